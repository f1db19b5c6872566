"""HIP graph capture of engine work (fhea_graph_*): a recorded sequence replayed on
new data at the same offsets gives the same words as running it eagerly, replay
after replay with eager work in between; host-synchronising calls are refused
while capturing."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-graphs]", *a, flush=True)
    sys.stdout.flush()


def _rand(rng, ctx, n):
    return np.concatenate([np.stack([rng.integers(0, q, ctx.N, dtype=np.uint64) for q in ctx.q]).reshape(-1)
                           for _ in range(n)])


def test_graph_replay_matches_eager():
    from anns_fhe_amd._lib import FheError
    from anns_fhe_amd.driver import FheContext, FHEHostDriver
    N, L, K, dnum = 4096, 4, 2, 2
    W = L * N
    ctx = FheContext(N, L, K, dnum, poly_words=64 * W, key_words=16 * 2 * dnum * (L + K) * N)
    try:
        drv = FHEHostDriver(ctx)
        rng = np.random.default_rng(3)
        s = ctx.keygen_secret(1)
        ksz = ctx.key_size_words(L)
        ctx.keygen_relin_key(s, L, 2, 0)
        drv.set_relin_key(0, L)
        for i, st in enumerate((1, 2)):
            g = ctx.galois_for_step(st)
            ctx.keygen_rotation_key(s, g, L, 3 + i, (1 + i) * ksz)
            drv.add_rotation_key(st, g, (1 + i) * ksz, L)
        B = 3
        mult = [dict(kind=1, limbs=L, a0=4 * z * W, a1=(4 * z + 1) * W, b0=(4 * z + 2) * W, b1=(4 * z + 3) * W,
                     d0=(20 + 2 * z) * W, d1=(21 + 2 * z) * W) for z in range(B)]
        rot = [dict(kind=2, limbs=L, key_limbs=L, a0=4 * z * W, a1=(4 * z + 1) * W, d0=(30 + 2 * z) * W,
                    d1=(31 + 2 * z) * W, galois=ctx.galois_for_step(1), evk=ksz) for z in range(B)]

        def work():
            ctx.submit(mult)
            ctx.submit(rot)
            drv.eval_rotate(0, W, 40 * W, 41 * W, 3, L)          # two steps: ping-pong buffer
            assert ctx.op(8, 2 * W, 3 * W, 42 * W, 0, L) == 0      # staged OP 8 (device copies)
            ctx.linear_sum([0, 4 * W], [W, 5 * W], np.array([[2 * W, 6 * W]], np.uint64), [44 * W], [45 * W], L,
                           bias=(8 * W, 9 * W))

        outs = (20 * W, 26 * W)
        ctx.write_poly(0, _rand(rng, ctx, 4 * B))
        work()  # eager once: workspaces reach their size
        ctx.sync()
        say("capture")
        ctx.graph_begin()
        with pytest.raises(FheError) as ei:
            ctx.read_poly(0, 8)
        assert ei.value.status == -8
        work()
        gid = ctx.graph_end()
        for rep in range(3):
            ctx.write_poly(0, _rand(rng, ctx, 4 * B))
            ctx.graph_launch(gid)
            ctx.sync()
            got = ctx.read_poly(outs[0], 26 * W - 20 * W + 0)
            got_rest = ctx.read_poly(outs[1], 20 * W)
            work()  # eager, same inputs and offsets
            ctx.sync()
            exp = ctx.read_poly(outs[0], 26 * W - 20 * W)
            exp_rest = ctx.read_poly(outs[1], 20 * W)
            say("replay", rep)
            assert np.array_equal(got, exp)
            assert np.array_equal(got_rest, exp_rest)
        ctx.graph_destroy(gid)
        with pytest.raises(FheError):
            ctx.graph_launch(gid)
        # a capture needing a larger workspace than allocated is refused, not recorded
        big = [dict(kind=1, limbs=L, a0=0, a1=W, b0=2 * W, b1=3 * W, d0=(46 + 2 * z) * W, d1=(47 + 2 * z) * W)
               for z in range(8)]
        ctx.graph_begin()
        with pytest.raises(FheError) as ei:
            ctx.submit(big)
        assert ei.value.status == -6
        ctx.graph_end()
    finally:
        ctx.close()
