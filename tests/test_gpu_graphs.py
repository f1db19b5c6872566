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


def test_graph_replay_after_workspace_growth():
    """A recorded graph keeps addressing the workspace it was captured with: eager
    work that needs a larger workspace (or a larger rotation ping-pong buffer)
    afterwards must not free it under the graph.  Capture at 2 items / level 2,
    grow both buffers eagerly (8 items, level 4), replay, compare with eager."""
    from anns_fhe_amd.driver import FheContext, FHEHostDriver
    N, L, K, dnum = 4096, 4, 2, 2
    W = L * N
    ctx = FheContext(N, L, K, dnum, poly_words=96 * W, key_words=8 * 2 * dnum * (L + K) * N)
    try:
        drv = FHEHostDriver(ctx)
        rng = np.random.default_rng(11)
        s = ctx.keygen_secret(4)
        ksz = ctx.key_size_words(L)
        ctx.keygen_relin_key(s, L, 5, 0)
        drv.set_relin_key(0, L)
        for i, st in enumerate((1, 2)):
            g = ctx.galois_for_step(st)
            ctx.keygen_rotation_key(s, g, L, 6 + i, (1 + i) * ksz)
            drv.add_rotation_key(st, g, (1 + i) * ksz, L)
        small = [dict(kind=1, limbs=2, a0=4 * z * W, a1=(4 * z + 1) * W, b0=(4 * z + 2) * W, b1=(4 * z + 3) * W,
                      d0=(40 + 2 * z) * W, d1=(41 + 2 * z) * W) for z in range(2)]
        big = [dict(kind=1, limbs=L, a0=4 * z * W, a1=(4 * z + 1) * W, b0=(4 * z + 2) * W, b1=(4 * z + 3) * W,
                    d0=(50 + 2 * z) * W, d1=(51 + 2 * z) * W) for z in range(8)]

        def small_work():
            ctx.submit(small)
            drv.eval_rotate(0, W, 70 * W, 71 * W, 3, 2)  # two steps at level 2

        ctx.write_poly(0, _rand(rng, ctx, 32))
        small_work()
        ctx.sync()
        say("capture small")
        ctx.graph_begin()
        small_work()
        gid = ctx.graph_end()
        say("grow workspaces")
        ctx.submit(big)
        drv.eval_rotate(0, W, 72 * W, 73 * W, 3, L)  # ping-pong buffer for level 4
        ctx.sync()
        for rep in range(2):
            ctx.write_poly(0, _rand(rng, ctx, 32))
            ctx.graph_launch(gid)
            ctx.sync()
            got = ctx.read_poly(40 * W, 4 * W)
            got_rot = ctx.read_poly(70 * W, 2 * W)
            small_work()
            ctx.sync()
            say("replay after growth", rep)
            assert np.array_equal(got, ctx.read_poly(40 * W, 4 * W))
            assert np.array_equal(got_rot, ctx.read_poly(70 * W, 2 * W))
        ctx.graph_destroy(gid)
        # with no graph left the replaced buffers are freed; eager work still runs
        ctx.submit(big)
        ctx.sync()
    finally:
        ctx.close()


def test_submit_refuses_cross_item_aliasing():
    """fhea_submit: an item may write over its own inputs, but an output that
    overlaps another item's input or output is FHEA_ERR_ARG (the group's items
    run in chunks and on two streams)."""
    from anns_fhe_amd._lib import FheError
    from anns_fhe_amd.driver import FheContext, FHEHostDriver
    N, L, K, dnum = 4096, 3, 1, 3
    W = L * N
    ctx = FheContext(N, L, K, dnum, poly_words=32 * W, key_words=2 * dnum * (L + K) * N + 8)
    try:
        drv = FHEHostDriver(ctx)
        s = ctx.keygen_secret(8)
        ctx.keygen_relin_key(s, L, 9, 0)
        drv.set_relin_key(0, L)
        rng = np.random.default_rng(12)
        data = _rand(rng, ctx, 8)
        ctx.write_poly(0, data)

        def item(z, d0, d1):
            return dict(kind=1, limbs=L, a0=4 * z * W, a1=(4 * z + 1) * W, b0=(4 * z + 2) * W, b1=(4 * z + 3) * W,
                        d0=d0, d1=d1)

        # in place: each item overwrites its own a0/a1
        ctx.submit([item(0, 0, W), item(1, 4 * W, 5 * W)])
        got = ctx.read_poly(0, 8 * W)
        ctx.write_poly(0, data)
        ctx.submit([item(0, 16 * W, 17 * W), item(1, 18 * W, 19 * W)])
        W1 = (L - 1) * N
        exp = ctx.read_poly(16 * W, 4 * W)
        assert np.array_equal(got[:W1], exp[:W1]) and np.array_equal(got[W:W + W1], exp[W:W + W1])
        assert np.array_equal(got[4 * W:4 * W + W1], exp[2 * W:2 * W + W1])
        for bad in ([item(0, 4 * W, 17 * W), item(1, 18 * W, 19 * W)],      # output over item 1's a0
                    [item(0, 16 * W, 17 * W), item(1, 16 * W + 8, 19 * W)],  # outputs overlap
                    [item(0, 16 * W, 16 * W + W1 - 8), item(1, 18 * W, 19 * W)]):  # d0/d1 of one item
            with pytest.raises(FheError) as ei:
                ctx.submit(bad)
            assert ei.value.status == -4
    finally:
        ctx.close()
