"""GPU parity at the headline sizes and on the integer path.

* The benchmark's own configuration: batched eval_mult_relin_rescale (work kind 1)
  at N=2^16, L=24, K=8, dnum=3, B=64 on two streams with key-switch chunks of 16
  and of 64 and on three streams; items at the chunk and lane boundaries against
  the oracle.
* eval_mult_plain_rescale (fhe_host_driver.cpp:308-329) at the C1 shape and at
  N=2^16, L=24.
* The integer kernels, which run when a prime is >= 2^50 or the chain spans a
  ratio >= 4: 60-bit primes, and the reference's own chain shape (a 60-bit q0,
  45-bit scaling primes, 60-bit special primes;
  integration_tools/generate_interactive_test_vectors.cpp:161-166), at N=2^13,
  the reference's N=2^14/L=11/K=4/dnum=3 and N=2^16/L=24/K=8/dnum=3: OP 1/2,
  OP 7, OP 8, eval_mult_relin_rescale and eval_rotate against the oracle.
All comparisons are bit-exact on every residue."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-fullsize]", *a, flush=True)
    sys.stdout.flush()


def rand_poly(rng, primes, N):
    return np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in primes])


def rand_key(rng, ctx, key_limbs):
    L, K, N = ctx.L, ctx.K, ctx.N
    kidx = list(range(key_limbs)) + [L + j for j in range(K)]
    return np.concatenate([rand_poly(rng, [ctx.primes[i] for i in kidx], N).reshape(-1)
                           for _ in range(2 * ctx.dnum)])


def test_bench_config_batched_hmult_vs_oracle(orc):
    """B=64 HMults sharing one relin key at N=2^16/L=24/K=8/dnum=3, two lanes
    (items 32..63 on the second stream) with chunks of 16 and of 64, three lanes
    (items 0..20, 21..41, 42..63) with chunks of 16, and two lanes with the inverse
    NTT split at 7 (the conversion kernels' inverse column pass one stage shorter than
    the forward one): items 0, 15, 16, 31, 32, 63 equal the oracle's
    eval_mult_relin_rescale."""
    from anns_fhe_amd import _lib
    from anns_fhe_amd.driver import FheContext, FHEHostDriver
    N, L, K, dnum, B = 65536, 24, 8, 3, 64
    W, W1 = L * N, (L - 1) * N
    out_base = 4 * B * W
    ctx = FheContext(N, L, K, dnum, poly_words=out_base + 2 * B * W1 + 8,
                     key_words=2 * dnum * (L + K) * N + 8)
    try:
        drv = FHEHostDriver(ctx)
        rng = np.random.default_rng(64)
        key = rand_key(rng, ctx, L)
        ctx.write_key(0, key)
        drv.set_relin_key(0, L)
        sample = (0, 15, 16, 31, 32, 63)
        inputs = {}
        say("writing", B, "ciphertext pairs")
        for z in range(B):
            d = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(4)])
            ctx.write_poly(4 * z * W, d)
            if z in sample:
                inputs[z] = d
        works = [dict(kind=1, limbs=L, a0=4 * z * W, a1=(4 * z + 1) * W, b0=(4 * z + 2) * W,
                      b1=(4 * z + 3) * W, d0=out_base + 2 * z * W1, d1=out_base + (2 * z + 1) * W1)
                 for z in range(B)]
        got = {}
        # (streams, chunk, inverse NTT split): split 7 makes the inverse column pass one
        # stage shorter than the forward one inside the fused conversion kernels
        schedules = ((2, 16, 0), (2, 64, 0), (3, 16, 0), (2, 16, 7))
        for streams, chunk, split_inv in schedules:
            ctx.set_tuning(_lib.TUNE_STREAMS, streams)
            ctx.set_tuning(_lib.TUNE_CHUNK, chunk)
            ctx.set_tuning(_lib.TUNE_NTT_SPLIT_INV, split_inv)
            say("submit streams", streams, "chunk", chunk, "split_inv", split_inv)
            ctx.write_poly(out_base, np.zeros(2 * B * W1, np.uint64))
            ctx.submit(works)
            ctx.sync()
            got[streams, chunk, split_inv] = {z: ctx.read_poly(out_base + 2 * z * W1, 2 * W1) for z in sample}
        ctx.set_tuning(_lib.TUNE_STREAMS, 1)
        ctx.set_tuning(_lib.TUNE_CHUNK, 16)
        ctx.set_tuning(_lib.TUNE_NTT_SPLIT_INV, 0)
        o = orc.OracleCtx(N, ctx.q, ctx.p, dnum)
        for z in sample:
            say("oracle item", z)
            pm = np.zeros(6 * W, np.uint64)
            pm[:4 * W] = inputs[z]
            assert o.eval_mult_relin_rescale(pm, key, 0, L, 0, W, 2 * W, 3 * W, 4 * W, 5 * W, L) == 0
            for sc in schedules:
                g = got[sc][z]
                assert np.array_equal(g[:W1], pm[4 * W:4 * W + W1]), (sc, z, "c0")
                assert np.array_equal(g[W1:], pm[5 * W:5 * W + W1]), (sc, z, "c1")
    finally:
        ctx.close()


@pytest.mark.parametrize("N,L,K,dnum", [(8192, 4, 1, 2), (65536, 24, 8, 3)])
def test_mult_plain_rescale_vs_oracle(orc, N, L, K, dnum):
    """fhea_eval_mult_plain_rescale: (c0 pt, c1 pt) then OP 8, also with the output
    over the input (d0 == a0)."""
    from anns_fhe_amd.driver import FheContext, FHEHostDriver
    W, W1 = L * N, (L - 1) * N
    ctx = FheContext(N, L, K, dnum, poly_words=8 * W, key_words=64)
    try:
        drv = FHEHostDriver(ctx)
        rng = np.random.default_rng(N + L)
        data = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(3)])
        data[:8] = ctx.q[0] - 1
        ctx.write_poly(0, data)
        say("mult_plain_rescale", N, L)
        drv.eval_mult_plain_rescale(0, W, 2 * W, 3 * W, 4 * W, L)
        got = ctx.read_poly(3 * W, 2 * W)
        o = orc.OracleCtx(N, ctx.q, ctx.p, dnum)
        pm = np.zeros(5 * W, np.uint64)
        pm[:3 * W] = data
        assert o.eval_mult_plain_rescale(pm, 0, W, 2 * W, 3 * W, 4 * W, L) == 0
        assert np.array_equal(got[:W1], pm[3 * W:3 * W + W1])
        assert np.array_equal(got[W:W + W1], pm[4 * W:4 * W + W1])
        drv.eval_mult_plain_rescale(0, W, 2 * W, 0, W, L)  # in place
        got = ctx.read_poly(0, 2 * W)
        assert np.array_equal(got[:W1], pm[3 * W:3 * W + W1])
        assert np.array_equal(got[W:W + W1], pm[4 * W:4 * W + W1])
    finally:
        ctx.close()


def _chain(orc, N, L, K, kind):
    """(q primes, p primes) for an integer-path parameter set."""
    if kind == "q60":
        ps = orc.gen_primes(2 * N, 60, L + K)
        return ps[:L], ps[L:]
    # the reference's chain shape: q0 of 60 bits, 45-bit scaling primes, 60-bit P
    big = orc.gen_primes(2 * N, 60, K + 1)
    return [big[0]] + orc.gen_primes(2 * N, 45, L - 1), big[1:]


@pytest.mark.parametrize("N,L,K,dnum,kind", [(8192, 4, 1, 2, "q60"), (8192, 4, 1, 2, "ref_chain"),
                                             (16384, 11, 4, 3, "ref_chain"), (65536, 24, 8, 3, "ref_chain")])
def test_integer_path_vs_oracle(orc, N, L, K, dnum, kind):
    from anns_fhe_amd.driver import FheContext, FHEHostDriver
    q, p = _chain(orc, N, L, K, kind)
    W, W1 = L * N, (L - 1) * N
    ksz = 2 * dnum * (L + K) * N
    ctx = FheContext(N, L, K, dnum, q_primes=q, p_primes=p, poly_words=16 * W, key_words=2 * ksz + 8)
    try:
        assert ctx.q == q and ctx.p == p
        drv = FHEHostDriver(ctx)
        o = orc.OracleCtx(N, q, p, dnum)
        rng = np.random.default_rng(N + L + len(kind))
        relin, rot = rand_key(rng, ctx, L), rand_key(rng, ctx, L)
        km = np.concatenate([relin, rot])
        ctx.write_key(0, km)
        drv.set_relin_key(0, L)
        g1 = ctx.galois_for_step(1)
        drv.add_rotation_key(1, g1, ksz, L)
        data = np.concatenate([rand_poly(rng, q, N).reshape(-1) for _ in range(4)])
        for i in range(L):
            data[i * N] = q[i] - 1
        ctx.write_poly(0, data)
        pm = np.zeros(16 * W, np.uint64)
        pm[:4 * W] = data
        say("int path", kind, N, L, "ntt/intt")
        for op in (1, 2):
            assert ctx.op(op, 0, 0, 4 * W, 0, L) == 0
            assert o.accel_op(pm, km, op, 0, 0, 4 * W, 0, L) == 0
            assert np.array_equal(ctx.read_poly(4 * W, W), pm[4 * W:5 * W]), op
        say("int path", kind, "op7")
        assert ctx.op(7, W, 0, 5 * W, 0, L, L) == 0
        assert o.accel_op(pm, km, 7, W, 0, 5 * W, 0, L, L) == 0
        assert np.array_equal(ctx.read_poly(5 * W, 2 * W), pm[5 * W:7 * W])
        say("int path", kind, "op8")
        assert ctx.op(8, 0, W, 7 * W, 0, L) == 0
        assert o.accel_op(pm, km, 8, 0, W, 7 * W, 0, L) == 0
        assert np.array_equal(ctx.read_poly(7 * W, 2 * W1), pm[7 * W:7 * W + 2 * W1])
        say("int path", kind, "hmult")
        drv.eval_mult_relin_rescale(0, W, 2 * W, 3 * W, 9 * W, 10 * W, L)
        assert o.eval_mult_relin_rescale(pm, km, 0, L, 0, W, 2 * W, 3 * W, 9 * W, 10 * W, L) == 0
        assert np.array_equal(ctx.read_poly(9 * W, W1), pm[9 * W:9 * W + W1])
        assert np.array_equal(ctx.read_poly(10 * W, W1), pm[10 * W:10 * W + W1])
        say("int path", kind, "rotate")
        drv.eval_rotate(9 * W, 10 * W, 11 * W, 12 * W, 1, L - 1)
        assert o.eval_rotate(pm, km, [(1, g1, ksz, L)], 9 * W, 10 * W, 11 * W, 12 * W, 1, L - 1) == 0
        assert np.array_equal(ctx.read_poly(11 * W, W1), pm[11 * W:11 * W + W1])
        assert np.array_equal(ctx.read_poly(12 * W, W1), pm[12 * W:12 * W + W1])
    finally:
        ctx.close()


@pytest.mark.parametrize("switch,cc_variant", [(0, 0), (0, 1), (0, 2), (0, 3), (0, "inv6"), (0, "inv8"), (128, 0), (256, 0), (64, 0), (512, 0),
                                               (512 | 128, 0), (1024, 0), (1024 | 128, 0), (2048, 0), (8192, 0), (4096, 0),
                                               (1 | 2 | 4 | 8 | 16 | 32, 0)])
def test_pipeline_switches_keep_words(orc, monkeypatch, switch, cc_variant):
    """FHEA_F64_OFF selects alternative kernel pipelines of the FP64 path (128: the
    P slots' ModDown INTT pass inside the fused inner product; 256: the slot-major
    tile order (instead of unit-major) for the passes whose prologue reads the
    rescale lift; 64: the inner product as its own kernel; 512: the INTT column
    pass, basis conversion and NTT column pass as three launches instead of the
    fused conversion kernel; 1024: the HMult tensor in full instead of folded into
    the key switch; 2048: the HMult's Q-slot sums stored and the ModDown + rescale
    row pass as its own launch instead of fused into the rows kernel; 8192: the
    P-slot sums (HMult and rotations) stored and the ModDown INTT row pass as its
    own launch;
    4096: rotations
    through the automorphism kernel instead of sigma_g applied while loading; 1..32: each component on the integer kernels), and
    FHEA_CC_VARIANT the fused conversion kernel's tile shape; "invK" sets the inverse
    NTT split to K (inv6: the conversion kernels' inverse column pass one stage shorter
    than the forward one; inv8: one longer, which leaves the conversions unfused).
    Every setting gives the oracle's words for a batched HMult (3 items, one lane and
    two) and a rotation at N=2^14, L=8, K=2, dnum=3 (two-pass transforms, 3 digits)."""
    from anns_fhe_amd import _lib
    from anns_fhe_amd.driver import FheContext, FHEHostDriver
    split_inv = 0
    if isinstance(cc_variant, str):
        split_inv, cc_variant = int(cc_variant[3:]), 0
    monkeypatch.setenv("FHEA_F64_OFF", str(switch))
    monkeypatch.setenv("FHEA_CC_VARIANT", str(cc_variant))
    N, L, K, dnum = 16384, 8, 2, 3
    W, W1 = L * N, (L - 1) * N
    ksz = 2 * dnum * (L + K) * N
    ctx = FheContext(N, L, K, dnum, poly_words=24 * W, key_words=2 * ksz + 8)
    try:
        if split_inv:
            ctx.set_tuning(_lib.TUNE_NTT_SPLIT_INV, split_inv)
        drv = FHEHostDriver(ctx)
        o = orc.OracleCtx(N, ctx.q, ctx.p, dnum)
        rng = np.random.default_rng(300 + switch + 7 * cc_variant + 13 * split_inv)
        relin, rot = rand_key(rng, ctx, L), rand_key(rng, ctx, L)
        km = np.concatenate([relin, rot])
        ctx.write_key(0, km)
        drv.set_relin_key(0, L)
        g1 = ctx.galois_for_step(1)
        drv.add_rotation_key(1, g1, ksz, L)
        B = 3
        data = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(4 * B)])
        ctx.write_poly(0, data)
        pm = np.zeros(24 * W, np.uint64)
        pm[:4 * B * W] = data
        works = [dict(kind=1, limbs=L, a0=4 * z * W, a1=(4 * z + 1) * W, b0=(4 * z + 2) * W,
                      b1=(4 * z + 3) * W, d0=(12 + 2 * z) * W, d1=(13 + 2 * z) * W) for z in range(B)]
        for streams in (1, 2):
            ctx.set_tuning(_lib.TUNE_STREAMS, streams)
            say("switch", switch, "cc_variant", cc_variant, "split_inv", split_inv, "streams", streams)
            ctx.submit(works)
            for z in range(B):
                assert o.eval_mult_relin_rescale(pm, km, 0, L, 4 * z * W, (4 * z + 1) * W, (4 * z + 2) * W,
                                                 (4 * z + 3) * W, (12 + 2 * z) * W, (13 + 2 * z) * W, L) == 0
                for d in (12 + 2 * z, 13 + 2 * z):
                    assert np.array_equal(ctx.read_poly(d * W, W1), pm[d * W:d * W + W1]), (switch, streams, z)
        say("switch", switch, "rotate")
        drv.eval_rotate(0, W, 20 * W, 21 * W, 1, L)
        assert o.eval_rotate(pm, km, [(1, g1, ksz, L)], 0, W, 20 * W, 21 * W, 1, L) == 0
        assert np.array_equal(ctx.read_poly(20 * W, W), pm[20 * W:21 * W])
        assert np.array_equal(ctx.read_poly(21 * W, W), pm[21 * W:22 * W])
        # in place (outputs over the inputs): sigma_g(c0) cannot be read while the
        # outputs are written, so this takes the automorphism-kernel route
        say("switch", switch, "rotate in place")
        drv.eval_automorphism_ks(g1, ksz, L, 20 * W, 21 * W, 20 * W, 21 * W, L)
        assert o.eval_automorphism_ks(pm, km, g1, ksz, L, 20 * W, 21 * W, 20 * W, 21 * W, L) == 0
        assert np.array_equal(ctx.read_poly(20 * W, 2 * W), pm[20 * W:22 * W])
        # c1 overwritten only: sigma_g read while loading, but the Q-slot sums stored
        # (the kernel that writes the outputs would read c1 through sigma_g)
        say("switch", switch, "rotate over c1")
        drv.eval_automorphism_ks(g1, ksz, L, 20 * W, 21 * W, 22 * W, 21 * W, L)
        assert o.eval_automorphism_ks(pm, km, g1, ksz, L, 20 * W, 21 * W, 22 * W, 21 * W, L) == 0
        assert np.array_equal(ctx.read_poly(20 * W, 3 * W), pm[20 * W:23 * W])
    finally:
        ctx.close()


@pytest.mark.parametrize("shape", ["p60", "q2big"])
def test_mixed_chain_levels_vs_oracle(orc, shape):
    """Chains with primes on both sides of 2^50 take the FP64 fused ModDown / rescale
    transforms on the limbs past a leading run of >= 2^50 Q primes and the integer
    kernels on that run.  p60: every Q prime 45-bit, P 60-bit (no run); q2big: q0, q1
    of 60 bits, then 45-bit primes (a run of two, so at levels 3 and 2 every output
    limb of the rescale is in the run and the integer path takes over).  At every
    level: OP 7, OP 8, eval_mult_relin_rescale and a rotation = oracle."""
    from anns_fhe_amd.driver import FheContext, FHEHostDriver
    N, L, K, dnum = 8192, 6, 2, 3
    if shape == "p60":
        big = orc.gen_primes(2 * N, 60, K)
        q, p = orc.gen_primes(2 * N, 45, L), big
    else:
        big = orc.gen_primes(2 * N, 60, K + 2)
        q, p = big[:2] + orc.gen_primes(2 * N, 45, L - 2), big[2:]
    W = L * N
    ksz = 2 * dnum * (L + K) * N
    ctx = FheContext(N, L, K, dnum, q_primes=q, p_primes=p, poly_words=16 * W, key_words=2 * ksz + 8)
    try:
        drv = FHEHostDriver(ctx)
        o = orc.OracleCtx(N, q, p, dnum)
        rng = np.random.default_rng(len(shape) + 77)
        relin, rot = rand_key(rng, ctx, L), rand_key(rng, ctx, L)
        km = np.concatenate([relin, rot])
        ctx.write_key(0, km)
        drv.set_relin_key(0, L)
        g1 = ctx.galois_for_step(1)
        drv.add_rotation_key(1, g1, ksz, L)
        data = np.concatenate([rand_poly(rng, q, N).reshape(-1) for _ in range(4)])
        for i in range(L):
            data[i * N] = q[i] - 1
        ctx.write_poly(0, data)
        pm = np.zeros(16 * W, np.uint64)
        pm[:4 * W] = data
        for l in (6, 4, 3, 2):
            l1 = (l - 1) * N
            say("mixed", shape, "level", l)
            assert ctx.op(7, W, 0, 5 * W, 0, l, L) == 0
            assert o.accel_op(pm, km, 7, W, 0, 5 * W, 0, l, L) == 0
            assert np.array_equal(ctx.read_poly(5 * W, l * N), pm[5 * W:5 * W + l * N]), ("op7", l)
            assert np.array_equal(ctx.read_poly(5 * W + l * N, l * N), pm[5 * W + l * N:5 * W + 2 * l * N]), ("op7", l)
            assert ctx.op(8, 0, W, 7 * W, 0, l) == 0
            assert o.accel_op(pm, km, 8, 0, W, 7 * W, 0, l) == 0
            assert np.array_equal(ctx.read_poly(7 * W, 2 * l1), pm[7 * W:7 * W + 2 * l1]), ("op8", l)
            drv.eval_mult_relin_rescale(0, W, 2 * W, 3 * W, 9 * W, 10 * W, l)
            assert o.eval_mult_relin_rescale(pm, km, 0, L, 0, W, 2 * W, 3 * W, 9 * W, 10 * W, l) == 0
            assert np.array_equal(ctx.read_poly(9 * W, l1), pm[9 * W:9 * W + l1]), ("hmult", l)
            assert np.array_equal(ctx.read_poly(10 * W, l1), pm[10 * W:10 * W + l1]), ("hmult", l)
            drv.eval_rotate(0, W, 11 * W, 12 * W, 1, l)
            assert o.eval_rotate(pm, km, [(1, g1, ksz, L)], 0, W, 11 * W, 12 * W, 1, l) == 0
            assert np.array_equal(ctx.read_poly(11 * W, l * N), pm[11 * W:11 * W + l * N]), ("rot", l)
            assert np.array_equal(ctx.read_poly(12 * W, l * N), pm[12 * W:12 * W + l * N]), ("rot", l)
    finally:
        ctx.close()


@pytest.mark.parametrize("chain", ["q50", "mixed"])
def test_largest_n_vs_oracle(orc, chain):
    """N = 2^17, the largest ring the tables accept: the transforms' two passes are
    2^8 x 2^9 (the fused key-switch row pass runs its 2^9 instance).  OP 7, OP 8,
    eval_mult_relin_rescale and a rotation = oracle, on 50-bit primes and on a mixed
    60/45-bit chain."""
    from anns_fhe_amd.driver import FheContext, FHEHostDriver
    N, L, K, dnum = 1 << 17, 4, 2, 2
    if chain == "q50":
        ps = orc.gen_primes(2 * N, 50, L + K)
        q, p = ps[:L], ps[L:]
    else:
        big = orc.gen_primes(2 * N, 60, K + 1)
        q, p = [big[0]] + orc.gen_primes(2 * N, 45, L - 1), big[1:]
    W, W1 = L * N, (L - 1) * N
    ksz = 2 * dnum * (L + K) * N
    ctx = FheContext(N, L, K, dnum, q_primes=q, p_primes=p, poly_words=16 * W, key_words=2 * ksz + 8)
    try:
        drv = FHEHostDriver(ctx)
        o = orc.OracleCtx(N, q, p, dnum)
        rng = np.random.default_rng(17 + len(chain))
        relin, rot = rand_key(rng, ctx, L), rand_key(rng, ctx, L)
        km = np.concatenate([relin, rot])
        ctx.write_key(0, km)
        drv.set_relin_key(0, L)
        g1 = ctx.galois_for_step(-1)
        drv.add_rotation_key(-1, g1, ksz, L)
        data = np.concatenate([rand_poly(rng, q, N).reshape(-1) for _ in range(4)])
        ctx.write_poly(0, data)
        pm = np.zeros(16 * W, np.uint64)
        pm[:4 * W] = data
        say("N=2^17", chain, "op7")
        assert ctx.op(7, W, 0, 5 * W, 0, L, L) == 0
        assert o.accel_op(pm, km, 7, W, 0, 5 * W, 0, L, L) == 0
        assert np.array_equal(ctx.read_poly(5 * W, 2 * W), pm[5 * W:7 * W])
        say("N=2^17", chain, "op8")
        assert ctx.op(8, 0, W, 7 * W, 0, L) == 0
        assert o.accel_op(pm, km, 8, 0, W, 7 * W, 0, L) == 0
        assert np.array_equal(ctx.read_poly(7 * W, 2 * W1), pm[7 * W:7 * W + 2 * W1])
        say("N=2^17", chain, "hmult")
        drv.eval_mult_relin_rescale(0, W, 2 * W, 3 * W, 9 * W, 10 * W, L)
        assert o.eval_mult_relin_rescale(pm, km, 0, L, 0, W, 2 * W, 3 * W, 9 * W, 10 * W, L) == 0
        assert np.array_equal(ctx.read_poly(9 * W, W1), pm[9 * W:9 * W + W1])
        assert np.array_equal(ctx.read_poly(10 * W, W1), pm[10 * W:10 * W + W1])
        say("N=2^17", chain, "rotate")
        drv.eval_rotate(0, W, 11 * W, 12 * W, -1, L)
        assert o.eval_rotate(pm, km, [(-1, g1, ksz, L)], 0, W, 11 * W, 12 * W, -1, L) == 0
        assert np.array_equal(ctx.read_poly(11 * W, W), pm[11 * W:12 * W])
        assert np.array_equal(ctx.read_poly(12 * W, W), pm[12 * W:13 * W])
    finally:
        ctx.close()


@pytest.mark.parametrize("N", [1024, 2048])
@pytest.mark.parametrize("chain", ["q50", "mixed"])
def test_smallest_n_vs_oracle(orc, N, chain):
    """The smallest rings (one-pass transforms of a whole limb, unfused key switch):
    eval_mult_relin_rescale and a rotation = oracle, on 50-bit and mixed chains."""
    from anns_fhe_amd.driver import FheContext, FHEHostDriver
    L, K, dnum = 3, 1, 3
    if chain == "q50":
        ps = orc.gen_primes(2 * N, 50, L + K)
        q, p = ps[:L], ps[L:]
    else:
        big = orc.gen_primes(2 * N, 60, K + 1)
        q, p = [big[0]] + orc.gen_primes(2 * N, 45, L - 1), big[1:]
    W, W1 = L * N, (L - 1) * N
    ksz = 2 * dnum * (L + K) * N
    ctx = FheContext(N, L, K, dnum, q_primes=q, p_primes=p, poly_words=16 * W, key_words=2 * ksz + 8)
    try:
        drv = FHEHostDriver(ctx)
        o = orc.OracleCtx(N, q, p, dnum)
        rng = np.random.default_rng(N + len(chain))
        relin, rot = rand_key(rng, ctx, L), rand_key(rng, ctx, L)
        km = np.concatenate([relin, rot])
        ctx.write_key(0, km)
        drv.set_relin_key(0, L)
        g = ctx.galois_for_step(4)
        drv.add_rotation_key(4, g, ksz, L)
        data = np.concatenate([rand_poly(rng, q, N).reshape(-1) for _ in range(4)])
        ctx.write_poly(0, data)
        pm = np.zeros(16 * W, np.uint64)
        pm[:4 * W] = data
        say("small N", N, chain)
        drv.eval_mult_relin_rescale(0, W, 2 * W, 3 * W, 9 * W, 10 * W, L)
        assert o.eval_mult_relin_rescale(pm, km, 0, L, 0, W, 2 * W, 3 * W, 9 * W, 10 * W, L) == 0
        assert np.array_equal(ctx.read_poly(9 * W, W1), pm[9 * W:9 * W + W1])
        assert np.array_equal(ctx.read_poly(10 * W, W1), pm[10 * W:10 * W + W1])
        drv.eval_rotate(0, W, 11 * W, 12 * W, 4, L)
        assert o.eval_rotate(pm, km, [(4, g, ksz, L)], 0, W, 11 * W, 12 * W, 4, L) == 0
        assert np.array_equal(ctx.read_poly(11 * W, W), pm[11 * W:12 * W])
        assert np.array_equal(ctx.read_poly(12 * W, W), pm[12 * W:13 * W])
    finally:
        ctx.close()


def test_batches_beyond_one_argument_block(orc):
    """A submit of more items than one staged argument block holds (4,096) runs in
    segments: 4,100 HMults and 4,100 rotations at N=2^13, L=3 sharing their inputs;
    items 0, 4095, 4096 and 4099 (either side of the segment boundary) equal the
    oracle's words."""
    from anns_fhe_amd.driver import FheContext, FHEHostDriver
    N, L, K, dnum, B = 8192, 3, 1, 3, 4100
    W, W1 = L * N, (L - 1) * N
    ksz = 2 * dnum * (L + K) * N
    out_m = 6 * W                 # HMult outputs: B pairs of L-1 limbs
    out_r = out_m + 2 * B * W1    # rotation outputs: B pairs of L limbs
    ctx = FheContext(N, L, K, dnum, poly_words=out_r + 2 * B * W + 8, key_words=2 * ksz + 8)
    try:
        drv = FHEHostDriver(ctx)
        o = orc.OracleCtx(N, ctx.q, ctx.p, dnum)
        rng = np.random.default_rng(4100)
        relin, rot = rand_key(rng, ctx, L), rand_key(rng, ctx, L)
        km = np.concatenate([relin, rot])
        ctx.write_key(0, km)
        drv.set_relin_key(0, L)
        g1 = ctx.galois_for_step(1)
        data = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(4)])
        ctx.write_poly(0, data)
        pm = np.zeros(6 * W + 2 * W, np.uint64)
        pm[:4 * W] = data
        say("submit", B, "HMults")
        ctx.submit([dict(kind=1, limbs=L, a0=0, a1=W, b0=2 * W, b1=3 * W,
                         d0=out_m + 2 * z * W1, d1=out_m + (2 * z + 1) * W1) for z in range(B)])
        say("submit", B, "rotations")
        ctx.submit([dict(kind=2, limbs=L, a0=0, a1=W, galois=g1, evk=ksz, key_limbs=L,
                         d0=out_r + 2 * z * W, d1=out_r + (2 * z + 1) * W) for z in range(B)])
        ctx.sync()
        assert o.eval_mult_relin_rescale(pm, km, 0, L, 0, W, 2 * W, 3 * W, 4 * W, 5 * W, L) == 0
        assert o.eval_automorphism_ks(pm, km, g1, ksz, L, 0, W, 6 * W, 7 * W, L) == 0
        for z in (0, 4095, 4096, B - 1):
            say("item", z)
            gm = ctx.read_poly(out_m + 2 * z * W1, 2 * W1)
            assert np.array_equal(gm[:W1], pm[4 * W:4 * W + W1]), ("hmult c0", z)
            assert np.array_equal(gm[W1:], pm[5 * W:5 * W + W1]), ("hmult c1", z)
            gr = ctx.read_poly(out_r + 2 * z * W, 2 * W)
            assert np.array_equal(gr, pm[6 * W:8 * W]), ("rotation", z)
    finally:
        ctx.close()


@pytest.mark.parametrize("case", ["same_words", "shifted_d0", "shifted_d1"])
def test_hmult_outputs_over_own_inputs(orc, case):
    """An item may write over its own inputs (fhea.h).  The Q-slot rows kernel reads
    an item's inputs while it writes its outputs, which is safe when an output lies
    on an input's own words (same_words: d0 = a0, d1 = a1) and not when it is shifted
    against it; then the pipeline stores the sums and runs the ModDown + rescale row
    pass on its own.  N=2^14, L=8, K=2, dnum=3; every case = oracle."""
    from anns_fhe_amd.driver import FheContext, FHEHostDriver
    N, L, K, dnum = 16384, 8, 2, 3
    W, W1 = L * N, (L - 1) * N
    ctx = FheContext(N, L, K, dnum, poly_words=8 * W, key_words=2 * dnum * (L + K) * N + 8)
    try:
        drv = FHEHostDriver(ctx)
        o = orc.OracleCtx(N, ctx.q, ctx.p, dnum)
        rng = np.random.default_rng({"same_words": 1, "shifted_d0": 2, "shifted_d1": 3}[case])
        km = rand_key(rng, ctx, L)
        ctx.write_key(0, km)
        drv.set_relin_key(0, L)
        data = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(4)])
        ctx.write_poly(0, data)
        pm = np.zeros(8 * W, np.uint64)
        pm[:4 * W] = data
        a0, a1, b0, b1 = 0, W, 2 * W, 3 * W
        d0, d1 = {"same_words": (a0, a1), "shifted_d0": (a0 + N, 5 * W), "shifted_d1": (5 * W, b1 + 3 * N)}[case]
        ctx.submit([dict(kind=1, limbs=L, a0=a0, a1=a1, b0=b0, b1=b1, d0=d0, d1=d1)])
        assert o.eval_mult_relin_rescale(pm, km, 0, L, a0, a1, b0, b1, d0, d1, L) == 0
        for d in (d0, d1):
            assert np.array_equal(ctx.read_poly(d, W1), pm[d:d + W1]), (case, d)
    finally:
        ctx.close()
