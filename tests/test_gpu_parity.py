"""GPU parity: the HIP engine (through the C-ABI) against the CPU oracle,
bit-exact on every RNS residue.  Mirrors the reference testbenches
(tb_ntt.cpp, tb_intt.cpp, tb_poly_arith.cpp, tb_automorphism.cpp,
tb_rescale.cpp, tb_keyswitch.cpp, tb_fhe_accel.cpp) and the main_sim.cpp
sequence, at runtime N/L/K/dnum instead of the compiled-in N=16384."""
import sys

import numpy as np
import pytest

from tests import ckks_ref as ck

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-parity]", *a, flush=True)
    sys.stdout.flush()


def rand_poly(rng, primes, N):
    return np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in primes])


_CTX = {}


def get_ctx(N, L, K, dnum, q_bits=50, poly_words=1 << 24, key_words=1 << 24):
    """q_bits: generated primes of that size, or "mixed": the reference's chain shape
    (60-bit q0 and P primes, 45-bit q1..), which runs the FP64 kernels on its 45-bit
    limbs and the integer kernels on its 60-bit ones."""
    from anns_fhe_amd.driver import FheContext
    key = (N, L, K, dnum, q_bits, poly_words, key_words)
    if key not in _CTX:
        say("ctx create", key)
        if q_bits == "mixed":
            from oracle import oracle as O
            big = O.gen_primes(2 * N, 60, K + 1)
            prm = dict(q_primes=[big[0]] + O.gen_primes(2 * N, 45, L - 1), p_primes=big[1:])
        else:
            prm = dict(q_bits=q_bits)
        _CTX[key] = FheContext(N, L, K, dnum, poly_words=poly_words, key_words=key_words, **prm)
    return _CTX[key]


def seed_of(q_bits):
    return q_bits if isinstance(q_bits, int) else 45


def get_orc(orc, ctx):
    return orc.OracleCtx(ctx.N, ctx.q, ctx.p, ctx.dnum)


def test_tables_match_oracle(orc):
    ctx = get_ctx(8192, 4, 1, 2)
    assert ctx.primes == orc.gen_primes(2 * 8192, 50, 5)
    o = get_orc(orc, ctx)
    assert np.array_equal(ctx.table(0), o.psi)
    assert np.array_equal(ctx.table(1), o.n_inv)
    assert np.array_equal(ctx.table(2), o.tw)
    assert np.array_equal(ctx.table(3), o.itw)


@pytest.mark.parametrize("q_bits", [50, 61])
@pytest.mark.parametrize("N,L", [(1024, 3), (2048, 2), (4096, 2), (8192, 4), (16384, 3), (32768, 2),
                                 (65536, 3), (131072, 2)])
def test_ntt_intt_parity(orc, N, L, q_bits):
    """q_bits 50: the FP64 transform; 61: the integer transform (lazy words up to 4q,
    the largest primes the tables accept)."""
    say("ntt", N, L, q_bits)
    ctx = get_ctx(N, L, 1, 1, q_bits=q_bits)
    o = get_orc(orc, ctx)
    rng = np.random.default_rng(42)
    a = rand_poly(rng, ctx.q, N)
    a[0, :4] = [0, 1, ctx.q[0] - 1, ctx.q[0] - 2]
    W = L * N
    ctx.write_poly(0, a)
    assert ctx.op(1, 0, 0, W, 0, L) == 0
    got = ctx.read_poly(W, W).reshape(L, N)
    assert np.array_equal(got, o.ntt_poly(a))
    # INTT back (in place) is the identity, INTT of random data matches
    assert ctx.op(2, W, 0, W, 0, L) == 0
    assert np.array_equal(ctx.read_poly(W, W).reshape(L, N), a)
    assert ctx.op(2, 0, 0, 2 * W, 0, L) == 0
    assert np.array_equal(ctx.read_poly(2 * W, W).reshape(L, N), o.intt_poly(a))


@pytest.mark.parametrize("N", [1024, 4096, 32768])
def test_ntt_mixed_chain_parity(orc, N):
    """A chain with primes on both sides of 2^50 (60-bit q0 and P, 45-bit q1..):
    the transforms of the 45-bit limbs run on the FP64 kernel and those of the
    60-bit limbs on the integer kernel, in launches over the same polynomials.
    Every limb matches the oracle; in-place INTT restores the input."""
    from anns_fhe_amd.driver import FheContext
    L = 4
    big = orc.gen_primes(2 * N, 60, 2)
    q, p = [big[0]] + orc.gen_primes(2 * N, 45, L - 1), big[1:]
    W = L * N
    ctx = FheContext(N, L, 1, 1, q_primes=q, p_primes=p, poly_words=4 * W, key_words=1 << 12)
    try:
        o = orc.OracleCtx(N, q, p, 1)
        rng = np.random.default_rng(N + 5)
        a = rand_poly(rng, q, N)
        for i in range(L):
            a[i, :2] = [q[i] - 1, 0]
        say("mixed ntt", N)
        ctx.write_poly(0, a)
        assert ctx.op(1, 0, 0, W, 0, L) == 0
        assert np.array_equal(ctx.read_poly(W, W).reshape(L, N), o.ntt_poly(a))
        assert ctx.op(2, W, 0, W, 0, L) == 0
        assert np.array_equal(ctx.read_poly(W, W).reshape(L, N), a)
        assert ctx.op(2, 0, 0, 2 * W, 0, L) == 0
        assert np.array_equal(ctx.read_poly(2 * W, W).reshape(L, N), o.intt_poly(a))
    finally:
        ctx.close()


def test_elementwise_parity(orc):
    N, L = 8192, 4
    ctx = get_ctx(N, L, 1, 2)
    o = get_orc(orc, ctx)
    rng = np.random.default_rng(1)
    W = L * N
    a, b = rand_poly(rng, ctx.q, N), rand_poly(rng, ctx.q, N)
    a[:, 0] = [q - 1 for q in ctx.q]
    b[:, 0] = [q - 1 for q in ctx.q]
    pm = np.zeros(8 * W, np.uint64)
    pm[:W], pm[W:2 * W] = a.reshape(-1), b.reshape(-1)
    ctx.write_poly(0, pm[:2 * W])
    for op in (3, 4, 5):
        say("binop", op)
        assert ctx.op(op, 0, W, 2 * W, 0, L) == 0
        ref = pm.copy()
        assert o.accel_op(ref, None, op, 0, W, 2 * W, 0, L) == 0
        assert np.array_equal(ctx.read_poly(2 * W, W), ref[2 * W:3 * W])
    # in place (dst == src_a)
    assert ctx.op(5, 0, W, 0, 0, L) == 0
    ref = pm.copy()
    o.accel_op(ref, None, 5, 0, W, 0, 0, L)
    assert np.array_equal(ctx.read_poly(0, W), ref[:W])


def test_automorphism_op6_parity(orc):
    N, L = 8192, 3
    ctx = get_ctx(N, L, 1, 1)
    o = get_orc(orc, ctx)
    rng = np.random.default_rng(2)
    W = L * N
    a = rand_poly(rng, ctx.q, N)
    for g in (5, pow(5, 7, 2 * N), 2 * N - 1):
        say("automorphism", g)
        packed = orc.pack_map(orc.galois_map(N, g))
        pm = np.zeros(4 * W, np.uint64)
        pm[:W] = a.reshape(-1)
        pm[W:W + N // 2] = packed
        ctx.write_poly(0, pm[:2 * W])
        assert ctx.op(6, 0, W, 2 * W, 0, L) == 0
        ref = pm.copy()
        assert o.accel_op(ref, None, 6, 0, W, 2 * W, 0, L) == 0
        assert np.array_equal(ctx.read_poly(2 * W, W), ref[2 * W:3 * W])
        # in place
        assert ctx.op(6, 0, W, 0, 0, L) == 0
        assert np.array_equal(ctx.read_poly(0, W), ref[2 * W:3 * W])


@pytest.mark.parametrize("L", [4, 2])
def test_rescale_op8_parity(orc, L):
    N = 8192
    ctx = get_ctx(N, 4, 1, 2)
    o = get_orc(orc, ctx)
    rng = np.random.default_rng(3 + L)
    W = L * N
    c0, c1 = rand_poly(rng, ctx.q[:L], N), rand_poly(rng, ctx.q[:L], N)
    pm = np.zeros(8 * W, np.uint64)
    pm[:W], pm[W:2 * W] = c0.reshape(-1), c1.reshape(-1)
    ctx.write_poly(0, pm[:2 * W])
    say("rescale", L)
    assert ctx.op(8, 0, W, 4 * W, 0, L) == 0
    ref = pm.copy()
    assert o.accel_op(ref, None, 8, 0, W, 4 * W, 0, L) == 0
    W1 = (L - 1) * N
    assert np.array_equal(ctx.read_poly(4 * W, 2 * W1), ref[4 * W:4 * W + 2 * W1])
    # output overlapping the inputs (dst == src_a)
    assert ctx.op(8, 0, W, 0, 0, L) == 0
    assert np.array_equal(ctx.read_poly(0, 2 * W1), ref[4 * W:4 * W + 2 * W1])


@pytest.mark.parametrize("sizeQl,key_limbs", [(6, 6), (5, 5), (3, 6), (1, 2)])
def test_keyswitch_op7_parity(orc, sizeQl, key_limbs):
    N, L, K, dnum = 4096, 6, 2, 3  # alpha = 2
    ctx = get_ctx(N, L, K, dnum)
    o = get_orc(orc, ctx)
    rng = np.random.default_rng(10 * sizeQl + key_limbs)
    W = sizeQl * N
    kidx = list(range(key_limbs)) + [L + j for j in range(K)]
    key = np.concatenate([rand_poly(rng, [ctx.primes[i] for i in kidx], N).reshape(-1)
                          for _ in range(2 * dnum)])
    evk = 1024
    ctx.write_key(evk, key)
    c = rand_poly(rng, ctx.q[:sizeQl], N)
    ctx.write_poly(0, c)
    say("keyswitch", sizeQl, key_limbs)
    assert ctx.op(7, 0, 0, 8 * W, evk, sizeQl, key_limbs) == 0
    km = np.zeros(evk + key.size, np.uint64)
    km[evk:] = key
    pm = np.zeros(12 * W, np.uint64)
    pm[:W] = c.reshape(-1)
    assert o.accel_op(pm, km, 7, 0, 0, 8 * W, evk, sizeQl, key_limbs) == 0
    assert np.array_equal(ctx.read_poly(8 * W, 2 * W), pm[8 * W:10 * W])
    # in place: ks0 over the input
    assert ctx.op(7, 0, 0, 0, evk, sizeQl, key_limbs) == 0
    assert np.array_equal(ctx.read_poly(0, 2 * W), pm[8 * W:10 * W])


def test_status_codes_gpu():
    ctx = get_ctx(8192, 4, 1, 2)
    N = ctx.N
    assert ctx.op(42, 0, 0, 0, 0, 2) == -1
    assert ctx.op(1, 3, 0, 0, 0, 2) == -3
    assert ctx.op(1, ctx.poly_words - N, 0, 0, 0, 2) == -2
    assert ctx.op(1, 0, 0, 0, 0, 0) == -4
    assert ctx.op(8, 0, 8 * N, 16 * N, 0, 1) == -4
    assert ctx.op(7, 0, 0, 8 * N, ctx.key_words - 8, 2) == -2
    assert ctx.op(99, 0, 0, 0, 0, 4) == 0


def _c1_setup(ctx, seed=42):
    """C1 keys: seeded ternary secret, relin key, rotation key for +1."""
    N, L = ctx.N, ctx.L
    s = ctx.keygen_secret(seed)
    relin = 0
    ctx.keygen_relin_key(s, L, seed + 1, relin)
    rot = ctx.key_size_words(L)
    g1 = ctx.galois_for_step(1)
    ctx.keygen_rotation_key(s, g1, L, seed + 2, rot)
    return s, relin, rot, g1


def test_c1_sequence_bit_exact(orc):
    """C1: N=2^13, L=4, K=1, dnum=2; eval_sub_plain -> eval_mult_relin_rescale
    -> eval_rotate(+1), uniform residues, bit-exact (main_sim.cpp CP2-CP5 order)."""
    from anns_fhe_amd.driver import FHEHostDriver
    N, L = 8192, 4
    ctx = get_ctx(N, L, 1, 2)
    o = get_orc(orc, ctx)
    drv = FHEHostDriver(ctx)
    s, relin, rot, g1 = _c1_setup(ctx)
    drv.set_relin_key(relin, L)
    drv.add_rotation_key(1, g1, rot, L)
    km = ctx.read_key(0, 2 * rot)
    rng = np.random.default_rng(42)
    W = L * N
    ct0, ct1, pt = rand_poly(rng, ctx.q, N), rand_poly(rng, ctx.q, N), rand_poly(rng, ctx.q, N)
    pm = np.zeros(12 * W, np.uint64)
    pm[:W], pm[W:2 * W], pm[2 * W:3 * W] = ct0.reshape(-1), ct1.reshape(-1), pt.reshape(-1)
    ctx.write_poly(0, pm[:3 * W])
    say("c1 sequence")
    drv.eval_sub_plain(0, W, 2 * W, 3 * W, 4 * W, L)
    drv.eval_mult_relin_rescale(3 * W, 4 * W, 3 * W, 4 * W, 5 * W, 6 * W, L)
    drv.eval_rotate(5 * W, 6 * W, 7 * W, 8 * W, 1, L - 1)
    got = ctx.read_poly(0, 9 * W)
    ref = pm.copy()
    assert o.eval_sub_plain(ref, 0, W, 2 * W, 3 * W, 4 * W, L) == 0
    assert np.array_equal(got[3 * W:5 * W], ref[3 * W:5 * W])
    assert o.eval_mult_relin_rescale(ref, km, relin, L, 3 * W, 4 * W, 3 * W, 4 * W, 5 * W, 6 * W, L) == 0
    W1 = (L - 1) * N
    assert np.array_equal(got[5 * W:5 * W + W1], ref[5 * W:5 * W + W1])
    assert np.array_equal(got[6 * W:6 * W + W1], ref[6 * W:6 * W + W1])
    assert o.eval_rotate(ref, km, [(1, g1, rot, L)], 5 * W, 6 * W, 7 * W, 8 * W, 1, L - 1) == 0
    assert np.array_equal(got[7 * W:7 * W + W1], ref[7 * W:7 * W + W1])
    assert np.array_equal(got[8 * W:8 * W + W1], ref[8 * W:8 * W + W1])


def test_rotate_decomposition_and_missing_key(orc):
    from anns_fhe_amd.driver import FHEHostDriver, FheError
    N, L = 4096, 3
    ctx = get_ctx(N, L, 1, 3)
    o = get_orc(orc, ctx)
    drv = FHEHostDriver(ctx)
    s = ctx.keygen_secret(5)
    ksz = ctx.key_size_words(L)
    keys = []
    for i, step in enumerate([1, 2, -1, -4]):
        g = ctx.galois_for_step(step)
        ctx.keygen_rotation_key(s, g, L, 100 + i, i * ksz)
        drv.add_rotation_key(step, g, i * ksz, L)
        keys.append((step, g, i * ksz, L))
    km = ctx.read_key(0, 4 * ksz)
    rng = np.random.default_rng(8)
    W = L * N
    c0, c1 = rand_poly(rng, ctx.q, N), rand_poly(rng, ctx.q, N)
    pm = np.zeros(8 * W, np.uint64)
    pm[:W], pm[W:2 * W] = c0.reshape(-1), c1.reshape(-1)
    for rot in (3, -5, 0):
        say("rotate", rot)
        ctx.write_poly(0, pm[:2 * W])
        drv.eval_rotate(0, W, 2 * W, 3 * W, rot, L)
        ref = pm.copy()
        assert o.eval_rotate(ref, km, keys, 0, W, 2 * W, 3 * W, rot, L) == 0
        assert np.array_equal(ctx.read_poly(2 * W, 2 * W), ref[2 * W:4 * W])
    with pytest.raises(FheError) as ei:
        drv.eval_rotate(0, W, 2 * W, 3 * W, 8, L)
    assert ei.value.status == -5


def test_batched_submit_matches_single_calls(orc):
    from anns_fhe_amd.driver import FHEHostDriver
    N, L = 4096, 4
    ctx = get_ctx(N, L, 2, 2)
    drv = FHEHostDriver(ctx)
    s = ctx.keygen_secret(9)
    ctx.keygen_relin_key(s, L, 10, 0)
    drv.set_relin_key(0, L)
    ksz = ctx.key_size_words(L)
    gs = [ctx.galois_for_step(k) for k in (1, 3, 7)]
    for i, g in enumerate(gs):
        ctx.keygen_rotation_key(s, g, L, 20 + i, (i + 1) * ksz)
    rng = np.random.default_rng(11)
    W = L * N
    n = 5
    data = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(2 * n)])
    ctx.write_poly(0, data)
    say("batched mult")
    works = [dict(kind=1, limbs=L, a0=2 * i * W, a1=(2 * i + 1) * W, b0=2 * ((i + 1) % n) * W,
                  b1=(2 * ((i + 1) % n) + 1) * W, d0=(20 + 2 * i) * W, d1=(21 + 2 * i) * W) for i in range(n)]
    ctx.submit(works)
    batched = ctx.read_poly(20 * W, 2 * n * W)
    for w in works:
        drv.eval_mult_relin_rescale(w["a0"], w["a1"], w["b0"], w["b1"], w["d0"], w["d1"], L)
    single = ctx.read_poly(20 * W, 2 * n * W)
    assert np.array_equal(batched, single)
    say("batched automorphism-ks")
    works = [dict(kind=2, limbs=L, key_limbs=L, a0=2 * i * W, a1=(2 * i + 1) * W, d0=(40 + 2 * i) * W,
                  d1=(41 + 2 * i) * W, galois=gs[i % 3], evk=((i % 3) + 1) * ksz) for i in range(n)]
    ctx.submit(works)
    batched = ctx.read_poly(40 * W, 2 * n * W)
    for w in works:
        drv.eval_automorphism_ks(w["galois"], w["evk"], L, w["a0"], w["a1"], w["d0"], w["d1"], L)
    single = ctx.read_poly(40 * W, 2 * n * W)
    assert np.array_equal(batched, single)
    # and one item against the oracle
    o = get_orc(orc, ctx)
    km = ctx.read_key(0, 4 * ksz)
    pm = np.zeros(50 * W, np.uint64)
    pm[:2 * n * W] = data
    w = works[1]
    assert o.eval_automorphism_ks(pm, km, w["galois"], w["evk"], L, w["a0"], w["a1"], w["d0"], w["d1"], L) == 0
    assert np.array_equal(single[2 * W:4 * W], pm[w["d0"]:w["d0"] + 2 * W])


def test_keygen_keys_decrypt(orc):
    """The engine's own keys: Dec(KS_relin(d)) ~= d * s^2 (product client pinned
    by the decryption invariant, evaluated in Python big integers)."""
    N, L = 1024, 3
    ctx = get_ctx(N, L, 1, 3)
    o = get_orc(orc, ctx)
    s = ctx.keygen_secret(77)
    assert set(np.unique(s)) <= {-1, 0, 1}
    ctx.keygen_relin_key(s, L, 78, 0)
    rng = np.random.default_rng(12)
    W = L * N
    d = rand_poly(rng, ctx.q, N)
    ctx.write_poly(0, d)
    assert ctx.op(7, 0, 0, 4 * W, 0, L, L) == 0
    ks = ctx.read_poly(4 * W, 2 * W).reshape(2, L, N)
    got = ck.decrypt_coeffs(o, s.astype(np.int64), ks[0], ks[1])
    from tests.bigint_ref import centered, crt, prod
    dc = o.intt_poly(d)
    Q = prod(ctx.q)
    dint = [crt([dc[i][j] for i in range(L)], ctx.q)[0] for j in range(N)]
    s2 = ck.negacyclic(s.astype(np.int64), s.astype(np.int64))
    # d * s^2 via two negacyclic products with exact integers
    exp = [centered(v, Q) for v in ck.negacyclic(dint, s2)]
    err = max(abs(centered(g - e, Q)) for g, e in zip(got, exp))
    assert err < 2 ** 24, err


def test_full_size_hmult_bit_exact(orc):
    """One eval_mult_relin_rescale at N=2^16, L=24, K=8, dnum=3 against the oracle."""
    N, L, K, dnum = 65536, 24, 8, 3
    W = L * N
    ctx = get_ctx(N, L, K, dnum, poly_words=8 * W, key_words=2 * dnum * (L + K) * N + 64)
    from anns_fhe_amd.driver import FHEHostDriver
    drv = FHEHostDriver(ctx)
    rng = np.random.default_rng(2024)
    kidx = list(range(L)) + [L + j for j in range(K)]
    key = np.concatenate([rand_poly(rng, [ctx.primes[i] for i in kidx], N).reshape(-1) for _ in range(2 * dnum)])
    ctx.write_key(0, key)
    drv.set_relin_key(0, L)
    data = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(4)])
    ctx.write_poly(0, data)
    say("full-size hmult gpu")
    drv.eval_mult_relin_rescale(0, W, 2 * W, 3 * W, 4 * W, 5 * W, L)
    W1 = (L - 1) * N
    got0 = ctx.read_poly(4 * W, W1)
    got1 = ctx.read_poly(5 * W, W1)
    say("full-size hmult oracle")
    o = get_orc(orc, ctx)
    pm = np.zeros(6 * W, np.uint64)
    pm[:4 * W] = data
    assert o.eval_mult_relin_rescale(pm, key, 0, L, 0, W, 2 * W, 3 * W, 4 * W, 5 * W, L) == 0
    assert np.array_equal(got0, pm[4 * W:4 * W + W1])
    assert np.array_equal(got1, pm[5 * W:5 * W + W1])


def test_cpp_host_driver_c1_sequence(orc, tmp_path):
    """The C++ FHEHostDriver over the C-ABI (anns_fhe_amd/c1_sequence) runs the
    main_sim flow; every output word must equal the oracle's."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "anns_fhe_amd", "c1_sequence")
    out = str(tmp_path / "c1.bin")
    say("c++ host driver")
    r = subprocess.run([exe, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, dtype=np.uint64)
    N, L, K, dnum = (int(x) for x in raw[:4])
    pos = 4
    primes = [int(x) for x in raw[pos:pos + L + K]]
    pos += L + K
    W, W1 = L * N, (L - 1) * N
    inp = raw[pos:pos + 3 * W]
    pos += 3 * W
    ksz = 2 * dnum * (L + K) * N
    km = raw[pos:pos + 2 * ksz].copy()
    pos += 2 * ksz
    sub, mult, rot = raw[pos:pos + 2 * W], raw[pos + 2 * W:pos + 2 * W + 2 * W1], raw[pos + 2 * W + 2 * W1:]
    assert primes == orc.gen_primes(2 * N, 50, L + K)
    o = orc.OracleCtx(N, primes[:L], primes[L:], dnum)
    pm = np.zeros(12 * W, np.uint64)
    pm[:3 * W] = inp
    assert o.eval_sub_plain(pm, 0, W, 2 * W, 3 * W, 4 * W, L) == 0
    assert np.array_equal(sub, pm[3 * W:5 * W])
    assert o.eval_mult_relin_rescale(pm, km, 0, L, 3 * W, 4 * W, 3 * W, 4 * W, 5 * W, 6 * W, L) == 0
    assert np.array_equal(mult[:W1], pm[5 * W:5 * W + W1])
    assert np.array_equal(mult[W1:], pm[6 * W:6 * W + W1])
    g1 = pow(5, 1, 2 * N)
    assert o.eval_rotate(pm, km, [(1, g1, ksz, L)], 5 * W, 6 * W, 7 * W, 8 * W, 1, L - 1) == 0
    assert np.array_equal(rot[:W1], pm[7 * W:7 * W + W1])
    assert np.array_equal(rot[W1:], pm[8 * W:8 * W + W1])


@pytest.mark.parametrize("q_bits", [50, 60, "mixed"])
def test_linear_sum_parity(orc, q_bits):
    """fhea_eval_linear_sum (FP64 path at 50-bit primes, integer path at 60-bit)
    against the oracle's OP 5 / OP 3 composition, with bias and per-output plaintexts."""
    from anns_fhe_amd._lib import NONE, FheError
    N, L = 4096, 3
    ctx = get_ctx(N, L, 1, 3, q_bits=q_bits)
    o = get_orc(orc, ctx)
    rng = np.random.default_rng(seed_of(q_bits))
    W = L * N
    R, G = 20, 11
    n_slots = 2 * R + G * R + G + 2 + 2 * G
    data = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(2 * R + G * R + G + 2)])
    data[:N] = ctx.q[0] - 1
    assert n_slots * W <= ctx.poly_words
    ctx.write_poly(0, data)
    ct0 = [2 * r * W for r in range(R)]
    ct1 = [(2 * r + 1) * W for r in range(R)]
    pt = np.array([[(2 * R + g * R + r) * W for r in range(R)] for g in range(G)], np.uint64)
    base = 2 * R + G * R
    ptb = [(base + g) * W if g % 3 else NONE for g in range(G)]
    bias = ((base + G) * W, (base + G + 1) * W)
    out0 = [(base + G + 2 + 2 * g) * W for g in range(G)]
    out1 = [(base + G + 3 + 2 * g) * W for g in range(G)]
    say("linear sum", q_bits)
    ctx.linear_sum(ct0, ct1, pt, out0, out1, L, ptb=ptb, bias=bias)
    got = ctx.read_poly(out0[0], 2 * G * W)
    pm = np.zeros(n_slots * W, np.uint64)
    pm[:data.size] = data
    assert o.linear_sum(pm, ct0, ct1, pt, out0, out1, L, ptb=ptb, bias=bias) == 0
    assert np.array_equal(got, pm[out0[0]:out0[0] + 2 * G * W])
    # no bias, one output, two limbs of the same words
    ctx.linear_sum(ct0, ct1, pt[:1], out0[:1], out1[:1], 2)
    assert o.linear_sum(pm, ct0, ct1, pt[:1], out0[:1], out1[:1], 2) == 0
    assert np.array_equal(ctx.read_poly(out0[0], 2 * N), pm[out0[0]:out0[0] + 2 * N])
    # an output overlapping an input is refused
    with pytest.raises(FheError) as ei:
        ctx.linear_sum(ct0, ct1, pt[:1], [ct0[3]], out1[:1], L)
    assert ei.value.status == -4


def test_submit_mult_relin_and_add_parity(orc):
    from anns_fhe_amd.driver import FHEHostDriver
    N, L = 4096, 4
    ctx = get_ctx(N, L, 2, 2)
    drv = FHEHostDriver(ctx)
    o = get_orc(orc, ctx)
    rng = np.random.default_rng(21)
    kidx = list(range(L)) + [L + j for j in range(ctx.K)]
    key = np.concatenate([rand_poly(rng, [ctx.primes[i] for i in kidx], N).reshape(-1) for _ in range(4)])
    ctx.write_key(0, key)
    drv.set_relin_key(0, L)
    W = L * N
    data = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(8)])
    ctx.write_poly(0, data)
    for limbs in (L, 2):
        works = [dict(kind=5, limbs=limbs, a0=0, a1=W, b0=2 * W, b1=3 * W, d0=8 * W, d1=9 * W),
                 dict(kind=5, limbs=limbs, a0=4 * W, a1=5 * W, b0=6 * W, b1=7 * W, d0=10 * W, d1=11 * W),
                 dict(kind=6, limbs=limbs, a0=8 * W, a1=9 * W, b0=10 * W, b1=11 * W, d0=12 * W, d1=13 * W)]
        ctx.submit(works)
        pm = np.zeros(14 * W, np.uint64)
        pm[:8 * W] = data
        assert o.eval_mult_relin(pm, key, 0, L, 0, W, 2 * W, 3 * W, 8 * W, 9 * W, limbs) == 0
        assert o.eval_mult_relin(pm, key, 0, L, 4 * W, 5 * W, 6 * W, 7 * W, 10 * W, 11 * W, limbs) == 0
        assert o.eval_add(pm, 8 * W, 9 * W, 10 * W, 11 * W, 12 * W, 13 * W, limbs) == 0
        for off in (8, 9, 10, 11, 12, 13):
            assert np.array_equal(ctx.read_poly(off * W, limbs * N), pm[off * W:off * W + limbs * N]), (limbs, off)


@pytest.mark.parametrize("q_bits,N", [(50, 4096), (60, 4096), ("mixed", 4096), (50, 16384)])
def test_hoisted_rotations_parity(orc, q_bits, N):
    """Work kind 7 (runs of rotations of the same input share one ModUp) against the
    oracle's hoisted rotations, two inputs in one submit, both arithmetic paths (at
    N=2^14 the fused ModDown row pass reads sigma_g(c0) itself)."""
    L = 4
    ctx = get_ctx(N, L, 2, 2, q_bits=q_bits)
    o = get_orc(orc, ctx)
    rng = np.random.default_rng(seed_of(q_bits) + 1)
    kidx = list(range(L)) + [L + j for j in range(ctx.K)]
    ksz = ctx.key_size_words(L)
    n_keys = 5
    km = np.concatenate([rand_poly(rng, [ctx.primes[i] for i in kidx], N).reshape(-1)
                         for _ in range(2 * ctx.dnum * n_keys)])
    ctx.write_key(0, km)
    W = L * N
    data = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(4)])
    ctx.write_poly(0, data)
    gal = [ctx.galois_for_step(s) for s in (1, 2, 7, -3, 11)]
    works = []
    for inp in range(2):
        for z in range(5 - inp * 2):
            k = len(works)
            works.append(dict(kind=7, limbs=L, key_limbs=L, a0=2 * inp * W, a1=(2 * inp + 1) * W,
                              d0=(8 + 2 * k) * W, d1=(9 + 2 * k) * W, galois=gal[z], evk=z * ksz))
    say("hoisted", q_bits, len(works))
    ctx.submit(works)
    got = ctx.read_poly(8 * W, 2 * len(works) * W)
    pm = np.zeros((8 + 2 * len(works)) * W, np.uint64)
    pm[:4 * W] = data
    for inp in range(2):
        ws = [w for w in works if w["a0"] == 2 * inp * W]
        assert o.eval_rotations_hoisted(pm, km, ws[0]["a0"], ws[0]["a1"], [w["galois"] for w in ws],
                                        [w["evk"] for w in ws], L, [w["d0"] for w in ws], [w["d1"] for w in ws],
                                        L) == 0
    assert np.array_equal(got, pm[8 * W:])
    # a run where a later input is an earlier output (rotation of a rotation): the
    # groups then run one after the other
    base = 8 + 2 * len(works)
    chain = [dict(kind=7, limbs=L, key_limbs=L, a0=0, a1=W, d0=(base + 2 * k) * W, d1=(base + 2 * k + 1) * W,
                  galois=gal[k], evk=k * ksz) for k in range(2)]
    chain += [dict(kind=7, limbs=L, key_limbs=L, a0=base * W, a1=(base + 1) * W, d0=(base + 4 + 2 * k) * W,
                   d1=(base + 5 + 2 * k) * W, galois=gal[2 + k], evk=(2 + k) * ksz) for k in range(2)]
    ctx.submit(chain)
    got = ctx.read_poly(base * W, 8 * W)
    pm = np.concatenate([pm, np.zeros(8 * W, np.uint64)])
    for grp in (chain[:2], chain[2:]):
        assert o.eval_rotations_hoisted(pm, km, grp[0]["a0"], grp[0]["a1"], [w["galois"] for w in grp],
                                        [w["evk"] for w in grp], L, [w["d0"] for w in grp],
                                        [w["d1"] for w in grp], L) == 0
    assert np.array_equal(got, pm[base * W:(base + 8) * W])


@pytest.mark.parametrize("q_bits", [50, 61])
def test_ntt_splits_and_tuning_knobs(orc, q_bits):
    """Every NTT split the tuner may choose (N=2^16: first pass of 6..10 stages) gives
    the oracle's transform word for word, on the FP64 (50-bit) and the integer
    (61-bit) kernels; invalid knob values are refused; chunk size does not change a
    batched result."""
    from anns_fhe_amd import _lib
    from anns_fhe_amd._lib import FheError
    from anns_fhe_amd import autotune as AT
    N, L = 65536, 3
    ctx = get_ctx(N, L, 1, 1, q_bits=q_bits)
    o = get_orc(orc, ctx)
    rng = np.random.default_rng(77)
    a = rand_poly(rng, ctx.q, N)
    W = L * N
    ctx.write_poly(0, a)
    fwd, inv = o.ntt_poly(a), o.intt_poly(a)
    for k in AT.ntt_splits(16):
        say("split", k, q_bits)
        ctx.set_tuning(_lib.TUNE_NTT_SPLIT_FWD, k)
        ctx.set_tuning(_lib.TUNE_NTT_SPLIT_INV, k)
        for w in (0, 2):
            ctx.set_tuning(_lib.TUNE_NTT_WG_PER_CU, w)
            ctx.ntt_batch(0, W, 1, L, inverse=False)
            assert np.array_equal(ctx.read_poly(W, W).reshape(L, N), fwd), (k, w)
            ctx.ntt_batch(0, 2 * W, 1, L, inverse=True)
            assert np.array_equal(ctx.read_poly(2 * W, W).reshape(L, N), inv), (k, w)
    for bad in ((_lib.TUNE_NTT_SPLIT_FWD, 5), (_lib.TUNE_NTT_SPLIT_INV, 11), (_lib.TUNE_NTT_WG_PER_CU, 5),
                (_lib.TUNE_CHUNK, 0), (99, 1)):
        with pytest.raises(FheError):
            ctx.set_tuning(*bad)
    for k in (_lib.TUNE_NTT_SPLIT_FWD, _lib.TUNE_NTT_SPLIT_INV, _lib.TUNE_NTT_WG_PER_CU):
        ctx.set_tuning(k, 0)
    # the tuner itself on a small batch: it checks every candidate against the default
    t = AT.tune_ntt(ctx, 3 * W, 1, L, reps=1)
    assert any(key.startswith("ntt_split_fwd") for key in t)
    for k in (_lib.TUNE_NTT_SPLIT_FWD, _lib.TUNE_NTT_SPLIT_INV, _lib.TUNE_NTT_WG_PER_CU):
        ctx.set_tuning(k, 0)


def test_workload_tuner_keeps_results(orc):
    """The workload tuner times a batched eval_mult_relin_rescale under every
    schedule candidate (NTT splits, workgroups per CU, chunk sizes), checks the
    output words each time, and the chosen schedule still gives the oracle's words."""
    from anns_fhe_amd import autotune as AT
    from anns_fhe_amd.driver import FHEHostDriver
    N, L = 16384, 4
    ctx = get_ctx(N, L, 2, 2)
    drv = FHEHostDriver(ctx)
    s = ctx.keygen_secret(21)
    ctx.keygen_relin_key(s, L, 22, 0)
    drv.set_relin_key(0, L)
    rng = np.random.default_rng(23)
    W, n = L * N, 6
    data = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(4 * n)])
    ctx.write_poly(0, data)
    works = [dict(kind=1, limbs=L, a0=4 * i * W, a1=(4 * i + 1) * W, b0=(4 * i + 2) * W, b1=(4 * i + 3) * W,
                  d0=(4 * n + 2 * i) * W, d1=(4 * n + 2 * i + 1) * W) for i in range(n)]
    say("tune")
    t = AT.tune(ctx, lambda: ctx.submit(works), check=lambda: ctx.read_poly(4 * n * W, 2 * n * W), reps=1)
    assert len(t) >= 4 + 2 * len(AT.ntt_splits(14))
    ctx.submit(works)
    got = ctx.read_poly(4 * n * W, 2 * W)
    o = get_orc(orc, ctx)
    pm = np.zeros((4 * n + 2) * W, np.uint64)
    pm[:4 * n * W] = data
    km = ctx.read_key(0, ctx.key_size_words(L))
    w = works[0]
    assert o.eval_mult_relin_rescale(pm, km, 0, L, w["a0"], w["a1"], w["b0"], w["b1"], 4 * n * W,
                                     (4 * n + 1) * W, L) == 0
    assert np.array_equal(got[:(L - 1) * N], pm[4 * n * W:4 * n * W + (L - 1) * N])
    assert np.array_equal(got[W:W + (L - 1) * N], pm[(4 * n + 1) * W:(4 * n + 1) * W + (L - 1) * N])
    assert any(k.startswith("streams=2") for k in t)  # the two-lane schedule was run and checked
    AT.apply(ctx, {"ntt_split_fwd": 0, "ntt_split_inv": 0, "ntt_wg_per_cu": 0, "chunk": 16, "streams": 1})


def test_async_reads_overlap_later_work():
    """fhea_write_polys puts back-to-back words into scattered ranges;
    fhea_read_polys_async: gathered reads on the copy stream equal the synchronous
    read, also while later kernels (NTTs on other words) run; tickets may be waited
    out of order; a ticket more than 16 reads old returns at once (it is complete),
    one not yet issued is refused."""
    from anns_fhe_amd._lib import FheError
    from anns_fhe_amd.driver import PinnedArray
    N, L = 8192, 3
    ctx = get_ctx(N, L, 1, 1)
    rng = np.random.default_rng(41)
    W = L * N
    data = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(8)])
    ctx.write_poly(0, data)
    ranges = [(0, 100), (3 * W + 17, W), (5 * W, 2 * W), (W + 8, 1)]
    ref = ctx.read_polys(ranges)
    words = rng.integers(0, 1 << 40, ref.size, dtype=np.uint64)
    ctx.write_polys(ranges, words)
    assert np.array_equal(ctx.read_polys(ranges), words)
    ctx.write_polys(ranges, ref)
    pin = PinnedArray(4 * ref.size)
    tickets = []
    for k in range(3):
        say("async read", k)
        tickets.append(ctx.read_polys_async(ranges, pin.array[k * ref.size:(k + 1) * ref.size]))
        ctx.ntt_batch(8 * W, 8 * W, 2, L)  # later work on words no read touches
    for k in (2, 0, 1):
        ctx.read_wait(tickets[k])
        assert np.array_equal(pin.array[k * ref.size:(k + 1) * ref.size], ref), k
    with pytest.raises(FheError):
        ctx.read_wait(tickets[-1] + 1)
    for _ in range(17):
        t = ctx.read_polys_async(ranges[:1], pin.array[3 * ref.size:])
    ctx.read_wait(t)
    ctx.read_wait(tickets[0])  # more than 16 reads old: complete already
    pin.close()


@pytest.mark.parametrize("q_bits", [50, 60, "mixed"])
def test_linear_sum_batch_and_rescale_scatter(orc, q_bits):
    """fhea_eval_linear_sum_batch (jobs of different G sharing plaintext blocks, with
    and without bias) equals the oracle's job-by-job sums; an output of one job
    overlapping an input of another is refused; fhea_rescale_scatter equals the
    oracle's rescale of every ciphertext."""
    from anns_fhe_amd._lib import NONE, FheError
    N, L = 4096, 3
    ctx = get_ctx(N, L, 1, 3, q_bits=q_bits)
    o = get_orc(orc, ctx)
    rng = np.random.default_rng(7 + seed_of(q_bits))
    W = L * N
    R, n_blocks = 9, 6
    n_in = 3 * 2 * R + n_blocks * R + n_blocks + 2
    data = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(n_in)])
    ctx.write_poly(0, data)
    blk = lambda b: [(3 * 2 * R + b * R + r) * W for r in range(R)]
    ptb_of = lambda b: (3 * 2 * R + n_blocks * R + b) * W if b % 2 else NONE
    bias = ((n_in - 2) * W, (n_in - 1) * W)
    out = n_in
    jobs = []
    for j, (blocks, bs) in enumerate((([0], bias), ([1, 2, 3, 4, 5], None), ([5, 4, 3, 2, 1, 0, 1, 2, 3, 4, 5], bias))):
        G = len(blocks)
        jobs.append(dict(ct0=[(2 * j * R + 2 * r) * W for r in range(R)],
                         ct1=[(2 * j * R + 2 * r + 1) * W for r in range(R)],
                         pt=np.array([blk(b) for b in blocks], np.uint64), ptb=[ptb_of(b) for b in blocks],
                         out0=[(out + 2 * g) * W for g in range(G)], out1=[(out + 2 * g + 1) * W for g in range(G)],
                         bias=bs))
        out += 2 * G
    say("linear sum batch", q_bits)
    ctx.linear_sum_batch(jobs, L)
    got = ctx.read_poly(n_in * W, (out - n_in) * W)
    pm = np.zeros((out + 40) * W, np.uint64)
    pm[:data.size] = data
    for jb in jobs:
        assert o.linear_sum(pm, jb["ct0"], jb["ct1"], jb["pt"], jb["out0"], jb["out1"], L, ptb=jb["ptb"],
                            bias=jb["bias"]) == 0
    assert np.array_equal(got, pm[n_in * W:out * W])
    bad = [dict(jobs[0]), dict(jobs[1])]
    bad[0]["out0"] = [jobs[1]["ct0"][4]]
    with pytest.raises(FheError) as ei:
        ctx.linear_sum_batch(bad, L)
    assert ei.value.status == -4
    # rescale of the 17 outputs (stored back to back) to scattered places
    n = (out - n_in) // 2
    W1 = (L - 1) * N
    dst = [(out + 1) * W + ((7 * z) % n) * 2 * W1 for z in range(n)]
    ctx.rescale_scatter(n_in * W, dst, L)
    for z in range(n):
        for k in range(2):
            a = pm[(n_in + 2 * z + k) * W:(n_in + 2 * z + k + 1) * W]
            assert np.array_equal(ctx.read_poly(dst[z] + k * W1, W1), o.rescale(a).reshape(-1)), (z, k)
    with pytest.raises(FheError) as ei:
        ctx.rescale_scatter(n_in * W, [dst[0], dst[0] + 8], L)
    assert ei.value.status == -4


@pytest.mark.parametrize("q_bits", [50, 60, "mixed"])
def test_linear_sum_groups_parity(orc, q_bits):
    """fhea_eval_linear_sum_groups: groups of 5x6, 1x3, 4x4 and 3x9 (query x block),
    queries and blocks shared between groups, absent biases and block biases; every
    pair equals the oracle's linear sum (at 50-bit primes FP64: the 5x6 group as one
    LDS unit, the others as 2x8 tiles, partial in both directions; integer 1x8 units
    at 60)."""
    from anns_fhe_amd._lib import NONE
    N, L = 4096, 3
    ctx = get_ctx(N, L, 1, 3, q_bits=q_bits)
    o = get_orc(orc, ctx)
    rng = np.random.default_rng(70 + seed_of(q_bits))
    W = L * N
    R, n_q, n_b = 7, 6, 9
    n_in = n_q * 2 * R + n_b * R + n_b + 2 * n_q
    data = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(n_in)])
    ctx.write_poly(0, data)
    ctq = lambda z, c: [(z * 2 * R + 2 * r + c) * W for r in range(R)]
    blk = lambda b: [(n_q * 2 * R + b * R + r) * W for r in range(R)]
    ptb = lambda b: (n_q * 2 * R + n_b * R + b) * W if b % 3 else NONE
    bias = lambda z: ((n_q * 2 * R + n_b * R + n_b + 2 * z) * W, (n_q * 2 * R + n_b * R + n_b + 2 * z + 1) * W)
    out = n_in
    groups, pairs = [], []
    for zs, bs in (([0, 1, 2, 3, 4], [0, 1, 2, 3, 4, 5]), ([5], [6, 0, 2]), ([2, 3, 5, 1], [3, 4, 5, 6]),
                   ([4, 0, 3], [8, 7, 6, 5, 4, 3, 2, 1, 0])):
        o0 = np.array([[(out + 2 * (i * len(bs) + g)) * W for g in range(len(bs))] for i in range(len(zs))], np.uint64)
        groups.append(dict(ct0=[ctq(z, 0) for z in zs], ct1=[ctq(z, 1) for z in zs], pt=[blk(b) for b in bs],
                           ptb=[ptb(b) for b in bs], bias=[bias(z) if z != 4 else (NONE, NONE) for z in zs],
                           out0=o0, out1=o0 + W))
        pairs += [(z, bs, o0[i], o0[i] + W) for i, z in enumerate(zs)]
        out += 2 * len(zs) * len(bs)
    say("linear sum groups", q_bits)
    ctx.linear_sum_groups(groups, L)
    got = ctx.read_poly(n_in * W, (out - n_in) * W)
    pm = np.zeros(out * W, np.uint64)
    pm[:data.size] = data
    for z, bs, o0, o1 in pairs:
        assert o.linear_sum(pm, ctq(z, 0), ctq(z, 1), np.array([blk(b) for b in bs], np.uint64), o0, o1, L,
                            ptb=[ptb(b) for b in bs], bias=bias(z) if z != 4 else None) == 0
    assert np.array_equal(got, pm[n_in * W:out * W])


@pytest.mark.parametrize("N", [8192, 16384])
def test_single_pass_limb_ntt(orc, N):
    """N = 2^13, 2^14 with the split knob at logN (one pass over the whole limb, the
    limb in LDS): NTT / INTT, the fused rescale (OP 8) and eval_mult_relin_rescale
    (fused ModDown + rescale ends) give the oracle's words."""
    from anns_fhe_amd import _lib
    from anns_fhe_amd.driver import FHEHostDriver
    L = 4
    lg = N.bit_length() - 1
    ctx = get_ctx(N, L, 2, 2)
    o = get_orc(orc, ctx)
    rng = np.random.default_rng(N + 5)
    W = L * N
    try:
        ctx.set_tuning(_lib.TUNE_NTT_SPLIT_FWD, lg)
        ctx.set_tuning(_lib.TUNE_NTT_SPLIT_INV, lg)
        a = rand_poly(rng, ctx.q, N)
        ctx.write_poly(0, a)
        say("single-pass ntt", N)
        assert ctx.op(1, 0, 0, W, 0, L) == 0
        assert np.array_equal(ctx.read_poly(W, W).reshape(L, N), o.ntt_poly(a))
        assert ctx.op(2, 0, 0, 2 * W, 0, L) == 0
        assert np.array_equal(ctx.read_poly(2 * W, W).reshape(L, N), o.intt_poly(a))
        # OP 8 rescale of a ciphertext (c0 at 0, c1 at 3W) -> 5W
        c1 = rand_poly(rng, ctx.q, N)
        ctx.write_poly(3 * W, c1)
        assert ctx.op(8, 0, 3 * W, 5 * W, 0, L) == 0
        W1 = (L - 1) * N
        assert np.array_equal(ctx.read_poly(5 * W, W1).reshape(L - 1, N), o.rescale(a.reshape(-1)))
        assert np.array_equal(ctx.read_poly(5 * W + W1, W1).reshape(L - 1, N), o.rescale(c1.reshape(-1)))
        # HMult with the fused ModDown + rescale
        drv = FHEHostDriver(ctx)
        s = ctx.keygen_secret(5)
        ctx.keygen_relin_key(s, L, 6, 0)
        drv.set_relin_key(0, L)
        ins = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(4)])
        ctx.write_poly(8 * W, ins)
        drv.eval_mult_relin_rescale(8 * W, 9 * W, 10 * W, 11 * W, 12 * W, 13 * W, L)
        got = ctx.read_poly(12 * W, 2 * W)
        pm = np.zeros(14 * W, np.uint64)
        pm[8 * W:12 * W] = ins
        km = ctx.read_key(0, ctx.key_size_words(L))
        assert o.eval_mult_relin_rescale(pm, km, 0, L, 8 * W, 9 * W, 10 * W, 11 * W, 12 * W, 13 * W, L) == 0
        assert np.array_equal(got[:W1], pm[12 * W:12 * W + W1])
        assert np.array_equal(got[W:W + W1], pm[13 * W:13 * W + W1])
    finally:
        ctx.set_tuning(_lib.TUNE_NTT_SPLIT_FWD, 0)
        ctx.set_tuning(_lib.TUNE_NTT_SPLIT_INV, 0)
