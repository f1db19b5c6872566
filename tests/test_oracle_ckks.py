"""Pins the oracle's CKKS-RNS ops (rescale, fast_bases_conv, approx_mod_down,
key_switch) and the composites against CRT closed forms and decryption
invariants (tb_rescale.cpp, tb_keyswitch.cpp, main_sim.cpp semantics)."""
import numpy as np
import pytest

from tests import ckks_ref as ck
from tests.bigint_ref import centered, crt, prod


def mkctx(orc, N, L, K, dnum, bits=50):
    ps = orc.gen_primes(2 * N, bits, L + K)
    return orc.OracleCtx(N, ps[:L], ps[L:], dnum)


@pytest.fixture(scope="module")
def ctx(orc):
    return mkctx(orc, 32, 5, 2, 3, bits=50)  # alpha = 2: digits (2, 2, 1)


def test_modup_tables_bigint(ctx):
    for sizeQl in (5, 4, 3):
        nparts = -(-sizeQl // ctx.alpha)
        for part in range(nparts):
            start = ctx.alpha * part
            sz = min(ctx.alpha, sizeQl - start)
            dig = ctx.q[start:start + sz]
            outp = [q for i, q in enumerate(ctx.q[:sizeQl]) if not (start <= i < start + sz)] + ctx.p
            inv, hat = ctx.modup_tables(sizeQl, part)
            Qp = prod(dig)
            for i, qi in enumerate(dig):
                assert int(inv[i]) == pow(Qp // qi, -1, qi)
                for j, pj in enumerate(outp):
                    assert int(hat[i, j]) == (Qp // qi) % pj


def test_moddown_and_rescale_tables_bigint(ctx):
    P = prod(ctx.p)
    for sizeQl in (5, 3):
        pinv, phinv, phat = ctx.moddown_tables(sizeQl)
        for i in range(sizeQl):
            assert int(pinv[i]) == pow(P, -1, ctx.q[i])
        for j, pj in enumerate(ctx.p):
            assert int(phinv[j]) == pow(P // pj, -1, pj)
            for i in range(sizeQl):
                assert int(phat[j, i]) == (P // pj) % ctx.q[i]
        # OpenFHE's definition: floor(Q' u / q_l) mod q_i, Q' = Q_l / q_l, u = Q'^-1 mod q_l
        qlql, qlinv = ctx.rescale_tables(sizeQl)
        ql = ctx.q[sizeQl - 1]
        Qp = prod(ctx.q[: sizeQl - 1])
        u = pow(Qp, -1, ql)
        for i in range(sizeQl - 1):
            assert int(qlql[i]) == ((Qp * u) // ql) % ctx.q[i]
            assert int(qlinv[i]) == pow(ql, -1, ctx.q[i])


def test_rescale_crt(ctx):
    rng = np.random.default_rng(5)
    for L in (5, 2):
        a = np.stack([rng.integers(0, q, ctx.N, dtype=np.uint64) for q in ctx.q[:L]])
        coeffs = ctx.intt_poly(a)
        ql = ctx.q[L - 1]
        expect = np.zeros((L - 1, ctx.N), np.uint64)
        for j in range(ctx.N):
            x, _ = crt([coeffs[i][j] for i in range(L)], ctx.q[:L])
            c = int(coeffs[L - 1][j])
            c = c - ql if c > ql // 2 else c  # rescale.cpp:46-54 centring
            y = (x - c) // ql
            assert (x - c) % ql == 0
            for i in range(L - 1):
                expect[i, j] = y % ctx.q[i]
        expect = ctx.ntt_poly(expect)
        assert np.array_equal(ctx.rescale(a), expect)


def test_fast_bases_conv_closed_form(orc, ctx):
    rng = np.random.default_rng(9)
    sizeQl, part = 5, 1
    start, sz = 2, 2
    dig = ctx.q[start:start + sz]
    outi = [0, 1, 4] + [ctx.L + j for j in range(ctx.K)]
    outp = [ctx.primes[i] for i in outi]
    inv, hat = ctx.modup_tables(sizeQl, part)
    x = np.stack([rng.integers(0, q, ctx.N, dtype=np.uint64) for q in dig])
    x[:, 0] = [q - 1 for q in dig]
    out = np.zeros((len(outp), ctx.N), np.uint64)
    idx_in = np.array([start, start + 1])
    orc.lib().orc_fast_bases_conv(
        ctx.N, orc.P64(x), orc.P64(out), sz, len(outp), orc.P64(np.array(dig, np.uint64)),
        orc.P64(np.array(outp, np.uint64)), orc.P64(inv), orc.P64(np.ascontiguousarray(hat.reshape(-1))),
        orc.P64(ctx.m_lo[idx_in].copy()), orc.P64(ctx.m_hi[idx_in].copy()), orc.P32(ctx.k[idx_in].copy()),
        orc.P64(ctx.m_lo[outi].copy()), orc.P64(ctx.m_hi[outi].copy()), orc.P32(ctx.k[outi].copy()))
    Qd = prod(dig)
    for r in range(ctx.N):
        s = sum(((int(x[i, r]) * pow(Qd // qi, -1, qi)) % qi) * (Qd // qi) for i, qi in enumerate(dig))
        for j, pj in enumerate(outp):
            assert int(out[j, r]) == s % pj


def test_approx_mod_down_closed_form(ctx):
    rng = np.random.default_rng(11)
    sizeQl = 4
    idxs = list(range(sizeQl)) + [ctx.L + j for j in range(ctx.K)]
    a = np.stack([rng.integers(0, ctx.primes[i], ctx.N, dtype=np.uint64) for i in idxs])
    got = ctx.approx_mod_down(a, sizeQl)
    xP = ctx.intt_poly(a[sizeQl:], [ctx.L + j for j in range(ctx.K)])
    P = prod(ctx.p)
    y = [sum(((int(xP[j, r]) * pow(P // pj, -1, pj)) % pj) * (P // pj) for j, pj in enumerate(ctx.p))
         for r in range(ctx.N)]
    for i in range(sizeQl):
        qi = ctx.q[i]
        t = ctx.ntt(np.array([v % qi for v in y], np.uint64), i)
        expect = [((int(a[i, r]) - int(t[r])) * pow(P, -1, qi)) % qi for r in range(ctx.N)]
        assert [int(v) for v in got[i]] == expect


def test_key_switch_noiseless_exact(ctx):
    """With s = 0 and no noise, b_p = [i in digit p] P s' so KS(d) = (d*s', .) exactly."""
    rng = np.random.default_rng(13)
    N = ctx.N
    for sizeQl in (5, 4):
        s_new = ck.ternary(rng, N)
        key = ck.make_ks_key(ctx, rng, np.zeros(N, np.int64), s_new, sizeQl, noise=False)
        d = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in ctx.q[:sizeQl]])
        o0, _ = ctx.key_switch(d, key, 0, sizeQl, sizeQl)
        sn = ck.small_to_limbs(ctx, s_new, range(sizeQl))
        for i in range(sizeQl):
            assert np.array_equal(o0[i], ck.mulmod(d[i], sn[i], ctx.q[i]))


@pytest.mark.parametrize("sizeQl,key_limbs", [(5, 5), (4, 5), (3, 3)])
def test_key_switch_decrypts(ctx, sizeQl, key_limbs):
    rng = np.random.default_rng(17 + sizeQl)
    N = ctx.N
    s = ck.ternary(rng, N)
    s_new = ck.ternary(rng, N)
    key = ck.make_ks_key(ctx, rng, s, s_new, key_limbs)
    d = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in ctx.q[:sizeQl]])
    o0, o1 = ctx.key_switch(d, key, 0, key_limbs, sizeQl)
    got = ck.decrypt_coeffs(ctx, s, o0, o1)
    # expected d * s_new (coefficient domain, mod Q_l)
    dc = ctx.intt_poly(d)
    Q = prod(ctx.q[:sizeQl])
    dint = [crt([dc[i][j] for i in range(sizeQl)], ctx.q[:sizeQl])[0] for j in range(N)]
    exp = [centered(v, Q) for v in ck.negacyclic(dint, s_new)]
    err = max(abs(centered(g - e, Q)) for g, e in zip(got, exp))
    assert err < 2 ** 16, err


def test_accel_op_status_codes(ctx):
    N = ctx.N
    pm = np.zeros(64 * N, np.uint64)
    assert ctx.accel_op(pm, None, 42, 0, 0, 0, 0, 2) == -1       # unknown opcode
    assert ctx.accel_op(pm, None, 1, 3, 0, 0, 0, 2) == -3        # unaligned offset
    assert ctx.accel_op(pm, None, 1, 63 * N, 0, 0, 0, 2) == -2   # out of range
    assert ctx.accel_op(pm, None, 1, 0, 0, 0, 0, 0) == -4        # zero limbs
    assert ctx.accel_op(pm, None, 8, 0, 8 * N, 16 * N, 0, 1) == -4  # rescale of 1 limb
    assert ctx.accel_op(pm, None, 1, 0, 0, 8 * N, 0, 2) == 0


def test_accel_ops_in_place_and_ntt_roundtrip(ctx):
    rng = np.random.default_rng(19)
    N, L = ctx.N, ctx.L
    pm = np.zeros(64 * N, np.uint64)
    a = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in ctx.q]).reshape(-1)
    pm[: L * N] = a
    assert ctx.accel_op(pm, None, 1, 0, 0, 0, 0, L) == 0            # in-place NTT
    assert np.array_equal(pm[: L * N].reshape(L, N), ctx.ntt_poly(a))
    assert ctx.accel_op(pm, None, 2, 0, 0, 8 * N, 0, L) == 0
    assert np.array_equal(pm[8 * N: 8 * N + L * N], a)


def test_composites_decrypt(ctx):
    """eval_sub_plain -> eval_mult_relin_rescale -> eval_rotate(+1), then Dec."""
    rng = np.random.default_rng(23)
    N, L = ctx.N, ctx.L
    s = ck.ternary(rng, N)
    s2 = ck.negacyclic(s, s)
    relin = ck.make_ks_key(ctx, rng, s, s2, L)
    g1 = pow(5, 1, 2 * N)
    g2 = pow(5, 2, 2 * N)
    rot1 = ck.make_ks_key(ctx, rng, s, ck.poly_automorph_coeffs(s, g1, N), L - 1)
    rot2 = ck.make_ks_key(ctx, rng, s, ck.poly_automorph_coeffs(s, g2, N), L - 1)
    km = np.concatenate([relin, rot1, rot2])
    off_rot1 = relin.size
    off_rot2 = relin.size + rot1.size
    delta = 1 << 30
    u = rng.integers(-3, 4, N)
    v = rng.integers(-3, 4, N)
    c0, c1 = ck.encrypt(ctx, rng, s, [delta * x for x in u], L)
    pt = ck.small_to_limbs(ctx, [delta * x for x in v], range(L))
    W = L * N
    pm = np.zeros(16 * W, np.uint64)
    pm[0:W] = c0.reshape(-1)
    pm[W:2 * W] = c1.reshape(-1)
    pm[2 * W:3 * W] = pt.reshape(-1)
    # diff = ct - pt  (u - v) * delta
    assert ctx.eval_sub_plain(pm, 0, W, 2 * W, 3 * W, 4 * W, L) == 0
    # sq = Rescale(diff * diff)
    assert ctx.eval_mult_relin_rescale(pm, km, 0, L, 3 * W, 4 * W, 3 * W, 4 * W, 5 * W, 6 * W, L) == 0
    W1 = (L - 1) * N
    r0 = pm[5 * W:5 * W + W1].reshape(L - 1, N).copy()
    r1 = pm[6 * W:6 * W + W1].reshape(L - 1, N).copy()
    w = [int(x) - int(y) for x, y in zip(u, v)]
    ww = ck.negacyclic(w, w)
    ql = ctx.q[L - 1]
    got = ck.decrypt_coeffs(ctx, s, r0, r1)
    exp = [delta * delta * x / ql for x in ww]
    assert max(abs(g - e) for g, e in zip(got, exp)) < 2 ** 12
    # rotate by +3 = steps (+2, +1) -> sigma_{5^3}
    keys = [(1, g1, off_rot1, L - 1), (2, g2, off_rot2, L - 1)]
    assert ctx.eval_rotate(pm, km, keys, 5 * W, 6 * W, 7 * W, 8 * W, 3, L - 1) == 0
    t0 = pm[7 * W:7 * W + W1].reshape(L - 1, N)
    t1 = pm[8 * W:8 * W + W1].reshape(L - 1, N)
    got_r = ck.decrypt_coeffs(ctx, s, t0, t1)
    exp_r = ck.poly_automorph_coeffs(got, pow(5, 3, 2 * N), N)
    assert max(abs(g - e) for g, e in zip(got_r, exp_r)) < 2 ** 12
    # a missing step key is an error (fhe_host_driver.cpp:336-339)
    assert ctx.eval_rotate(pm, km, keys, 5 * W, 6 * W, 7 * W, 8 * W, 4, L - 1) == -5


def test_linear_sum_is_op5_op3_composition(orc):
    """orc_linear_sum equals OP 5 (poly_mul) and OP 3 (poly_add) applied term by term
    through the op dispatcher, with bias and per-output plaintexts."""
    N, L = 1024, 2
    q = orc.gen_primes(2 * N, 50, 3)
    o = orc.OracleCtx(N, q[:L], q[L:], 1)
    rng = np.random.default_rng(8)
    W = L * N
    R, G = 3, 2
    pm = np.zeros(40 * W, np.uint64)
    for k in range(20):
        pm[k * W:(k + 1) * W] = np.concatenate([rng.integers(0, p, N, dtype=np.uint64) for p in q[:L]])
    ct0, ct1 = [0, 2 * W, 4 * W], [W, 3 * W, 5 * W]
    pt = np.array([[6 * W, 7 * W, 8 * W], [9 * W, 10 * W, 11 * W]], np.uint64)
    ptb = [12 * W, (1 << 64) - 1]
    bias = (13 * W, 14 * W)
    out0, out1 = [20 * W, 22 * W], [21 * W, 23 * W]
    assert o.linear_sum(pm, ct0, ct1, pt, out0, out1, L, ptb=ptb, bias=bias) == 0
    ref = pm.copy()
    T = 30 * W
    for g in range(G):
        a0, a1 = 24 * W, 25 * W
        ref[a0:a0 + W] = ref[bias[0]:bias[0] + W]
        ref[a1:a1 + W] = ref[bias[1]:bias[1] + W]
        if ptb[g] != (1 << 64) - 1:
            assert o.accel_op(ref, None, 3, a0, ptb[g], a0, 0, L) == 0
        for r in range(R):
            for c, a in ((ct0[r], a0), (ct1[r], a1)):
                assert o.accel_op(ref, None, 5, c, int(pt[g, r]), T, 0, L) == 0
                assert o.accel_op(ref, None, 3, a, T, a, 0, L) == 0
        assert np.array_equal(pm[out0[g]:out0[g] + W], ref[a0:a0 + W])
        assert np.array_equal(pm[out1[g]:out1[g] + W], ref[a1:a1 + W])


def test_mult_relin_then_rescale_is_mult_relin_rescale(orc):
    N, L, K = 1024, 3, 2
    q = orc.gen_primes(2 * N, 50, L + K)
    o = orc.OracleCtx(N, q[:L], q[L:], 3)
    rng = np.random.default_rng(12)
    W = L * N
    km = np.concatenate([rng.integers(0, p, N, dtype=np.uint64) for _ in range(2 * 3) for p in q])
    pm = np.zeros(12 * W, np.uint64)
    for k in range(4):
        pm[k * W:(k + 1) * W] = np.concatenate([rng.integers(0, p, N, dtype=np.uint64) for p in q[:L]])
    assert o.eval_mult_relin(pm, km, 0, L, 0, W, 2 * W, 3 * W, 4 * W, 5 * W, L) == 0
    r0 = o.rescale(pm[4 * W:5 * W]).reshape(-1)
    r1 = o.rescale(pm[5 * W:6 * W]).reshape(-1)
    assert o.eval_mult_relin_rescale(pm, km, 0, L, 0, W, 2 * W, 3 * W, 6 * W, 7 * W, L) == 0
    W1 = (L - 1) * N
    assert np.array_equal(r0, pm[6 * W:6 * W + W1])
    assert np.array_equal(r1, pm[7 * W:7 * W + W1])


def test_hoisted_rotations_pinned(orc):
    """Hoisted rotations (one ModUp, permuted digits): with g = 1 they equal the
    pinned automorphism + key switch word for word; with real keys every output
    decrypts to the rotated slots."""
    from anns_fhe_amd.client import CkksClient
    N, L, K, dnum = 1024, 3, 2, 2
    cl = CkksClient(N, L, K, dnum)
    o = orc.OracleCtx(N, cl.q, cl.p, dnum)
    s = cl.keygen_secret(3)
    pk = cl.keygen_public(s, 4)
    ksz = cl.key_size_words(L)
    steps = [1, 3, -2, 100]
    km = np.zeros((len(steps) + 1) * ksz, np.uint64)
    gal = [cl.galois_for_step(st) for st in steps] + [1]
    for z, g in enumerate(gal):
        km[z * ksz:(z + 1) * ksz] = cl.rotation_key(s, g, L, 10 + z)
    W = L * N
    n = N // 2
    x = np.linspace(-1, 1, n)
    pm = np.zeros(16 * W, np.uint64)
    pm[:2 * W] = cl.encrypt_values(pk, x, L, 2.0 ** 40, 5)
    outs0 = [(2 + 2 * z) * W for z in range(len(gal))]
    outs1 = [(3 + 2 * z) * W for z in range(len(gal))]
    assert o.eval_rotations_hoisted(pm, km, 0, W, gal, [z * ksz for z in range(len(gal))], L, outs0, outs1, L) == 0
    # identity element: hoisting changes nothing
    ref = pm.copy()
    assert o.eval_automorphism_ks(ref, km, 1, len(steps) * ksz, L, 0, W, 14 * W, 15 * W, L) == 0
    assert np.array_equal(pm[outs0[-1]:outs0[-1] + W], ref[14 * W:15 * W])
    assert np.array_equal(pm[outs1[-1]:outs1[-1] + W], ref[15 * W:16 * W])
    for z, st in enumerate(steps):
        ct = np.concatenate([pm[outs0[z]:outs0[z] + W], pm[outs1[z]:outs1[z] + W]])
        got = cl.decrypt_values(s, ct, L, 2.0 ** 40, n)
        assert np.max(np.abs(got - np.roll(x, -st))) < 1e-6, st


def test_oracle_calls_on_threads_equal_sequential_calls():
    """The CPU baseline times oracle HMults on several host threads at once (one
    ciphertext pair per thread, one shared context and key): the words equal those
    of the same calls made one after another."""
    import concurrent.futures as cf
    from anns_fhe_amd.client import CkksClient
    from oracle import oracle as O
    O.build()
    N, L, K, D = 1 << 11, 3, 2, 2
    ck = CkksClient(N, L, K, D)
    o = O.OracleCtx(N, ck.q, ck.p, D)
    key = ck.relin_key(ck.keygen_secret(3), L, 5)
    W = L * N
    rng = np.random.default_rng(1)
    bufs = []
    for _ in range(6):
        pm = np.zeros(6 * W, np.uint64)
        pm[:4 * W] = np.concatenate([rng.integers(0, q, N, dtype=np.uint64) for _ in range(4) for q in ck.q])
        bufs.append(pm)
    seq = [b.copy() for b in bufs]
    for b in seq:
        assert o.eval_mult_relin_rescale(b, key, 0, L, 0, W, 2 * W, 3 * W, 4 * W, 5 * W, L) == 0
    par = [b.copy() for b in bufs]
    with cf.ThreadPoolExecutor(6) as ex:
        st = list(ex.map(lambda b: o.eval_mult_relin_rescale(b, key, 0, L, 0, W, 2 * W, 3 * W, 4 * W, 5 * W, L),
                         par))
    assert st == [0] * 6
    assert all(np.array_equal(a, b) for a, b in zip(seq, par))
