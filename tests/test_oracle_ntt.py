"""Pins the oracle's NTT, INTT and automorphism (tb_ntt.cpp, tb_intt.cpp,
tb_automorphism.cpp semantics) against closed forms."""
import numpy as np
import pytest

from tests.bigint_ref import intt_closed_form, negacyclic_mul, ntt_closed_form


def mkctx(orc, N, L, bits=50, K=1, dnum=1):
    q = orc.gen_primes(2 * N, bits, L + K)
    return orc.OracleCtx(N, q[:L], q[L:], dnum)


@pytest.mark.parametrize("N,bits", [(16, 30), (64, 50), (256, 60)])
def test_ntt_closed_form(orc, N, bits):
    ctx = mkctx(orc, N, 2, bits)
    rng = np.random.default_rng(42)
    for idx in range(2):
        q = ctx.primes[idx]
        a = rng.integers(0, q, N, dtype=np.uint64)
        A = ctx.ntt(a, idx)
        assert [int(x) for x in A] == ntt_closed_form(a, int(ctx.psi[idx]), q)
        back = ctx.intt(A, idx)
        assert np.array_equal(back, a)
        assert [int(x) for x in ctx.intt(a, idx)] == intt_closed_form(a, int(ctx.psi[idx]), q)


def test_foreign_roots_closed_form(orc):
    """orc_ctx_set_psi: tables from another primitive root (psi^3, as a directory
    dumped by another library may hold) still give the closed-form transform with
    that root; a root that is not a primitive 2N-th one is refused."""
    N = 64
    q = orc.gen_primes(2 * N, 50, 2)
    base = orc.OracleCtx(N, q[:1], q[1:], 1)
    psi3 = [pow(int(w), 3, p) for w, p in zip(base.psi, q)]
    ctx = orc.OracleCtx(N, q[:1], q[1:], 1, psi=psi3)
    assert [int(x) for x in ctx.psi] == psi3
    rng = np.random.default_rng(3)
    for idx in range(2):
        a = rng.integers(0, q[idx], N, dtype=np.uint64)
        assert [int(x) for x in ctx.ntt(a, idx)] == ntt_closed_form(a, psi3[idx], q[idx])
        assert [int(x) for x in ctx.intt(a, idx)] == intt_closed_form(a, psi3[idx], q[idx])
    bad = [pow(psi3[0], 2, q[0]), psi3[1]]
    with pytest.raises(AssertionError):
        orc.OracleCtx(N, q[:1], q[1:], 1, psi=bad)


def test_ntt_convolution_theorem(orc):
    N = 1024
    ctx = mkctx(orc, N, 1, 55)
    q = ctx.primes[0]
    rng = np.random.default_rng(7)
    a = rng.integers(0, q, N, dtype=np.uint64)
    b = rng.integers(0, 1 << 20, N, dtype=np.uint64)
    prod = ctx.intt(np.array([(int(x) * int(y)) % q for x, y in zip(ctx.ntt(a, 0), ctx.ntt(b, 0))],
                             dtype=np.uint64), 0)
    assert [int(x) for x in prod] == negacyclic_mul(a, b, q)


def test_ntt_edge_inputs(orc):
    N = 4096
    ctx = mkctx(orc, N, 1, 61)
    q = ctx.primes[0]
    for a in [np.zeros(N, np.uint64), np.full(N, q - 1, np.uint64)]:
        assert np.array_equal(ctx.intt(ctx.ntt(a, 0), 0), a)
    # delta at 0 maps to all-ones in every slot
    d = np.zeros(N, np.uint64)
    d[0] = 1
    assert np.all(ctx.ntt(d, 0) == 1)


@pytest.mark.parametrize("N", [16, 256])
def test_automorphism_is_x_to_xg(orc, N):
    """NTT(a(X^g)) == automorphism(NTT(a), map_g): coefficient-domain identity."""
    ctx = mkctx(orc, N, 1, 50)
    q = ctx.primes[0]
    rng = np.random.default_rng(3)
    a = rng.integers(0, q, N, dtype=np.uint64)
    for g in [5, pow(5, 3, 2 * N), pow(5, N // 2 - 1, 2 * N), 2 * N - 1]:
        b = [0] * N
        for j in range(N):
            e = (j * g) % (2 * N)
            if e < N:
                b[e] = (b[e] + int(a[j])) % q
            else:
                b[e - N] = (b[e - N] - int(a[j])) % q
        expect = ctx.ntt(np.array(b, dtype=np.uint64), 0)
        packed = orc.pack_map(orc.galois_map(N, g))
        out = np.zeros(N, np.uint64)
        orc.lib().orc_automorphism(orc.P64(ctx.ntt(a, 0)), orc.P64(out), orc.P64(packed), N)
        assert np.array_equal(out, expect)
