"""CPU checks of the FP64 modular arithmetic the HIP kernels use for primes < 2^50.

The FP64 kernels (anns_fhe_amd/csrc/fhea_f64.hpp, fhea_ntt_f64.hip, k_fbc_f64,
k_ks_inner_f64) must give the same residues as the integer path.  Their
correctness rests on magnitude bounds (every intermediate an integer below 2^53).
Here the same operation sequences run in exact binary64 emulation
(tests/f64_emu.py) and are compared with the oracle and with big-integer closed
forms, on random and adversarial inputs, with the bounds asserted at every step.
"""
import random
from fractions import Fraction

import numpy as np
import pytest

import f64_emu as E
from oracle import oracle as O

TWO52 = 2 ** 52


def _primes(bits, n, count):
    return [int(p) for p in O.gen_primes(2 * n, bits, count)]


@pytest.mark.parametrize("bits", [50, 45, 36])
def test_mulmod_f64_congruence_and_bound(bits):
    """r = v*w - k*q exactly, and |r| <= q (1/2 + |v| 2^-52 (1 + 2^-52)) for |v| <= 8q."""
    rng = random.Random(bits)
    for q in _primes(bits, 1024, 3):
        qd = float(q)
        cases = []
        for _ in range(3000):
            v = rng.randint(-8 * q, 8 * q)
            w = rng.randrange(q)
            cases.append((v, w))
        for v in (8 * q, -8 * q, 8 * q - 1, 4 * q, -4 * q, q - 1, 1, 0):
            for w in (q - 1, q - 2, (q - 1) // 2, 1, 0):
                cases.append((v, w))
        for v, w in cases:
            vd, wd = float(v), float(w)
            assert vd == v and wd == w
            r = E.mulmod_f64(vd, wd, wd / qd, qd)
            assert r == int(r)
            assert (int(r) - v * w) % q == 0
            bound = Fraction(q) * (Fraction(1, 2) + Fraction(abs(v), TWO52) * (1 + Fraction(1, TWO52)))
            assert abs(int(r)) <= bound, (q, v, w, r)


@pytest.mark.parametrize("bits", [50, 40])
def test_mulmod_vv_f64_bound(bits):
    """Canonical a, b: a*b mod q exact with |r| <= 0.9 q."""
    rng = random.Random(7 + bits)
    for q in _primes(bits, 1024, 2):
        qd, qinv = float(q), 1.0 / float(q)
        pairs = [(rng.randrange(q), rng.randrange(q)) for _ in range(4000)]
        pairs += [(q - 1, q - 1), (q - 1, 1), (0, q - 1), ((q - 1) // 2, q - 1)]
        for a, b in pairs:
            r = E.mulmod_vv_f64(float(a), float(b), qd, qinv)
            assert (int(r) - a * b) % q == 0
            assert abs(r) <= 0.9 * q


def test_reduce_f64_bound():
    rng = random.Random(3)
    for q in _primes(50, 1024, 2) + _primes(33, 1024, 1):
        qd, qinv = float(q), 1.0 / float(q)
        xs = [rng.randint(-8 * q, 8 * q) for _ in range(4000)] + [8 * q, -8 * q, 0, q, -q, q // 2]
        for x in xs:
            r = E.reduce_f64(float(x), qd, qinv)
            assert (int(r) - x) % q == 0
            assert abs(r) <= q / 2 + 2


def _adversarial_inputs(rng, n, q):
    return [
        [rng.randrange(q) for _ in range(n)],
        [q - 1] * n,
        [(q - 1) if (j & 1) else 0 for j in range(n)],
        [(q - 1) // 2 + (j & 1) for j in range(n)],
    ]


@pytest.mark.parametrize("n", [16, 64, 256])
@pytest.mark.parametrize("round_stages", [4, 3, 2])
def test_ntt_f64_matches_oracle(n, round_stages):
    """The FP64 forward transform (radix-2^r rounds, one reduction per round) and
    inverse equal the oracle's NTT/INTT word for word; no intermediate exceeds 8q."""
    rng = random.Random(n * 10 + round_stages)
    q = _primes(50, n, 2)
    o = O.OracleCtx(n, q[:1], q[1:], 1)
    tr = E.Trace()
    for idx in range(2):
        p = o.primes[idx]
        for a in _adversarial_inputs(rng, n, p):
            got = E.ntt_forward_f64(a, o.tw[idx], p, round_stages, tr)
            ref = o.ntt(np.array(a, np.uint64), idx)
            assert got == [int(v) for v in ref]
            back = E.ntt_inverse_f64(got, o.itw[idx], int(o.n_inv[idx]), p, tr)
            assert back == a
    assert tr.max_ratio["mulmod_in"] < 8.0
    assert tr.max_ratio["butterfly_out"] < 8.0
    assert tr.max_ratio["reduce_in"] < 8.0


@pytest.mark.parametrize("n,k1", [(256, 8), (1024, 10), (8192, 6), (8192, 7)])
def test_ntt_f64_kernel_rounds_factorized_twiddles(n, k1):
    """The forward kernel's own pass/round structure with its twiddles formed as
    products tw[base] * tw[t] (one load per stage): equal to the oracle's NTT word
    for word, every intermediate below 8q, every formed twiddle below 0.75q."""
    rng = random.Random(n + k1)
    q = _primes(50, n, 2)
    o = O.OracleCtx(n, q[:1], q[1:], 1)
    tr = E.Trace()
    for idx in range(2):
        p = o.primes[idx]
        inputs = _adversarial_inputs(rng, n, p) if n <= 1024 else [
            [p - 1] * n, [rng.randrange(p) for _ in range(n)]]
        for a in inputs:
            got = E.ntt_forward_f64_kernel(a, o.tw[idx], p, k1, tr)
            assert got == [int(v) for v in o.ntt(np.array(a, np.uint64), idx)]
    assert tr.max_ratio["butterfly_out"] < 8.0
    assert tr.max_ratio["reduce_in"] < 8.0
    assert tr.max_ratio["twiddle"] < 0.75


@pytest.mark.parametrize("n,k1", [(256, 8), (1024, 10), (8192, 6), (8192, 7)])
def test_intt_f64_kernel_factorized_first_round(n, k1):
    """The inverse with the first row round's twiddles formed as products equals
    the oracle's INTT word for word; differences and products stay far below 8q."""
    rng = random.Random(3 * n + k1)
    q = _primes(50, n, 2)
    o = O.OracleCtx(n, q[:1], q[1:], 1)
    tr = E.Trace()
    for idx in range(2):
        p = o.primes[idx]
        inputs = _adversarial_inputs(rng, n, p) if n <= 1024 else [
            [p - 1] * n, [rng.randrange(p) for _ in range(n)]]
        for a in inputs:
            got = E.ntt_inverse_f64_kernel(a, o.itw[idx], int(o.n_inv[idx]), p, k1, tr=tr)
            assert got == [int(v) for v in o.intt(np.array(a, np.uint64), idx)]
    assert tr.max_ratio["inv_diff"] < 4.0
    assert tr.max_ratio["mulmod_out"] < 2.0
    assert tr.max_ratio["twiddle"] < 0.75


def test_ntt_f64_largest_prime_below_2_50():
    """N = 1024 at the largest NTT prime below 2^50 (the tightest case for 2^53)."""
    n = 1024
    p = _primes(50, n, 1)
    o = O.OracleCtx(n, p, [], 1)
    rng = random.Random(11)
    tr = E.Trace()
    a = [rng.randrange(p[0]) for _ in range(n)]
    got = E.ntt_forward_f64(a, o.tw[0], p[0], 4, tr)
    assert got == [int(v) for v in o.ntt(np.array(a, np.uint64), 0)]
    assert E.ntt_inverse_f64(got, o.itw[0], int(o.n_inv[0]), p[0], tr) == a
    assert tr.max_ratio["butterfly_out"] < 8.0


@pytest.mark.parametrize("in_bits,out_bits", [(50, 50), (50, 49), (49, 50)])
def test_fbc_f64_matches_closed_form(in_bits, out_bits):
    """Basis conversion as k_fbc_f64 computes it equals
    sum_i [x_i qhat_i^-1]_{q_i} (qhat_i mod p_j) mod p_j (integer y_i, not its residue),
    for input/output prime ratios below 2 (the FP64 path admits ratios below 4)."""
    rng = random.Random(in_bits * 100 + out_bits)
    n_in, n_out = 8, 12
    qs = _primes(in_bits, 64, n_in)
    ps = [p for p in _primes(out_bits, 64, n_in + n_out) if p not in qs][:n_out]
    assert max(qs + ps) / min(qs + ps) < 4
    Q = 1
    for q in qs:
        Q *= q
    qhat = [Q // q for q in qs]
    qhat_inv = [pow(h % q, -1, q) for h, q in zip(qhat, qs)]
    qhat_mod_p = [[h % p for p in ps] for h in qhat]
    tr = E.Trace()
    for trial in range(300):
        if trial == 0:
            x = [q - 1 for q in qs]
        else:
            x = [rng.randrange(q) for q in qs]
        got = E.fast_bases_conv_f64(x, qs, qhat_inv, qhat_mod_p, ps, tr)
        ys = [(xi * hi) % q for xi, hi, q in zip(x, qhat_inv, qs)]
        exp = [sum(y * hm[j] for y, hm in zip(ys, qhat_mod_p)) % p for j, p in enumerate(ps)]
        assert got == exp
    assert tr.max_ratio["fbc_acc"] < 8.0


def test_fbc_centred_intermediate_is_wrong():
    """The planted mistake: using the centred product [x qhat^-1] in (-q, q) instead of
    the canonical one changes the converted value (so the test above can fail)."""
    qs = _primes(50, 64, 3)
    ps = [p for p in _primes(50, 64, 6) if p not in qs][:2]
    Q = qs[0] * qs[1] * qs[2]
    qhat = [Q // q for q in qs]
    qhat_inv = [pow(h % q, -1, q) for h, q in zip(qhat, qs)]
    rng = random.Random(5)
    differs = 0
    for _ in range(50):
        x = [rng.randrange(q) for q in qs]
        canon = [(xi * hi) % q for xi, hi, q in zip(x, qhat_inv, qs)]
        centred = [c - q if c > q // 2 else c for c, q in zip(canon, qs)]
        for p in ps:
            a = sum(c * (h % p) for c, h in zip(canon, qhat)) % p
            b = sum(c * (h % p) for c, h in zip(centred, qhat)) % p
            differs += a != b
    assert differs > 0
