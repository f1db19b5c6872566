"""Pins the oracle's L0 arithmetic and table generators against big-int definitions.

Mirrors tb_mod_arith.cpp:32-62 (per prime: random (a, b) pairs, add/sub/mul
compared bit-for-bit), plus the edge cases the reference never tested.
"""
import random

import numpy as np
import pytest

from tests.bigint_ref import bitrev, is_probable_prime


PRIME_BITS = [30, 40, 45, 50, 55, 60, 61]


@pytest.fixture(scope="module")
def primes(orc):
    out = []
    for b in PRIME_BITS:
        out += orc.gen_primes(2 * 4096, b, 2)
    return out


def test_barrett_constants(orc, primes):
    for q in primes:
        lo, hi, k = orc.barrett(q)
        assert k == 2 * q.bit_length() + 3
        assert (hi << 64 | lo) == (1 << k) // q


def test_mod_arith_random_and_edges(orc, primes):
    L = orc.lib()
    rng = random.Random(42)
    for q in primes:
        lo, hi, k = orc.barrett(q)
        cases = [(0, 0), (0, q - 1), (q - 1, q - 1), (1, q - 1), (q - 1, 1), (q // 2, q // 2 + 1), (1, 1)]
        cases += [(rng.randrange(q), rng.randrange(q)) for _ in range(1000)]
        for a, b in cases:
            assert L.orc_mod_add(a, b, q) == (a + b) % q
            assert L.orc_mod_sub(a, b, q) == (a - b) % q
            assert L.orc_mod_mul(a, b, q, lo, hi, k) == (a * b) % q
            # reduction use (b = 1) with x < 2^64, as key_switch.cpp:51 / rescale.cpp:41 use it
        for x in [0, 1, q, 2 * q - 1, (1 << 62) - 1, rng.randrange(1 << 62)]:
            assert L.orc_mod_mul(x, 1, q, lo, hi, k) == x % q


def test_inverse_and_pow(orc, primes):
    L = orc.lib()
    rng = random.Random(1)
    for q in primes:
        for _ in range(20):
            a = rng.randrange(1, q)
            assert L.orc_inv_mod(a, q) == pow(a, -1, q)
            e = rng.randrange(1 << 40)
            assert L.orc_pow_mod(a, e, q) == pow(a, e, q)


def test_gen_primes_rule(orc):
    # q == 1 mod 2N, q < 2^bits, prime, strictly descending, and no prime of that
    # form is skipped (checked exhaustively at 20 bits).
    two_n = 2 * 64
    ps = orc.gen_primes(two_n, 20, 6)
    assert ps == sorted(ps, reverse=True)
    expected = [q for q in range((1 << 20) - 1, 0, -1) if q % two_n == 1 and is_probable_prime(q)][:6]
    assert ps == expected
    for bits in (50, 60):
        ps = orc.gen_primes(2 * 65536, bits, 8)
        assert all(q % (2 * 65536) == 1 and q < (1 << bits) and is_probable_prime(q) for q in ps)
        assert len(set(ps)) == 8
    skip = orc.gen_primes(two_n, 20, 2)
    rest = orc.gen_primes(two_n, 20, 3, skip=skip)
    assert rest == expected[2:5]


def test_min_primitive_root(orc):
    L = orc.lib()
    for two_n in (16, 64, 256):
        for q in orc.gen_primes(two_n, 20, 3):
            psi = L.orc_min_primitive_root(q, two_n)
            assert pow(psi, two_n // 2, q) == q - 1
            # brute force minimality: no smaller x is a primitive two_n-th root
            for x in range(2, psi):
                assert not (pow(x, two_n // 2, q) == q - 1)


def test_twiddles(orc):
    N = 64
    q = orc.gen_primes(2 * N, 40, 1)[0]
    psi = orc.lib().orc_min_primitive_root(q, 2 * N)
    tw = np.zeros(N, np.uint64)
    itw = np.zeros(N, np.uint64)
    orc.lib().orc_twiddles(q, psi, N, orc.P64(tw), orc.P64(itw))
    for j in range(N):
        assert int(tw[j]) == pow(psi, bitrev(j, 6), q)
        assert int(itw[j]) == pow(pow(psi, -1, q), bitrev(j, 6), q)
