"""Python big-integer closed forms used to pin the oracle.

These are the mathematical definitions (O(N^2) transforms, CRT reconstruction,
exact rational identities), not restatements of the oracle's loops: a wrong
sign, index or transposition in the oracle makes them disagree.
"""
import random

import numpy as np


def bitrev(x, bits):
    r = 0
    for _ in range(bits):
        r = (r << 1) | (x & 1)
        x >>= 1
    return r


def is_probable_prime(n, rounds=24, seed=7):
    """Miller-Rabin with random bases (independent of the oracle's fixed bases)."""
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    rng = random.Random(seed)
    for _ in range(rounds):
        a = rng.randrange(2, n - 1)
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def ntt_closed_form(a, psi, q):
    """A[i] = sum_j a_j psi^((2 br(i) + 1) j) mod q  (natural in, bit-reversed out)."""
    N = len(a)
    lg = N.bit_length() - 1
    out = []
    for i in range(N):
        e = 2 * bitrev(i, lg) + 1
        w = pow(psi, e, q)
        acc, wp = 0, 1
        for j in range(N):
            acc += int(a[j]) * wp
            wp = wp * w % q
        out.append(acc % q)
    return out


def intt_closed_form(A, psi, q):
    """Inverse of ntt_closed_form: a_j = N^-1 sum_i A_i psi^(-(2br(i)+1) j)."""
    N = len(A)
    lg = N.bit_length() - 1
    ninv = pow(N, -1, q)
    pinv = pow(psi, -1, q)
    out = [0] * N
    for i in range(N):
        e = 2 * bitrev(i, lg) + 1
        w = pow(pinv, e, q)
        wp = 1
        for j in range(N):
            out[j] += int(A[i]) * wp
            wp = wp * w % q
    return [x * ninv % q for x in out]


def negacyclic_mul(a, b, q):
    """Schoolbook a*b mod (X^N + 1, q)."""
    N = len(a)
    a = [int(x) for x in a]
    b = [int(x) for x in b]
    out = [0] * N
    for i in range(N):
        ai = a[i]
        if ai == 0:
            continue
        for j in range(N):
            k = i + j
            if k < N:
                out[k] += ai * b[j]
            else:
                out[k - N] -= ai * b[j]
    return [x % q for x in out]


def crt(residues, primes):
    """x in [0, prod(primes)) with x == r_i mod p_i."""
    M = 1
    for p in primes:
        M *= p
    x = 0
    for r, p in zip(residues, primes):
        Mi = M // p
        x += int(r) * Mi * pow(Mi, -1, p)
    return x % M, M


def prod(xs):
    r = 1
    for x in xs:
        r *= int(x)
    return r


def rand_poly(rng, primes, N):
    return np.stack([rng.integers(0, q, size=N, dtype=np.uint64) for q in primes]).astype(np.uint64)


def centered(x, m):
    x %= m
    return x - m if x > m // 2 else x
