"""Small-N CKKS helpers in Python for the oracle's decryption invariants.

Keys follow the hybrid key-switching definition (OpenFHE KeySwitchHYBRID):
digit p of a key for s' under s is  b_p = -a_p s + e_p + [i in digit p] (P mod q_i) s'
over the Q limbs of the key level and all P limbs, a_p uniform.  Layout in the
key memory: a at evk as [dnum][key_limbs+K][N], then b (key_switch.cpp:352-381).
Only for tiny N (Python integers); the product has its own C++ key generator.
"""
import numpy as np

from tests.bigint_ref import centered, crt, prod


def mulmod(a, b, q):
    return np.array([(int(x) * int(y)) % q for x, y in zip(a, b)], dtype=np.uint64)


def small_to_limbs(ctx, coeffs, idxs):
    """Signed small coefficients -> NTT form on the given prime indices."""
    out = []
    for i in idxs:
        q = ctx.primes[i]
        out.append(ctx.ntt(np.array([int(c) % q for c in coeffs], dtype=np.uint64), i))
    return np.stack(out)


def ternary(rng, N):
    return rng.integers(-1, 2, N)


def gaussian(rng, N, sigma=3.2):
    return np.rint(rng.normal(0, sigma, N)).astype(np.int64)


def key_index_primes(ctx, key_limbs):
    return list(range(key_limbs)) + [ctx.L + j for j in range(ctx.K)]


def make_ks_key(ctx, rng, s, s_new, key_limbs, noise=True):
    """Return the flat key words for switching s_new -> s at level key_limbs."""
    N, K, dnum, alpha = ctx.N, ctx.K, ctx.dnum, ctx.alpha
    idxs = key_index_primes(ctx, key_limbs)
    s_e = small_to_limbs(ctx, s, idxs)
    sn_e = small_to_limbs(ctx, s_new, idxs)
    P = prod(ctx.p)
    A = np.zeros((dnum, key_limbs + K, N), np.uint64)
    B = np.zeros((dnum, key_limbs + K, N), np.uint64)
    for p in range(dnum):
        e = small_to_limbs(ctx, gaussian(rng, N) if noise else np.zeros(N, np.int64), idxs)
        for r, i in enumerate(idxs):
            q = ctx.primes[i]
            a = rng.integers(0, q, N, dtype=np.uint64)
            A[p, r] = a
            b = (-(np.array([int(x) for x in mulmod(a, s_e[r], q)], dtype=object)) + e[r].astype(object)) % q
            if i < key_limbs and alpha * p <= i < alpha * (p + 1):
                b = (b + np.array([int(x) for x in sn_e[r]], dtype=object) * (P % q)) % q
            B[p, r] = np.array([int(x) for x in b], dtype=np.uint64)
    return np.concatenate([A.reshape(-1), B.reshape(-1)])


def encrypt(ctx, rng, s, m_coeffs, limbs):
    """(c0, c1) with c0 + c1 s = m + e over Q limbs 0..limbs-1, eval domain."""
    N = ctx.N
    idxs = list(range(limbs))
    s_e = small_to_limbs(ctx, s, idxs)
    m_e = small_to_limbs(ctx, m_coeffs, idxs)
    e_e = small_to_limbs(ctx, gaussian(rng, N), idxs)
    c0 = np.zeros((limbs, N), np.uint64)
    c1 = np.zeros((limbs, N), np.uint64)
    for i in idxs:
        q = ctx.primes[i]
        a = rng.integers(0, q, N, dtype=np.uint64)
        c1[i] = a
        t = np.array([(-(int(x)) + int(y) + int(z)) % q for x, y, z in zip(mulmod(a, s_e[i], q), m_e[i], e_e[i])],
                     dtype=np.uint64)
        c0[i] = t
    return c0, c1


def decrypt_coeffs(ctx, s, c0, c1):
    """Centered integer coefficients of c0 + c1 s (CRT over the limbs present)."""
    limbs = c0.shape[0]
    idxs = list(range(limbs))
    s_e = small_to_limbs(ctx, s, idxs)
    rows = []
    for i in idxs:
        q = ctx.primes[i]
        d = np.array([(int(x) + int(y)) % q for x, y in zip(c0[i], mulmod(c1[i], s_e[i], q))], dtype=np.uint64)
        rows.append(ctx.intt(d, i))
    qs = ctx.primes[:limbs]
    out = []
    for j in range(ctx.N):
        x, M = crt([rows[i][j] for i in idxs], qs)
        out.append(centered(x, M))
    return out


def poly_automorph_coeffs(m, g, N):
    out = [0] * N
    for j in range(N):
        e = (j * g) % (2 * N)
        if e < N:
            out[e] += int(m[j])
        else:
            out[e - N] -= int(m[j])
    return out


def negacyclic(a, b):
    N = len(a)
    out = [0] * N
    for i in range(N):
        for j in range(N):
            k = i + j
            if k < N:
                out[k] += int(a[i]) * int(b[j])
            else:
                out[k - N] -= int(a[i]) * int(b[j])
    return out
