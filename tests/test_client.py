"""CPU tests of the CKKS client (anns_fhe_amd/client.py over libfhea's host-only
client entry points).

Pinned by definitions, not by the implementation's own FFT:
* encoding/decoding against the canonical embedding evaluated directly,
  z_j = m(zeta^(5^j mod 2N)) / scale with zeta = exp(i pi / N), from the
  plaintext's coefficients recovered by the oracle's INTT and big-integer CRT;
* encryption/decryption round trips within the noise bound;
* homomorphic semantics through the oracle (which the GPU matches bit for bit):
  Dec(HMult(Enc(x), Enc(y))) = x * y, Dec(rotate(Enc(x), r)) = x shifted left by r.
"""
import numpy as np
import pytest

from tests.bigint_ref import centered, crt, prod


def _client(N, L, K, dnum, **kw):
    from anns_fhe_amd.client import CkksClient
    return CkksClient(N, L, K, dnum, **kw)


def _coeffs(orc, cl, pt, limbs):
    """Centred integer coefficients of a plaintext (oracle INTT + CRT)."""
    o = orc.OracleCtx(cl.N, cl.q, cl.p, cl.dnum)
    c = o.intt_poly(pt.reshape(limbs, cl.N), list(range(limbs)))
    Q = prod(cl.q[:limbs])
    return [centered(crt([c[i][k] for i in range(limbs)], cl.q[:limbs])[0], Q) for k in range(cl.N)]


def _embed(coeffs, N, n):
    """Direct canonical embedding: m(zeta^(5^j)) for j < n."""
    k = np.arange(N)
    e = np.array([pow(5, j, 2 * N) for j in range(n)])
    zeta = np.exp(1j * np.pi * np.outer(e, k) / N)
    return zeta @ np.array(coeffs, dtype=np.float64)


def test_client_primes_match_device_rules(orc):
    cl = _client(1024, 3, 2, 3)
    assert cl.primes == orc.gen_primes(2 * 1024, 50, 5)


@pytest.mark.parametrize("N,n", [(256, 128), (1024, 512), (1024, 16)])
def test_encode_is_the_canonical_embedding(orc, N, n):
    cl = _client(N, 3, 1, 3)
    rng = np.random.default_rng(N + n)
    z = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    scale = 2.0 ** 40
    pt = cl.encode(z, 3, scale)
    m = _coeffs(orc, cl, pt, 3)
    got = _embed(m, N, N // 2) / scale
    zz = np.tile(z, N // 2 // n)  # replicated with period n
    assert np.max(np.abs(got - zz)) < N * 2.0 ** -40
    # conjugate slots make m real: the embedding at zeta^(-5^j) is the conjugate
    assert np.all(np.abs(m) < 2 ** 62)


def test_decode_is_the_canonical_embedding(orc):
    N, L = 512, 3
    cl = _client(N, L, 1, 3)
    o = orc.OracleCtx(N, cl.q, cl.p, 3)
    rng = np.random.default_rng(3)
    m = [int(v) for v in rng.integers(-2 ** 45, 2 ** 45, N)]
    limbs = np.stack([np.array([v % q for v in m], np.uint64) for q in cl.q])
    pt = o.ntt_poly(limbs).reshape(-1)
    scale = 2.0 ** 30
    got = cl.decode(pt, L, scale, N // 2, complex_out=True)
    exp = _embed(m, N, N // 2) / scale
    assert np.max(np.abs(got - exp)) < 1e-6 * np.max(np.abs(exp))


def test_decode_centres_large_negative_values():
    """Coefficients near -Q/2 .. Q/2 across many limbs (Garner lift, sign)."""
    N, L = 64, 6
    cl = _client(N, L, 1, 2)
    scale = 2.0 ** 50
    z = np.linspace(-3.0, 3.0, N // 2)
    pt = cl.encode(z, L, scale)
    assert np.max(np.abs(cl.decode(pt, L, scale, N // 2) - z)) < 1e-9
    # the same message at 1 limb and at 6 limbs decodes the same
    pt1 = cl.encode(z, 1, 2.0 ** 30)
    assert np.max(np.abs(cl.decode(pt1, 1, 2.0 ** 30, N // 2) - z)) < 1e-6


def test_encrypt_decrypt_round_trip():
    N, L = 2048, 4
    cl = _client(N, L, 2, 2)
    s = cl.keygen_secret(5)
    pk = cl.keygen_public(s, 6)
    rng = np.random.default_rng(9)
    z = rng.uniform(-4, 4, N // 2)
    scale = 2.0 ** 40
    for limbs in (L, 2):
        ct = cl.encrypt_values(pk, z, limbs, scale, seed=10 + limbs)
        got = cl.decrypt_values(s, ct, limbs, scale, N // 2)
        assert np.max(np.abs(got - z)) < 2.0 ** -20
    # a different secret does not decrypt
    s2 = cl.keygen_secret(99)
    bad = cl.decrypt_values(s2, ct, 2, scale, N // 2)
    assert np.max(np.abs(bad - z)) > 1.0


def test_batch_decrypt_decode_equals_single_calls():
    """decrypt_values_batch (secret transformed once, several host threads) gives the
    words of decrypt_values on each ciphertext; wrong sizes are refused."""
    N, L = 1024, 3
    cl = _client(N, L, 1, 3)
    s = cl.keygen_secret(11)
    pk = cl.keygen_public(s, 12)
    rng = np.random.default_rng(13)
    se = cl.secret_eval(s)
    for limbs in (1, 3):
        zs = [rng.uniform(-2, 2, N // 2) for _ in range(5)]
        cts = np.stack([cl.encrypt_values(pk, z, limbs, 2.0 ** 35, seed=20 + k) for k, z in enumerate(zs)])
        single = np.stack([cl.decrypt_values(s, c, limbs, 2.0 ** 35, N // 2) for c in cts])
        for threads in (1, 3, 0):
            got = cl.decrypt_values_batch(se, cts, limbs, 2.0 ** 35, N // 2, threads=threads)
            assert np.array_equal(got, single), (limbs, threads)
        assert np.max(np.abs(single - np.stack(zs))) < 2.0 ** -15
        z = cl.decrypt_values_batch(se, cts[1], limbs, 2.0 ** 35, N // 2, complex_out=True)
        assert z.shape == (1, N // 2) and np.max(np.abs(z[0].real - zs[1])) < 2.0 ** -15
    with pytest.raises(ValueError):
        cl.decrypt_values_batch(se, cts[:, :-1], 3, 2.0 ** 35, N // 2)
    with pytest.raises(ValueError):
        cl.decrypt_values_batch(cl.secret_eval(s, 1), cts, 3, 2.0 ** 35, N // 2)


@pytest.mark.parametrize("N,q_bits", [(4096, 50), (16384, 49), (1024, 60)])
def test_fp64_host_transforms_equal_integer_ones(orc, N, q_bits):
    """The CPU transforms on exact FP64 products (AVX2+FMA, primes below 2^50) give
    the words of the 64-bit integer ones: the secret's evaluation form (forward),
    the plaintexts of encode (forward), the ciphertexts of encrypt (forward and the
    FP64 products) and decrypt+decode (inverse) agree exactly between the two
    modes, and the forward form equals the oracle's NTT."""
    L = 3
    cl = _client(N, L, 1, 3, q_bits=q_bits)
    o = orc.OracleCtx(N, cl.q, cl.p, 3)
    s = cl.keygen_secret(31)
    pk = cl.keygen_public(s, 32)
    rng = np.random.default_rng(33)
    z = rng.uniform(-2, 2, N // 2)
    out = {}
    for mode in (1, 0):
        cl.set_host_ntt(mode)
        se = cl.secret_eval(s)
        pt = cl.encode(z, L, 2.0 ** 35)
        ct = cl.encrypt(pk, pt, L, 34)
        out[mode] = (se, pt, ct, cl.decrypt_values_batch(se, ct, L, 2.0 ** 35, N // 2))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    sc = np.stack([np.array([int(v) % q for v in s], np.uint64) for q in cl.q])
    assert np.array_equal(out[0][0].reshape(L, N), o.ntt_poly(sc))
    assert np.max(np.abs(out[0][3][0] - z)) < 2.0 ** -15
    with pytest.raises(Exception):
        cl.set_host_ntt(2)


def test_hmult_and_rotation_semantics_through_oracle(orc):
    """Client keys + oracle evaluation: products and rotations decrypt to the
    expected slot vectors (the GPU path equals the oracle word for word)."""
    N, L, K, dnum = 2048, 4, 2, 2
    cl = _client(N, L, K, dnum)
    o = orc.OracleCtx(N, cl.q, cl.p, dnum)
    s = cl.keygen_secret(1)
    pk = cl.keygen_public(s, 2)
    ksz = cl.key_size_words(L)
    steps = [1, 2, 4, 8, -1, -2, -4, -8]
    km = np.zeros((2 + len(steps)) * ksz, np.uint64)
    km[:ksz] = cl.relin_key(s, L, 3)
    keys = []
    for i, st in enumerate(steps):
        g = cl.galois_for_step(st)
        km[(1 + i) * ksz:(2 + i) * ksz] = cl.rotation_key(s, g, L, 10 + i)
        keys.append((st, g, (1 + i) * ksz, L))
    rng = np.random.default_rng(4)
    n = N // 2
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    scale = float(cl.q[L - 1])  # rescale by q_{L-1} brings scale^2 back to scale
    W = L * N
    pm = np.zeros(8 * W, np.uint64)
    pm[:2 * W] = cl.encrypt_values(pk, x, L, scale, 20)
    pm[2 * W:4 * W] = cl.encrypt_values(pk, y, L, scale, 21)
    assert o.eval_mult_relin_rescale(pm, km, 0, L, 0, W, 2 * W, 3 * W, 4 * W, 5 * W, L) == 0
    W1 = (L - 1) * N
    ct = np.concatenate([pm[4 * W:4 * W + W1], pm[5 * W:5 * W + W1]])
    got = cl.decrypt_values(s, ct, L - 1, scale * scale / cl.q[L - 1], n)
    assert np.max(np.abs(got - x * y)) < 1e-6
    for r in (1, 5, -3, 12, -15):
        assert o.eval_rotate(pm, km, keys, 0, W, 6 * W, 7 * W, r, L) == 0
        rot = np.concatenate([pm[6 * W:7 * W], pm[7 * W:8 * W]])
        got = cl.decrypt_values(s, rot, L, scale, n)
        assert np.max(np.abs(got - np.roll(x, -r))) < 1e-6, r


def test_client_errors():
    from anns_fhe_amd._lib import FheError
    cl = _client(256, 2, 1, 1)
    with pytest.raises(FheError):
        cl.encode(np.ones(3), 2, 2.0 ** 20)  # n not a power of two
    with pytest.raises(FheError):
        cl.encode(np.ones(4), 3, 2.0 ** 20)  # more limbs than L
    with pytest.raises(FheError):
        cl.encode(np.full(4, np.inf), 2, 2.0 ** 20)  # not finite


def test_large_scale_encoding_and_batch():
    """Scales beyond 2^62 (a product scale of two 2^45 factors) encode exactly;
    the threaded batch encoder equals one-by-one encoding."""
    N, L = 1024, 3
    cl = _client(N, L, 1, 3)
    rng = np.random.default_rng(2)
    z = rng.uniform(0, 500, 64)
    scale = 2.0 ** 90
    pt = cl.encode(z, L, scale)
    assert np.max(np.abs(cl.decode(pt, L, scale, 64) - z)) < 1e-9
    vals = rng.uniform(-2, 2, (7, 128))
    batch = cl.encode_batch(vals, 2, 2.0 ** 45, threads=3)
    for k in range(7):
        assert np.array_equal(batch[k], cl.encode(vals[k], 2, 2.0 ** 45))


def test_key_noise_is_rounded_gaussian(orc):
    """The public key's noise e = b + a s (taken to coefficients by the oracle's
    inverse NTT; every limb gives the same small integers) follows round(N(0, 3.2^2)):
    mean 0, variance 3.2^2 + 1/12, |e| <= 40, and the frequency of each |e| = k,
    k <= 10, within 5 binomial standard deviations of 2^[k>0] (Phi((k+1/2)/3.2) -
    Phi((k-1/2)/3.2))."""
    import math
    N, L = 4096, 2
    cl = _client(N, L, 1, 2, q_bits=50)
    o = orc.OracleCtx(N, cl.q, cl.p, 2)
    es = []
    for seed in range(8):
        s = cl.keygen_secret(100 + seed)
        pk = cl.keygen_public(s, 200 + seed).reshape(2, L, N)
        se = cl.secret_eval(s, L).reshape(L, N)
        limbs = []
        for i in range(L):
            q = int(cl.q[i])
            ee = np.array([(int(b) + int(a) * int(x)) % q for b, a, x in zip(pk[0][i], pk[1][i], se[i])], np.uint64)
            c = o.intt(ee, i).astype(object)
            limbs.append(np.array([int(v) - q if int(v) > q // 2 else int(v) for v in c], np.int64))
        assert np.array_equal(limbs[0], limbs[1])
        es.append(limbs[0])
    e = np.concatenate(es)
    n = e.size
    assert abs(e.mean()) < 0.1
    assert abs(e.var() - (3.2 ** 2 + 1.0 / 12.0)) < 0.4
    assert np.abs(e).max() <= 40
    Phi = lambda x: 0.5 * math.erfc(-x / math.sqrt(2.0))
    for k in range(11):
        p = (Phi((k + 0.5) / 3.2) - Phi((k - 0.5) / 3.2)) * (2 if k else 1)
        got = int((np.abs(e) == k).sum())
        assert abs(got - n * p) <= 5 * math.sqrt(n * p * (1 - p)) + 2, (k, got, n * p)
