"""Full-size parity for the batched workloads on one GPU.

* Batched NTT/INTT: 256 polynomials x 24 limbs at N=2^16 (50-bit primes
  q = 1 mod 2^17), forward then inverse (ntt.cpp:4-80 per limb).  Checked word
  for word against the oracle on sampled polynomials, on every output word
  through a linearity checksum (sum over polynomials of NTT(a_p) equals
  NTT(sum_p a_p)), and by the round trip (INTT(NTT(a)) == a on all 6,144 limbs).
* Batched hybrid key switching: 64 ciphertexts at N=2^16, L=24, K=8, dnum=3,
  each through sigma_g then KS with its own resident key, g = 5^k mod 2N with
  k seeded uniform in [1, N/2) (fhe_host_driver.cpp:331-359 with one explicit
  Galois element).  The batch equals per-ciphertext calls word for word, and
  sampled ciphertexts equal the oracle.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, L, K, DNUM = 1 << 16, 24, 8, 3


def say(*a):
    print("[gpu-configs]", *a, flush=True)
    sys.stdout.flush()


def _uniform(rng, primes, shape_lead):
    """Array of shape shape_lead + (len(primes), N), limb i uniform in [0, q_i)."""
    out = np.empty(shape_lead + (len(primes), N), np.uint64)
    for i, q in enumerate(primes):
        out[..., i, :] = rng.integers(0, q, shape_lead + (N,), dtype=np.uint64)
    return out


def test_batched_ntt_256x24_full_size(orc):
    from anns_fhe_amd.driver import FheContext
    P = 256
    T = P * L * N  # words per batch (3 GiB)
    ctx = FheContext(N, L, 1, 1, q_bits=50, poly_words=2 * T + 8, key_words=8)
    try:
        assert all(q % (2 * N) == 1 and q < 2 ** 50 for q in ctx.q)
        rng = np.random.default_rng(42)
        say("generate 256 x 24 limbs")
        a = _uniform(rng, ctx.q, (P,))
        a[0, :, 0] = [q - 1 for q in ctx.q]
        a[0, :, 1] = 0
        ctx.write_poly(0, a)
        say("forward")
        ctx.ntt_batch(0, T, P, L, inverse=False)
        fwd = ctx.read_poly(T, T).reshape(P, L, N)
        o = orc.OracleCtx(N, ctx.q, ctx.p, 1)
        for p in (0, 97, P - 1):
            say("oracle poly", p)
            assert np.array_equal(fwd[p], o.ntt_poly(a[p]))
        say("linearity checksum over all polynomials")
        qs = np.array(ctx.q, np.uint64)[:, None]
        sum_fwd = fwd.sum(axis=0, dtype=np.uint64) % qs  # 256 * 2^50 < 2^64
        sum_in = a.sum(axis=0, dtype=np.uint64) % qs
        assert np.array_equal(sum_fwd, o.ntt_poly(sum_in))
        del fwd
        say("inverse (in place)")
        ctx.ntt_batch(T, T, P, L, inverse=True)
        back = ctx.read_poly(T, T).reshape(P, L, N)
        assert np.array_equal(back, a)
    finally:
        ctx.close()


def test_batched_keyswitch_64_cts_distinct_keys(orc):
    from anns_fhe_amd.driver import FheContext
    n_ct = 64
    W = L * N
    ksz = 2 * DNUM * (L + K) * N  # 96 MiB per key
    ctx = FheContext(N, L, K, DNUM, q_bits=50, poly_words=3 * n_ct * 2 * W + 8, key_words=n_ct * ksz + 8)
    try:
        assert ctx.key_size_words(L) == ksz
        rng = np.random.default_rng(42)
        ks = rng.integers(1, N // 2, n_ct)
        gal = [pow(5, int(k), 2 * N) for k in ks]
        say("keys: 64 resident, 8 distinct contents")
        kidx = list(range(L)) + [L + j for j in range(K)]
        kprimes = [ctx.primes[i] for i in kidx]
        blobs = [_uniform(rng, kprimes, (2 * DNUM,)).reshape(-1) for _ in range(8)]
        for z in range(n_ct):
            ctx.write_key(z * ksz, blobs[z % 8])
        say("ciphertexts")
        cts = _uniform(rng, ctx.q, (n_ct, 2))
        ctx.write_poly(0, cts)
        out_b, out_s = 2 * n_ct * W, 4 * n_ct * W
        works = [dict(kind=2, limbs=L, key_limbs=L, a0=2 * z * W, a1=(2 * z + 1) * W, d0=out_b + 2 * z * W,
                      d1=out_b + (2 * z + 1) * W, galois=gal[z], evk=z * ksz) for z in range(n_ct)]
        say("batched submit")
        ctx.submit(works)
        batched = ctx.read_poly(out_b, 2 * n_ct * W)
        say("single calls")
        for z, w in enumerate(works):
            st = ctx.lib.fhea_eval_automorphism_ks(ctx.h, w["galois"], w["evk"], L, w["a0"], w["a1"],
                                                   out_s + 2 * z * W, out_s + (2 * z + 1) * W, L)
            assert st == 0
        single = ctx.read_poly(out_s, 2 * n_ct * W)
        assert np.array_equal(batched, single)
        o = orc.OracleCtx(N, ctx.q, ctx.p, DNUM)
        for z in (5, 42):
            say("oracle ct", z)
            pm = np.zeros(4 * W, np.uint64)
            pm[:2 * W] = cts[z].reshape(-1)
            km = blobs[z % 8]
            assert o.eval_automorphism_ks(pm, km, gal[z], 0, L, 0, W, 2 * W, 3 * W, L) == 0
            assert np.array_equal(batched[2 * z * W:2 * (z + 1) * W], pm[2 * W:4 * W])
        # the same 64 rotations recorded as one HIP graph (as the C3 bench runs them),
        # replayed on new ciphertexts with eager work in between: every replay equals
        # the eager result, and one item of the last replay equals the oracle's
        say("graph capture")
        ctx.graph_begin()
        ctx.submit(works)
        gid = ctx.graph_end()
        for rep in range(3):
            cts = _uniform(rng, ctx.q, (n_ct, 2))
            ctx.write_poly(0, cts)
            ctx.graph_launch(gid)
            ctx.sync()
            replay = ctx.read_poly(out_b, 2 * n_ct * W)
            ctx.submit(works)
            say("graph replay", rep)
            assert np.array_equal(replay, ctx.read_poly(out_b, 2 * n_ct * W)), rep
        z = 17
        pm = np.zeros(4 * W, np.uint64)
        pm[:2 * W] = cts[z].reshape(-1)
        assert o.eval_automorphism_ks(pm, blobs[z % 8], gal[z], 0, L, 0, W, 2 * W, 3 * W, L) == 0
        assert np.array_equal(replay[2 * z * W:2 * (z + 1) * W], pm[2 * W:4 * W])
        ctx.graph_destroy(gid)
    finally:
        ctx.close()


@pytest.mark.parametrize("chain", ["q50", "ref"])
def test_full_size_ckks_semantics_on_gpu(orc, chain):
    """Client-encrypted vectors at N=2^16, L=24, K=8, dnum=3 through the GPU:
    Dec(HMult(Enc(x), Enc(y))) = x*y and Dec(rotate(Enc(x), r)) = x shifted by r.
    q50: 50-bit primes (FP64 kernels); ref: the reference's chain shape, 60-bit q0
    and P with 45-bit q1..q23 (integer kernels, FP64 transforms on 45-bit limbs)."""
    from anns_fhe_amd.client import CkksClient
    from anns_fhe_amd.driver import FheContext, FHEHostDriver
    W = L * N
    ksz = 2 * DNUM * (L + K) * N
    steps = [1, 2, 4, -1]
    primes = dict(q_bits=50)
    if chain == "ref":
        big = orc.gen_primes(2 * N, 60, K + 1)
        primes = dict(q_primes=[big[0]] + orc.gen_primes(2 * N, 45, L - 1), p_primes=big[1:])
    ctx = FheContext(N, L, K, DNUM, poly_words=8 * W + 8, key_words=(1 + len(steps)) * ksz + 8, **primes)
    try:
        cl = CkksClient(0, 0, 0, 0, ctx=ctx)
        assert cl.primes == ctx.primes
        drv = FHEHostDriver(ctx)
        s = cl.keygen_secret(11)
        pk = cl.keygen_public(s, 12)
        say("keys")
        ctx.write_key(0, cl.relin_key(s, L, 13))
        drv.set_relin_key(0, L)
        for i, st in enumerate(steps):
            g = cl.galois_for_step(st)
            ctx.write_key((1 + i) * ksz, cl.rotation_key(s, g, L, 20 + i))
            drv.add_rotation_key(st, g, (1 + i) * ksz, L)
        rng = np.random.default_rng(5)
        n = N // 2
        x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
        scale = float(ctx.q[L - 1])
        say(chain, "encrypt")
        ctx.write_poly(0, cl.encrypt_values(pk, x, L, scale, 30))
        ctx.write_poly(2 * W, cl.encrypt_values(pk, y, L, scale, 31))
        drv.eval_mult_relin_rescale(0, W, 2 * W, 3 * W, 4 * W, 5 * W, L)
        W1 = (L - 1) * N
        ct = np.concatenate([ctx.read_poly(4 * W, W1), ctx.read_poly(5 * W, W1)])
        say("decrypt product")
        got = cl.decrypt_values(s, ct, L - 1, scale * scale / ctx.q[L - 1], n)
        assert np.max(np.abs(got - x * y)) < 1e-6
        for r in (1, 7, -1):
            drv.eval_rotate(0, W, 6 * W, 7 * W, r, L)
            rot = np.concatenate([ctx.read_poly(6 * W, W), ctx.read_poly(7 * W, W)])
            got = cl.decrypt_values(s, rot, L, scale, n)
            assert np.max(np.abs(got - np.roll(x, -r))) < 1e-6, r
    finally:
        ctx.close()
