"""fhea_eval_mult_relin_rescale_noalloc (the entry point that allocates nothing,
its scratch in a static device array) against fhea_eval_mult_relin_rescale,
both called through the C-ABI: bit-identical words on two calls in a row, the
second at a smaller level, and FHEA_ERR_SCRATCH for work beyond the array."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[no-alloc]", *a, flush=True)
    sys.stdout.flush()


def rand_poly(rng, primes, N):
    return np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in primes])


@pytest.mark.parametrize("chain", ["q50", "q60_45"])
def test_noalloc_equals_allocating_entry_point(orc, chain):
    """q50: every prime below 2^50 (FP64 kernels); q60_45: a 60-bit q0 and P with
    45-bit q1.. (integer kernels, FP64 transforms for the 45-bit limbs)."""
    from anns_fhe_amd import _lib
    from anns_fhe_amd.driver import FheContext
    lib = _lib.load()
    N, L, K, dnum = 8192, 4, 1, 2
    W = L * N
    primes = {}
    if chain == "q60_45":
        big = orc.gen_primes(2 * N, 60, K + 1)
        primes = dict(q_primes=[big[0]] + orc.gen_primes(2 * N, 45, L - 1), p_primes=big[1:])
    ctx = FheContext(N, L, K, dnum, poly_words=16 * W, key_words=2 * dnum * (L + K) * N + 8, **primes)
    s = ctx.keygen_secret(5)
    ctx.keygen_relin_key(s, L, 6, 0)
    assert lib.fhea_set_relin_key(ctx.h, 0, L) == 0
    rng = np.random.default_rng(99)
    data = np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(4)])
    ctx.write_poly(0, data)
    for limbs in (4, 3):  # second call: a smaller input (one level down)
        say(chain, "level", limbs)
        W1 = (limbs - 1) * N
        args = (0, W, 2 * W, 3 * W)  # the first `limbs` limbs of each input polynomial
        assert lib.fhea_eval_mult_relin_rescale(ctx.h, *args, 4 * W, 5 * W, limbs) == 0
        ref = ctx.read_poly(4 * W, 2 * W)
        ctx.write_poly(6 * W, np.zeros(2 * W, np.uint64))
        st = lib.fhea_eval_mult_relin_rescale_noalloc(ctx.h, *args, 6 * W, 7 * W, limbs)
        assert st == 0, lib.fhea_status_string(st)
        got = ctx.read_poly(6 * W, 2 * W)
        assert np.array_equal(got[:W1], ref[:W1]), limbs
        assert np.array_equal(got[W:W + W1], ref[W:W + W1]), limbs
        if limbs == L:
            o = orc.OracleCtx(N, ctx.q, ctx.p, dnum)
            pm = np.zeros(6 * W, np.uint64)
            pm[:4 * W] = data
            km = ctx.read_key(0, ctx.key_size_words(L))
            assert o.eval_mult_relin_rescale(pm, km, 0, L, *args, 4 * W, 5 * W, limbs) == 0
            assert np.array_equal(got[:W1], pm[4 * W:4 * W + W1])
            assert np.array_equal(got[W:W + W1], pm[5 * W:5 * W + W1])
    ctx.close()


def test_noalloc_rejects_work_beyond_the_scratch():
    """At N=2^15, L=24, K=8, dnum=3 one pair at the top level needs 320 N words
    (more than the array): FHEA_ERR_SCRATCH, nothing launched.  Eight levels
    down (112 N words) the same context runs and equals the allocating call."""
    from anns_fhe_amd import _lib
    from anns_fhe_amd.driver import FheContext
    lib = _lib.load()
    N, L, K, dnum = 32768, 24, 8, 3
    W = L * N
    ctx = FheContext(N, L, K, dnum, poly_words=8 * W, key_words=2 * dnum * (L + K) * N + 8)
    rng = np.random.default_rng(7)
    kidx = list(range(L)) + [L + j for j in range(K)]
    key = np.concatenate([rand_poly(rng, [ctx.primes[i] for i in kidx], N).reshape(-1) for _ in range(2 * dnum)])
    ctx.write_key(0, key)
    assert lib.fhea_set_relin_key(ctx.h, 0, L) == 0
    say("too large")
    st = lib.fhea_eval_mult_relin_rescale_noalloc(ctx.h, 0, W, 2 * W, 3 * W, 4 * W, 5 * W, L)
    assert st == _lib.ERR_SCRATCH
    l = 8
    ctx.write_poly(0, np.concatenate([rand_poly(rng, ctx.q, N).reshape(-1) for _ in range(4)]))
    say("level", l)
    assert lib.fhea_eval_mult_relin_rescale(ctx.h, 0, W, 2 * W, 3 * W, 4 * W, 5 * W, l) == 0
    st = lib.fhea_eval_mult_relin_rescale_noalloc(ctx.h, 0, W, 2 * W, 3 * W, 6 * W, 7 * W, l)
    assert st == 0, lib.fhea_status_string(st)
    W1 = (l - 1) * N
    ref, got = ctx.read_poly(4 * W, 2 * W), ctx.read_poly(6 * W, 2 * W)
    assert np.array_equal(got[:W1], ref[:W1])
    assert np.array_equal(got[W:W + W1], ref[W:W + W1])
    ctx.close()
