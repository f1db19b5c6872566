"""Serving loop on the GPU: two parameter sets (N=2^14/L=8 and N=2^16/L=24)
served together from one Poisson arrival stream with n_probe uniform in 1..8.
Every result equals the plaintext ADC search, and equals the same request served
alone (identical decrypted distances)."""
import sys

import numpy as np
import pytest

from anns_fhe_amd import ivfpq as IV
from anns_fhe_amd import serving as SV

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-serving]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("k_special", [(4, 8), (2, 2)])
def test_serving_mixed_params_poisson(k_special):
    """k_special: special primes of the two sets; (2, 2) is the bench's serving
    configuration (key switching at 2 limbs needs no more)."""
    x, _ = IV.make_gaussian_mixture(20000, dim=128, n_centres=32, sigma=0.1, seed=42)
    ix = IV.IvfPqIndex(n_list=32, M=8, K=256).train(x, train_size=20000, iters=8, seed=0).add(x)
    say("param sets", k_special)
    small = SV.ParamSet.create("n14", 1 << 14, 8, k_special[0], 2, ix, max_slots=8, seed=1)
    large = SV.ParamSet.create("n16", 1 << 16, 24, k_special[1], 3, ix, max_slots=4, seed=2)
    try:
        rng = np.random.default_rng(3)
        n = 24
        arr = SV.poisson_arrivals(40.0, n, seed=4)
        reqs = [SV.Request(rid=i, pset="n14" if rng.random() < 0.5 else "n16",
                           query=x[rng.integers(len(x))] + rng.normal(0, 0.05, 128).astype(np.float32),
                           n_probe=int(rng.integers(1, 9)), arrival=float(arr[i])) for i in range(n)]
        budget = small.static_bytes + large.static_bytes + 6 * small.slot_bytes + 3 * large.slot_bytes
        loop = SV.ServingLoop([small, large], budget, max_batch=8, client_workers=8)
        say("serve")
        out = loop.run(reqs)
        st = SV.latency_stats(out)
        say("stats", st)
        ps = IV.PlaintextSearcher(ix)
        for r in out:
            pids, pdist, pprobes = ps.search(r.query, r.n_probe, 10)
            assert list(r.probes) == list(pprobes), r.rid
            assert set(r.ids) == set(pids), r.rid
            assert np.allclose(r.dist, pdist, atol=1e-3), r.rid
        say("alone")
        for r in out[:4]:
            solo = SV.Request(rid=r.rid, pset=r.pset, query=r.query, n_probe=r.n_probe, arrival=0.0)
            SV.ServingLoop([small, large], budget, max_batch=1, client_workers=1).run([solo])
            assert np.array_equal(solo.ids, r.ids)
            assert np.array_equal(solo.dist, r.dist)
    finally:
        small.close()
        large.close()


def test_slots_sized_from_device_memory():
    """ParamSet.slots_for_memory: the query slots a share of the device's free HBM
    (hipMemGetInfo through fhea_mem_info) holds; more share, more slots, capped."""
    from anns_fhe_amd import _lib
    free, total = _lib.mem_info(0)
    say("device memory", free, total)
    assert 0 < free <= total
    x, _ = IV.make_gaussian_mixture(20000, dim=128, n_centres=8, sigma=0.1, seed=1)
    ix = IV.IvfPqIndex(n_list=8, M=8, K=256).train(x, train_size=20000, iters=4, seed=0).add(x)
    s1 = SV.ParamSet.slots_for_memory(1 << 14, 8, 4, 2, ix, hbm_share=0.01, slot_cap=10 ** 6)
    s2 = SV.ParamSet.slots_for_memory(1 << 14, 8, 4, 2, ix, hbm_share=0.02, slot_cap=10 ** 6)
    assert 1 <= s1 < s2
    w1, kw, _ = SV.ParamSet.sizes(1 << 14, 8, 4, 2, ix, s1)
    assert 8 * (w1 + kw) <= 0.01 * free + 8 * kw
    assert SV.ParamSet.slots_for_memory(1 << 14, 8, 4, 2, ix, hbm_share=0.5, slot_cap=48) == 48
