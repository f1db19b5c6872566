"""Serving loop on the CPU (oracle-backed contexts): admission by slots and HBM
budget, first-come-first-served order, continuous batching across two parameter
sets, and results identical alone and in the mix."""
import numpy as np
import pytest

from anns_fhe_amd import ivfpq as IV
from anns_fhe_amd import serving as SV


@pytest.fixture(scope="module")
def tiny():
    x, _ = IV.make_gaussian_mixture(1200, dim=128, n_centres=4, sigma=0.1, seed=8)
    ix = IV.IvfPqIndex(n_list=4, M=8, K=256).train(x, train_size=1200, iters=5, seed=3).add(x)
    return x, ix


def _oracle_set(orc, name, N, L, K, dnum, ix, slots, seed, group=None):
    from anns_fhe_amd.client import CkksClient
    from tests.oracle_ctx import OracleContext
    ck = CkksClient(N, L, K, dnum)
    words, kw, max_g = SV.ParamSet.sizes(N, L, K, dnum, ix, slots, max_probe=2)
    ctx = OracleContext(orc, N, L, K, dnum, ck.q, ck.p, words, kw)
    return SV.ParamSet(name, ctx, ck, ix, slots, max_g, seed=seed, group=group)


def test_serving_loop_mixed_sets_match_alone(orc, tiny):
    x, ix = tiny
    a = _oracle_set(orc, "small", 4096, 4, 2, 2, ix, 2, seed=1)
    b = _oracle_set(orc, "large", 8192, 5, 2, 3, ix, 2, seed=2)
    rng = np.random.default_rng(0)
    qs = [x[rng.integers(len(x))] + rng.normal(0, 0.05, 128).astype(np.float32) for _ in range(5)]
    reqs = [SV.Request(rid=i, pset="small" if i % 2 == 0 else "large", query=qs[i], n_probe=1 + i % 2,
                       arrival=0.0) for i in range(5)]
    # budget admits the static state plus three slots in all
    budget = a.static_bytes + b.static_bytes + 2 * a.slot_bytes + b.slot_bytes
    loop = SV.ServingLoop([a, b], budget, max_batch=4, client_workers=2)
    out = loop.run(reqs)
    assert all(r.t_done > 0 for r in out)
    ps = IV.PlaintextSearcher(ix)
    for r in out:
        pids, pdist, pprobes = ps.search(r.query, r.n_probe, 10)
        assert list(r.probes) == list(pprobes)
        assert set(r.ids) == set(pids)
        assert np.allclose(r.dist, pdist, atol=1e-3)
    # the same requests one at a time give identical decrypted distances
    for r in out[:3]:
        solo = SV.Request(rid=r.rid, pset=r.pset, query=r.query, n_probe=r.n_probe, arrival=0.0)
        SV.ServingLoop([a, b], budget, max_batch=1, client_workers=1).run([solo])
        assert np.array_equal(solo.ids, r.ids)
        assert np.array_equal(solo.dist, r.dist)


def test_admission_respects_budget_and_order(orc, tiny):
    x, ix = tiny
    a = _oracle_set(orc, "small", 4096, 4, 2, 2, ix, 3, seed=1)
    with pytest.raises(ValueError):
        SV.ServingLoop([a], a.static_bytes - 1)
    rng = np.random.default_rng(1)
    reqs = [SV.Request(rid=i, pset="small", query=x[rng.integers(len(x))], n_probe=1, arrival=0.0)
            for i in range(4)]
    # one slot's worth of budget: requests run strictly one after another, in arrival order
    loop = SV.ServingLoop([a], a.static_bytes + a.slot_bytes, max_batch=4, client_workers=1)
    out = loop.run(reqs)
    admits = [r.t_admit for r in out]
    dones = [r.t_done for r in out]
    assert admits == sorted(admits)
    for i in range(1, 4):
        assert out[i].t_admit >= out[i - 1].events["stage3"]


def test_replica_sets_share_a_group(orc, tiny):
    """Two replicas of one parameter set (same seed) serve the requests of their group:
    both take requests, and every result equals the plaintext search and the same
    request served by one replica alone; replicas with different keys are refused."""
    x, ix = tiny
    a1 = _oracle_set(orc, "small_a", 4096, 4, 2, 2, ix, 2, seed=1, group="small")
    a2 = _oracle_set(orc, "small_b", 4096, 4, 2, 2, ix, 2, seed=1, group="small")
    rng = np.random.default_rng(5)
    qs = [x[rng.integers(len(x))] + rng.normal(0, 0.05, 128).astype(np.float32) for _ in range(4)]
    reqs = [SV.Request(rid=i, pset="small", query=qs[i], n_probe=1 + i % 2, arrival=0.0) for i in range(4)]
    budget = a1.static_bytes + a2.static_bytes + 4 * a1.slot_bytes
    out = SV.ServingLoop([a1, a2], budget, max_batch=2, client_workers=2).run(reqs)
    assert {r.replica for r in out} == {"small_a", "small_b"}
    ps = IV.PlaintextSearcher(ix)
    for r in out:
        pids, pdist, pprobes = ps.search(r.query, r.n_probe, 10)
        assert list(r.probes) == list(pprobes)
        assert set(r.ids) == set(pids)
        solo = SV.Request(rid=r.rid, pset="small", query=r.query, n_probe=r.n_probe, arrival=0.0)
        SV.ServingLoop([a1], a1.static_bytes + a1.slot_bytes, max_batch=1, client_workers=1).run([solo])
        assert np.array_equal(solo.ids, r.ids) and np.array_equal(solo.dist, r.dist)
    other = _oracle_set(orc, "small_c", 4096, 4, 2, 2, ix, 2, seed=9, group="small")
    with pytest.raises(ValueError):
        SV.ServingLoop([a1, other], budget)
