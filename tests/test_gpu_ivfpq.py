"""Encrypted IVF-PQ on the GPU.

* The interactive protocol on the engine produces the same ciphertext words as
  the same protocol on the oracle-backed context (N=2^12, bit-exact).
* At the encrypted-search shape (N=2^14, L=8, K=4, dnum=2; 128-dim Gaussian
  mixture, n_list=32, M=8, K=256, n_probe=4): decrypted distances equal the
  plaintext ADC distances within the CKKS error, the top-10 ids equal the
  plaintext search's, and recall@10 against exact k-NN is unchanged.
"""
import sys

import numpy as np
import pytest

from anns_fhe_amd import ivfpq as IV

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-ivfpq]", *a, flush=True)
    sys.stdout.flush()


def _queries(x, B, seed):
    rng = np.random.default_rng(seed)
    return [x[rng.integers(len(x))] + rng.normal(0, 0.05, x.shape[1]).astype(np.float32) for _ in range(B)]


def test_protocol_words_equal_oracle(orc):
    from anns_fhe_amd.client import CkksClient
    from anns_fhe_amd.driver import FheContext
    from tests.oracle_ctx import OracleContext
    x, _ = IV.make_gaussian_mixture(4000, dim=128, n_centres=4, sigma=0.1, seed=3)
    ix = IV.IvfPqIndex(n_list=4, M=8, K=256).train(x, train_size=4000, iters=8, seed=1).add(x)
    N, L, K, dnum, lt = 4096, 4, 2, 2, 3
    n_blocks = 1 + sum(-(-len(l) // (N // 2)) for l in ix.lists)
    words = IV.FheIvfPqServer.arena_words(N, L, lt, 128, n_blocks, 2, 4)
    ctx = FheContext(N, L, K, dnum, poly_words=words, key_words=128 * 2 * dnum * (lt + K) * N + 8)
    try:
        ck = CkksClient(0, 0, 0, 0, ctx=ctx)
        octx = OracleContext(orc, N, L, K, dnum, ck.q, ck.p, words, 128 * ck.key_size_words(lt) + 8)
        queries = _queries(x, 2, 4)
        results = []
        for c in (ctx, octx):
            client = IV.FheIvfPqClient(ck, dim=128, log_scale=45, lt=lt, seed=7)
            server = IV.FheIvfPqServer(c, ck, ix, lt=lt, log_scale=45, max_queries=2, max_blocks_per_query=4)
            rel, rots = client.evaluation_keys()
            server.load_keys(rel, rots)
            say("protocol on", type(c).__name__)
            res, _ = IV.execute_interactive_queries(server, client, queries, n_probe=2, top_k=10)
            words_out = [c.read_poly(server._offs(z)["res"], 2 * (lt - 1) * N * 2) for z in range(2)]
            norms = [c.read_poly(server._offs(z)["norm"], 2 * lt * N) for z in range(2)]
            results.append((res, words_out, norms))
        (rg, wg, ng), (ro, wo, no) = results
        for z in range(2):
            assert np.array_equal(ng[z], no[z])
            assert np.array_equal(wg[z], wo[z])
            assert np.array_equal(rg[z][0], ro[z][0])
    finally:
        ctx.close()


def test_encrypted_search_recall_c4_shape():
    from anns_fhe_amd.client import CkksClient
    from anns_fhe_amd.driver import FheContext
    N, L, K, dnum, lt = 1 << 14, 8, 4, 2, 2
    say("dataset + index")
    x, _ = IV.make_gaussian_mixture(40000, dim=128, n_centres=32, sigma=0.1, seed=42)
    ix = IV.IvfPqIndex(n_list=32, M=8, K=256).train(x, train_size=20000, iters=10, seed=0).add(x)
    B, n_probe = 8, 4
    n_blocks = 1 + sum(-(-len(l) // (N // 2)) for l in ix.lists)
    words = IV.FheIvfPqServer.arena_words(N, L, lt, 128, n_blocks, B, 8)
    ctx = FheContext(N, L, K, dnum, poly_words=words, key_words=128 * 2 * dnum * (lt + K) * N + 8)
    try:
        ck = CkksClient(0, 0, 0, 0, ctx=ctx)
        client = IV.FheIvfPqClient(ck, dim=128, log_scale=40, lt=lt, seed=11)
        say("encode database")
        server = IV.FheIvfPqServer(ctx, ck, ix, lt=lt, log_scale=40, max_queries=B, max_blocks_per_query=8)
        say("keys")
        rel, rots = client.evaluation_keys()
        server.load_keys(rel, rots)
        queries = _queries(x, B, 9)
        say("queries")
        res, t = IV.execute_interactive_queries(server, client, queries, n_probe=n_probe, top_k=10)
        say("timing", {k: round(v, 3) for k, v in t.items()})
        ps = IV.PlaintextSearcher(ix)
        rec_fhe, rec_plain = [], []
        for q, (ids, dist, probes) in zip(queries, res):
            pids, pdist, pprobes = ps.search(q, n_probe, 10)
            assert list(probes) == list(pprobes)
            assert np.allclose(dist, pdist, atol=1e-3), np.max(np.abs(dist - pdist))
            assert set(ids) == set(pids)
            truth = IV.exact_knn(x, q, 10)
            rec_fhe.append(IV.recall_at(ids, truth))
            rec_plain.append(IV.recall_at(pids, truth))
        say("recall@10", np.mean(rec_fhe), np.mean(rec_plain))
        assert rec_fhe == rec_plain
    finally:
        ctx.close()


def test_stage1_graph_replay_equals_eager():
    """Stage 1 recorded as a HIP graph and replayed for new queries in the same slots
    gives the same coarse-distance ciphertexts as the eager launches."""
    from anns_fhe_amd.client import CkksClient
    from anns_fhe_amd.driver import FheContext
    x, _ = IV.make_gaussian_mixture(4000, dim=128, n_centres=4, sigma=0.1, seed=3)
    ix = IV.IvfPqIndex(n_list=4, M=8, K=256).train(x, train_size=4000, iters=6, seed=1).add(x)
    N, L, K, dnum, lt = 8192, 5, 2, 3, 2
    n_blocks = 1 + sum(-(-len(l) // (N // 2)) for l in ix.lists)
    words = IV.FheIvfPqServer.arena_words(N, L, lt, 128, n_blocks, 3, 4)
    ctx = FheContext(N, L, K, dnum, poly_words=words, key_words=128 * 2 * dnum * (lt + K) * N + 8)
    try:
        ck = CkksClient(0, 0, 0, 0, ctx=ctx)
        client = IV.FheIvfPqClient(ck, dim=128, lt=lt, seed=5)
        server = IV.FheIvfPqServer(ctx, ck, ix, lt=lt, max_queries=3, max_blocks_per_query=4)
        rel, rots = client.evaluation_keys()
        server.load_keys(rel, rots)
        slots = [0, 1, 2]
        server.use_graphs = True
        for rep in range(3):
            qs = _queries(x, 3, 20 + rep)
            for z, q in zip(slots, qs):
                server.upload(z, client.encrypt_query(q))
            server.stage1_coarse_centroid_distance(slots)   # rep 0: eager + capture; later: replay
            got = [server.read_distances(z, 1) for z in slots]
            server.use_graphs = False
            server.stage1_coarse_centroid_distance(slots)   # eager on the same inputs
            exp = [server.read_distances(z, 1) for z in slots]
            server.use_graphs = True
            say("stage-1 replay", rep)
            for a, b in zip(got, exp):
                assert np.array_equal(a, b)
        assert len(server.graphs) == 1
    finally:
        ctx.close()
