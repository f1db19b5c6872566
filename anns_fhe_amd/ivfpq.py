"""IVF-PQ approximate nearest-neighbour search, in the clear and encrypted.

Plaintext side (the reference's software_golden_model: data_loader.py,
ivf_quantizer.py, product_quantizer.py, searcher.py:26-87, and its C++ twin
fhe_core/plaintext_searcher.cpp): a seeded Gaussian-mixture dataset, IVF
(k-means coarse quantiser) + PQ (per-subspace k-means on residuals) training,
encoding to uint8 codes, and ADC search.  This is also the recall baseline.

Encrypted side, interactive mode (cpp_host_model/src/interactive_query_manager.cpp:85-274,
software_golden_model/fhe_core/fhe_searcher.cpp:129-230,377-409): the client
encrypts its query, the server evaluates squared distances to the coarse
centroids, the client decrypts them and picks n_probe clusters, the server
evaluates the distances of every candidate in those clusters, the client
decrypts them and keeps the top k.

The distance of candidate x_s (the PQ reconstruction c + r_hat, so that
||q - x_s||^2 is exactly the ADC distance sum_m ||(q - c)_m - r_hat_m||^2) is
evaluated as
    ||q||^2 - 2 <q, x_s> + ||x_s||^2
with the query replicated over the slots with period D:
  * ||q||^2 in every slot: relinearised q*q, then rotate-and-sum with strides
    1, 2, ..., D/2 (the reference's coarse tree sum, interactive_query_manager.cpp:111-124);
  * -2 <q, x_s> for S = N/2 candidates at once by the diagonal method:
    sum_{r<D} rot_r(q) (.) pt_r with pt_r[s] = -2 x_s[(s + r) mod D];
    the D-1 rotations of the query are computed once per query and shared by
    every candidate block (where the reference runs a 7-rotation tree sum, a
    mask multiply and a position rotation per block of B = S/D candidates);
  * ||x_s||^2 as a plaintext.
Everything runs at `lt` limbs (the query is encrypted at L limbs and its upper
limbs dropped); the result is rescaled once and returned at lt-1 limbs.
"""
import time

import numpy as np

from ._lib import NONE


# ---------------------------------------------------------------------------
# data and index
# ---------------------------------------------------------------------------
def make_gaussian_mixture(n, dim=128, n_centres=32, sigma=0.1, seed=42):
    """n float32 vectors from n_centres isotropic Gaussians (centres uniform in [-1, 1]^dim)."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-1.0, 1.0, (n_centres, dim)).astype(np.float32)
    labels = rng.integers(0, n_centres, n)
    x = centres[labels] + rng.normal(0.0, sigma, (n, dim)).astype(np.float32)
    return x.astype(np.float32), centres


def _nearest(x, c, chunk=65536):
    """Index of the nearest row of c for every row of x (squared L2)."""
    cn = (c.astype(np.float64) ** 2).sum(1)
    out = np.empty(len(x), np.int64)
    for i in range(0, len(x), chunk):
        xb = x[i:i + chunk].astype(np.float64)
        d = cn[None, :] - 2.0 * xb @ c.T.astype(np.float64)
        out[i:i + chunk] = np.argmin(d, axis=1)
    return out


def kmeans(x, k, iters=20, seed=0):
    """Lloyd's k-means, seeded initialisation from distinct data points."""
    rng = np.random.default_rng(seed)
    c = x[rng.choice(len(x), k, replace=False)].astype(np.float64)
    for _ in range(iters):
        a = _nearest(x, c)
        sums = np.zeros_like(c)
        np.add.at(sums, a, x.astype(np.float64))
        cnt = np.bincount(a, minlength=k)
        empty = cnt == 0
        c[~empty] = sums[~empty] / cnt[~empty, None]
        if empty.any():
            c[empty] = x[rng.choice(len(x), int(empty.sum()), replace=False)]
    return c.astype(np.float32)


class IvfPqIndex:
    """IVF (n_list centroids) + PQ (M subquantisers of K centroids) over residuals."""

    def __init__(self, n_list=32, M=8, K=256):
        self.n_list, self.M, self.K = n_list, M, K

    def train(self, x, train_size=65536, iters=20, seed=0):
        rng = np.random.default_rng(seed)
        t = x[rng.choice(len(x), min(train_size, len(x)), replace=False)]
        self.dim = x.shape[1]
        self.sub = self.dim // self.M
        self.centroids = kmeans(t, self.n_list, iters, seed)
        res = t - self.centroids[_nearest(t, self.centroids)]
        self.codebooks = np.stack([kmeans(res[:, m * self.sub:(m + 1) * self.sub], self.K, iters, seed + 1 + m)
                                   for m in range(self.M)])
        return self

    def add(self, x):
        self.assign = _nearest(x, self.centroids).astype(np.int32)
        res = x - self.centroids[self.assign]
        self.codes = np.stack([_nearest(res[:, m * self.sub:(m + 1) * self.sub], self.codebooks[m])
                               for m in range(self.M)], axis=1).astype(np.uint8)
        self.lists = [np.flatnonzero(self.assign == c) for c in range(self.n_list)]
        return self

    def reconstruct(self, ids):
        ids = np.asarray(ids)
        r = np.concatenate([self.codebooks[m][self.codes[ids, m]] for m in range(self.M)], axis=1)
        return self.centroids[self.assign[ids]] + r


class PlaintextSearcher:
    """ADC search over an IvfPqIndex (searcher.py:26-87, plaintext_searcher.cpp)."""

    def __init__(self, index):
        self.ix = index

    def coarse(self, q):
        return ((self.ix.centroids - q) ** 2).sum(1)

    def search(self, q, n_probe, top_k):
        ix = self.ix
        probes = np.argsort(self.coarse(q), kind="stable")[:n_probe]
        cand, dist = [], []
        for c in probes:
            ids = ix.lists[c]
            qr = q - ix.centroids[c]
            lut = np.stack([((ix.codebooks[m] - qr[m * ix.sub:(m + 1) * ix.sub]) ** 2).sum(1)
                            for m in range(ix.M)])  # [M][K]
            d = np.zeros(len(ids), np.float64)
            for m in range(ix.M):
                d += lut[m][ix.codes[ids, m]]
            cand.append(ids)
            dist.append(d)
        cand = np.concatenate(cand) if cand else np.zeros(0, np.int64)
        dist = np.concatenate(dist) if dist else np.zeros(0)
        order = topk_stable(dist, top_k)
        return cand[order], dist[order], probes


def topk_stable(d, k):
    """np.argsort(d, kind="stable")[:k] without sorting all of d: the k smallest by
    partition, ties broken by index."""
    d = np.asarray(d)
    if k >= len(d):
        return np.argsort(d, kind="stable")[:k]
    if k <= 0:
        return np.zeros(0, np.int64)
    t = np.partition(d, k - 1)[k - 1]
    below = np.flatnonzero(d < t)
    cand = np.concatenate([below, np.flatnonzero(d == t)[:k - len(below)]])
    return cand[np.argsort(d[cand], kind="stable")]


def exact_knn(x, q, k):
    """The first k of a stable argsort of the squared distances (float64) from q to
    the rows of x.  Pass x as float64 to avoid a conversion per call."""
    return topk_stable(((np.asarray(x, np.float64) - np.asarray(q, np.float64)) ** 2).sum(1), k)


def recall_at(found, truth):
    return len(set(int(i) for i in found) & set(int(i) for i in truth)) / max(1, len(truth))


# ---------------------------------------------------------------------------
# encrypted, interactive
# ---------------------------------------------------------------------------
class FheIvfPqClient:
    """Client of the interactive protocol: keys, query encryption, decryption of
    distances, cluster selection and top-k (fhe_context_manager.cpp:457-512,
    interactive_query_manager.cpp:166-174,276-290)."""

    def __init__(self, ckks, dim=128, log_scale=40, lt=2, seed=None):
        """seed None (default): keys and encryptions draw from the secure generator;
        an integer makes them reproducible (tests, replicas sharing keys), not secure."""
        self.ck = ckks
        self.D = dim
        self.scale = 2.0 ** log_scale
        self.lt = lt
        self.seed0 = seed
        self.s = ckks.keygen_secret(self.next_seed(0))
        self.s_eval = ckks.secret_eval(self.s)  # decryption reuses the transformed secret
        self.pk = ckks.keygen_public(self.s, self.next_seed(1))
        self.seed = 100

    def next_seed(self, k):
        """The seed of draw k: None (secure) without a base seed, else base + k."""
        return None if self.seed0 is None else self.seed0 + k

    def evaluation_keys(self, threads=16):
        """Relinearisation key and rotation keys for r = 1..D-1, all at lt limbs
        (generated on host threads; each key has its own seed, so the result does
        not depend on the thread count)."""
        import concurrent.futures as cf
        rel = self.ck.relin_key(self.s, self.lt, self.next_seed(self.seed))

        def one(r):
            return r, self.ck.rotation_key(self.s, self.ck.galois_for_step(r), self.lt, self.next_seed(self.seed + r))

        with cf.ThreadPoolExecutor(max(1, threads)) as ex:
            rots = dict(ex.map(one, range(1, self.D)))
        return rel, rots

    def encrypt_query(self, q, limbs=None):
        limbs = limbs or self.ck.L
        self.seed += 1
        return self.ck.encrypt_values(self.pk, np.asarray(q, np.float64), limbs, self.scale,
                                      self.next_seed(self.seed))

    def result_scale(self, q_last):
        return self.scale * self.scale / q_last

    def decrypt_distances(self, ct, limbs, scale, n):
        return self.decrypt_distances_batch(ct, limbs, scale, n)[0]

    def decrypt_distances_batch(self, cts, limbs, scale, n, threads=1):
        """Slot values of several result ciphertexts (count x 2*limbs*N words): (count, n)."""
        return self.ck.decrypt_values_batch(self.s_eval, cts, limbs, scale, n, threads)


class _HostArray:
    """Pageable stand-in for PinnedArray (contexts without a device)."""

    def __init__(self, words):
        self.words = words
        self.array = np.zeros(words, np.uint64)

    def close(self):
        self.array = None


class FheIvfPqServer:
    """Server side: the encoded database resident in HBM, keys, and the batched
    homomorphic stages.  Word layout of the poly arena:
        [plaintext blocks | query slots]
    block b: D plaintexts pt_r (lt limbs each) then ptb (||x_s||^2);
    query slot: ct (L limbs x 2), D-1 rotations, norm, tmp, outputs, rescaled outputs."""

    def __init__(self, ctx, ckks, index, lt=2, log_scale=40, max_queries=64, max_blocks_per_query=None,
                 threads=0, log=None):
        self.ctx, self.ck, self.ix = ctx, ckks, index
        self.N, self.L, self.lt = ctx.N, ctx.L, lt
        self.S = ctx.N // 2
        self.D = index.dim
        assert self.S % self.D == 0 and (self.D & (self.D - 1)) == 0
        self.scale = 2.0 ** log_scale
        self.log = log or (lambda *a: None)
        Wl = lt * self.N
        # ---- blocks: 0 = coarse centroids, then every cluster in chunks of S candidates
        self.blocks = [("coarse", np.arange(index.n_list))]
        self.cluster_blocks = []
        for c in range(index.n_list):
            ids = index.lists[c]
            bl = []
            for o in range(0, max(1, len(ids)), self.S):
                if len(ids) == 0:
                    break
                bl.append(len(self.blocks))
                self.blocks.append((c, ids[o:o + self.S]))
            self.cluster_blocks.append(bl)
        self.block_words = (self.D + 1) * Wl
        if max_blocks_per_query is None:
            per = sorted((len(b) for b in self.cluster_blocks), reverse=True)
            max_blocks_per_query = max(1, sum(per[:8]))
        self.max_g = max_blocks_per_query
        self.pt_words = len(self.blocks) * self.block_words
        self.ct_words = 2 * self.L * self.N
        # per query slot: the ciphertext, its rotations, norm, scratch and its results;
        # then one output area shared by the slots of a call (packed in job order)
        self.slot_words = (self.ct_words + (self.D - 1) * 2 * Wl + 4 * Wl + self.max_g * 2 * (lt - 1) * self.N)
        self.out_base = self.pt_words + max_queries * self.slot_words
        self.max_queries = max_queries
        self._pinned = None  # page-locked read buffers (engine contexts)
        self._bufs = {}
        self._copy_pool = None
        self.use_graphs = False
        self.max_graphs = 16
        self.graphs = {}
        need = self.out_base + max_queries * self.max_g * 2 * Wl
        assert need <= ctx.poly_words, (need, ctx.poly_words)
        self._encode_blocks(threads)

    @staticmethod
    def arena_words(N, L, lt, D, n_blocks, max_queries, max_g):
        Wl = lt * N
        slot = 2 * L * N + (D - 1) * 2 * Wl + 4 * Wl + max_g * 2 * (lt - 1) * N
        return n_blocks * (D + 1) * Wl + max_queries * (slot + max_g * 2 * Wl) + 64

    def _encode_blocks(self, threads):
        D, S, lt = self.D, self.S, self.lt
        t0 = time.perf_counter()
        s = np.arange(S)
        for b, (_, ids) in enumerate(self.blocks):
            if b == 0:
                x = self.ix.centroids.astype(np.float64)
            else:
                x = self.ix.reconstruct(ids).astype(np.float64)
            n = len(x)
            vals = np.zeros((D + 1, S), np.float64)
            for r in range(D):
                vals[r, :n] = -2.0 * x[s[:n], (s[:n] + r) % D]
            vals[D, :n] = (x ** 2).sum(1)
            pts = self.ck.encode_batch(vals[:D], lt, self.scale, threads)
            ptb = self.ck.encode_batch(vals[D:], lt, self.scale * self.scale, threads)
            base = b * self.block_words
            self.ctx.write_poly(base, pts)
            self.ctx.write_poly(base + D * lt * self.N, ptb)
        self.log(f"encoded {len(self.blocks)} blocks in {time.perf_counter() - t0:.1f} s")

    # ---- keys
    def load_keys(self, relin_key, rot_keys, key_base=0):
        ksz = relin_key.size
        self.ctx.write_key(key_base, relin_key)
        self.ctx.lib.fhea_set_relin_key(self.ctx.h, key_base, self.lt)
        self.rot_off = {}
        for k, (r, key) in enumerate(sorted(rot_keys.items())):
            off = key_base + (1 + k) * ksz
            self.ctx.write_key(off, key)
            self.rot_off[r] = off

    # ---- offsets inside a query slot
    def slot_base(self, z):
        return self.pt_words + z * self.slot_words

    def _offs(self, z):
        N, L, lt, D = self.N, self.L, self.lt, self.D
        Wl = lt * N
        b = self.slot_base(z)
        o = {"q0": b, "q1": b + L * N}
        o["rot"] = b + self.ct_words  # rotation r (1..D-1) at rot + (r-1)*2*Wl
        o["norm"] = o["rot"] + (D - 1) * 2 * Wl
        o["tmp"] = o["norm"] + 2 * Wl
        o["res"] = o["tmp"] + 2 * Wl  # result g at res + g*2*(lt-1)*N
        return o

    def upload(self, z, ct):
        self.ctx.write_poly(self.slot_base(z), ct)

    def upload_batch(self, items):
        """Query ciphertexts [(z, ct), ...] into their slots: staged back to back in a
        page-locked buffer, then one gathered write."""
        if not items or not hasattr(self.ctx, "write_polys"):
            for z, ct in items:
                self.upload(z, ct)
            return
        W = self.ct_words
        buf = self._result_buffer(len(items) * W, which="up")
        for k, (z, ct) in enumerate(items):
            a = np.asarray(ct, np.uint64).reshape(-1)
            assert a.size == W, (a.size, W)
            buf[k * W:(k + 1) * W] = a
        self.ctx.write_polys([(self.slot_base(z), W) for z, _ in items], buf[:len(items) * W])

    def prepare(self, slots):
        """Rotations of every query by 1..D-1 (hoisted: one ModUp per query) and the
        encrypted ||q||^2 (all at lt limbs)."""
        N, lt, D = self.N, self.lt, self.D
        Wl = lt * N
        works = []
        for z in slots:
            o = self._offs(z)
            for r in range(1, D):
                d = o["rot"] + (r - 1) * 2 * Wl
                works.append(dict(kind=7, limbs=lt, key_limbs=lt, a0=o["q0"], a1=o["q1"], d0=d, d1=d + Wl,
                                  galois=self.ck.galois_for_step(r), evk=self.rot_off[r]))
        self.ctx.submit(works)
        self.ctx.submit([dict(kind=5, limbs=lt, a0=self._offs(z)["q0"], a1=self._offs(z)["q1"],
                              b0=self._offs(z)["q0"], b1=self._offs(z)["q1"], d0=self._offs(z)["norm"],
                              d1=self._offs(z)["norm"] + Wl) for z in slots])
        stride = 1
        while stride < D:
            rot, add = [], []
            for z in slots:
                o = self._offs(z)
                rot.append(dict(kind=2, limbs=lt, key_limbs=lt, a0=o["norm"], a1=o["norm"] + Wl, d0=o["tmp"],
                                d1=o["tmp"] + Wl, galois=self.ck.galois_for_step(stride), evk=self.rot_off[stride]))
                add.append(dict(kind=6, limbs=lt, a0=o["norm"], a1=o["norm"] + Wl, b0=o["tmp"], b1=o["tmp"] + Wl,
                                d0=o["norm"], d1=o["norm"] + Wl))
            self.ctx.submit(rot)
            self.ctx.submit(add)
            stride *= 2

    def distances(self, z, block_ids):
        self.distances_batch([(z, block_ids)])

    def distances_batch(self, jobs):
        """Encrypted squared distances, jobs = [(z, block_ids)]: slot z's rescaled
        ciphertexts (lt-1 limbs) are left at o['res'] + g*2*(lt-1)*N."""
        self._distances_groups([([z], list(b), [0]) for z, b in jobs], {z: len(b) for z, b in jobs})

    def _distances_groups(self, groups, sizes):
        """groups: [(slots, block_ids, starts)] - every slot of a group needs every block
        of it, starts[i] = index of block_ids[0] in the result list of slots[i];
        sizes: {z: number of results}.  One launch per unit shape for all groups."""
        N, lt, D = self.N, self.lt, self.D
        Wl = lt * N
        first, k = {}, 0
        for z, G in sizes.items():
            assert 0 < G <= self.max_g
            first[z] = k
            k += G
        assert k <= self.max_queries * self.max_g
        rot = (np.arange(1, D, dtype=np.uint64) - 1) * (2 * Wl)
        lin = []
        for zs, block_ids, starts in groups:
            os_ = [self._offs(z) for z in zs]
            b = np.asarray(block_ids, np.uint64) * self.block_words
            ng = len(block_ids)
            out0 = np.array([[self.out_base + (first[z] + s0 + g) * 2 * Wl for g in range(ng)]
                             for z, s0 in zip(zs, starts)], np.uint64)
            lin.append(dict(
                ct0=np.array([np.concatenate([[o["q0"]], o["rot"] + rot]) for o in os_], np.uint64),
                ct1=np.array([np.concatenate([[o["q1"]], o["rot"] + rot + Wl]) for o in os_], np.uint64),
                bias=[(o["norm"], o["norm"] + Wl) for o in os_],
                pt=b[:, None] + np.arange(D, dtype=np.uint64)[None, :] * Wl, ptb=b + D * Wl,
                out0=out0, out1=out0 + Wl))
        self.ctx.linear_sum_groups(lin, lt)
        res = np.concatenate([self._offs(z)["res"] + np.arange(G, dtype=np.uint64) * (2 * (lt - 1) * N)
                              for z, G in sizes.items()])
        self.ctx.rescale_scatter(self.out_base, res, lt)

    def _result_buffer(self, words, which="main"):
        """Page-locked read buffer `which` ("main": stage 3 and gathered reads; "read":
        reads queued beside them), at least `words` long (page-locking is slow: a
        buffer grows by half again each time)."""
        cur = self._bufs.get(which)
        if cur is None or cur.words < words:
            size = max(words, self.max_queries * 4 * 2 * (self.lt - 1) * self.N)
            if cur is not None:
                size = max(size, cur.words * 3 // 2)
                cur.close()
            if hasattr(self.ctx, "graph_begin"):  # a device context: page-locked
                from .driver import PinnedArray
                cur = PinnedArray(size)
            else:
                cur = _HostArray(size)
            self._bufs[which] = cur
        return cur.array

    def read_distances_batch(self, items, copy=True):
        """{z: (G, 2*(lt-1)*N) array} for items [(z, G), ...]: one gathered read into
        page-locked memory, then copies out on host threads.  copy=False returns views
        of the page-locked buffer instead, valid until the next read."""
        n = 2 * (self.lt - 1) * self.N
        if not hasattr(self.ctx, "read_polys"):
            return {z: self.read_distances(z, G) for z, G in items}
        ranges = [(self._offs(z)["res"], G * n) for z, G in items]
        total = sum(c for _, c in ranges)
        buf = self.ctx.read_polys(ranges, out=self._result_buffer(total))
        at, parts = 0, []
        for z, G in items:
            parts.append((z, at, G))
            at += G * n
        return self._hand_out(buf, parts, copy)

    def _hand_out(self, buf, parts, copy):
        """{z: (G, n) results} from [(z, word offset in buf, G)]: views of buf, or copies
        made on host threads."""
        n = 2 * (self.lt - 1) * self.N
        if not copy:
            return {z: buf[a:a + G * n].reshape(G, n) for z, a, G in parts}
        if self._copy_pool is None:
            import concurrent.futures as cf
            self._copy_pool = cf.ThreadPoolExecutor(8)
        return dict(self._copy_pool.map(lambda p: (p[0], buf[p[1]:p[1] + p[2] * n].reshape(p[2], n).copy()), parts))

    def read_distances(self, z, G):
        n = 2 * (self.lt - 1) * self.N
        if self._pinned is None and hasattr(self.ctx, "graph_begin"):
            from .driver import PinnedArray
            self._pinned = PinnedArray(self.max_g * n)
        if self._pinned is not None:
            buf = self.ctx.read_poly(self._offs(z)["res"], G * n, out=self._pinned.array)
            return buf.reshape(G, n).copy()
        return self.ctx.read_poly(self._offs(z)["res"], G * n).reshape(G, n)

    # ---- stages (interactive_query_manager.cpp:85-274)
    def stage1_coarse_centroid_distance(self, slots):
        """Rotations, norm and coarse distances of the queries in `slots`.  With
        use_graphs, the launch sequence of a slot set is recorded once as a HIP graph
        (after one eager run) and replayed for later queries in the same slots."""
        key = tuple(slots)
        gid = self.graphs.get(key) if self.use_graphs else None
        if gid is not None:
            self.ctx.graph_launch(gid)
            return
        coarse = ([(list(slots), [0], [0] * len(slots))], {z: 1 for z in slots})
        self.prepare(slots)
        self._distances_groups(*coarse)
        if self.use_graphs and len(self.graphs) < self.max_graphs:
            self.ctx.sync()
            self.ctx.graph_begin()
            try:
                self.prepare(slots)
                self._distances_groups(*coarse)
            finally:
                self.graphs[key] = self.ctx.graph_end()

    def stage3_fine_candidate_batch_distance(self, z, clusters):
        return self.stage3_fine_candidate_distances([(z, clusters)])[z]

    def stage3_fine_candidate_distances(self, probes):
        """Candidate distances of several queries in one launch: probes is a list of
        (z, clusters); returns {z: candidate block ids}."""
        blocks = {z: [b for c in clusters for b in self.cluster_blocks[c]] for z, clusters in probes}
        self._stage3_launch(probes, blocks)
        return blocks

    def stage3_and_read(self, probes, chunks=2, copy=True):
        """Stage 3 and the read of its results: ({z: candidate block ids}, {z: (G, n)
        results}).  The queries go in `chunks` parts of about equal result counts, each
        launched and then read on the copy stream, so that a part's read-back overlaps
        the next part's sums (a cluster's plaintext words are then read once per part).
        copy=False returns views of the page-locked buffer, valid until the next read."""
        if chunks <= 1 or not hasattr(self.ctx, "read_polys_async"):
            blocks = self.stage3_fine_candidate_distances(probes)
            return blocks, self.read_distances_batch([(z, len(b)) for z, b in blocks.items() if b], copy=copy)
        pending = self.stage3_submit(probes, chunks)
        return pending["blocks"], self.collect(pending, copy)

    def stage3_submit(self, probes, chunks=2):
        """stage3_and_read without the wait: launches the parts, queues their reads and
        returns a pending record for collect()."""
        n = 2 * (self.lt - 1) * self.N
        blocks = {z: [b for c in clusters for b in self.cluster_blocks[c]] for z, clusters in probes}
        total = sum(len(b) for b in blocks.values())
        buf = self._result_buffer(total * n)
        parts, cur, acc = [], [], 0
        for z, clusters in probes:
            cur.append((z, clusters))
            acc += len(blocks[z])
            if acc * chunks >= total * (len(parts) + 1) and len(parts) < chunks - 1:
                parts.append(cur)
                cur = []
        if cur:
            parts.append(cur)
        where, tickets, at = [], [], 0
        for part in parts:
            self._stage3_launch(part, {z: blocks[z] for z, _ in part})
            items = [(z, len(blocks[z])) for z, _ in part if blocks[z]]
            if not items:
                continue
            ranges = [(self._offs(z)["res"], G * n) for z, G in items]
            cnt = sum(c for _, c in ranges)
            tickets.append(self.ctx.read_polys_async(ranges, buf[at:at + cnt]))
            for z, G in items:
                where.append((z, at, G))
                at += G * n
        return {"blocks": blocks, "buf": buf, "where": where, "tickets": tickets}

    def read_distances_submit(self, items):
        """Queues the read of results [(z, G), ...] (e.g. stage 1's) into a buffer of its
        own on the copy stream; returns a pending record for collect()."""
        n = 2 * (self.lt - 1) * self.N
        ranges = [(self._offs(z)["res"], G * n) for z, G in items]
        buf = self._result_buffer(sum(c for _, c in ranges), which="read")
        where, at = [], 0
        for z, G in items:
            where.append((z, at, G))
            at += G * n
        tickets = [self.ctx.read_polys_async(ranges, buf)] if ranges else []
        return {"buf": buf, "where": where, "tickets": tickets}

    def collect(self, pending, copy=True):
        """Waits for a pending record's reads: {z: (G, n) results} (views of the buffer,
        valid until its next use, or copies)."""
        for tk in pending["tickets"]:
            self.ctx.read_wait(tk)
        return self._hand_out(pending["buf"], pending["where"], copy)

    def _stage3_launch(self, probes, blocks):
        # one group per probed cluster: the queries probing it x its blocks
        members = {}
        for z, clusters in probes:
            s0 = 0
            for c in clusters:
                if self.cluster_blocks[c]:
                    members.setdefault(c, ([], []))
                    members[c][0].append(z)
                    members[c][1].append(s0)
                s0 += len(self.cluster_blocks[c])
        groups = [(zs, self.cluster_blocks[c], st) for c, (zs, st) in members.items()]
        if groups:
            self._distances_groups(groups, {z: len(b) for z, b in blocks.items() if b})


def execute_interactive_queries(server, client, queries, n_probe, top_k, client_threads=16, stage3_chunks=8):
    """Runs the whole protocol for a batch of queries; returns per query
    (ids, distances, probes) and a timing record.  Client-side work (encryption,
    decryption, selection) runs on `client_threads` host threads, as independent
    clients would; stage 3 goes in `stage3_chunks` parts whose result reads overlap
    the next part's evaluation."""
    import concurrent.futures as cf
    ck, N = client.ck, server.N
    B = len(queries)
    assert B <= server.max_queries
    lt1 = server.lt - 1
    q_last = ck.q[server.lt - 1]
    scale = client.result_scale(q_last)
    pool = cf.ThreadPoolExecutor(max(1, client_threads))
    t = {}
    t0 = time.perf_counter()
    seeds = [client.next_seed(client.seed + 1 + z) for z in range(B)]
    client.seed += B
    cts = list(pool.map(lambda zq: ck.encrypt_values(client.pk, np.asarray(zq[1], np.float64), ck.L, client.scale,
                                                      seeds[zq[0]]), enumerate(queries)))
    server.upload_batch(list(enumerate(cts)))
    t["client_encrypt_upload"] = time.perf_counter() - t0
    slots = list(range(B))
    t1 = time.perf_counter()
    server.stage1_coarse_centroid_distance(slots)
    cmap = server.read_distances_batch([(z, 1) for z in slots], copy=False)
    coarse = [cmap[z][0] for z in slots]
    t["server_stage1"] = time.perf_counter() - t1
    # stage 2 (client): decrypt the coarse distances, pick n_probe clusters
    t2 = time.perf_counter()

    def choose(z):
        d = client.decrypt_distances(coarse[z], lt1, scale, server.S)[:server.ix.n_list]
        return np.argsort(d, kind="stable")[:n_probe]

    probes = list(pool.map(choose, slots))
    t["client_stage2"] = time.perf_counter() - t2
    t3 = time.perf_counter()
    # views of the read buffer: every client is done with them before the next read
    bmap, fmap = server.stage3_and_read([(z, probes[z]) for z in slots], chunks=stage3_chunks, copy=False)
    blocks = [bmap[z] for z in slots]
    fine = [fmap.get(z) for z in slots]
    t["server_stage3"] = time.perf_counter() - t3
    t4 = time.perf_counter()

    def top(z):
        ids, dist = [], []
        dec = client.decrypt_distances_batch(fine[z], lt1, scale, server.S) if blocks[z] else None
        for g, b in enumerate(blocks[z]):
            cand = server.blocks[b][1]
            ids.append(cand)
            dist.append(dec[g][:len(cand)])
        ids = np.concatenate(ids) if ids else np.zeros(0, np.int64)
        dist = np.concatenate(dist) if dist else np.zeros(0)
        order = topk_stable(dist, top_k)
        return ids[order], dist[order], probes[z]

    out = list(pool.map(top, slots))
    t["client_stage4"] = time.perf_counter() - t4
    pool.shutdown()
    return out, t
