"""The reference's interactive query sequence, op by op over FHEHostDriver.

Mirrors InteractiveQueryManager (cpp_host_model/src/interactive_query_manager.cpp:85-274)
and the stage sequence of integration_tools/generate_interactive_test_vectors.cpp:400-560,
generalised to n_probe >= 1 clusters and to more than one slot-ciphertext of
candidates, with the validity mask covering batch_count slots
(generate_interactive_test_vectors.cpp:521-523):

  stage 1   diff = q_rep - pt_centroids; sq = Rescale(diff * diff);
            tree sum over the D dimensions with strides 1, 2, ..., D/2
  stage 1b  for every cluster i: Rescale(sq * mask_i), rotate by i*(D-1), accumulate
            -> the n_list coarse distances in slots 0..n_list-1
  stage 2   (client) decrypt, pick n_probe clusters
  stage 3   for every batch of B = S/D candidates: diff = q_dim - pt_batch;
            sq = Rescale(diff * diff); tree sum with strides B, 2B, ..., (D/2)B;
            Rescale(sq * mask(batch_count)); rotate by -offset; accumulate
  stage 4   (client) decrypt, top-k

Every call is one single-ciphertext composite of FHEHostDriver (no batching
across ciphertexts or queries): this is the plain per-op GPU path the serving
loop is measured against.
"""
import time

import numpy as np


def _pow2_steps(n):
    out, k = [], 1
    while k <= n:
        out.append(k)
        k <<= 1
    return out


class InteractiveQueryManager:
    def __init__(self, ctx, drv, ckks, index, log_scale=50, threads=0):
        self.ctx, self.drv, self.ck, self.ix = ctx, drv, ckks, index
        self.N, self.L = ctx.N, ctx.L
        self.S = ctx.N // 2
        self.D = index.dim
        self.B = self.S // self.D
        assert index.n_list * self.D <= self.S
        self.scale = 2.0 ** log_scale
        self.W = self.L * self.N
        W = self.W
        # fixed regions (the reference's SLOT_* / SCRATCH_* offsets, fhe_host_driver.h:63-69)
        names = ["q_rep", "q_rep1", "q_dim", "q_dim1", "pt_centroids", "diff0", "diff1", "sq0", "sq1",
                 "s0", "s1", "t0", "t1", "acc0", "acc1"]
        self.off = {n: k * W for k, n in enumerate(names)}
        self.n_fixed = len(names)
        self.mask_base = self.n_fixed * W  # n_list coarse masks, then B+1 batch masks (by count)
        self.batch_base = self.mask_base + (index.n_list + self.B + 1) * W
        self.threads = threads
        self._encode_static()

    @staticmethod
    def arena_words(N, L, n_list, n_batches, n_out):
        S, W = N // 2, L * N
        B = S // 128
        return (15 + n_list + B + 1 + n_batches + 2 * n_out) * W + 64

    def rotation_steps(self):
        """Signed powers of two covering every rotation the stages use."""
        m = max(self.D, self.ix.n_list * (self.D - 1), self.S - 1)
        p = _pow2_steps(m)
        return p + [-k for k in p]

    def _encode_static(self):
        ck, D, S, L = self.ck, self.D, self.S, self.L
        c = np.zeros(S)
        for i in range(self.ix.n_list):
            c[i * D:(i + 1) * D] = self.ix.centroids[i]
        self.ctx.write_poly(self.off["pt_centroids"], ck.encode(c, L, self.scale))
        # masks multiply ciphertexts at L-1 (stage 1b) and L-1 (stage 3) limbs; their scale is
        # the prime the following rescale drops, so the ciphertext scale is kept
        ms = float(ck.q[L - 2])
        for i in range(self.ix.n_list):
            m = np.zeros(S)
            m[i * D] = 1.0
            self.ctx.write_poly(self.mask_base + i * self.W, ck.encode(m, L - 1, ms))
        for cnt in range(self.B + 1):
            m = np.zeros(S)
            m[:cnt] = 1.0
            self.ctx.write_poly(self.mask_base + (self.ix.n_list + cnt) * self.W, ck.encode(m, L - 1, ms))
        # candidate batches: dimension-major pt_batch[d*B + j] = x_hat_j[d]
        self.batches = []  # per cluster: list of (offset, ids)
        vals, where = [], []
        for cl in range(self.ix.n_list):
            ids = self.ix.lists[cl]
            bl = []
            for o in range(0, len(ids), self.B):
                b_ids = ids[o:o + self.B]
                x = self.ix.reconstruct(b_ids).astype(np.float64)
                v = np.zeros(S)
                for d in range(D):
                    v[d * self.B:d * self.B + len(b_ids)] = x[:, d]
                vals.append(v)
                off = self.batch_base + len(where) * self.W
                where.append(off)
                bl.append((off, b_ids))
            self.batches.append(bl)
        if vals:
            pts = ck.encode_batch(np.stack(vals), L, self.scale, self.threads)
            for k, off in enumerate(where):
                self.ctx.write_poly(off, pts[k])
        self.out_base = self.batch_base + len(where) * self.W

    def load_keys(self, client_secret, seed=500):
        """Relinearisation key and the signed power-of-two rotation keys, at L limbs."""
        ksz = self.ck.key_size_words(self.L)
        self.ctx.write_key(0, self.ck.relin_key(client_secret, self.L, seed))
        self.drv.set_relin_key(0, self.L)
        self.n_rot_keys = len(self.rotation_steps())
        for k, st in enumerate(self.rotation_steps()):
            g = self.ck.galois_for_step(st)
            off = (1 + k) * ksz
            self.ctx.write_key(off, self.ck.rotation_key(client_secret, g, self.L, seed + 1 + k))
            self.drv.add_rotation_key(st, g, off, self.L)

    def upload_query(self, ct_rep, ct_dim):
        self.ctx.write_poly(self.off["q_rep"], ct_rep)
        self.ctx.write_poly(self.off["q_dim"], ct_dim)

    # ---- stages
    def _tree_sum(self, a0, a1, first_stride, limbs):
        o, drv = self.off, self.drv
        stride = first_stride
        while stride < self.D * first_stride:
            drv.eval_rotate(a0, a1, o["t0"], o["t1"], stride, limbs)
            drv.eval_add(a0, a1, o["t0"], o["t1"], a0, a1, limbs)
            stride *= 2

    def stage1_coarse_centroid_distance(self):
        o, drv, L = self.off, self.drv, self.L
        drv.eval_sub_plain(o["q_rep"], o["q_rep"] + self.W, o["pt_centroids"], o["diff0"], o["diff1"], L)
        drv.eval_mult_relin_rescale(o["diff0"], o["diff1"], o["diff0"], o["diff1"], o["sq0"], o["sq1"], L)
        self._tree_sum(o["sq0"], o["sq1"], 1, L - 1)

    def stage1b_distance_compaction(self):
        o, drv, L, D = self.off, self.drv, self.L, self.D
        for i in range(self.ix.n_list):
            drv.eval_mult_plain_rescale(o["sq0"], o["sq1"], self.mask_base + i * self.W, o["s0"], o["s1"], L - 1)
            tgt = (o["acc0"], o["acc1"]) if i == 0 else (o["t0"], o["t1"])
            if i == 0:
                self.ctx.lib.fhea_copy_poly(self.ctx.h, tgt[0], o["s0"], (L - 2) * self.N)
                self.ctx.lib.fhea_copy_poly(self.ctx.h, tgt[1], o["s1"], (L - 2) * self.N)
            else:
                drv.eval_rotate(o["s0"], o["s1"], tgt[0], tgt[1], i * (D - 1), L - 2)
                drv.eval_add(o["acc0"], o["acc1"], tgt[0], tgt[1], o["acc0"], o["acc1"], L - 2)

    def read_coarse(self):
        n = (self.L - 2) * self.N
        o = self.off
        return np.concatenate([self.ctx.read_poly(o["acc0"], n), self.ctx.read_poly(o["acc1"], n)])

    def stage3_fine_candidate_batch_distance(self, clusters):
        """Returns [(ciphertext words, candidate ids)] for every output ciphertext."""
        o, drv, L, S = self.off, self.drv, self.L, self.S
        n = (L - 2) * self.N
        outs, cur_ids, offset = [], [], 0
        k_out = 0

        def out_at(k):
            return self.out_base + 2 * k * self.W, self.out_base + (2 * k + 1) * self.W

        for cl in clusters:
            for pt_off, b_ids in self.batches[cl]:
                cnt = len(b_ids)
                if offset + cnt > S:
                    outs.append((k_out, cur_ids))
                    k_out, cur_ids, offset = k_out + 1, [], 0
                drv.eval_sub_plain(o["q_dim"], o["q_dim"] + self.W, pt_off, o["diff0"], o["diff1"], L)
                drv.eval_mult_relin_rescale(o["diff0"], o["diff1"], o["diff0"], o["diff1"], o["sq0"], o["sq1"], L)
                self._tree_sum(o["sq0"], o["sq1"], self.B, L - 1)
                drv.eval_mult_plain_rescale(o["sq0"], o["sq1"], self.mask_base + (self.ix.n_list + cnt) * self.W,
                                            o["s0"], o["s1"], L - 1)
                a0, a1 = out_at(k_out)
                if offset == 0:
                    self.ctx.lib.fhea_copy_poly(self.ctx.h, a0, o["s0"], n)
                    self.ctx.lib.fhea_copy_poly(self.ctx.h, a1, o["s1"], n)
                else:
                    drv.eval_rotate(o["s0"], o["s1"], o["t0"], o["t1"], -offset, L - 2)
                    drv.eval_add(a0, a1, o["t0"], o["t1"], a0, a1, L - 2)
                cur_ids.extend(int(i) for i in b_ids)
                offset += cnt
        if cur_ids:
            outs.append((k_out, cur_ids))
        res = []
        for k, ids in outs:
            a0, a1 = out_at(k)
            res.append((np.concatenate([self.ctx.read_poly(a0, n), self.ctx.read_poly(a1, n)]), np.array(ids)))
        return res


def run_reference_query(qm, client_ck, s, pk, q, n_probe, top_k, seed):
    """One query through the op-by-op reference sequence; returns ids, distances, probes, seconds."""
    ck, D, S, L = client_ck, qm.D, qm.S, qm.L
    t0 = time.perf_counter()
    rep = np.tile(np.asarray(q, np.float64), S // D)
    dim = np.repeat(np.asarray(q, np.float64), qm.B)
    qm.upload_query(ck.encrypt_values(pk, rep, L, qm.scale, seed), ck.encrypt_values(pk, dim, L, qm.scale, seed + 1))
    qm.stage1_coarse_centroid_distance()
    qm.stage1b_distance_compaction()
    scale2 = qm.scale * qm.scale / ck.q[L - 1]
    coarse = ck.decrypt_values(s, qm.read_coarse(), L - 2, scale2, S)[:qm.ix.n_list]
    probes = np.argsort(coarse, kind="stable")[:n_probe]
    outs = qm.stage3_fine_candidate_batch_distance(probes)
    ids, dist = [], []
    for ct, cand in outs:
        ids.append(cand)
        dist.append(ck.decrypt_values(s, ct, L - 2, scale2, S)[:len(cand)])
    ids = np.concatenate(ids) if ids else np.zeros(0, np.int64)
    dist = np.concatenate(dist) if dist else np.zeros(0)
    order = np.argsort(dist, kind="stable")[:top_k]
    return ids[order], dist[order], probes, time.perf_counter() - t0
