"""A CPU stand-in for anns_fhe_amd.driver.FheContext backed by the oracle.

Same method names and word-offset semantics (poly arena, key arena, submit work
items, linear_sum, rescale_batch), so host-side protocol code written against
FheContext runs unchanged on the oracle: the GPU run and this run must produce
the same words.  Test infrastructure only.
"""
import numpy as np

NONE = (1 << 64) - 1


class OracleContext:
    def __init__(self, orc, N, L, K, dnum, q_primes, p_primes, poly_words, key_words):
        self.o = orc.OracleCtx(N, q_primes, p_primes, dnum)
        self.N, self.L, self.K, self.dnum = N, L, K, dnum
        self.q, self.p = list(q_primes), list(p_primes)
        self.primes = self.q + self.p
        self.pm = np.zeros(poly_words, np.uint64)
        self.km = np.zeros(key_words, np.uint64)
        self.poly_words, self.key_words = poly_words, key_words
        self.relin = None
        self.rot = {}
        self.lib = self  # host code calls ctx.lib.fhea_set_relin_key / fhea_copy_poly (ctx.h, ...)
        self.h = None

    def fhea_set_relin_key(self, _h, off, limbs):
        self.relin = (off, limbs)
        return 0

    def fhea_copy_poly(self, _h, dst, src, count):
        self.pm[dst:dst + count] = self.pm[src:src + count].copy()
        return 0

    def write_poly(self, off, data):
        a = np.ascontiguousarray(data, dtype=np.uint64).reshape(-1)
        self.pm[off:off + a.size] = a

    def read_poly(self, off, count):
        return self.pm[off:off + count].copy()

    def read_polys(self, ranges, out=None):
        total = sum(c for _, c in ranges)
        a = np.zeros(total, np.uint64) if out is None else out
        at = 0
        for off, c in ranges:
            a[at:at + c] = self.pm[off:off + c]
            at += c
        return a[:total]

    def write_polys(self, ranges, data):
        a = np.ascontiguousarray(data, dtype=np.uint64).reshape(-1)
        at = 0
        for off, c in ranges:
            self.pm[off:off + c] = a[at:at + c]
            at += c

    def read_polys_async(self, ranges, out):
        # the oracle runs every call to completion: the read is done at once
        self.read_polys(ranges, out)
        self._tickets = getattr(self, "_tickets", 0) + 1
        return self._tickets - 1

    def read_wait(self, ticket):
        assert 0 <= ticket < getattr(self, "_tickets", 0)

    def write_key(self, off, data):
        a = np.ascontiguousarray(data, dtype=np.uint64).reshape(-1)
        self.km[off:off + a.size] = a

    def submit(self, works):
        i = 0
        while i < len(works):
            w = works[i]
            if w["kind"] == 7:  # runs of the same input share one ModUp
                j = i
                while j < len(works) and works[j]["kind"] == 7 and works[j]["a0"] == w["a0"] and \
                        works[j]["a1"] == w["a1"]:
                    j += 1
                run = works[i:j]
                st = self.o.eval_rotations_hoisted(self.pm, self.km, w["a0"], w["a1"], [r["galois"] for r in run],
                                                   [r["evk"] for r in run], w.get("key_limbs", w["limbs"]),
                                                   [r["d0"] for r in run], [r["d1"] for r in run], w["limbs"])
                assert st == 0, st
                i = j
                continue
            self._one(w)
            i += 1

    def _one(self, w):
        k, l = w["kind"], w["limbs"]
        if k == 2:
            st = self.o.eval_automorphism_ks(self.pm, self.km, w["galois"], w["evk"], w.get("key_limbs", l),
                                             w["a0"], w["a1"], w["d0"], w["d1"], l)
        elif k == 5:
            off, kl = self.relin
            st = self.o.eval_mult_relin(self.pm, self.km, off, kl, w["a0"], w["a1"], w["b0"], w["b1"], w["d0"],
                                        w["d1"], l)
        elif k == 6:
            st = self.o.eval_add(self.pm, w["a0"], w["a1"], w["b0"], w["b1"], w["d0"], w["d1"], l)
        else:
            raise ValueError(k)
        assert st == 0, (k, st)

    def linear_sum(self, ct0, ct1, pt, out0, out1, limbs, ptb=None, bias=None):
        assert self.o.linear_sum(self.pm, ct0, ct1, pt, out0, out1, limbs, ptb=ptb, bias=bias) == 0

    def linear_sum_batch(self, jobs, limbs):
        # no output overlaps an input, so jobs run one after the other give the words
        # of one launch
        for j in jobs:
            self.linear_sum(j["ct0"], j["ct1"], j["pt"], j["out0"], j["out1"], limbs, ptb=j.get("ptb"),
                            bias=j.get("bias"))

    def linear_sum_groups(self, groups, limbs):
        for g in groups:
            R = np.asarray(g["ct0"]).shape[-1]
            c0 = np.asarray(g["ct0"], np.uint64).reshape(-1, R)
            c1 = np.asarray(g["ct1"], np.uint64).reshape(-1, R)
            pt = np.asarray(g["pt"], np.uint64).reshape(-1, R)
            o0 = np.asarray(g["out0"], np.uint64).reshape(c0.shape[0], pt.shape[0])
            o1 = np.asarray(g["out1"], np.uint64).reshape(c0.shape[0], pt.shape[0])
            bias = g.get("bias")
            for i in range(c0.shape[0]):
                self.linear_sum(c0[i], c1[i], pt, o0[i], o1[i], limbs, ptb=g.get("ptb"),
                                bias=tuple(np.asarray(bias, np.uint64).reshape(-1, 2)[i]) if bias is not None else None)

    def rescale_batch(self, src, dst, n, limbs):
        W1 = (limbs - 1) * self.N
        self.rescale_scatter(src, [dst + z * 2 * W1 for z in range(n)], limbs)

    def rescale_scatter(self, src, dst, limbs):
        W, W1 = limbs * self.N, (limbs - 1) * self.N
        for z, d in enumerate(dst):
            for k in range(2):
                a = self.pm[src + z * 2 * W + k * W:src + z * 2 * W + (k + 1) * W]
                self.pm[d + k * W1:d + (k + 1) * W1] = self.o.rescale(a).reshape(-1)


class OracleDriver:
    """FHEHostDriver's composites (fhe_host_driver.cpp:248-404) on an OracleContext."""

    def __init__(self, octx):
        self.c = octx
        self.o = octx.o

    def set_relin_key(self, evk, key_limbs=0):
        self.c.relin = (evk, key_limbs)

    def add_rotation_key(self, step, galois, evk, key_limbs=0):
        self.c.rot[step] = (galois, evk, key_limbs)

    def eval_add(self, a0, a1, b0, b1, d0, d1, limbs):
        assert self.o.eval_add(self.c.pm, a0, a1, b0, b1, d0, d1, limbs) == 0

    def eval_sub_plain(self, a0, a1, pt, d0, d1, limbs):
        assert self.o.eval_sub_plain(self.c.pm, a0, a1, pt, d0, d1, limbs) == 0

    def eval_mult_relin_rescale(self, a0, a1, b0, b1, d0, d1, in_limbs):
        off, kl = self.c.relin
        assert self.o.eval_mult_relin_rescale(self.c.pm, self.c.km, off, kl, a0, a1, b0, b1, d0, d1, in_limbs) == 0

    def eval_mult_plain_rescale(self, a0, a1, pt, d0, d1, in_limbs):
        assert self.o.eval_mult_plain_rescale(self.c.pm, a0, a1, pt, d0, d1, in_limbs) == 0

    def eval_rotate(self, in0, in1, d0, d1, total_rot, limbs):
        keys = [(st, g, off, kl) for st, (g, off, kl) in sorted(self.c.rot.items())]
        assert self.o.eval_rotate(self.c.pm, self.c.km, keys, in0, in1, d0, d1, total_rot, limbs) == 0
