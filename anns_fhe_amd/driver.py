"""Host-side mirror of the reference's FHEHostDriver over the C-ABI.

FheContext owns a libfhea context (parameters, device tables, the poly and key
arenas in HBM).  FHEHostDriver keeps the reference's method names, argument
order and meaning (cpp_host_model/src/fhe_host_driver.h:71-111): offsets are
64-bit word offsets into the poly arena ("poly_gmem") or key arena
("key_gmem").  Errors raise FheError, as the reference raises runtime_error /
out_of_range (fhe_host_driver.cpp:336-339, sim_memory_bus.cpp:13-15).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import FheError, check  # noqa: F401  (re-exported)


def _p64(a):
    return a.ctypes.data_as(_lib.u64p)


def _seed(seed):
    """None -> FHEA_SEED_RANDOM (secure generator); an integer -> reproducible stream."""
    if seed is None:
        return _lib.SEED_RANDOM
    if not 0 <= int(seed) < _lib.SEED_RANDOM:
        raise ValueError("seed must be None or an integer in [0, 2^64 - 1)")
    return int(seed)


class PinnedArray:
    """A uint64 numpy array in page-locked host memory (fhea_host_alloc): arena reads
    and writes through it run at full DMA speed instead of through a bounce buffer."""

    def __init__(self, words):
        self.lib = _lib.load()
        self.words = int(words)
        self.ptr = self.lib.fhea_host_alloc(8 * max(1, self.words))
        if not self.ptr:
            raise FheError(-6, "fhea_host_alloc")
        self.array = np.ctypeslib.as_array(C.cast(self.ptr, _lib.u64p), shape=(max(1, self.words),))[:self.words]

    def close(self):
        if getattr(self, "ptr", None):
            self.array = None
            self.lib.fhea_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FheContext:
    """Parameters + device-resident tables + arenas (replaces OP 99 and SimMemoryBus)."""

    def __init__(self, N, L, K, dnum, q_bits=50, q_primes=None, p_primes=None,
                 poly_words=1 << 23, key_words=1 << 23, device=0, psi=None):
        self.lib = _lib.load()
        prm = _lib.fhea_params()
        prm.N, prm.L, prm.K, prm.dnum, prm.q_bits, prm.device = N, L, K, dnum, q_bits, device
        self._qa = np.array(q_primes, np.uint64) if q_primes is not None else None
        self._pa = np.array(p_primes, np.uint64) if p_primes is not None else None
        prm.q_primes = _p64(self._qa) if self._qa is not None else C.cast(None, _lib.u64p)
        prm.p_primes = _p64(self._pa) if self._pa is not None else C.cast(None, _lib.u64p)
        prm.poly_words, prm.key_words = poly_words, key_words
        self._psi = np.array(psi, np.uint64) if psi is not None else None
        prm.psi = _p64(self._psi) if self._psi is not None else C.cast(None, _lib.u64p)
        h = C.c_void_p()
        check(self.lib.fhea_ctx_create(C.byref(prm), C.byref(h)), "fhea_ctx_create")
        self.h = h
        info = (C.c_uint32 * 5)()
        check(self.lib.fhea_ctx_info(self.h, info), "fhea_ctx_info")
        self.N, self.L, self.K, self.dnum, self.alpha = list(info)
        pr = np.zeros(self.L + self.K, np.uint64)
        check(self.lib.fhea_ctx_primes(self.h, _p64(pr)), "fhea_ctx_primes")
        self.primes = [int(x) for x in pr]
        self.q = self.primes[: self.L]
        self.p = self.primes[self.L:]
        self.poly_words = int(self.lib.fhea_poly_words(self.h))
        self.key_words = int(self.lib.fhea_key_words(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.fhea_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def table(self, which):
        T = self.L + self.K
        n = T if which in (0, 1) else T * self.N
        out = np.zeros(n, np.uint64)
        check(self.lib.fhea_ctx_table(self.h, which, _p64(out), n), "fhea_ctx_table")
        return out if which in (0, 1) else out.reshape(T, self.N)

    # ---- arenas (SimMemoryBus::write_poly / read_poly / write_key / read_key)
    def write_poly(self, off, data):
        a = np.ascontiguousarray(data, dtype=np.uint64).reshape(-1)
        check(self.lib.fhea_write_poly(self.h, off, _p64(a), a.size), "write_poly")

    def read_poly(self, off, count, out=None):
        """Arena words [off, off+count) into a new array, or into `out` (e.g. a
        PinnedArray's array) when given."""
        a = np.zeros(count, np.uint64) if out is None else out
        assert a.dtype == np.uint64 and a.size >= count and a.flags["C_CONTIGUOUS"]
        check(self.lib.fhea_read_poly(self.h, off, _p64(a), count), "read_poly")
        return a if out is None else a[:count]

    def read_polys(self, ranges, out=None):
        """Several arena ranges [(off, count), ...] back to back into one array (or
        into `out`, e.g. a PinnedArray's array), with one synchronisation."""
        offs = np.array([r[0] for r in ranges], np.uint64)
        cnts = np.array([r[1] for r in ranges], np.uint64)
        total = int(cnts.sum())
        a = np.zeros(total, np.uint64) if out is None else out
        assert a.dtype == np.uint64 and a.size >= total and a.flags["C_CONTIGUOUS"]
        check(self.lib.fhea_read_polys(self.h, len(ranges), _p64(offs), _p64(cnts), _p64(a)), "read_polys")
        return a[:total]

    def write_polys(self, ranges, data):
        """`data` (the ranges' words back to back, ideally page-locked) into the arena
        ranges [(off, count), ...], with one synchronisation."""
        offs = np.array([r[0] for r in ranges], np.uint64)
        cnts = np.array([r[1] for r in ranges], np.uint64)
        a = np.ascontiguousarray(data, dtype=np.uint64).reshape(-1)
        assert a.size >= int(cnts.sum())
        check(self.lib.fhea_write_polys(self.h, len(ranges), _p64(offs), _p64(cnts), _p64(a)), "write_polys")

    def read_polys_async(self, ranges, out):
        """read_polys into `out` (page-locked, e.g. a PinnedArray's array or a slice
        of it), returned at once: the copies run on the context's copy stream after
        the work submitted so far.  Returns a ticket for read_wait; `out` must stay
        alive and the ranges unwritten until then."""
        offs = np.array([r[0] for r in ranges], np.uint64)
        cnts = np.array([r[1] for r in ranges], np.uint64)
        assert out.dtype == np.uint64 and out.size >= int(cnts.sum()) and out.flags["C_CONTIGUOUS"]
        t = C.c_uint32()
        check(self.lib.fhea_read_polys_async(self.h, len(ranges), _p64(offs), _p64(cnts), _p64(out), C.byref(t)),
              "read_polys_async")
        return t.value

    def read_wait(self, ticket):
        check(self.lib.fhea_read_wait(self.h, ticket), "read_wait")

    def write_key(self, off, data):
        a = np.ascontiguousarray(data, dtype=np.uint64).reshape(-1)
        check(self.lib.fhea_write_key(self.h, off, _p64(a), a.size), "write_key")

    def read_key(self, off, count):
        a = np.zeros(count, np.uint64)
        check(self.lib.fhea_read_key(self.h, off, _p64(a), count), "read_key")
        return a

    # ---- raw op code (fhe_accel_top)
    def op(self, op_code, src_a, src_b, dst, evk=0, limbs=1, key_limbs=0):
        return self.lib.fhea_op(self.h, op_code, src_a, src_b, dst, evk, limbs, key_limbs)

    def submit(self, works):
        arr = (_lib.fhea_work * len(works))()
        for i, w in enumerate(works):
            for k, v in w.items():
                setattr(arr[i], k, v)
        check(self.lib.fhea_submit(self.h, arr, len(works)), "fhea_submit")

    def ntt_batch(self, src, dst, npolys, limbs, inverse=False):
        check(self.lib.fhea_ntt_batch(self.h, src, dst, npolys, limbs, 1 if inverse else 0), "ntt_batch")

    def linear_sum(self, ct0, ct1, pt, out0, out1, limbs, ptb=None, bias=None):
        """out_g = bias + ptb[g] + sum_r ct_r (.) pt[g][r]; pt is a G x R array of offsets."""
        pt = np.ascontiguousarray(pt, dtype=np.uint64)
        G, R = pt.shape
        c0 = np.ascontiguousarray(ct0, dtype=np.uint64)
        c1 = np.ascontiguousarray(ct1, dtype=np.uint64)
        o0 = np.ascontiguousarray(out0, dtype=np.uint64)
        o1 = np.ascontiguousarray(out1, dtype=np.uint64)
        assert c0.size == R and c1.size == R and o0.size == G and o1.size == G
        pb = np.ascontiguousarray(ptb, dtype=np.uint64) if ptb is not None else None
        b0, b1 = bias if bias is not None else (_lib.NONE, _lib.NONE)
        check(self.lib.fhea_eval_linear_sum(self.h, _p64(c0), _p64(c1), R, _p64(pt),
                                            _p64(pb) if pb is not None else C.cast(None, _lib.u64p), G, b0, b1,
                                            _p64(o0), _p64(o1), limbs), "fhea_eval_linear_sum")

    def linear_sum_batch(self, jobs, limbs):
        """One launch for several linear sums.  jobs: list of dicts with ct0, ct1
        (R offsets each), pt (G x R offsets), out0, out1 (G each) and optional ptb (G)
        and bias (pair of offsets)."""
        R = len(jobs[0]["ct0"])
        c0 = np.concatenate([np.asarray(j["ct0"], np.uint64) for j in jobs])
        c1 = np.concatenate([np.asarray(j["ct1"], np.uint64) for j in jobs])
        pts = [np.asarray(j["pt"], np.uint64).reshape(-1, R) for j in jobs]
        G = np.array([p.shape[0] for p in pts], np.uint32)
        pt = np.ascontiguousarray(np.concatenate(pts).reshape(-1))
        o0 = np.concatenate([np.asarray(j["out0"], np.uint64) for j in jobs])
        o1 = np.concatenate([np.asarray(j["out1"], np.uint64) for j in jobs])
        assert c0.size == R * len(jobs) and c1.size == c0.size and o0.size == G.sum() and o1.size == G.sum()
        nil = C.cast(None, _lib.u64p)
        pb = None
        if any(j.get("ptb") is not None for j in jobs):
            pb = np.concatenate([np.asarray(j["ptb"], np.uint64) if j.get("ptb") is not None
                                 else np.full(g, _lib.NONE, np.uint64) for j, g in zip(jobs, G)])
        b0 = b1 = None
        if any(j.get("bias") is not None for j in jobs):
            b0 = np.array([j["bias"][0] if j.get("bias") is not None else _lib.NONE for j in jobs], np.uint64)
            b1 = np.array([j["bias"][1] if j.get("bias") is not None else _lib.NONE for j in jobs], np.uint64)
        check(self.lib.fhea_eval_linear_sum_batch(
            self.h, len(jobs), R, _p64(c0), _p64(c1), G.ctypes.data_as(_lib.u32p), _p64(pt),
            _p64(pb) if pb is not None else nil, _p64(b0) if b0 is not None else nil,
            _p64(b1) if b1 is not None else nil, _p64(o0), _p64(o1), limbs), "fhea_eval_linear_sum_batch")

    def linear_sum_groups(self, groups, limbs):
        """One launch per unit shape for groups of queries sharing blocks.  groups: list
        of dicts with ct0, ct1 (nq x R offsets), pt (ng x R), out0, out1 (nq x ng) and
        optional bias (nq pairs) and ptb (ng)."""
        R = np.asarray(groups[0]["ct0"]).shape[-1]
        nq = np.array([np.asarray(g["ct0"]).reshape(-1, R).shape[0] for g in groups], np.uint32)
        ng = np.array([np.asarray(g["pt"]).reshape(-1, R).shape[0] for g in groups], np.uint32)
        cat = lambda key: np.ascontiguousarray(np.concatenate([np.asarray(g[key], np.uint64).reshape(-1)
                                                               for g in groups]))
        c0, c1, pt, o0, o1 = cat("ct0"), cat("ct1"), cat("pt"), cat("out0"), cat("out1")
        assert c0.size == R * nq.sum() and pt.size == R * ng.sum() and o0.size == (nq.astype(np.int64) * ng).sum()
        nil = C.cast(None, _lib.u64p)
        pb = b0 = b1 = None
        if any(g.get("ptb") is not None for g in groups):
            pb = np.concatenate([np.asarray(g["ptb"], np.uint64) if g.get("ptb") is not None
                                 else np.full(n, _lib.NONE, np.uint64) for g, n in zip(groups, ng)])
        if any(g.get("bias") is not None for g in groups):
            bb = np.concatenate([np.asarray(g["bias"], np.uint64).reshape(-1, 2) if g.get("bias") is not None
                                 else np.full((n, 2), _lib.NONE, np.uint64) for g, n in zip(groups, nq)])
            b0, b1 = np.ascontiguousarray(bb[:, 0]), np.ascontiguousarray(bb[:, 1])
        check(self.lib.fhea_eval_linear_sum_groups(
            self.h, len(groups), R, nq.ctypes.data_as(_lib.u32p), ng.ctypes.data_as(_lib.u32p), _p64(c0), _p64(c1),
            _p64(b0) if b0 is not None else nil, _p64(b1) if b1 is not None else nil, _p64(pt),
            _p64(pb) if pb is not None else nil, _p64(o0), _p64(o1), limbs), "fhea_eval_linear_sum_groups")

    def rescale_batch(self, src, dst, n, limbs):
        check(self.lib.fhea_rescale_batch(self.h, src, dst, n, limbs), "fhea_rescale_batch")

    def rescale_scatter(self, src, dst, limbs):
        """Rescale len(dst) ciphertexts stored back to back at src; ciphertext z to dst[z]."""
        d = np.ascontiguousarray(dst, dtype=np.uint64)
        check(self.lib.fhea_rescale_scatter(self.h, src, _p64(d), d.size, limbs), "fhea_rescale_scatter")

    # ---- HIP graphs: record a fixed sequence once, replay it on new data at the same offsets
    def graph_begin(self):
        check(self.lib.fhea_graph_begin(self.h), "fhea_graph_begin")

    def graph_end(self):
        gid = C.c_uint32()
        check(self.lib.fhea_graph_end(self.h, C.byref(gid)), "fhea_graph_end")
        return gid.value

    def graph_launch(self, gid):
        check(self.lib.fhea_graph_launch(self.h, gid), "fhea_graph_launch")

    def graph_destroy(self, gid):
        check(self.lib.fhea_graph_destroy(self.h, gid), "fhea_graph_destroy")

    def set_tuning(self, knob, value):
        check(self.lib.fhea_ctx_set_tuning(self.h, knob, value), "fhea_ctx_set_tuning")

    def get_tuning(self, knob):
        v = C.c_uint32()
        check(self.lib.fhea_ctx_get_tuning(self.h, knob, C.byref(v)), "fhea_ctx_get_tuning")
        return v.value

    def sync(self):
        check(self.lib.fhea_sync(self.h), "fhea_sync")

    def timer_start(self):
        check(self.lib.fhea_timer_start(self.h), "fhea_timer_start")

    def timer_stop(self):
        ms = C.c_float()
        check(self.lib.fhea_timer_stop(self.h, C.byref(ms)), "fhea_timer_stop")
        return ms.value

    # ---- client key generation (CPU)
    def keygen_secret(self, seed=None):
        """seed None: the secure generator; an integer: a reproducible stream (tests)."""
        s = np.zeros(self.N, np.int8)
        check(self.lib.fhea_keygen_secret(self.h, _seed(seed), s.ctypes.data_as(_lib.i8p)), "keygen_secret")
        return s

    def key_size_words(self, key_limbs):
        return int(self.lib.fhea_key_size_words(self.h, key_limbs))

    def galois_for_step(self, step):
        return int(self.lib.fhea_galois_for_step(self.h, step))

    def keygen_relin_key(self, s, key_limbs, seed, evk):
        check(self.lib.fhea_keygen_relin_key(self.h, s.ctypes.data_as(_lib.i8p), key_limbs, _seed(seed), evk),
              "keygen_relin_key")

    def keygen_rotation_key(self, s, galois, key_limbs, seed, evk):
        check(self.lib.fhea_keygen_rotation_key(self.h, s.ctypes.data_as(_lib.i8p), galois, key_limbs, _seed(seed),
                                                evk), "keygen_rotation_key")

    def keygen_switch_key(self, s, s_new, key_limbs, seed, evk):
        sn = np.ascontiguousarray(s_new, dtype=np.int64)
        check(self.lib.fhea_keygen_switch_key(self.h, s.ctypes.data_as(_lib.i8p), sn.ctypes.data_as(_lib.i64p),
                                              key_limbs, _seed(seed), evk), "keygen_switch_key")


class FHEHostDriver:
    """cpp_host_model/src/fhe_host_driver.h:55-123 over libfhea."""

    def __init__(self, ctx: FheContext):
        self.ctx = ctx
        self.lib = ctx.lib
        self.h = ctx.h
        self.N = ctx.N

    # ---- key material (initialize_context_from_directory, fhe_host_driver.cpp:105-172)
    def set_relin_key(self, evk, key_limbs=0):
        check(self.lib.fhea_set_relin_key(self.h, evk, key_limbs), "set_relin_key")

    def add_rotation_key(self, step, galois, evk, key_limbs=0):
        check(self.lib.fhea_add_rotation_key(self.h, step, galois, evk, key_limbs), "add_rotation_key")

    def load_context_to_hardware(self):
        check(self.ctx.op(_lib.OP_LOAD_CONTEXT, 0, 0, 0, 0, self.ctx.L), "load_context_to_hardware")

    # ---- primitive wrappers (fhe_host_driver.cpp:210-242)
    def ntt(self, src, dst, active_limbs):
        check(self.ctx.op(_lib.OP_NTT, src, 0, dst, 0, active_limbs), "ntt")

    def intt(self, src, dst, active_limbs):
        check(self.ctx.op(_lib.OP_INTT, src, 0, dst, 0, active_limbs), "intt")

    def poly_add(self, a, b, dst, active_limbs):
        check(self.ctx.op(_lib.OP_POLY_ADD, a, b, dst, 0, active_limbs), "poly_add")

    def poly_sub(self, a, b, dst, active_limbs):
        check(self.ctx.op(_lib.OP_POLY_SUB, a, b, dst, 0, active_limbs), "poly_sub")

    def poly_mul_ntt(self, a, b, dst, active_limbs):
        check(self.ctx.op(_lib.OP_POLY_MUL_NTT, a, b, dst, 0, active_limbs), "poly_mul_ntt")

    def automorphism(self, src, auto_map, dst, active_limbs, galois_elt=0):
        check(self.ctx.op(_lib.OP_AUTOMORPHISM, src, auto_map, dst, 0, active_limbs), "automorphism")

    def rescale(self, in_c0, in_c1, out_c0, out_c1, active_limbs):
        # OP 8 writes c1' at out_c0 + (L-1)N, as the reference does (out_c1 unused)
        check(self.ctx.op(_lib.OP_RESCALE, in_c0, in_c1, out_c0, 0, active_limbs), "rescale")

    def key_switch(self, src_c1, evk, dst_c0, dst_c1, size_ql, key_limbs=0):
        # OP 7 writes ks1 at dst_c0 + L*N (dst_c1 unused, fhe_host_driver.cpp:236-239)
        check(self.ctx.op(_lib.OP_KEY_SWITCH, src_c1, 0, dst_c0, evk, size_ql, key_limbs), "key_switch")

    # ---- composites (fhe_host_driver.cpp:248-404)
    def eval_add(self, a0, a1, b0, b1, d0, d1, limbs):
        check(self.lib.fhea_eval_add(self.h, a0, a1, b0, b1, d0, d1, limbs), "eval_add")

    def eval_sub_plain(self, a0, a1, pt, d0, d1, limbs):
        check(self.lib.fhea_eval_sub_plain(self.h, a0, a1, pt, d0, d1, limbs), "eval_sub_plain")

    def eval_mult_relin_rescale(self, a0, a1, b0, b1, d0, d1, in_limbs):
        check(self.lib.fhea_eval_mult_relin_rescale(self.h, a0, a1, b0, b1, d0, d1, in_limbs),
              "eval_mult_relin_rescale")

    def eval_mult_plain_rescale(self, a0, a1, pt, d0, d1, in_limbs):
        check(self.lib.fhea_eval_mult_plain_rescale(self.h, a0, a1, pt, d0, d1, in_limbs),
              "eval_mult_plain_rescale")

    def eval_rotate(self, in0, in1, d0, d1, total_rot, limbs):
        check(self.lib.fhea_eval_rotate(self.h, in0, in1, d0, d1, total_rot, limbs), "eval_rotate")

    def eval_automorphism_ks(self, galois, evk, key_limbs, in0, in1, d0, d1, limbs):
        check(self.lib.fhea_eval_automorphism_ks(self.h, galois, evk, key_limbs, in0, in1, d0, d1, limbs),
              "eval_automorphism_ks")
