"""ctypes wrapper around oracle/liboracle.so (the CPU restatement).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker / the timed CPU baseline.
The product (anns_fhe_amd) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int32)


_REF_SRC = "/root/reference/hardware_fpga_model/hls/src"
REF_AUTOMORPHISM = os.path.join(_HERE, "_ref", "libref_automorphism.so")


def build():
    """Compile the oracle with gcc (oracle/Makefile) and, where the reference's
    sources are present, the part of them that builds here (oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    if os.path.isdir(_REF_SRC):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def ref_automorphism():
    """The reference's own automorphism() (hardware_fpga_model/hls/src/automorphism.cpp:3-12,
    N fixed at 16384), compiled unmodified into oracle/_ref; None when not built."""
    if not os.path.exists(REF_AUTOMORPHISM):
        return None
    f = getattr(C.CDLL(REF_AUTOMORPHISM), "_Z12automorphismPKmPmS0_")  # automorphism(const u64*, u64*, const u64*)
    f.restype = None
    f.argtypes = [u64p, u64p, u64p]
    return f


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_mod_add": (C.c_uint64, [C.c_uint64] * 3),
            "orc_mod_sub": (C.c_uint64, [C.c_uint64] * 3),
            "orc_mod_mul": (C.c_uint64, [C.c_uint64] * 5 + [C.c_uint32]),
            "orc_barrett_constants": (None, [C.c_uint64, u64p, u64p, u32p]),
            "orc_inv_mod": (C.c_uint64, [C.c_uint64, C.c_uint64]),
            "orc_pow_mod": (C.c_uint64, [C.c_uint64] * 3),
            "orc_is_prime": (C.c_int, [C.c_uint64]),
            "orc_gen_primes": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32, u64p, C.c_uint32, u64p]),
            "orc_min_primitive_root": (C.c_uint64, [C.c_uint64, C.c_uint32]),
            "orc_bit_reverse": (C.c_uint32, [C.c_uint32, C.c_uint32]),
            "orc_twiddles": (None, [C.c_uint64, C.c_uint64, C.c_uint32, u64p, u64p]),
            "orc_ntt_forward": (None, [u64p, C.c_uint32, u64p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32]),
            "orc_intt_inverse": (None, [u64p, C.c_uint32, u64p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32]),
            "orc_automorphism": (None, [u64p, u64p, u64p, C.c_uint32]),
            "orc_galois_map": (None, [C.c_uint32, C.c_uint64, u32p]),
            "orc_ctx_create": (C.c_void_p, [C.c_uint32] * 4 + [u64p, u64p]),
            "orc_ctx_destroy": (None, [C.c_void_p]),
            "orc_ctx_set_psi": (C.c_int, [C.c_void_p, u64p]),
            "orc_ctx_alpha": (C.c_uint32, [C.c_void_p]),
            "orc_ctx_tables": (None, [C.c_void_p, u64p, u64p, u64p, u64p, u32p, u64p, u64p]),
            "orc_modup_tables": (C.c_uint32, [C.c_void_p, C.c_uint32, C.c_uint32, u64p, u64p]),
            "orc_moddown_tables": (None, [C.c_void_p, C.c_uint32, u64p, u64p, u64p]),
            "orc_rescale_tables": (None, [C.c_void_p, C.c_uint32, u64p, u64p]),
            "orc_rescale": (None, [C.c_void_p, u64p, u64p, C.c_uint32]),
            "orc_fast_bases_conv": (None, [C.c_uint32, u64p, u64p, C.c_uint32, C.c_uint32, u64p, u64p, u64p, u64p,
                                           u64p, u64p, u32p, u64p, u64p, u32p]),
            "orc_approx_mod_down": (None, [C.c_void_p, u64p, u64p, C.c_uint32]),
            "orc_key_switch": (None, [C.c_void_p, u64p, u64p, C.c_uint64, C.c_uint32, u64p, u64p, C.c_uint32]),
            "orc_accel_op": (C.c_int, [C.c_void_p, u64p, C.c_uint64, u64p, C.c_uint64, C.c_uint32,
                                       C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]),
            "orc_eval_add": (C.c_int, [C.c_void_p, u64p, C.c_uint64] + [C.c_uint64] * 6 + [C.c_uint32]),
            "orc_eval_sub_plain": (C.c_int, [C.c_void_p, u64p, C.c_uint64] + [C.c_uint64] * 5 + [C.c_uint32]),
            "orc_eval_mult_relin_rescale": (C.c_int, [C.c_void_p, u64p, C.c_uint64, u64p, C.c_uint64, C.c_uint64,
                                                      C.c_uint32] + [C.c_uint64] * 6 + [C.c_uint32]),
            "orc_eval_mult_plain_rescale": (C.c_int, [C.c_void_p, u64p, C.c_uint64] + [C.c_uint64] * 5 + [C.c_uint32]),
            "orc_eval_rotate": (C.c_int, [C.c_void_p, u64p, C.c_uint64, u64p, C.c_uint64, C.c_uint32, i32p, u64p,
                                          u64p, u32p] + [C.c_uint64] * 4 + [C.c_int32, C.c_uint32]),
            "orc_eval_automorphism_ks": (C.c_int, [C.c_void_p, u64p, C.c_uint64, u64p, C.c_uint64, C.c_uint64,
                                                   C.c_uint64, C.c_uint32] + [C.c_uint64] * 4 + [C.c_uint32]),
            "orc_eval_mult_relin": (C.c_int, [C.c_void_p, u64p, C.c_uint64, u64p, C.c_uint64, C.c_uint64,
                                              C.c_uint32] + [C.c_uint64] * 6 + [C.c_uint32]),
            "orc_eval_rotations_hoisted": (C.c_int, [C.c_void_p, u64p, C.c_uint64, u64p, C.c_uint64, C.c_uint64,
                                                     C.c_uint64, C.c_uint32, u64p, u64p, C.c_uint32, u64p, u64p,
                                                     C.c_uint32]),
            "orc_linear_sum": (C.c_int, [C.c_void_p, u64p, C.c_uint64, u64p, u64p, C.c_uint32, u64p, u64p,
                                         C.c_uint32, C.c_uint64, C.c_uint64, u64p, u64p, C.c_uint32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def P64(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u64p)


def P32(a):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u32p)


def PI32(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(i32p)


def _null64():
    return C.cast(None, u64p)


def gen_primes(two_n, bits, count, skip=()):
    out = np.zeros(count, dtype=np.uint64)
    sk = np.array(list(skip) or [0], dtype=np.uint64)
    n = lib().orc_gen_primes(two_n, bits, count, P64(sk), len(skip), P64(out))
    assert n == count, "not enough primes"
    return [int(x) for x in out]


def barrett(q):
    lo, hi, k = C.c_uint64(), C.c_uint64(), C.c_uint32()
    lib().orc_barrett_constants(q, C.byref(lo), C.byref(hi), C.byref(k))
    return lo.value, hi.value, k.value


def galois_map(N, g):
    m = np.zeros(N, dtype=np.uint32)
    lib().orc_galois_map(N, g, P32(m))
    return m


def pack_map(map_u32):
    m = map_u32.astype(np.uint64)
    return (m[0::2] | (m[1::2] << np.uint64(32))).astype(np.uint64)


class OracleCtx:
    """Parameter block (N, L Q-primes, K P-primes, dnum) plus its tables."""

    def __init__(self, N, q_primes, p_primes, dnum, psi=None):
        """psi: optional primitive 2N-th roots (Q then P) instead of the minimal ones."""
        self.N, self.L, self.K, self.dnum = N, len(q_primes), len(p_primes), dnum
        self.q = [int(x) for x in q_primes]
        self.p = [int(x) for x in p_primes]
        qa = np.array(self.q, dtype=np.uint64)
        pa = np.array(self.p or [0], dtype=np.uint64)
        self.ptr = lib().orc_ctx_create(N, self.L, self.K, dnum, P64(qa), P64(pa))
        assert self.ptr, "orc_ctx_create failed"
        if psi is not None:
            ps = np.array([int(x) for x in psi], dtype=np.uint64)
            assert lib().orc_ctx_set_psi(self.ptr, P64(ps)) == 0, "not primitive 2N-th roots"
        self.alpha = lib().orc_ctx_alpha(self.ptr)
        T = self.L + self.K
        self.psi = np.zeros(T, np.uint64)
        self.n_inv = np.zeros(T, np.uint64)
        self.m_lo = np.zeros(T, np.uint64)
        self.m_hi = np.zeros(T, np.uint64)
        self.k = np.zeros(T, np.uint32)
        self.tw = np.zeros(T * N, np.uint64)
        self.itw = np.zeros(T * N, np.uint64)
        lib().orc_ctx_tables(self.ptr, P64(self.psi), P64(self.n_inv), P64(self.m_lo), P64(self.m_hi),
                             P32(self.k), P64(self.tw), P64(self.itw))
        self.tw = self.tw.reshape(T, N)
        self.itw = self.itw.reshape(T, N)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_ctx_destroy(self.ptr)
            self.ptr = None

    @property
    def primes(self):
        return self.q + self.p

    # ---- single-limb transforms
    def ntt(self, a, idx):
        a = np.ascontiguousarray(a, dtype=np.uint64).copy()
        lib().orc_ntt_forward(P64(a), self.N, P64(np.ascontiguousarray(self.tw[idx])), self.primes[idx],
                              int(self.m_lo[idx]), int(self.m_hi[idx]), int(self.k[idx]))
        return a

    def intt(self, a, idx):
        a = np.ascontiguousarray(a, dtype=np.uint64).copy()
        lib().orc_intt_inverse(P64(a), self.N, P64(np.ascontiguousarray(self.itw[idx])), int(self.n_inv[idx]),
                               self.primes[idx], int(self.m_lo[idx]), int(self.m_hi[idx]), int(self.k[idx]))
        return a

    def ntt_poly(self, a, idxs=None):
        a = a.reshape(-1, self.N)
        idxs = range(a.shape[0]) if idxs is None else idxs
        return np.stack([self.ntt(a[r], i) for r, i in enumerate(idxs)])

    def intt_poly(self, a, idxs=None):
        a = a.reshape(-1, self.N)
        idxs = range(a.shape[0]) if idxs is None else idxs
        return np.stack([self.intt(a[r], i) for r, i in enumerate(idxs)])

    # ---- tables
    def modup_tables(self, sizeQl, part):
        inv = np.zeros(64, np.uint64)
        hat = np.zeros(64 * (sizeQl + self.K), np.uint64)
        n_out = lib().orc_modup_tables(self.ptr, sizeQl, part, P64(inv), P64(hat))
        a = self.alpha
        nparts = -(-sizeQl // a)
        sz = sizeQl - a * part if part == nparts - 1 else a
        return inv[:sz].copy(), hat[: sz * n_out].reshape(sz, n_out).copy()

    def moddown_tables(self, sizeQl):
        pinv = np.zeros(sizeQl, np.uint64)
        phinv = np.zeros(max(self.K, 1), np.uint64)
        phat = np.zeros(max(self.K, 1) * sizeQl, np.uint64)
        lib().orc_moddown_tables(self.ptr, sizeQl, P64(pinv), P64(phinv), P64(phat))
        return pinv, phinv[: self.K], phat[: self.K * sizeQl].reshape(self.K, sizeQl)

    def rescale_tables(self, sizeQl):
        a = np.zeros(sizeQl, np.uint64)
        b = np.zeros(sizeQl, np.uint64)
        lib().orc_rescale_tables(self.ptr, sizeQl, P64(a), P64(b))
        return a[: sizeQl - 1], b[: sizeQl - 1]

    # ---- L2
    def rescale(self, a):
        L = a.size // self.N
        a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1)
        out = np.zeros((L - 1) * self.N, np.uint64)
        lib().orc_rescale(self.ptr, P64(a), P64(out), L)
        return out.reshape(L - 1, self.N)

    def approx_mod_down(self, a, sizeQl):
        a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1)
        out = np.zeros(sizeQl * self.N, np.uint64)
        lib().orc_approx_mod_down(self.ptr, P64(a), P64(out), sizeQl)
        return out.reshape(sizeQl, self.N)

    def key_switch(self, c_in, key_mem, evk_off, key_limbs, sizeQl):
        c_in = np.ascontiguousarray(c_in, dtype=np.uint64).reshape(-1)
        o0 = np.zeros(sizeQl * self.N, np.uint64)
        o1 = np.zeros(sizeQl * self.N, np.uint64)
        lib().orc_key_switch(self.ptr, P64(c_in), P64(key_mem), evk_off, key_limbs, P64(o0), P64(o1), sizeQl)
        return o0.reshape(sizeQl, self.N), o1.reshape(sizeQl, self.N)

    # ---- L3 / L5 over word-addressed memories
    def accel_op(self, poly_mem, key_mem, op, src_a, src_b, dst, evk, limbs, key_limbs=0):
        km = key_mem if key_mem is not None else np.zeros(8, np.uint64)
        return lib().orc_accel_op(self.ptr, P64(poly_mem), poly_mem.size, P64(km), km.size, op, src_a, src_b,
                                  dst, evk, limbs, key_limbs)

    def eval_add(self, pm, a0, a1, b0, b1, d0, d1, limbs):
        return lib().orc_eval_add(self.ptr, P64(pm), pm.size, a0, a1, b0, b1, d0, d1, limbs)

    def eval_sub_plain(self, pm, a0, a1, pt, d0, d1, limbs):
        return lib().orc_eval_sub_plain(self.ptr, P64(pm), pm.size, a0, a1, pt, d0, d1, limbs)

    def eval_mult_relin_rescale(self, pm, km, relin_evk, key_limbs, a0, a1, b0, b1, d0, d1, limbs):
        return lib().orc_eval_mult_relin_rescale(self.ptr, P64(pm), pm.size, P64(km), km.size, relin_evk,
                                                 key_limbs, a0, a1, b0, b1, d0, d1, limbs)

    def eval_mult_plain_rescale(self, pm, a0, a1, pt, d0, d1, limbs):
        return lib().orc_eval_mult_plain_rescale(self.ptr, P64(pm), pm.size, a0, a1, pt, d0, d1, limbs)

    def eval_rotate(self, pm, km, keys, in0, in1, d0, d1, total_rot, limbs):
        """keys: list of (step, galois, evk_offset, key_limbs)."""
        steps = np.array([k[0] for k in keys] or [0], np.int32)
        gal = np.array([k[1] for k in keys] or [0], np.uint64)
        offs = np.array([k[2] for k in keys] or [0], np.uint64)
        kls = np.array([k[3] for k in keys] or [0], np.uint32)
        return lib().orc_eval_rotate(self.ptr, P64(pm), pm.size, P64(km), km.size, len(keys), PI32(steps),
                                     P64(gal), P64(offs), P32(kls), in0, in1, d0, d1, total_rot, limbs)

    def eval_automorphism_ks(self, pm, km, galois, evk, key_limbs, in0, in1, d0, d1, limbs):
        return lib().orc_eval_automorphism_ks(self.ptr, P64(pm), pm.size, P64(km), km.size, galois, evk,
                                              key_limbs, in0, in1, d0, d1, limbs)

    def eval_mult_relin(self, pm, km, relin_evk, key_limbs, a0, a1, b0, b1, d0, d1, limbs):
        return lib().orc_eval_mult_relin(self.ptr, P64(pm), pm.size, P64(km), km.size, relin_evk, key_limbs,
                                         a0, a1, b0, b1, d0, d1, limbs)

    def linear_sum(self, pm, ct0, ct1, pt, out0, out1, limbs, ptb=None, bias=None):
        """pt: G x R offsets; ptb: G offsets or None; bias: (b0, b1) or None."""
        none = (1 << 64) - 1
        pt = np.ascontiguousarray(pt, dtype=np.uint64)
        G, R = pt.shape
        c0 = np.ascontiguousarray(ct0, dtype=np.uint64)
        c1 = np.ascontiguousarray(ct1, dtype=np.uint64)
        o0 = np.ascontiguousarray(out0, dtype=np.uint64)
        o1 = np.ascontiguousarray(out1, dtype=np.uint64)
        pb = np.ascontiguousarray(ptb if ptb is not None else [none] * G, dtype=np.uint64)
        b0, b1 = bias if bias is not None else (none, none)
        return lib().orc_linear_sum(self.ptr, P64(pm), pm.size, P64(c0), P64(c1), R, P64(pt), P64(pb), G, b0, b1,
                                    P64(o0), P64(o1), limbs)

    def eval_rotations_hoisted(self, pm, km, in0, in1, galois, evk, key_limbs, d0, d1, limbs):
        g = np.ascontiguousarray(galois, dtype=np.uint64)
        e = np.ascontiguousarray(evk, dtype=np.uint64)
        o0 = np.ascontiguousarray(d0, dtype=np.uint64)
        o1 = np.ascontiguousarray(d1, dtype=np.uint64)
        return lib().orc_eval_rotations_hoisted(self.ptr, P64(pm), pm.size, P64(km), km.size, in0, in1, g.size,
                                                P64(g), P64(e), key_limbs, P64(o0), P64(o1), limbs)
