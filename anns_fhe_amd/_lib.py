"""ctypes binding of libfhea.so (the C-ABI declared in include/fhea.h).

The library is loaded from this package directory and nowhere else.  There is
no fallback: if libfhea.so is missing or cannot be loaded, every entry point
raises FheLibraryError.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfhea.so")

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
i8p = C.POINTER(C.c_int8)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)
vp = C.c_void_p

STATUS = {
    0: "ok", -1: "unknown op code", -2: "offset or size outside the arena",
    -3: "offset not a multiple of 8 words", -4: "invalid argument", -5: "missing evaluation key",
    -6: "device allocation failed", -7: "HIP runtime error", -8: "invalid context",
    -9: "work exceeds the static scratch area of the no-allocation entry point",
}
ERR_SCRATCH = -9
NOALLOC_SCRATCH_WORDS = 8 << 20

OP_NTT, OP_INTT, OP_POLY_ADD, OP_POLY_SUB, OP_POLY_MUL_NTT = 1, 2, 3, 4, 5
OP_AUTOMORPHISM, OP_KEY_SWITCH, OP_RESCALE, OP_LOAD_CONTEXT = 6, 7, 8, 99
WORK_MULT_RELIN_RESCALE, WORK_AUTOMORPHISM_KS, WORK_NTT, WORK_INTT = 1, 2, 3, 4
WORK_MULT_RELIN, WORK_ADD, WORK_ROTATE_HOISTED = 5, 6, 7
NONE = (1 << 64) - 1
SEED_RANDOM = (1 << 64) - 1  # FHEA_SEED_RANDOM: the secure (ChaCha20) source
TUNE_NTT_SPLIT_FWD, TUNE_NTT_SPLIT_INV, TUNE_NTT_WG_PER_CU, TUNE_CHUNK, TUNE_STREAMS = 1, 2, 3, 4, 5


class FheLibraryError(RuntimeError):
    pass


class FheError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__(f"{what}: {STATUS.get(status, status)} ({status})")


class fhea_params(C.Structure):
    _fields_ = [("N", C.c_uint32), ("L", C.c_uint32), ("K", C.c_uint32), ("dnum", C.c_uint32),
                ("q_bits", C.c_uint32), ("device", C.c_uint32), ("q_primes", u64p), ("p_primes", u64p),
                ("poly_words", C.c_uint64), ("key_words", C.c_uint64), ("psi", u64p)]


class fhea_work(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("limbs", C.c_uint32), ("a0", C.c_uint64), ("a1", C.c_uint64),
                ("b0", C.c_uint64), ("b1", C.c_uint64), ("d0", C.c_uint64), ("d1", C.c_uint64),
                ("galois", C.c_uint64), ("evk", C.c_uint64), ("key_limbs", C.c_uint32),
                ("reserved", C.c_uint32)]


# name -> (restype, argtypes); must match include/fhea.h
SIGNATURES = {
    "fhea_ctx_create": (C.c_int, [C.POINTER(fhea_params), C.POINTER(vp)]),
    "fhea_ctx_destroy": (C.c_int, [vp]),
    "fhea_ctx_info": (C.c_int, [vp, u32p]),
    "fhea_ctx_primes": (C.c_int, [vp, u64p]),
    "fhea_ctx_table": (C.c_int, [vp, C.c_int, u64p, C.c_uint64]),
    "fhea_ctx_set_stream": (C.c_int, [vp, vp]),
    "fhea_ctx_stream": (vp, [vp]),
    "fhea_status_string": (C.c_char_p, [C.c_int]),
    "fhea_write_poly": (C.c_int, [vp, C.c_uint64, u64p, C.c_uint64]),
    "fhea_read_poly": (C.c_int, [vp, C.c_uint64, u64p, C.c_uint64]),
    "fhea_read_polys": (C.c_int, [vp, C.c_uint32, u64p, u64p, u64p]),
    "fhea_write_polys": (C.c_int, [vp, C.c_uint32, u64p, u64p, u64p]),
    "fhea_read_polys_async": (C.c_int, [vp, C.c_uint32, u64p, u64p, u64p, C.POINTER(C.c_uint32)]),
    "fhea_read_wait": (C.c_int, [vp, C.c_uint32]),
    "fhea_write_key": (C.c_int, [vp, C.c_uint64, u64p, C.c_uint64]),
    "fhea_read_key": (C.c_int, [vp, C.c_uint64, u64p, C.c_uint64]),
    "fhea_copy_poly": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_uint64]),
    "fhea_ctx_set_tuning": (C.c_int, [vp, C.c_int, C.c_uint32]),
    "fhea_ctx_get_tuning": (C.c_int, [vp, C.c_int, u32p]),
    "fhea_mem_info": (C.c_int, [C.c_uint32, u64p, u64p]),
    "fhea_probe_valu_rate": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_double)]),
    "fhea_host_alloc": (vp, [C.c_uint64]),
    "fhea_host_free": (C.c_int, [vp]),
    "fhea_poly_arena": (vp, [vp]),
    "fhea_key_arena": (vp, [vp]),
    "fhea_poly_words": (C.c_uint64, [vp]),
    "fhea_key_words": (C.c_uint64, [vp]),
    "fhea_op": (C.c_int, [vp, C.c_uint32] + [C.c_uint64] * 4 + [C.c_uint32, C.c_uint32]),
    "fhea_eval_add": (C.c_int, [vp] + [C.c_uint64] * 6 + [C.c_uint32]),
    "fhea_eval_sub_plain": (C.c_int, [vp] + [C.c_uint64] * 5 + [C.c_uint32]),
    "fhea_set_relin_key": (C.c_int, [vp, C.c_uint64, C.c_uint32]),
    "fhea_eval_mult_relin_rescale": (C.c_int, [vp] + [C.c_uint64] * 6 + [C.c_uint32]),
    "fhea_eval_mult_relin_rescale_noalloc": (C.c_int, [vp] + [C.c_uint64] * 6 + [C.c_uint32]),
    "fhea_eval_mult_plain_rescale": (C.c_int, [vp] + [C.c_uint64] * 5 + [C.c_uint32]),
    "fhea_add_rotation_key": (C.c_int, [vp, C.c_int32, C.c_uint64, C.c_uint64, C.c_uint32]),
    "fhea_eval_rotate": (C.c_int, [vp] + [C.c_uint64] * 4 + [C.c_int32, C.c_uint32]),
    "fhea_eval_automorphism_ks": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_uint32] + [C.c_uint64] * 4
                                  + [C.c_uint32]),
    "fhea_submit": (C.c_int, [vp, C.POINTER(fhea_work), C.c_uint32]),
    "fhea_ntt_batch": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int]),
    "fhea_sync": (C.c_int, [vp]),
    "fhea_graph_begin": (C.c_int, [vp]),
    "fhea_graph_end": (C.c_int, [vp, u32p]),
    "fhea_graph_launch": (C.c_int, [vp, C.c_uint32]),
    "fhea_graph_destroy": (C.c_int, [vp, C.c_uint32]),
    "fhea_rescale_batch": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]),
    "fhea_rescale_scatter": (C.c_int, [vp, C.c_uint64, u64p, C.c_uint32, C.c_uint32]),
    "fhea_eval_linear_sum": (C.c_int, [vp, u64p, u64p, C.c_uint32, u64p, u64p, C.c_uint32, C.c_uint64,
                                       C.c_uint64, u64p, u64p, C.c_uint32]),
    "fhea_eval_linear_sum_batch": (C.c_int, [vp, C.c_uint32, C.c_uint32, u64p, u64p, u32p, u64p, u64p, u64p, u64p,
                                             u64p, u64p, C.c_uint32]),
    "fhea_eval_linear_sum_groups": (C.c_int, [vp, C.c_uint32, C.c_uint32, u32p, u32p, u64p, u64p, u64p, u64p,
                                              u64p, u64p, u64p, u64p, C.c_uint32]),
    "fhea_timer_start": (C.c_int, [vp]),
    "fhea_timer_stop": (C.c_int, [vp, C.POINTER(C.c_float)]),
    "fhea_keygen_secret": (C.c_int, [vp, C.c_uint64, i8p]),
    "fhea_keygen_switch_key": (C.c_int, [vp, i8p, i64p, C.c_uint32, C.c_uint64, C.c_uint64]),
    "fhea_keygen_relin_key": (C.c_int, [vp, i8p, C.c_uint32, C.c_uint64, C.c_uint64]),
    "fhea_keygen_rotation_key": (C.c_int, [vp, i8p, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64]),
    "fhea_key_size_words": (C.c_uint64, [vp, C.c_uint32]),
    "fhea_galois_for_step": (C.c_uint64, [vp, C.c_int32]),
    "fhea_client_create": (C.c_int, [C.POINTER(fhea_params), C.POINTER(vp)]),
    "fhea_client_from_ctx": (C.c_int, [vp, C.POINTER(vp)]),
    "fhea_client_destroy": (C.c_int, [vp]),
    "fhea_client_info": (C.c_int, [vp, u32p]),
    "fhea_client_primes": (C.c_int, [vp, u64p]),
    "fhea_client_table": (C.c_int, [vp, C.c_int, u64p, C.c_uint64]),
    "fhea_client_keygen_secret": (C.c_int, [vp, C.c_uint64, i8p]),
    "fhea_client_switch_key": (C.c_int, [vp, i8p, i64p, C.c_uint32, C.c_uint64, u64p]),
    "fhea_client_relin_key": (C.c_int, [vp, i8p, C.c_uint32, C.c_uint64, u64p]),
    "fhea_client_rotation_key": (C.c_int, [vp, i8p, C.c_uint64, C.c_uint32, C.c_uint64, u64p]),
    "fhea_encode": (C.c_int, [vp, f64p, f64p, C.c_uint32, C.c_uint32, C.c_double, u64p]),
    "fhea_encode_batch": (C.c_int, [vp, f64p, f64p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, u64p,
                                    C.c_uint32]),
    "fhea_decode": (C.c_int, [vp, u64p, C.c_uint32, C.c_double, C.c_uint32, f64p, f64p]),
    "fhea_chacha20_block": (C.c_int, [u32p, C.c_uint32, u32p, u32p]),
    "fhea_keygen_public": (C.c_int, [vp, i8p, C.c_uint64, u64p]),
    "fhea_encrypt": (C.c_int, [vp, u64p, u64p, C.c_uint32, C.c_uint64, u64p]),
    "fhea_decrypt": (C.c_int, [vp, i8p, u64p, C.c_uint32, u64p]),
    "fhea_client_set_host_ntt": (C.c_int, [vp, C.c_int]),
    "fhea_secret_eval": (C.c_int, [vp, i8p, C.c_uint32, u64p]),
    "fhea_decrypt_decode_batch": (C.c_int, [vp, u64p, u64p, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, f64p,
                                            f64p, C.c_uint32]),
}

_lib = None


def load():
    """Load libfhea.so from this package; raise if it is absent (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FheLibraryError(f"{LIB_PATH} is missing: build it with `python -m anns_fhe_amd.build`")
    try:
        lib = C.CDLL(LIB_PATH)
    except OSError as e:
        raise FheLibraryError(f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def mem_info(device=0):
    """(free, total) bytes of device memory (fhea_mem_info)."""
    lib = load()
    f, t = C.c_uint64(), C.c_uint64()
    check(lib.fhea_mem_info(device, C.byref(f), C.byref(t)), "fhea_mem_info")
    return f.value, t.value


def probe_valu_rate(kind, iters=4096, device=0):
    """Lane operations per second of one VALU instruction kind (fhea_probe_valu_rate):
    0 v_mul_lo_u32, 1 v_mul_hi_u32, 2 v_mad_u64_u32, 3 v_fma_f64."""
    lib = load()
    r = C.c_double()
    check(lib.fhea_probe_valu_rate(device, kind, iters, C.byref(r)), "fhea_probe_valu_rate")
    return r.value


def check(status, what):
    if status != 0:
        raise FheError(status, what)
