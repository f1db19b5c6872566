"""CPU-side checks of the C-ABI library: it loads, exports every symbol that
include/fhea.h declares, and fails with a status (not a crash) without a GPU."""
import os
import re

import pytest

from anns_fhe_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "fhea.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fhea_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding covers the same set
    assert sorted(_lib.SIGNATURES) == syms


def test_status_strings():
    lib = _lib.load()
    for st, msg in _lib.STATUS.items():
        assert lib.fhea_status_string(st).decode() == msg


def test_null_context_is_an_error_not_a_crash():
    lib = _lib.load()
    assert lib.fhea_op(None, 1, 0, 0, 0, 0, 1, 0) == -8
    assert lib.fhea_sync(None) == -8
    assert lib.fhea_ctx_destroy(None) == -8


def test_bad_params_rejected_before_device_use():
    import ctypes as C
    lib = _lib.load()
    h = C.c_void_p()
    for N, L, K, dnum, bits in [(1000, 4, 1, 2, 50), (8192, 0, 1, 1, 50), (8192, 4, 0, 2, 50),
                                (8192, 4, 1, 5, 50), (8192, 4, 1, 2, 62)]:
        p = _lib.fhea_params(N=N, L=L, K=K, dnum=dnum, q_bits=bits)
        assert lib.fhea_ctx_create(C.byref(p), C.byref(h)) == -4


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_context_without_gpu_reports_device_error():
    from anns_fhe_amd.driver import FheContext, FheError
    with pytest.raises(FheError) as ei:
        FheContext(8192, 4, 1, 2)
    assert ei.value.status == -7
