"""The oracle's OP 6 gather against the reference's own automorphism()
(hardware_fpga_model/hls/src/automorphism.cpp:3-12), compiled unmodified from its
source into oracle/_ref by oracle/Makefile (the one reference kernel file that
builds without Vitis ap_int.h).  N is fixed at 16384 there.  Pins the packed-map
convention (two u32 per word, low half = even index) with random permutations and
the closed-form Galois maps the engine uses."""
import numpy as np
import pytest

from oracle import oracle as O

N = 16384


@pytest.fixture(scope="module")
def ref():
    f = O.ref_automorphism()
    if f is None:
        pytest.skip("oracle/_ref not built (the reference sources are absent)")
    return f


def _oracle_gather(a, packed):
    out = np.zeros(N, np.uint64)
    O.lib().orc_automorphism(O.P64(a), O.P64(out), O.P64(packed), N)
    return out


def _ref_gather(ref, a, packed):
    out = np.zeros(N, np.uint64)
    ref(O.P64(a), O.P64(out), O.P64(packed))
    return out


@pytest.mark.parametrize("kind", ["random", "g5", "g5^77", "conj", "identity"])
def test_oracle_gather_equals_reference(ref, kind):
    rng = np.random.default_rng(len(kind))
    a = rng.integers(0, 1 << 63, N, dtype=np.uint64)
    if kind == "random":
        m = rng.permutation(N).astype(np.uint32)
    elif kind == "identity":
        m = np.arange(N, dtype=np.uint32)
    else:
        g = {"g5": 5, "g5^77": pow(5, 77, 2 * N), "conj": 2 * N - 1}[kind]
        m = O.galois_map(N, g)
    packed = O.pack_map(m)
    got = _oracle_gather(a, packed)
    assert np.array_equal(got, _ref_gather(ref, a, packed))
    assert np.array_equal(got, a[m])  # out[i] = a[map[i]]
