"""The oracle's OP 6 gather against the committed known-answer vectors of the reference's
automorphism() (hardware_fpga_model/hls/src/automorphism.cpp:3-12, N = 16384), made by
tests/golden/make_ref_automorphism.py from oracle/_ref.  Runs without the reference
sources; a map read with its two packed halves swapped (the plausible convention slip)
must miss every digest."""
import importlib.util
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
_spec = importlib.util.spec_from_file_location("make_ref_automorphism",
                                               os.path.join(HERE, "golden", "make_ref_automorphism.py"))
G = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(G)
with open(os.path.join(HERE, "golden", "ref_automorphism_n16384.json")) as f:
    FIX = json.load(f)


def _oracle_gather(a, packed):
    out = np.zeros(G.N, np.uint64)
    O.lib().orc_automorphism(O.P64(a), O.P64(out), O.P64(packed), G.N)
    return out


@pytest.mark.parametrize("kind", sorted(FIX["vectors"]))
def test_oracle_gather_matches_reference_vectors(kind):
    v = FIX["vectors"][kind]
    a, packed = G.inputs(kind)
    assert G.sha(packed) == v["map_sha256"], "input generation changed"
    out = _oracle_gather(a, packed)
    assert [int(out[i]) for i in v["sample_pos"]] == v["sample_words"]
    assert G.sha(out) == v["out_sha256"]


@pytest.mark.parametrize("kind", ["random", "g5", "conj"])
def test_swapped_map_halves_miss_the_vectors(kind):
    v = FIX["vectors"][kind]
    a, packed = G.inputs(kind)
    swapped = ((packed >> np.uint64(32)) | (packed << np.uint64(32))).astype(np.uint64)
    assert G.sha(_oracle_gather(a, swapped)) != v["out_sha256"]
