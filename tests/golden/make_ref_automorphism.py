"""Writes ref_automorphism_n16384.json: known-answer vectors of the reference's own
automorphism() (hardware_fpga_model/hls/src/automorphism.cpp:3-12, N fixed at 16384),
run from oracle/_ref, where oracle/Makefile compiles it unmodified from the reference
sources.  The reference ships no golden vectors for this path, so these are made here,
once, and committed; tests/test_golden_ref.py checks the oracle's OP 6 gather against
them, also where the reference sources (and so oracle/_ref) are absent.

Each vector: a seeded input polynomial (uniform 63-bit words) and a packed gather map
(two u32 indices per word, low half = even index); stored are the seed, the SHA-256 of
the packed map and of the reference's output words (little-endian u64), and 16 of the
output words in full.

    python tests/golden/make_ref_automorphism.py      (needs oracle/_ref built: make -C oracle ref)
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

N = 16384
OUT = os.path.join(HERE, "ref_automorphism_n16384.json")
# kind -> (seed, Galois element or None for a random permutation or the identity)
KINDS = {"random": (11, None), "identity": (12, None), "g5": (13, 5), "g5^77": (14, pow(5, 77, 2 * N)),
         "conj": (15, 2 * N - 1)}
SAMPLE = list(range(8)) + list(range(N - 8, N))


def inputs(kind):
    """The input words and the packed map of one vector (deterministic in the seed)."""
    seed, g = KINDS[kind]
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 1 << 63, N, dtype=np.uint64)
    if kind == "random":
        m = rng.permutation(N).astype(np.uint32)
    elif kind == "identity":
        m = np.arange(N, dtype=np.uint32)
    else:
        m = O.galois_map(N, g)
    return a, O.pack_map(m)


def sha(x):
    return hashlib.sha256(np.ascontiguousarray(x, dtype="<u8").tobytes()).hexdigest()


def main():
    ref = O.ref_automorphism()
    if ref is None:
        sys.exit("oracle/_ref is not built (make -C oracle ref)")
    vec = {}
    for kind in KINDS:
        a, packed = inputs(kind)
        out = np.zeros(N, np.uint64)
        ref(O.P64(a), O.P64(out), O.P64(packed))
        vec[kind] = {"seed": KINDS[kind][0], "galois": KINDS[kind][1], "map_sha256": sha(packed),
                     "out_sha256": sha(out), "sample_pos": SAMPLE, "sample_words": [int(out[i]) for i in SAMPLE]}
    with open(OUT, "w") as f:
        json.dump({"N": N, "source": "hardware_fpga_model/hls/src/automorphism.cpp:3-12", "vectors": vec}, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
