"""Reference-format directories (the files integration_tools/
generate_interactive_test_vectors.cpp:246-326 writes) loaded into the engine.

* The C++ FHEHostDriver (anns_fhe_amd/dir_sequence, main_sim.cpp:33-50 over the
  C-ABI) reads primes, roots, Barrett records, N^-1, the relinearisation key and
  the rotation-key directory, then runs eval_sub_plain -> eval_mult_relin_rescale
  -> eval_rotate(+1) and one rotation by -1 whose map file is written in the
  opposite Galois convention; every word equals the oracle built with the
  directory's own roots.
* formats.context_from_dir gives an engine context whose transform is the
  directory's."""
import os
import subprocess
import sys

import numpy as np
import pytest

from anns_fhe_amd import formats as F

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print("[gpu-formats]", *a, flush=True)
    sys.stdout.flush()


def _dir_with_foreign_roots(tmp_path, N, L, K, dnum):
    from anns_fhe_amd.client import CkksClient
    base = CkksClient(N, L, K, dnum)
    psi3 = [pow(int(w), 3, q) for w, q in zip(base.table(0), base.primes)]
    cl = CkksClient(N, L, K, dnum, q_primes=base.q, p_primes=base.p, psi=psi3)
    d = str(tmp_path / "tv")
    F.write_context_tables(d, N, cl.q, cl.p, psi3)
    F.write_ks_tables(d, cl.q, cl.p, L, cl.alpha)
    return cl, psi3, d


def test_cpp_driver_runs_reference_directory(orc, tmp_path):
    N, L, K, dnum = 8192, 4, 1, 2
    W, W1 = L * N, (L - 1) * N
    cl, psi3, d = _dir_with_foreign_roots(tmp_path, N, L, K, dnum)
    s = cl.keygen_secret(31)
    F.write_relin_key(d, cl.relin_key(s, L, 32), dnum, L, K, N)
    g1, gm1 = cl.galois_for_step(1), cl.galois_for_step(-1)
    gm1_inv = pow(gm1, -1, 2 * N)
    # rotation keys at the level below the top (the reference stores them level-trimmed);
    # step -1's map file holds sigma of gm1^-1 (the other convention), its key matches that map
    keys = {1: (g1, cl.rotation_key(s, g1, L - 1, 33), L - 1, F.galois_map(N, g1)),
            -1: (gm1, cl.rotation_key(s, gm1_inv, L - 1, 34), L - 1, F.galois_map(N, gm1_inv))}
    F.write_rotation_keys(d, keys, dnum, K, N)
    rng = np.random.default_rng(77)
    inp = [np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in cl.q]).reshape(-1) for _ in range(3)]
    for name, arr in zip(("ct_query_c0", "ct_query_c1", "pt_query"), inp):
        F.write_poly(os.path.join(d, name + ".bin"), arr)
    out = tmp_path / "out"
    out.mkdir()
    say("dir_sequence")
    r = subprocess.run([os.path.join(ROOT, "anns_fhe_amd", "dir_sequence"), d, str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    meta = {}
    for line in (out / "rotation_keys.txt").read_text().splitlines():
        f = line.split()
        if f[0] == "relin":
            relin_off, relin_limbs = int(f[1]), int(f[2])
        else:
            meta[int(f[0])] = dict(auto_idx=int(f[1]), galois=int(f[2]), map_off=int(f[3]), evk=int(f[4]),
                                   kl=int(f[5]))
    assert relin_off == 0 and relin_limbs == L
    assert meta[1]["galois"] == g1 and meta[-1]["galois"] == gm1_inv and meta[-1]["auto_idx"] == gm1
    assert sorted(m["map_off"] for m in meta.values()) == [0x580000, 0x580000 + N]
    assert all(m["kl"] == L - 1 and m["evk"] % 512 == 0 for m in meta.values())
    ksz_top, ksz_rot = 2 * dnum * (L + K) * N, 2 * dnum * (L - 1 + K) * N
    km = np.zeros(max(m["evk"] for m in meta.values()) + ksz_rot, np.uint64)
    km[:ksz_top] = F.read_relin_key(d, dnum)
    back = F.read_rotation_keys(d, dnum, K, N)
    for st, m in meta.items():
        km[m["evk"]:m["evk"] + ksz_rot] = back[st][1]
    o = orc.OracleCtx(N, cl.q, cl.p, dnum, psi=psi3)
    pm = np.zeros(12 * W, np.uint64)
    for k in range(3):
        pm[k * W:(k + 1) * W] = inp[k]
    rd = lambda n: np.fromfile(out / f"{n}.bin", np.uint64)  # noqa: E731
    assert o.eval_sub_plain(pm, 0, W, 2 * W, 3 * W, 4 * W, L) == 0
    assert np.array_equal(rd("sub_c0"), pm[3 * W:4 * W]) and np.array_equal(rd("sub_c1"), pm[4 * W:5 * W])
    assert o.eval_mult_relin_rescale(pm, km, 0, L, 3 * W, 4 * W, 3 * W, 4 * W, 5 * W, 6 * W, L) == 0
    assert np.array_equal(rd("mult_c0"), pm[5 * W:5 * W + W1])
    assert np.array_equal(rd("mult_c1"), pm[6 * W:6 * W + W1])
    assert o.eval_rotate(pm, km, [(1, g1, meta[1]["evk"], L - 1)], 5 * W, 6 * W, 7 * W, 8 * W, 1, L - 1) == 0
    assert np.array_equal(rd("rot_c0"), pm[7 * W:7 * W + W1])
    assert np.array_equal(rd("rot_c1"), pm[8 * W:8 * W + W1])
    assert o.eval_automorphism_ks(pm, km, gm1_inv, meta[-1]["evk"], L - 1, 5 * W, 6 * W, 9 * W, 10 * W, L - 1) == 0
    assert np.array_equal(rd("rotm1_c0"), pm[9 * W:9 * W + W1])
    assert np.array_equal(rd("rotm1_c1"), pm[10 * W:10 * W + W1])
    # a directory whose twiddles were tampered with is refused before any op runs
    tw = np.fromfile(os.path.join(d, "twiddles.bin"), np.uint64)
    tw[3] ^= np.uint64(1)
    tw.tofile(os.path.join(d, "twiddles.bin"))
    r = subprocess.run([os.path.join(ROOT, "anns_fhe_amd", "dir_sequence"), d, str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 1 and "twiddles" in r.stderr


def test_context_from_dir_runs_the_directorys_transform(orc, tmp_path):
    N, L, K, dnum = 4096, 3, 1, 3
    cl, psi3, d = _dir_with_foreign_roots(tmp_path, N, L, K, dnum)
    ctx = F.context_from_dir(d, dnum, poly_words=8 * L * N, key_words=64)
    try:
        o = orc.OracleCtx(N, cl.q, cl.p, dnum, psi=psi3)
        rng = np.random.default_rng(5)
        a = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in cl.q])
        W = L * N
        ctx.write_poly(0, a)
        say("context_from_dir ntt")
        assert ctx.op(1, 0, 0, W, 0, L) == 0
        assert np.array_equal(ctx.read_poly(W, W).reshape(L, N), o.ntt_poly(a))
        assert ctx.op(2, W, 0, 2 * W, 0, L) == 0
        assert np.array_equal(ctx.read_poly(2 * W, W).reshape(L, N), a)
    finally:
        ctx.close()
