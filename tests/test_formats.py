"""The reference's on-disk formats (anns_fhe_amd/formats.py): tables, keys, maps.

Round trips through files, agreement with the oracle's tables (Barrett records,
ModUp / ModDown / rescale constants, Galois maps), and a parameter block built
from a directory with its own roots of unity reproducing that directory's
twiddles (the path by which vectors dumped by another library are loaded)."""
import numpy as np
import pytest

from anns_fhe_amd import formats as F


def _client(**kw):
    from anns_fhe_amd.client import CkksClient
    return CkksClient(1024, 4, 2, 2, **kw)


def test_barrett_records_match_oracle(orc):
    for q in orc.gen_primes(2 * 1024, 50, 3) + orc.gen_primes(2 * 1024, 60, 2) + orc.gen_primes(2 * 1024, 31, 2):
        m_lo, m_hi, k = orc.barrett(q)
        assert F.barrett_record(q) == ((int(m_hi) << 64) | int(m_lo), int(k))


def test_context_tables_round_trip(tmp_path):
    cl = _client()
    F.write_context_tables(str(tmp_path), cl.N, cl.q, cl.p, [int(x) for x in cl.table(0)])
    t = F.read_context_tables(str(tmp_path))
    F.check_context_tables(t)
    assert t["N"] == cl.N and t["q"] == cl.q and t["p"] == cl.p
    assert np.array_equal(t["tw"], cl.table(2))
    assert np.array_equal(t["itw"], cl.table(3))
    assert t["n_inv"] == [int(x) for x in cl.table(1)]
    assert t["psi"] == [int(x) for x in cl.table(0)]
    # 20-byte unpadded records
    assert (tmp_path / "barrett_constants.bin").stat().st_size == 20 * (cl.L + cl.K)
    # a corrupted twiddle file is detected
    tw = np.fromfile(tmp_path / "twiddles.bin", np.uint64)
    tw[1] += 1
    tw.tofile(tmp_path / "twiddles.bin")
    with pytest.raises(ValueError):
        F.check_context_tables(F.read_context_tables(str(tmp_path)))


def test_foreign_roots_reproduce_foreign_twiddles(tmp_path):
    """A directory made with other primitive roots (psi^3) loads into a parameter block
    whose tables equal the files, and CKKS still round-trips with those tables."""
    base = _client()
    psi3 = [pow(int(w), 3, q) for w, q in zip(base.table(0), base.primes)]
    F.write_context_tables(str(tmp_path), base.N, base.q, base.p, psi3)
    t = F.read_context_tables(str(tmp_path))
    assert t["psi"] == psi3
    cl = _client(q_primes=t["q"], p_primes=t["p"], psi=t["psi"])
    assert np.array_equal(cl.table(2), t["tw"])
    assert not np.array_equal(cl.table(2), base.table(2))
    z = np.linspace(-1, 1, 512)
    s = cl.keygen_secret(1)
    pk = cl.keygen_public(s, 2)
    ct = cl.encrypt_values(pk, z, 4, 2.0 ** 40, 3)
    assert np.max(np.abs(cl.decrypt_values(s, ct, 4, 2.0 ** 40, 512) - z)) < 1e-6
    # a root that is not primitive is refused
    from anns_fhe_amd._lib import FheError
    bad = list(psi3)
    bad[0] = pow(bad[0], 2, base.primes[0])
    with pytest.raises(FheError):
        _client(q_primes=t["q"], p_primes=t["p"], psi=bad)


@pytest.mark.parametrize("sizeQl", [4, 3, 1])
def test_ks_and_rescale_tables_match_oracle(orc, tmp_path, sizeQl):
    cl = _client()
    o = orc.OracleCtx(cl.N, cl.q, cl.p, cl.dnum)
    parts, pinv, phatinv, phatmodq = F.ks_tables(cl.q, cl.p, sizeQl, cl.alpha)
    for j, (inv, hat) in enumerate(parts):
        oinv, ohat = o.modup_tables(sizeQl, j)
        assert [int(x) for x in oinv] == inv
        assert np.array_equal(np.array(hat, dtype=object).astype(np.uint64), ohat)
    opinv, ophinv, ophat = o.moddown_tables(sizeQl)
    assert [int(x) for x in opinv] == pinv
    assert [int(x) for x in ophinv] == phatinv
    assert np.array_equal(np.array(phatmodq, dtype=object).astype(np.uint64), ophat)
    if sizeQl > 1:
        a, b = F.rescale_tables(cl.q, sizeQl)
        oa, ob = o.rescale_tables(sizeQl)
        assert [int(x) for x in oa] == a and [int(x) for x in ob] == b
        F.write_ks_tables(str(tmp_path), cl.q, cl.p, sizeQl, cl.alpha)
        assert np.array_equal(np.fromfile(tmp_path / "qlInvModq.bin", np.uint64), ob)
        assert np.array_equal(np.fromfile(tmp_path / "tv_keyswitch_PartQlHatModp_0.bin", np.uint64),
                              o.modup_tables(sizeQl, 0)[1].reshape(-1))


def test_keys_and_maps_round_trip(orc, tmp_path):
    cl = _client()
    s = cl.keygen_secret(4)
    rel = cl.relin_key(s, cl.L, 5)
    F.write_relin_key(str(tmp_path), rel, cl.dnum, cl.L, cl.K, cl.N)
    assert np.array_equal(F.read_relin_key(str(tmp_path), cl.dnum), rel)
    keys = {}
    for st in (1, -2, 8):
        g = cl.galois_for_step(st)
        keys[st] = (g, cl.rotation_key(s, g, 3, 10 + st), 3, F.galois_map(cl.N, g))
    F.write_rotation_keys(str(tmp_path), keys, cl.dnum, cl.K, cl.N)
    back = F.read_rotation_keys(str(tmp_path), cl.dnum, cl.K, cl.N)
    assert sorted(back) == sorted(keys)
    for st, (g, key, kl, amap) in keys.items():
        g2, key2, kl2, amap2 = back[st]
        assert g2 == g and kl2 == kl and np.array_equal(key2, key) and np.array_equal(amap2, amap)
        # the closed-form map equals the oracle's, and its inverse is recognised
        assert np.array_equal(amap, orc.galois_map(cl.N, g))
        assert F.map_convention(amap, cl.N, g) == "same"
        assert F.map_convention(F.galois_map(cl.N, pow(g, -1, 2 * cl.N)), cl.N, g) == "inverse"
