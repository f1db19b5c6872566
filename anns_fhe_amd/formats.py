"""The reference's on-disk formats, read and written (all little-endian).

Files as integration_tools/generate_interactive_test_vectors.cpp:46-84,246-326 and
integration_tools/test_vector_generator.cpp:300-325,440-494 write them, and as
cpp_host_model/src/fhe_host_driver.cpp:18-172 reads them:

  rns_primes_Q.bin, rns_primes_P.bin   uint64[L], uint64[K]
  barrett_constants.bin                 (L+K) records {u64 m_lo, u64 m_hi, u32 k}, 20 B, unpadded;
                                        k = 2*bitlen(q) + 3, m = floor(2^k / q)
  n_inv.bin                             uint64[L+K]
  twiddles.bin, inv_twiddles.bin        uint64[(L+K)*N], tw[i*N + j] = psi_i^br(j)
  evalkey_mult_{a,b}_{d}.bin            one key digit, all Q then P limbs (limb-major)
  rotkey_step{s}_{a,b}_{d}.bin          the same at a lower level (sizeQl Q limbs, then P)
  auto_map_step{s}.bin                  uint32[N]
  rotation_keys_meta.txt                lines "step=%d auto_idx=%u"
  tv_keyswitch_PartQlHatInvModq_{j}.bin, tv_keyswitch_PartQlHatModp_{j}.bin (row-major),
  tv_keyswitch_PInvModq.bin, tv_keyswitch_PHatInvModp.bin, tv_keyswitch_PHatModq.bin (row-major by P),
  QlQlInvModqlDivqlModq.bin, qlInvModq.bin
  *_c0.bin / *_c1.bin                   ciphertext polynomials, uint64[limbs*N]

A directory written by OpenFHE elsewhere can be loaded here: its primes and its
roots of unity (psi_i = twiddles[i*N + N/2]) go into the context's parameter
block, so the engine's tables and transforms are the ones the files assume.
"""
import os
import re

import numpy as np

BARRETT_DTYPE = np.dtype([("m_lo", "<u8"), ("m_hi", "<u8"), ("k", "<u4")])  # 20-byte records


def _w(path, arr, dtype="<u8"):
    np.ascontiguousarray(arr, dtype=dtype).tofile(path)


def _r(path, dtype="<u8"):
    return np.fromfile(path, dtype=dtype)


def barrett_record(q):
    """(m, k) as generate_interactive_test_vectors.cpp:248-253 computes them."""
    k = 2 * int(q).bit_length() + 3
    return (1 << k) // int(q), k


def bit_reverse(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


# ---- polynomials / ciphertexts
def write_poly(path, words):
    _w(path, words)


def read_poly(path, limbs=None, N=None):
    a = _r(path)
    if limbs is not None and N is not None:
        assert a.size == limbs * N, (path, a.size, limbs, N)
    return a


def write_ciphertext(prefix, c0, c1):
    _w(prefix + "_c0.bin", c0)
    _w(prefix + "_c1.bin", c1)


def read_ciphertext(prefix):
    return _r(prefix + "_c0.bin"), _r(prefix + "_c1.bin")


# ---- context tables
def write_context_tables(out_dir, N, q_primes, p_primes, psi):
    """The table files of one parameter set (psi: root per prime, Q then P)."""
    os.makedirs(out_dir, exist_ok=True)
    primes = [int(x) for x in q_primes] + [int(x) for x in p_primes]
    _w(os.path.join(out_dir, "rns_primes_Q.bin"), [int(x) for x in q_primes])
    _w(os.path.join(out_dir, "rns_primes_P.bin"), [int(x) for x in p_primes])
    rec = np.zeros(len(primes), BARRETT_DTYPE)
    for i, q in enumerate(primes):
        m, k = barrett_record(q)
        rec[i] = (m & (2 ** 64 - 1), m >> 64, k)
    rec.tofile(os.path.join(out_dir, "barrett_constants.bin"))
    _w(os.path.join(out_dir, "n_inv.bin"), [pow(N, -1, q) for q in primes])
    bits = N.bit_length() - 1
    br = [bit_reverse(j, bits) for j in range(N)]
    tw = np.zeros((len(primes), N), np.uint64)
    itw = np.zeros((len(primes), N), np.uint64)
    for i, q in enumerate(primes):
        w, wi = int(psi[i]), pow(int(psi[i]), -1, q)
        pw = [1] * N
        pwi = [1] * N
        for e in range(1, N):
            pw[e] = pw[e - 1] * w % q
            pwi[e] = pwi[e - 1] * wi % q
        tw[i] = [pw[b] for b in br]
        itw[i] = [pwi[b] for b in br]
    _w(os.path.join(out_dir, "twiddles.bin"), tw)
    _w(os.path.join(out_dir, "inv_twiddles.bin"), itw)


def read_context_tables(in_dir):
    """Tables of a directory: primes, Barrett records, N^-1, twiddles and the roots."""
    q = [int(x) for x in _r(os.path.join(in_dir, "rns_primes_Q.bin"))]
    p = [int(x) for x in _r(os.path.join(in_dir, "rns_primes_P.bin"))]
    T = len(q) + len(p)
    tw = _r(os.path.join(in_dir, "twiddles.bin"))
    N = tw.size // T
    tw = tw.reshape(T, N)
    itw = _r(os.path.join(in_dir, "inv_twiddles.bin")).reshape(T, N)
    rec = np.fromfile(os.path.join(in_dir, "barrett_constants.bin"), BARRETT_DTYPE)
    barrett = [((int(r["m_hi"]) << 64) | int(r["m_lo"]), int(r["k"])) for r in rec]
    n_inv = [int(x) for x in _r(os.path.join(in_dir, "n_inv.bin"))]
    psi = [int(tw[i, N // 2]) for i in range(T)]  # tw[br(1)] = psi^1
    return {"N": N, "q": q, "p": p, "tw": tw, "itw": itw, "barrett": barrett, "n_inv": n_inv, "psi": psi}


def check_context_tables(t):
    """Consistency of a loaded directory with the definitions (raises ValueError)."""
    N = t["N"]
    for i, q in enumerate(t["q"] + t["p"]):
        if (q - 1) % (2 * N):
            raise ValueError(f"prime {q} is not 1 mod 2N")
        if pow(t["psi"][i], N, q) != q - 1:
            raise ValueError(f"twiddles of prime {q} do not come from a primitive 2N-th root")
        if t["barrett"][i] != barrett_record(q):
            raise ValueError(f"Barrett record of prime {q} differs")
        if t["n_inv"][i] * N % q != 1:
            raise ValueError(f"n_inv of prime {q} is wrong")
        if int(t["tw"][i, 1]) != pow(t["psi"][i], N // 2, q):
            raise ValueError(f"twiddle order of prime {q} is not bit-reversed")


def context_from_dir(in_dir, dnum, poly_words, key_words, device=0):
    """An engine context whose primes and roots are the directory's."""
    from .driver import FheContext
    t = read_context_tables(in_dir)
    check_context_tables(t)
    ctx = FheContext(t["N"], len(t["q"]), len(t["p"]), dnum, q_primes=t["q"], p_primes=t["p"],
                     psi=t["psi"], poly_words=poly_words, key_words=key_words, device=device)
    if not (np.array_equal(ctx.table(2), t["tw"]) and np.array_equal(ctx.table(3), t["itw"])):
        raise ValueError("engine twiddles differ from the directory's")
    return ctx


# ---- key-switching and rescale tables (OpenFHE definitions, big integers)
def _prod(xs):
    r = 1
    for x in xs:
        r *= int(x)
    return r


def ks_tables(q, p, sizeQl, alpha):
    """Per digit j: PartQlHatInvModq_j, PartQlHatModp_j (rows: digit limbs, columns: the
    complement basis = other Q limbs ascending then P); and PInvModq, PHatInvModp, PHatModq."""
    q = [int(x) for x in q[:sizeQl]]
    p = [int(x) for x in p]
    nparts = -(-sizeQl // alpha)
    parts = []
    for j in range(nparts):
        idx = list(range(j * alpha, min((j + 1) * alpha, sizeQl)))
        Qj = _prod(q[i] for i in idx)
        compl = [q[i] for i in range(sizeQl) if i not in idx] + p
        inv = [pow(Qj // q[i] % q[i], -1, q[i]) for i in idx]
        hat = [[(Qj // q[i]) % m for m in compl] for i in idx]
        parts.append((inv, hat))
    P = _prod(p)
    pinv = [pow(P % qi, -1, qi) for qi in q]
    phatinv = [pow(P // pk % pk, -1, pk) for pk in p]
    phatmodq = [[(P // pk) % qi for qi in q] for pk in p]
    return parts, pinv, phatinv, phatmodq


def rescale_tables(q, sizeQl):
    """(QlQlInvModqlDivqlModq, qlInvModq) for dropping q_{sizeQl-1}: with Q' = Q_l / q_l and
    u = Q'^-1 mod q_l, the first is floor(Q' u / q_l) mod q_i, the second q_l^-1 mod q_i."""
    q = [int(x) for x in q[:sizeQl]]
    ql = q[-1]
    Qp = _prod(q[:-1])
    u = pow(Qp % ql, -1, ql)
    a = [(Qp * u // ql) % qi for qi in q[:-1]]
    b = [pow(ql % qi, -1, qi) for qi in q[:-1]]
    return a, b


def write_ks_tables(out_dir, q, p, sizeQl, alpha):
    os.makedirs(out_dir, exist_ok=True)
    parts, pinv, phatinv, phatmodq = ks_tables(q, p, sizeQl, alpha)
    for j, (inv, hat) in enumerate(parts):
        _w(os.path.join(out_dir, f"tv_keyswitch_PartQlHatInvModq_{j}.bin"), inv)
        _w(os.path.join(out_dir, f"tv_keyswitch_PartQlHatModp_{j}.bin"), np.array(hat, dtype=object).reshape(-1).astype(np.uint64))
    _w(os.path.join(out_dir, "tv_keyswitch_PInvModq.bin"), pinv)
    _w(os.path.join(out_dir, "tv_keyswitch_PHatInvModp.bin"), phatinv)
    _w(os.path.join(out_dir, "tv_keyswitch_PHatModq.bin"), np.array(phatmodq, dtype=object).reshape(-1).astype(np.uint64))
    a, b = rescale_tables(q, sizeQl)
    _w(os.path.join(out_dir, "QlQlInvModqlDivqlModq.bin"), a)
    _w(os.path.join(out_dir, "qlInvModq.bin"), b)


# ---- keys
def write_switch_key(out_dir, stem, key, dnum, key_limbs, K, N):
    """Splits an arena-layout key ([a_0..a_{dnum-1} | b_0..]) into {stem}_{a,b}_{d}.bin files."""
    R = (key_limbs + K) * N
    key = np.asarray(key, np.uint64).reshape(2, dnum, R)
    for d in range(dnum):
        _w(os.path.join(out_dir, f"{stem}_a_{d}.bin"), key[0, d])
        _w(os.path.join(out_dir, f"{stem}_b_{d}.bin"), key[1, d])


def read_switch_key(in_dir, stem, dnum):
    """{stem}_{a,b}_{d}.bin -> one arena-layout key (a digits, then b digits)."""
    a = [_r(os.path.join(in_dir, f"{stem}_a_{d}.bin")) for d in range(dnum)]
    b = [_r(os.path.join(in_dir, f"{stem}_b_{d}.bin")) for d in range(dnum)]
    return np.concatenate(a + b)


def write_relin_key(out_dir, key, dnum, L, K, N):
    write_switch_key(out_dir, "evalkey_mult", key, dnum, L, K, N)


def read_relin_key(in_dir, dnum):
    return read_switch_key(in_dir, "evalkey_mult", dnum)


def write_rotation_keys(out_dir, keys, dnum, K, N):
    """keys: {step: (galois, key, key_limbs, map_u32 or None)}; writes the meta file too."""
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "rotation_keys_meta.txt"), "w") as f:
        for step in sorted(keys):
            g, key, kl, amap = keys[step]
            write_switch_key(out_dir, f"rotkey_step{step}", key, dnum, kl, K, N)
            if amap is not None:
                _w(os.path.join(out_dir, f"auto_map_step{step}.bin"), amap, "<u4")
            f.write(f"step={step} auto_idx={g}\n")


def read_rotation_keys(in_dir, dnum, K, N):
    """-> {step: (galois, key, key_limbs, map_u32 or None)} from the meta file."""
    out = {}
    with open(os.path.join(in_dir, "rotation_keys_meta.txt")) as f:
        for line in f:
            m = re.match(r"\s*step=(-?\d+)\s+auto_idx=(\d+)", line)
            if not m:
                continue
            step, g = int(m.group(1)), int(m.group(2))
            key = read_switch_key(in_dir, f"rotkey_step{step}", dnum)
            kl = key.size // (2 * dnum * N) - K
            mp = os.path.join(in_dir, f"auto_map_step{step}.bin")
            amap = np.fromfile(mp, "<u4") if os.path.exists(mp) else None
            out[step] = (g, key, kl, amap)
    return out


def galois_map(N, g):
    """Closed-form evaluation-domain map of sigma_g in bit-reversed order:
    map[i] = br(((g (2 br(i) + 1)) mod 2N - 1) / 2) (the engine's convention)."""
    bits = N.bit_length() - 1
    return np.array([bit_reverse(((g * (2 * bit_reverse(i, bits) + 1)) % (2 * N) - 1) // 2, bits)
                     for i in range(N)], np.uint32)


def map_convention(amap, N, g):
    """How a dumped map relates to the engine's: 'same', 'inverse' (sigma_{g^-1}) or None."""
    amap = np.asarray(amap, np.uint32)
    if np.array_equal(amap, galois_map(N, g)):
        return "same"
    if np.array_equal(amap, galois_map(N, pow(g, -1, 2 * N))):
        return "inverse"
    return None
