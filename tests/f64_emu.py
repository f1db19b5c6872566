"""Exact IEEE-754 binary64 emulation of the engine's FP64 modular arithmetic.

Mirrors anns_fhe_amd/csrc/fhea_f64.hpp operation by operation so that the FP64
kernels' arithmetic can be checked on the CPU:

* a * b, a + b, a - b, a / b on Python floats are correctly rounded binary64
  operations (round to nearest, ties to even), as on the GPU with contraction off;
* fma(a, b, c) is computed exactly with integers/fractions and rounded once;
* rint is round-half-to-even (the default rounding mode's rint).

`trace` collects the largest |value| seen at each checked point so that tests can
assert the magnitude bounds the kernels rely on (|x| < 2^53 keeps integers exact).
"""
from fractions import Fraction

TWO53 = float(2 ** 53)


def fma(a, b, c):
    if a == int(a) and b == int(b) and c == int(c):
        return float(int(a) * int(b) + int(c))  # int -> float rounds to nearest even
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def rint(x):
    return float(round(x))


class Trace:
    def __init__(self):
        self.max_ratio = {}

    def see(self, key, x, q):
        assert abs(x) < TWO53, (key, x)
        assert x == int(x), (key, x)
        r = abs(x) / q
        if r > self.max_ratio.get(key, 0.0):
            self.max_ratio[key] = r


NULL = Trace()


def mulmod_f64(v, w, wq, q, tr=NULL):
    tr.see("mulmod_in", v, q)
    k = rint(v * wq)
    hi = v * w
    lo = fma(v, w, -hi)
    r = fma(-k, q, hi) + lo
    tr.see("mulmod_out", r, q)
    return r


def mulmod_vv_f64(a, b, q, qinv, tr=NULL):
    hi = a * b
    lo = fma(a, b, -hi)
    k = rint(hi * qinv)
    r = fma(-k, q, hi) + lo
    tr.see("mulmod_vv_out", r, q)
    return r


def reduce_f64(x, q, qinv, tr=NULL):
    tr.see("reduce_in", x, q)
    r = fma(-rint(x * qinv), q, x)
    tr.see("reduce_out", r, q)
    return r


def f64_to_canon(x, q):
    y = x + q if x < 0.0 else x
    assert 0.0 <= y < q and y == int(y), (x, q)
    return int(y)


def bit_reverse(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


def ntt_forward_f64(a, tw, q, round_stages=4, tr=NULL):
    """Forward negacyclic NTT as the FP64 kernel computes it: Cooley-Tukey stages
    (W = tw[m + i]) in radix-2^round_stages rounds, signed residues, one
    reduce_f64 per element at the end of every round, canonical output."""
    n = len(a)
    qd, qinv = float(q), 1.0 / float(q)
    x = [float(v) for v in a]
    logn = n.bit_length() - 1
    m, t, st = 1, n // 2, 0
    while m < n:
        for i in range(m):
            w = float(tw[m + i])
            wq = w / qd
            for j in range(2 * i * t, 2 * i * t + t):
                rr = mulmod_f64(x[j + t], w, wq, qd, tr)
                u = x[j]
                x[j], x[j + t] = u + rr, u - rr
                tr.see("butterfly_out", x[j], qd)
                tr.see("butterfly_out", x[j + t], qd)
        st += 1
        if st % round_stages == 0 or st == logn:
            x = [reduce_f64(v, qd, qinv, tr) for v in x]
        m, t = 2 * m, t // 2
    return [f64_to_canon(v, qd) for v in x]


def kernel_rounds(logn, k1, log_tile=12):
    """The forward FP64 kernel's rounds: a list of rounds, each a list of
    (global bit b, window low bit P_global) in stage order.  logn <= log_tile: one
    pass over all bits; otherwise a first pass over the top k1 bits (columns) and a
    second over the low logn - k1 bits (rows).  Round geometry as in the kernel:
    4 stages per round, window [P, P+4) with P = top - 3 (or 0)."""
    passes = [(logn, 0)] if logn <= log_tile else [(k1, logn - k1), (logn - k1, 0)]
    rounds = []
    for logr, shift in passes:
        nr = (logr + 3) // 4
        for rd in range(nr):
            ns = 4 if rd + 1 < nr else logr - 4 * rd
            top = logr - 1 - 4 * rd
            p = top - 3 if top >= 3 else 0
            rounds.append([(top - st + shift, p + shift) for st in range(ns)])
    return rounds


def ntt_forward_f64_kernel(a, tw, q, k1=None, tr=NULL):
    """Forward NTT exactly as the FP64 kernel computes it: the kernel's passes and
    radix-16 rounds (kernel_rounds), a reduction at every round end, and each
    stage's twiddles tw[base | t] (t < 2^(P+3-b), base with those bits clear)
    formed as mulmod(tw[base], tw[t]) with w/q taken as w * fl(1/q)."""
    n = len(a)
    logn = n.bit_length() - 1
    k1 = logn // 2 if k1 is None else k1
    qd, qinv = float(q), 1.0 / float(q)
    pairs = [(float(w), float(w) / qd) for w in tw]
    x = [float(v) for v in a]
    for rnd in kernel_rounds(logn, k1):
        for b, p in rnd:
            tb = p + 3 - b          # bits of the group index taken from the window
            m = n >> (b + 1)
            for j in range(n):
                if j >> b & 1:
                    continue
                i = j >> (b + 1)
                t = i & ((1 << tb) - 1)
                w0, wq0 = pairs[m + i - t]
                if t:
                    c, cq = pairs[t]
                    w = mulmod_f64(w0, c, cq, qd, tr)
                    tr.see("twiddle", w, qd)
                    wq = w * qinv
                else:
                    w, wq = w0, wq0
                rr = mulmod_f64(x[j + (1 << b)], w, wq, qd, tr)
                u = x[j]
                x[j], x[j + (1 << b)] = u + rr, u - rr
                tr.see("butterfly_out", x[j], qd)
                tr.see("butterfly_out", x[j + (1 << b)], qd)
        x = [reduce_f64(v, qd, qinv, tr) for v in x]
    return [f64_to_canon(v, qd) for v in x]


def ntt_inverse_f64_kernel(a, itw, n_inv, q, k1=None, log_tile=12, tr=NULL):
    """Inverse NTT as the FP64 kernel computes it: Gentleman-Sande stages (sum
    reduced, difference times W = itw[m + i]), bits in ascending order; in the
    first round along rows (bits 0..3, window [0, 4)) each stage's twiddles
    itw[base | t] (t < 2^(3-b)) are formed as mulmod(itw[base], itw[t]) with w/q
    taken as w * fl(1/q); then the N^-1 scaling and the final reduction."""
    n = len(a)
    logn = n.bit_length() - 1
    k1 = logn // 2 if k1 is None else k1
    rows_bits = logn if logn <= log_tile else logn - k1
    qd, qinv = float(q), 1.0 / float(q)
    pairs = [(float(w), float(w) / qd) for w in itw]
    x = [float(v) for v in a]
    for b in range(logn):
        m = n >> (b + 1)
        fact_bits = 3 - b if (b < 4 and rows_bits >= 4) else 0
        for j in range(n):
            if j >> b & 1:
                continue
            i = j >> (b + 1)
            t = i & ((1 << fact_bits) - 1)
            w0, wq0 = pairs[m + i - t]
            if t:
                c, cq = pairs[t]
                w = mulmod_f64(w0, c, cq, qd, tr)
                tr.see("twiddle", w, qd)
                wq = w * qinv
            else:
                w, wq = w0, wq0
            u, v = x[j], x[j + (1 << b)]
            tr.see("inv_diff", u - v, qd)
            x[j] = reduce_f64(u + v, qd, qinv, tr)
            x[j + (1 << b)] = mulmod_f64(u - v, w, wq, qd, tr)
    ni = float(n_inv)
    niq = ni / qd
    out = []
    for v in x:
        v = mulmod_f64(v, ni, niq, qd, tr)
        out.append(f64_to_canon(reduce_f64(v, qd, qinv, tr), qd))
    return out


def ntt_inverse_f64(a, itw, n_inv, q, tr=NULL):
    """Inverse as the FP64 kernel computes it: Gentleman-Sande stages with the
    sum reduced every stage and the difference multiplied by W = itw[m + i],
    then the N^-1 scaling and a final reduction."""
    n = len(a)
    qd, qinv = float(q), 1.0 / float(q)
    x = [float(v) for v in a]
    m, t = n // 2, 1
    while m >= 1:
        for i in range(m):
            w = float(itw[m + i])
            wq = w / qd
            for j in range(2 * i * t, 2 * i * t + t):
                u, v = x[j], x[j + t]
                x[j] = reduce_f64(u + v, qd, qinv, tr)
                x[j + t] = mulmod_f64(u - v, w, wq, qd, tr)
        m, t = m // 2, 2 * t
    ni = float(n_inv)
    niq = ni / qd
    out = []
    for v in x:
        v = mulmod_f64(v, ni, niq, qd, tr)
        out.append(f64_to_canon(reduce_f64(v, qd, qinv, tr), qd))
    return out


def fast_bases_conv_f64(x, in_primes, qhat_inv, qhat_mod_p, out_primes, tr=NULL):
    """One coefficient of the FP64 basis conversion: y_i = [x_i qhat_i^-1]_{q_i}
    made canonical, then sum_i y_i (qhat_i mod p_j) with a reduction every four
    terms (the kernel's k_fbc_f64)."""
    ys = []
    for i, qi in enumerate(in_primes):
        w = float(qhat_inv[i])
        t = mulmod_f64(float(x[i]), w, w / float(qi), float(qi), tr)
        ys.append(t + float(qi) if t < 0.0 else t)
    out = []
    for j, pj in enumerate(out_primes):
        pd, pinv = float(pj), 1.0 / float(pj)
        acc = 0.0
        for i in range(len(in_primes)):
            h = float(qhat_mod_p[i][j])
            acc += mulmod_f64(ys[i], h, h / pd, pd, tr)
            tr.see("fbc_acc", acc, pd)
            if (i & 3) == 3:
                acc = reduce_f64(acc, pd, pinv, tr)
        out.append(f64_to_canon(reduce_f64(acc, pd, pinv, tr), pd))
    return out
