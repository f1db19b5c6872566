// fhea_conv_cols.hip -- fast basis conversion fused between the column passes of
// the transforms on either side of it (FP64 path, two-pass N).
//
// The key switch's ModUp is INTT -> fast_bases_conv -> NTT on every digit
// (key_switch.cpp:229-329), its ModDown INTT(P) -> conversion P -> Q -> NTT
// (key_switch.cpp:385-427).  With the two-pass transform of a limb viewed as
// [2^LOGR rows][2^logS columns], the INTT's second pass runs down the columns
// (stages on row bits 0 .. LOGR-1) and the NTT's first pass also runs down the
// columns (row bits LOGR-1 .. 0), and the conversion is position-wise.  So one
// workgroup that owns a tile of 2^LOGC adjacent columns (all their rows) can run
// the column pass of every input limb of a digit, keep the converted inputs
// [x_i qhat_i^-1]_{q_i} of its positions in registers, and produce each output
// limb's conversion and forward column pass from them: the converted limbs never
// go through HBM in coefficient form, and each input word is read once instead of
// once per output.  The rows passes before (INTT) and after (NTT) stay as they are.
//
// Layout.  TPB = 2^(LOGC+LOGR-W) threads; a thread holds E = 2^W elements of the
// tile: its column c = tid mod 2^LOGC and the rows whose index differs only in the
// W-bit window of the current round ("radix-E rounds", exchanges through LDS).
// The inverse's last window and the forward's first window are both the top W row
// bits, so the converted inputs sit in the registers where the forward column pass
// starts.  Every global access is a run of 2^LOGC consecutive words per row.
//
// Arithmetic: the FP64 exact products of fhea_f64.hpp.  Input limb i is scaled by
// N^-1 qhat_i^-1 mod q_i (the INTT's N^-1 and the conversion's first factor in one
// product), centred, and moved into [0, q_i): the conversion's input integer.  The
// conversion is k_fbc_f64's sum of the terms mulmod_f64(y_i, qhat_i mod p_j), each
// |term| <= q/2 + 1, so at most CC_MAX_IN = 8 of them stay below 4q + 8 < 2^52 and
// add exactly; the forward column pass starts from the centred sum.
// Every step is exact mod p_j, so the words equal INTT -> fast_bases_conv -> NTT.
#include "fhea_f64.hpp"
#include "fhea_kernels.hpp"
#include "fhea_ntt_tile.hpp"

namespace fhea {


// Values every lane holds equally (read from LDS with a uniform address) moved to
// scalar registers: the FP64 instructions take them as their SGPR operand and
// addresses built from them stay scalar.
__device__ __forceinline__ u32 uniform_u32(u32 v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uniform_f64(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const u32 lo = __builtin_amdgcn_readfirstlane((u32)b), hi = __builtin_amdgcn_readfirstlane((u32)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Round RD of a LOGR-stage column transform in W-bit windows.  Forward stages run
// on row bits LOGR-1 .. 0, inverse stages on 0 .. LOGR-1; P is the window's low
// bit (a short last round keeps a full-width window, shifted).
template <bool INV, int LOGR, int W, int RD>
struct RoundW {
    static constexpr int NR = (LOGR + W - 1) / W;
    static constexpr int TOP = LOGR - 1 - W * RD;
    static constexpr int NS = INV ? ((W * RD + W <= LOGR) ? W : LOGR - W * RD) : ((TOP + 1 >= W) ? W : TOP + 1);
    static constexpr int P = INV ? ((W * RD + W <= LOGR) ? W * RD : LOGR - W) : ((TOP >= W - 1) ? TOP - (W - 1) : 0);
    static constexpr int BIT0 = INV ? W * RD : TOP;
};

// tile element of the thread's element 0 in a round whose window starts at tile bit PP
// (PP = LOGC + P: the column bits sit below the row bits)
template <int W>
__device__ __forceinline__ u32 elem_base(u32 tid, u32 PP) {
    return (tid & ((1u << PP) - 1)) | ((tid >> PP) << (PP + W));
}

// Butterfly stages of one round on the thread's E elements.  j0 = global word
// index (within the limb) of element 0; the stage on row bit b uses twiddle
// tw[(N >> k) | (j >> k)], k = b + logS + 1 (ntt.cpp:4-80): the thread's 2^(W-1-wb)
// distinct (w, w/q) pairs of a stage are loaded (at most 2 at W = 2).
template <bool INV, int LOGR, int W, int RD>
__device__ __forceinline__ void round_w(double (&x)[1 << W], const double *__restrict__ tw, double q, double qinv,
                                        u32 logN, u32 logS, u32 j0) {
    using R = RoundW<INV, LOGR, W, RD>;
    constexpr int E = 1 << W;
    const double2 *tw2 = reinterpret_cast<const double2 *>(tw);
#pragma unroll
    for (int st = 0; st < R::NS; ++st) {
        const int b = INV ? R::BIT0 + st : R::BIT0 - st;
        const int wb = b - R::P;
        const u32 k = (u32)b + logS + 1;
        const u32 wbase = ((1u << logN) >> k) | (j0 >> k);
        constexpr int MAXP = E / 2;
        double w[MAXP], wq[MAXP];
        const int np = E >> (wb + 1);
        const double2 *tb = tw2 + wbase;  // wbase has its low bits clear: + t
#pragma unroll
        for (int t = 0; t < MAXP; ++t) {
            if (t < np) {
                const double2 pr = tb[t];
                w[t] = pr.x;
                wq[t] = pr.y;
            }
        }
#pragma unroll
        for (int u = 0; u < E; ++u) {
            if (u & (1 << wb)) continue;
            const int u2 = u | (1 << wb);
            const double ww = w[u >> (wb + 1)], wwq = wq[u >> (wb + 1)];
            if (!INV) {
                const double rr = mulmod_f64(x[u2], ww, wwq, q);
                const double a0 = x[u];
                x[u] = a0 + rr;
                x[u2] = a0 - rr;
            } else {
                // the sum is centred at the round's last stage only: from |x| <= q at
                // the round's start (canonical inputs, or centred values and products
                // <= q/2 + 1) it stays below 2^W q <= 8q, and a difference below 8q is
                // multiplied back to <= q/2 + 1 (canonical twiddles)
                const double a0 = x[u], b0 = x[u2];
                x[u] = (st == R::NS - 1) ? reduce_f64(a0 + b0, q, qinv) : a0 + b0;
                x[u2] = mulmod_f64(a0 - b0, ww, wwq, q);
            }
        }
    }
    // forward values are centred after every second round and after the last: from
    // |x| <= q/2 + 2 a stage grows the bound B to 1.25 B + q/2 (canonical twiddles,
    // mulmod_f64's bound), so four stages stay below 4.1q and six below 7.53q < 8q
    if (!INV && ((RD & 1) || RD == R::NR - 1)) {
#pragma unroll
        for (int u = 0; u < E; ++u) x[u] = reduce_f64(x[u], q, qinv);
    }
}

// Exchange between round RDP and round RDN (same direction) through one of the two
// LDS buffers; the buffers alternate, so one barrier per exchange suffices (a
// buffer is written again only after the barrier of the exchange that followed
// its last reads).
// Tile element e sits at cc_pad(e) = e + 4 floor(e / 16): element bits 4 and 5 move
// the address by 4 and 8 words mod 16, so in the exchanges whose lanes differ in
// element bits {0, 1, 4, 5} (window at row bit 0 with 4 columns) as in those whose
// lanes differ in bits {0, 1, 2, 3}, every 16 lanes of a 64-bit access fall on
// distinct LDS banks.  cc_pad is additive over disjoint bits, so element u sits at
// cc_pad(eb) + cc_pad(u << P): a per-thread base and a constant, which becomes the
// LDS instruction's immediate offset.
__device__ __forceinline__ constexpr u32 cc_pad(u32 e) { return e + ((e >> 4) << 2); }

template <bool INV, int LOGR, int W, int LOGC, int RDP, int RDN>
__device__ __forceinline__ void exch(double (&x)[1 << W], double *buf, u32 tid) {
    constexpr int E = 1 << W;
    constexpr u32 PP = LOGC + RoundW<INV, LOGR, W, RDP>::P, PN = LOGC + RoundW<INV, LOGR, W, RDN>::P;
    double *bp = buf + cc_pad(elem_base<W>(tid, PP));
    const double *bn = buf + cc_pad(elem_base<W>(tid, PN));
#pragma unroll
    for (int u = 0; u < E; ++u) bp[cc_pad((u32)u << PP)] = x[u];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < E; ++u) x[u] = bn[cc_pad((u32)u << PN)];
}

// global word index (within the limb) of tile element e: row e >> LOGC, column c0 + (e mod 2^LOGC)
template <int LOGC>
__device__ __forceinline__ u32 col_gidx(u32 e, u32 logS, u32 c0) {
    return ((e >> LOGC) << logS) + c0 + (e & ((1u << LOGC) - 1));
}

// All rounds of a column transform after round 0 (exchanges included); `xb` is the
// running LDS buffer index.  The transform's view of the tile: LOGR row bits above
// LOGCV column bits, over a limb of 2^logSV columns (its twiddles); the words' places
// in the limb follow the tile's own geometry (LOGCG column bits, 2^logSG columns),
// which differs from the view when the inverse column pass is one stage shorter
// than the forward one (then element bit LOGCG is a column bit of the view).
template <bool INV, int LOGR, int W, int LOGCV, int LOGCG, int RD>
__device__ __forceinline__ void rounds_from(double (&x)[1 << W], double *lds, u32 &xb, u32 tid,
                                            const double *__restrict__ tw, double q, double qinv, u32 logN,
                                            u32 logSV, u32 logSG, u32 c0) {
    constexpr int NR = RoundW<INV, LOGR, W, 0>::NR;
    if constexpr (RD < NR) {
        constexpr u32 TILE = 1u << (LOGCV + LOGR);
        constexpr u32 PADT = TILE + TILE / 4;  // cc_pad's span
        exch<INV, LOGR, W, LOGCV, RD - 1, RD>(x, lds + xb * PADT, tid);
        xb ^= 1u;
        constexpr u32 PN = LOGCV + RoundW<INV, LOGR, W, RD>::P;
        round_w<INV, LOGR, W, RD>(x, tw, q, qinv, logN, logSV, col_gidx<LOGCG>(elem_base<W>(tid, PN), logSG, c0));
        rounds_from<INV, LOGR, W, LOGCV, LOGCG, RD + 1>(x, lds, xb, tid, tw, q, qinv, logN, logSV, logSG, c0);
    }
}

// Work item = (polynomial, digit, tile of 2^LOGC columns).  With LOGC < 4 the tiles
// of a 16-column block are items 8 apart, so on a persistent grid of a multiple of
// 8 * 2^(4-LOGC) workgroups (dispatched round-robin over the 8 XCDs) they run at the
// same time on one XCD and share its L2 lines (128-byte rows).
// WV > 0: register budget of WV waves per SIMD (the compiler's choice otherwise)
// LOGRI: stages of the inverse column pass (LOGR, or LOGR - 1 when the inverse's
// first pass is the longer one): the tile's 2^LOGC columns of 2^LOGR rows are then,
// for the inverse, 2^(LOGC+1) columns of 2^LOGRI rows (columns c0.. and c0 + 2^logS..)
template <int LOGR, int LOGRI, int W, int LOGC, bool MDRS, int WV>
__global__ __launch_bounds__(((1u << (LOGC + LOGR)) >> W)) __attribute__((amdgpu_waves_per_eu(WV > 0 ? WV : 1))) void k_conv_cols_f64(
    ConvColsArgs a, const double *__restrict__ twf, const double *__restrict__ itwf,
    const ModParams *__restrict__ mp, u32 logN, u32 n_work) {
    constexpr int E = 1 << W;
    constexpr u32 TILE = 1u << (LOGC + LOGR);
    constexpr u32 PADT = TILE + TILE / 4;  // cc_pad's span
    static_assert(LOGRI == LOGR || LOGRI == LOGR - 1, "inverse column pass of LOGR or LOGR - 1 stages");
    constexpr int LOGCI = LOGC + (LOGR - LOGRI);                     // the inverse's column bits
    constexpr int NR = RoundW<false, LOGR, W, 0>::NR;
    constexpr int NRI = RoundW<true, LOGRI, W, 0>::NR;
    constexpr u32 PI0 = LOGCI + RoundW<true, LOGRI, W, 0>::P;     // input layout (inverse round 0)
    constexpr u32 PC = LOGC + RoundW<false, LOGR, W, 0>::P;       // conversion layout
    constexpr u32 PO = LOGC + RoundW<false, LOGR, W, NR - 1>::P;  // output layout (forward last round)
    static_assert(LOGCI + RoundW<true, LOGRI, W, NRI - 1>::P == PC, "windows must meet");
    static_assert(W <= 3, "forward values are centred every 2W <= 6 stages (bound 7.53q < 8q)");
    static_assert(CC_MAX_IN <= 8, "conversion sums of at most 8 terms add exactly without centring");
    __shared__ double lds[2 * PADT];
    // the unit's constants, staged once per digit: conversion pairs (qhat_i mod p_j,
    // / p_j) as [i][j] with zero pairs for i >= n_in, per output j its prime index and
    // (p_j, 1/p_j), per input i (q_i, 1/q_i) and the scaling pair, per output j the
    // ModDown + rescale pairs.  Read with uniform addresses (LDS broadcast) instead of
    // one dependent global load per conversion term.
    __shared__ double2 s_hat[CC_MAX_IN * CC_MAX_OUT];
    __shared__ double2 s_oq[CC_MAX_OUT];
    __shared__ double2 s_md[2 * CC_MAX_OUT];
    __shared__ double2 s_iq[2 * CC_MAX_IN];
    __shared__ u32 s_oprime[CC_MAX_OUT], s_iprime[CC_MAX_IN];
    const u32 tid = threadIdx.x;
    const u32 logS = logN - LOGR;
    const u32 logSI = logN - LOGRI;  // the inverse view's columns
    const u32 ntiles = 1u << (logS - LOGC);
    const u32 ebi = elem_base<W>(tid, PI0);
    constexpr u32 LG = (LOGC < 4) ? 4 - LOGC : 0;  // log2 tiles per 16-column block
    const u32 n_round = (n_work + (8u << LG) - 1) & ~((8u << LG) - 1);
    u32 xb = 0, staged = ~0u;
    for (u32 w = blockIdx.x; w < n_round; w += gridDim.x) {
        const u32 tt = (w & ~((8u << LG) - 1)) | ((w & 7u) << LG) | ((w >> 3) & ((1u << LG) - 1));
        if (tt >= n_work) continue;
        const u32 tI = tt % ntiles, unit = tt / ntiles;
        const u32 d = unit % a.ndig, poly = unit / a.ndig;
        const u32 c0 = tI << LOGC;
        const ConvDesc D = a.descs[d];
        const u32 n_in = D.n_in, n_out = D.n_out;
        if (d != staged) {
            __syncthreads();  // the previous digit's constants are no longer read
            for (u32 k = tid; k < CC_MAX_IN * n_out; k += blockDim.x) {
                const u32 i = k / n_out, j = k % n_out;
                s_hat[i * CC_MAX_OUT + j] =
                    i < n_in ? make_double2(D.hatf[2 * k], D.hatf[2 * k + 1]) : make_double2(0.0, 0.0);
            }
            for (u32 j = tid; j < n_out; j += blockDim.x) {
                const u32 pj = D.out_prime[j];
                s_oprime[j] = pj;
                s_oq[j] = make_double2(mp[pj].qd, mp[pj].qinv);
                if (MDRS && j + 1 < n_out) {
                    s_md[2 * j] = make_double2(a.cst_in[2 * j], a.cst_in[2 * j + 1]);
                    s_md[2 * j + 1] = make_double2(a.cst_pq[2 * j], a.cst_pq[2 * j + 1]);
                }
            }
            for (u32 i = tid; i < n_in; i += blockDim.x) {
                const u32 pi = D.in_prime[i];
                s_iprime[i] = pi;
                s_iq[2 * i] = make_double2(mp[pi].qd, mp[pi].qinv);
                s_iq[2 * i + 1] = make_double2(D.in_ncstf[2 * i], D.in_ncstf[2 * i + 1]);
            }
            __syncthreads();
            staged = d;
        }
        const auto *src = gptr(a.in + (size_t)poly * a.in_pstride + ((size_t)a.in_first[d] << logN));
        const u32 jin = col_gidx<LOGC>(ebi, logS, c0);
        const u32 rstep_in = 1u << (PI0 - LOGC);  // row distance of consecutive elements (in rows)
        // ---- column INTT pass of every input limb; y[i][u]: [x_i N^-1 qhat_i^-1]_{q_i}
        double y[CC_MAX_IN][E];
        double raw[E];
#pragma unroll
        for (int u = 0; u < E; ++u) raw[u] = src[jin + (((u32)u * rstep_in) << logS)];
#pragma unroll
        for (int i = 0; i < (int)CC_MAX_IN; ++i) {
            if ((u32)i < n_in) {
                double x[E];
#pragma unroll
                for (int u = 0; u < E; ++u) x[u] = raw[u];
                const u32 tl = opaque_tid(tid);
                const u32 jl = col_gidx<LOGC>(elem_base<W>(tl, PI0), logS, c0);
                if ((u32)i + 1 < n_in) {
                    const auto *s1 = src + ((size_t)(i + 1) << logN);
#pragma unroll
                    for (int u = 0; u < E; ++u) raw[u] = s1[jl + (((u32)u * rstep_in) << logS)];
                }
                const u32 pi = uniform_u32(s_iprime[i]);
                const double2 qq = s_iq[2 * i], cc = s_iq[2 * i + 1];
                const double q = uniform_f64(qq.x), qinv = uniform_f64(qq.y);
                const double *tw = itwf + ((size_t)pi << (logN + 1));
                round_w<true, LOGRI, W, 0>(x, tw, q, qinv, logN, logSI, jl);
                rounds_from<true, LOGRI, W, LOGCI, LOGC, 1>(x, lds, xb, tl, tw, q, qinv, logN, logSI, logS, c0);
                const double cs = uniform_f64(cc.x), csq = uniform_f64(cc.y);
#pragma unroll
                for (int u = 0; u < E; ++u) {
                    const double v = reduce_f64(mulmod_f64(x[u], cs, csq, q), q, qinv);
                    y[i][u] = v < 0.0 ? v + q : v;
                }
            } else {
#pragma unroll
                for (int u = 0; u < E; ++u) y[i][u] = 0.0;
            }
        }
        // ---- ModDown + rescale: the rescaled limb's ModDown result in coefficient form
        // at these positions, by the linearity of the transforms (mod q_{l-1}):
        //   v = INTT(d_{l-1}) + P^-1 (INTT(acc_{l-1}) - conv_{l-1}),
        // from the rows-pass intermediates of acc_{l-1} and d_{l-1} (without the d term
        // when d is already in acc: lift_d = 0), and its centred lift
        // (rescale.cpp:46-54): the words of the ModDown written out, transformed back and lifted
        double cl[E];
        if (MDRS) {
            const u32 pl = a.lift_prime;
            const double q = mp[pl].qd, qinv = mp[pl].qinv;
            const double *tw = itwf + ((size_t)pl << (logN + 1));
            const auto *lp = gptr(a.lift_in + (size_t)poly * a.lift_stride);
            // lift_cst: (N^-1, /q), (P^-1 N^-1, /q), (P^-1, /q) mod q_{l-1}
            const double ni = a.lift_cst[0], niq = a.lift_cst[1], pni = a.lift_cst[2], pniq = a.lift_cst[3];
            const double pi_ = a.lift_cst[4], piq = a.lift_cst[5];
            // one transform at a time, each folded into the partial sum vs (|terms| <= 0.82q)
            double vs[E], x[E];
#pragma unroll
            for (int u = 0; u < E; ++u) x[u] = lp[jin + (((u32)u * rstep_in) << logS)];
            round_w<true, LOGRI, W, 0>(x, tw, q, qinv, logN, logSI, jin);
            rounds_from<true, LOGRI, W, LOGCI, LOGC, 1>(x, lds, xb, opaque_tid(tid), tw, q, qinv, logN, logSI, logS, c0);
#pragma unroll
            for (int u = 0; u < E; ++u) vs[u] = mulmod_f64(x[u], pni, pniq, q);
            if (a.lift_d) {
#pragma unroll
                for (int u = 0; u < E; ++u) x[u] = lp[(1u << logN) + jin + (((u32)u * rstep_in) << logS)];
                round_w<true, LOGRI, W, 0>(x, tw, q, qinv, logN, logSI, jin);
                rounds_from<true, LOGRI, W, LOGCI, LOGC, 1>(x, lds, xb, opaque_tid(tid), tw, q, qinv, logN, logSI, logS,
                                                           c0);
#pragma unroll
                for (int u = 0; u < E; ++u) vs[u] += mulmod_f64(x[u], ni, niq, q);
            }
            // conv_{l-1}
#pragma unroll
            for (int u = 0; u < E; ++u) x[u] = 0.0;
            const u32 lj = a.lift_j;
            double2 hv[CC_MAX_IN];
#pragma unroll
            for (int i = 0; i < (int)CC_MAX_IN; ++i) {
                const double2 hh = s_hat[i * CC_MAX_OUT + lj];
                hv[i] = make_double2(uniform_f64(hh.x), uniform_f64(hh.y));
            }
#pragma unroll
            for (int i = 0; i < (int)CC_MAX_IN; ++i) {
                if ((u32)i < n_in) {
                    const double h = hv[i].x, hq = hv[i].y;
#pragma unroll
                    for (int u = 0; u < E; ++u) x[u] += mulmod_f64(y[i][u], h, hq, q);
                }
            }
            const double hq_l = 0.5 * q;
#pragma unroll
            for (int u = 0; u < E; ++u) {
                // |vs| <= 1.64q, |conv term| <= 0.63q: centred, then canonical, then lifted
                const double cv = reduce_f64(x[u], q, qinv);
                double v = reduce_f64(vs[u] - mulmod_f64(cv, pi_, piq, q), q, qinv);
                v = v < 0.0 ? v + q : v;
                cl[u] = v > hq_l ? v - q : v;
            }
        }
        // ---- every output limb: conversion, then the forward column pass
        const u32 rstep_o = 1u << (PO - LOGC);
        const u32 jend = (a.j0 + a.nj < n_out) ? a.j0 + a.nj : n_out;
        double *dst_base = a.out + (size_t)poly * a.out_pstride + (size_t)d * a.out_dstride;
        for (u32 j = a.j0; j < jend; ++j) {
            const u32 pj = uniform_u32(s_oprime[j]);
            const double2 qq = s_oq[j];
            const double q = uniform_f64(qq.x), qinv = uniform_f64(qq.y);
            double2 hv[CC_MAX_IN];
#pragma unroll
            for (int i = 0; i < (int)CC_MAX_IN; ++i) {
                const double2 hh = s_hat[i * CC_MAX_OUT + j];
                hv[i] = make_double2(uniform_f64(hh.x), uniform_f64(hh.y));
            }
            double x[E];
#pragma unroll
            for (int u = 0; u < E; ++u) x[u] = 0.0;
#pragma unroll
            for (int i = 0; i < (int)CC_MAX_IN; ++i) {
                if ((u32)i < n_in) {
                    const double h = hv[i].x, hq = hv[i].y;
#pragma unroll
                    for (int u = 0; u < E; ++u) x[u] += mulmod_f64(y[i][u], h, hq, q);
                }
            }
#pragma unroll
            for (int u = 0; u < E; ++u) x[u] = reduce_f64(x[u], q, qinv);
            if (MDRS) {
                // c*T_j - conv_j*(P q_l)^-1 (output j is Q slot j)
                const double2 tt2 = s_md[2 * j], pq2 = s_md[2 * j + 1];
                const double T = uniform_f64(tt2.x), Tq = uniform_f64(tt2.y);
                const double pq = uniform_f64(pq2.x), pqq = uniform_f64(pq2.y);
#pragma unroll
                for (int u = 0; u < E; ++u)
                    x[u] = reduce_f64(mulmod_f64(cl[u], T, Tq, q) - mulmod_f64(x[u], pq, pqq, q), q, qinv);
            }
            const double *tw = twf + ((size_t)pj << (logN + 1));
            const u32 tl = opaque_tid(tid);
            round_w<false, LOGR, W, 0>(x, tw, q, qinv, logN, logS, col_gidx<LOGC>(elem_base<W>(tl, PC), logS, c0));
            rounds_from<false, LOGR, W, LOGC, LOGC, 1>(x, lds, xb, tl, tw, q, qinv, logN, logS, logS, c0);
            const u32 jo = col_gidx<LOGC>(elem_base<W>(tl, PO), logS, c0);
            auto *dst = gptr_w(dst_base + ((size_t)j << logN));
#pragma unroll
            for (int u = 0; u < E; ++u) dst[jo + (((u32)u * rstep_o) << logS)] = x[u];
        }
    }
}

template <int LOGR, int LOGRI, int W, int LOGC, bool MDRS, int WV>
static void launch_cc_inst(const DevTables &t, const ConvColsArgs &a, hipStream_t s) {
    constexpr u32 TPB = (1u << (LOGC + LOGR)) >> W;
    const u32 ntiles = 1u << (t.logN - LOGR - LOGC);
    const u32 n_work = a.npolys * a.ndig * ntiles;
    static int per_cu = 0;
    if (per_cu == 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_conv_cols_f64<LOGR, LOGRI, W, LOGC, MDRS, WV>, TPB,
                                                         0) != hipSuccess ||
            n <= 0)
            n = 1;
        per_cu = n;
    }
    // persistent grid: all resident, a multiple of 8 * (tiles per 16-column block)
    constexpr u32 GR = 8u << ((LOGC < 4) ? 4 - LOGC : 0);
    u32 nwg = t.max_wg / 4 * (u32)per_cu;
    const u32 need = (n_work + GR - 1) & ~(GR - 1);
    if (nwg > need) nwg = need;
    nwg = nwg >= GR ? (nwg & ~(GR - 1)) : nwg;
    hipLaunchKernelGGL((k_conv_cols_f64<LOGR, LOGRI, W, LOGC, MDRS, WV>), dim3(nwg), dim3(TPB), 0, s, a, t.twf,
                       t.itwf, t.mp, t.logN, n_work);
}

// Column-pass lengths of the two-pass FP64 transforms (the first pass of the forward
// NTT, the second of the inverse), lr forward and lri inverse; lr = 0 where the
// kernel does not apply (one-pass transforms).
static void cols_logs(const DevTables &t, u32 &lr, u32 &lri) {
    const u32 lg = t.logN;
    const u32 kf = t.k1_fwd ? t.k1_fwd : lg / 2, ki = t.k1_inv ? t.k1_inv : lg / 2;
    lr = lri = 0;
    if ((lg == 13 || lg == 14) && (kf == lg || ki == lg)) return;  // one-pass transforms
    lr = kf;
    lri = ki;
}

bool conv_cols_fusable(const DevTables &t, u32 max_in) {
    if (!t.use_f64 || (t.f64_off & (1 | 2 | 512)) || t.logN <= 12 || max_in > CC_MAX_IN) return false;
    if (t.L + t.K > CC_MAX_OUT) return false;  // outputs per digit staged in LDS
    u32 lr, lri;
    cols_logs(t, lr, lri);
    // the inverse column pass as long as the forward one, or one stage shorter
    return lr >= 6 && lr <= 9 && t.logN - lr >= 4 && (lri == lr || (lri + 1 == lr && lri >= 6));
}

// Variants (DevTables::cc_var, FHEA_CC_VARIANT): 0 radix-4 rounds on 4-column
// tiles (256 threads at 2^8 rows) with a register budget of 3 waves per SIMD (131 /
// 118 registers with and without the lift); 1 8-column tiles (512 threads) at the
// compiler's register choice; 2 8-column tiles within 4 waves per SIMD; 3 4-column
// tiles within 4 waves per SIMD (127 / 118 registers).  Measured on MI355X (HMult at
// 2^16, B = 64, one lane): variant 0 8,621-8,707 ops/s, 3 8,433-8,469; budgets of 2
// waves or the compiler's default (127 registers each) 8,537-8,595.  The budget
// changes the instruction schedule more than the occupancy: in the build before
// cc_pad the lift instance took 127 registers under both 3 and 4 and ran at 8,627
// and 8,443 ops/s.  An inverse column pass one stage shorter than the forward one
// takes variant 0's shape.
template <bool MDRS>
static void launch_cc_var(const DevTables &t, const ConvColsArgs &a, u32 lr, u32 lri, hipStream_t s) {
#define FHEA_CC(LR)                                                                   \
    case LR:                                                                          \
        if (t.cc_var == 3)                                                            \
            launch_cc_inst<LR, LR, 2, 2, MDRS, 4>(t, a, s);                           \
        else if (t.cc_var == 1)                                                       \
            launch_cc_inst<LR, LR, 2, 3, MDRS, 0>(t, a, s);                           \
        else if (t.cc_var == 2)                                                       \
            launch_cc_inst<LR, LR, 2, 3, MDRS, 4>(t, a, s);                           \
        else                                                                          \
            launch_cc_inst<LR, LR, 2, 2, MDRS, 3>(t, a, s);                           \
        break;
    if (lri + 1 == lr) {
        switch (lr) {
        case 7: launch_cc_inst<7, 6, 2, 2, MDRS, 3>(t, a, s); break;
        case 8: launch_cc_inst<8, 7, 2, 2, MDRS, 3>(t, a, s); break;
        case 9: launch_cc_inst<9, 8, 2, 2, MDRS, 3>(t, a, s); break;
        default: break;
        }
        return;
    }
    switch (lr) {
        FHEA_CC(6)
        FHEA_CC(7)
        FHEA_CC(8)
        FHEA_CC(9)
    default:
        break;
    }
#undef FHEA_CC
}

void launch_conv_cols_f64(const DevTables &t, const ConvColsArgs &a, hipStream_t s) {
    u32 lr, lri;
    cols_logs(t, lr, lri);
    if (a.lift_in)
        launch_cc_var<true>(t, a, lr, lri, s);
    else
        launch_cc_var<false>(t, a, lr, lri, s);
}

}  // namespace fhea
