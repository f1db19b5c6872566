// fhea_ntt.hip -- batched negacyclic NTT / INTT on 64-bit integers (primes < 2^61).
//
// Computes exactly ntt_forward / intt_inverse of the reference
// (hardware_fpga_model/hls/src/ntt.cpp:4-80): Cooley-Tukey with W = tw[m+i]
// (natural in, bit-reversed out) and Gentleman-Sande with W = itw[m+i] followed
// by the N^-1 scaling.  This is the transform of chains the FP64 kernels cannot
// hold exactly (a prime >= 2^50, such as the reference's 60-bit first prime).
//
// Decomposition and tiling are those of the FP64 kernel (fhea_ntt_tile.hpp):
// N <= 4096 in one pass; larger N in two passes over [2^k1 rows][2^k2 columns]
// (forward: COLS then ROWS; inverse: ROWS then COLS), a 4096-word tile per
// workgroup, 16 words per thread in registers, up to four radix-2 stages per
// register round, rounds exchanging through LDS, a persistent grid that loads
// the next tile while the current one computes.
//
// Arithmetic (D. Harvey, "Faster arithmetic for number-theoretic transforms",
// J. Symbolic Comput. 60 (2014)): twiddles carry Shoup companions
// ws = floor(w 2^64 / q), and w*v - floor(v*ws / 2^64)*q lies in [0, 2q) for any
// 64-bit v (mul_shoup_lazy: one 64-bit high product and two low products).
//   forward  x in [0, 4q): a = x[u] folded into [0, 2q), t = w x[u2] in [0, 2q),
//            x[u] = a + t and x[u2] = a - t + 2q, both in [0, 4q);
//   inverse  x in [0, 2q): x[u] = a + b folded into [0, 2q),
//            x[u2] = w (a - b + 2q) in [0, 2q).
// 4q < 2^63 for q < 2^61.  Words stay lazy between the two passes and leave the
// transform canonical ([0, q)); modular arithmetic is exact, so the output words
// equal those of the reference's fully reducing butterflies.
#include "fhea_kernels.hpp"
#include "fhea_ntt_tile.hpp"

namespace fhea {

static constexpr u32 I_TPB = 256;
static constexpr u32 I_LOG_TILE = 12;

struct TileGeomI {
    u32 logN, logS, logC, b0, c0;
    const ulonglong2 *tw;  // (w, ws) pairs of the tile's prime
    u64 q;
    __device__ __forceinline__ u32 gidx(u32 e, u32 logR) const { return tile_gidx(e, logR, logS, logC, b0, c0); }
};

// The stages of round RD on the thread's 16 words; j0 = global index of its
// window's first word.  Stage bit b pairs words 2^wb apart in the window
// (wb = b - P) and needs the 2^(3-wb) twiddles tw[wbase | t].
template <bool INV, int LOGR, int RD>
__device__ __forceinline__ void round_arith_i(u64 (&x)[16], const TileGeomI &g, u32 j0) {
    using R = Round<INV, LOGR, RD>;
    const u64 q = g.q, q2 = g.q << 1;
#pragma unroll
    for (int st = 0; st < R::NS; ++st) {
        const int b = INV ? R::BIT0 + st : R::BIT0 - st;
        const int wb = b - R::P;
        const u32 k = (u32)b + g.logS + 1;
        const u32 wbase = ((1u << g.logN) >> k) | (j0 >> k);
        u64 w[8], ws[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            if (t < (8 >> wb)) {
                const ulonglong2 pr = g.tw[wbase | (u32)t];
                w[t] = pr.x;
                ws[t] = pr.y;
            }
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            if (u & (1 << wb)) continue;
            const int u2 = u | (1 << wb);
            const u64 ww = w[u >> (wb + 1)], wws = ws[u >> (wb + 1)];
            if (!INV) {
                u64 a = x[u];
                a = a >= q2 ? a - q2 : a;
                const u64 t = mul_shoup_lazy(x[u2], ww, wws, q);
                x[u] = a + t;
                x[u2] = a - t + q2;
            } else {
                const u64 a = x[u], bb = x[u2];
                const u64 s = a + bb;
                x[u2] = mul_shoup_lazy(a - bb + q2, ww, wws, q);
                x[u] = s >= q2 ? s - q2 : s;
            }
        }
    }
}

// round RD after round RD-1: exchange the words through LDS into round RD's
// windows, then run its stages
template <bool INV, int LOGR, int RD>
__device__ __forceinline__ void next_round_i(u64 (&x)[16], u64 *lds, const TileGeomI &g, u32 tid, bool active) {
    using Rp = Round<INV, LOGR, RD - 1>;
    using Rn = Round<INV, LOGR, RD>;
    const u32 Pp = g.logC + Rp::P, Pn = g.logC + Rn::P;
    const u32 ebp = (tid & ((1u << Pp) - 1)) | ((tid >> Pp) << (Pp + 4));
    const u32 ebn = (tid & ((1u << Pn) - 1)) | ((tid >> Pn) << (Pn + 4));
    __syncthreads();
    if (active) {
#pragma unroll
        for (int u = 0; u < 16; ++u) lds[lds_pad(ebp | ((u32)u << Pp))] = x[u];
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int u = 0; u < 16; ++u) x[u] = lds[lds_pad(ebn | ((u32)u << Pn))];
        round_arith_i<INV, LOGR, RD>(x, g, g.gidx(ebn, LOGR));
    }
}

// word leaving a pass: lazy between passes, canonical (and scaled by N^-1 after
// the inverse) at the transform's end
__device__ __forceinline__ u64 finish_word(u64 v, int out_canon, int scale, u64 ni, u64 nish, u64 q) {
    if (scale) return mul_shoup(v, ni, nish, q);
    if (!out_canon) return v;
    const u64 q2 = q << 1;
    v = v >= q2 ? v - q2 : v;
    return v >= q ? v - q : v;
}

template <bool INV, int LOGR, bool COLS>
__global__ __launch_bounds__(I_TPB) void k_ntt64i(const u64 *src, u64 *dst, u64 src_stride, u64 dst_stride,
                                                  u64 src_slot, u64 dst_slot, const u32 *__restrict__ prime_map,
                                                  const u64 *__restrict__ twtab, const u64 *__restrict__ ninv,
                                                  const ModParams *__restrict__ mp, u32 logN, u32 logT,
                                                  int out_canon, int scale, u32 tiles_per_limb, u32 n_tiles,
                                                  u32 npolys, const u64 *__restrict__ pmask) {
    __shared__ u64 lds[I_TPB * 16 + I_TPB];  // one tile, padded
    constexpr int NR = (LOGR + 3) / 4;
    using R0 = Round<INV, LOGR, 0>;
    using RL = Round<INV, LOGR, NR - 1>;
    // direct global <-> register transfers when a thread's words form runs of
    // >= 8 consecutive words across the wave (every COLS window, ROWS windows at
    // bit >= 3); otherwise staged through LDS
    constexpr bool DIRECT_IN = COLS || R0::P >= 3;
    constexpr bool DIRECT_OUT = COLS || RL::P >= 3;
    const u32 tid = threadIdx.x;
    const u32 T = 1u << logT;
    const bool active = tid < (T >> 4);
    TileGeomI g;
    g.logN = logN;
    g.logC = COLS ? (logT - LOGR) : 0;
    g.logS = COLS ? (logN - LOGR) : 0;
    const u32 logNB = logT - LOGR - g.logC;
    const u32 ncgl = g.logS - g.logC;  // log2 of column groups per limb
    const u32 P0 = g.logC + R0::P;
    const u32 eb0 = (tid & ((1u << P0) - 1)) | ((tid >> P0) << (P0 + 4));

    auto geom = [&](const TileId &id, TileGeomI &gg) {
        gg.c0 = (id.tI & ((1u << ncgl) - 1)) << gg.logC;
        gg.b0 = (id.tI >> ncgl) << logNB;
    };
    auto issue = [&](const TileId &id, u64 (&r)[16]) {
        TileGeomI gg = g;
        geom(id, gg);
        const u64 *s = src + (size_t)id.poly * src_stride + (size_t)id.slot * src_slot;
        if (DIRECT_IN) {
            if (active) {
                const u32 jb = gg.gidx(eb0, LOGR);
#pragma unroll
                for (int u = 0; u < 16; ++u) r[u] = s[jb + ((u32)u << (R0::P + gg.logS))];
            }
        } else {
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const u32 e = tid + (u32)k * I_TPB;
                if (e < T) r[k] = s[gg.gidx(e, LOGR)];
            }
        }
    };

    TileId cur = find_tile(blockIdx.x, n_tiles, tiles_per_limb, npolys, prime_map, nullptr, pmask);
    u64 raw[16];
    if (cur.tile < n_tiles) issue(cur, raw);
    while (cur.tile < n_tiles) {
        const TileId nxt = find_tile(cur.tile + gridDim.x, n_tiles, tiles_per_limb, npolys, prime_map, nullptr, pmask);
        geom(cur, g);
        g.q = mp[cur.pidx].q;
        g.tw = reinterpret_cast<const ulonglong2 *>(twtab) + ((size_t)cur.pidx << logN);
        u64 *d = dst + (size_t)cur.poly * dst_stride + (size_t)cur.slot * dst_slot;
        u64 x[16];
        if (DIRECT_IN) {
#pragma unroll
            for (int u = 0; u < 16; ++u) x[u] = raw[u];
        } else {
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const u32 e = tid + (u32)k * I_TPB;
                if (e < T) lds[lds_pad(e)] = raw[k];
            }
        }
        // this tile's words are in x or LDS: the next tile's loads go out now
        if (nxt.tile < n_tiles) issue(nxt, raw);
        if (!DIRECT_IN) {
            __syncthreads();
            if (active) {
#pragma unroll
                for (int u = 0; u < 16; ++u) x[u] = lds[lds_pad(eb0 | ((u32)u << P0))];
            }
        }
        if (active) round_arith_i<INV, LOGR, 0>(x, g, g.gidx(eb0, LOGR));
        if constexpr (NR > 1) next_round_i<INV, LOGR, 1>(x, lds, g, tid, active);
        if constexpr (NR > 2) next_round_i<INV, LOGR, 2>(x, lds, g, tid, active);
        if constexpr (NR > 3) next_round_i<INV, LOGR, 3>(x, lds, g, tid, active);

        const u64 q = g.q;
        const u64 ni = scale ? ninv[2 * cur.pidx] : 0, nish = scale ? ninv[2 * cur.pidx + 1] : 0;
        const u32 P = g.logC + RL::P;
        const u32 eb = (tid & ((1u << P) - 1)) | ((tid >> P) << (P + 4));
        if (DIRECT_OUT) {
            if (active) {
                const u32 jb = g.gidx(eb, LOGR);
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    d[jb + ((u32)u << (RL::P + g.logS))] = finish_word(x[u], out_canon, scale, ni, nish, q);
            }
        } else {
            __syncthreads();
            if (active) {
#pragma unroll
                for (int u = 0; u < 16; ++u) lds[lds_pad(eb | ((u32)u << P))] = x[u];
            }
            __syncthreads();
            for (u32 e = tid; e < T; e += I_TPB)
                d[g.gidx(e, LOGR)] = finish_word(lds[lds_pad(e)], out_canon, scale, ni, nish, q);
        }
        __syncthreads();  // LDS reads of this tile are done before the next tile writes
        cur = nxt;
    }
}

template <bool INV, bool COLS>
static void pass_i(const DevTables &t, int logR, u32 tiles_per_limb, u32 npolys, u32 limbs, hipStream_t s,
                   const u64 *src, u64 *dst, u64 ss, u64 ds, u64 sl, u64 dl, const u32 *pm, u32 logT, int oc,
                   int sc, const u64 *pmask) {
    const u64 *tw = INV ? t.itw : t.tw;
    const u32 ntiles = tiles_per_limb * npolys * limbs;
    // persistent grid: the workgroups that can be resident at once (at most 4 per CU)
#define FHEA_NTTI_CASE(LR)                                                                                  \
    case LR: {                                                                                              \
        static int per_cu = 0;                                                                              \
        if (per_cu == 0) {                                                                                  \
            int n = 0;                                                                                      \
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_ntt64i<INV, LR, COLS>, I_TPB, 0) !=      \
                    hipSuccess ||                                                                           \
                n <= 0)                                                                                     \
                n = 4;                                                                                      \
            per_cu = n < 4 ? n : 4;                                                                         \
        }                                                                                                   \
        const u32 use_cu = (t.wg_per_cu && t.wg_per_cu < (u32)per_cu) ? t.wg_per_cu : (u32)per_cu;          \
        const u32 cap = t.max_wg / 4 * use_cu;                                                              \
        const u32 nwg = ntiles < cap ? ntiles : cap;                                                        \
        hipLaunchKernelGGL((k_ntt64i<INV, LR, COLS>), dim3(nwg), dim3(I_TPB), 0, s, src, dst, ss, ds, sl, dl, \
                           pm, tw, t.ninv, t.mp, t.logN, logT, oc, sc, tiles_per_limb, ntiles, npolys, pmask); \
        break;                                                                                              \
    }
    switch (logR) {
        FHEA_NTTI_CASE(6)
        FHEA_NTTI_CASE(7)
        FHEA_NTTI_CASE(8)
        FHEA_NTTI_CASE(9)
        FHEA_NTTI_CASE(10)
        FHEA_NTTI_CASE(11)
        FHEA_NTTI_CASE(12)
    default:
        break;
    }
#undef FHEA_NTTI_CASE
}

// the integer transform of `limbs` slots; pmask restricts it to a prime subset
static void ntt_int(const DevTables &t, bool inverse, const u64 *src, u64 *dst, u64 src_stride, u64 dst_stride,
                    u64 sl, u64 dl, u32 npolys, u32 limbs, const u32 *prime_map, hipStream_t s,
                    const u64 *pmask) {
    const u32 lg = t.logN;
    if (lg <= I_LOG_TILE) {
        // the whole limb is one tile
        if (inverse)
            pass_i<true, false>(t, lg, 1, npolys, limbs, s, src, dst, src_stride, dst_stride, sl, dl, prime_map, lg,
                                1, 1, pmask);
        else
            pass_i<false, false>(t, lg, 1, npolys, limbs, s, src, dst, src_stride, dst_stride, sl, dl, prime_map,
                                 lg, 1, 0, pmask);
        return;
    }
    // column stages k1 as tuned (DevTables::k1_fwd / k1_inv) when both passes fit
    // the instantiated sub-transform lengths 2^6 .. 2^12, else log2(N)/2
    const u32 k1s = inverse ? t.k1_inv : t.k1_fwd;
    const u32 k1 = (k1s >= 6 && k1s <= 12 && lg - k1s >= 6 && lg - k1s <= 12) ? k1s : lg / 2;
    const u32 k2 = lg - k1, tpl = 1u << (lg - I_LOG_TILE);
    if (!inverse) {
        // stages 0..k1-1 over columns (stride 2^k2), then k1..lg-1 over rows
        pass_i<false, true>(t, k1, tpl, npolys, limbs, s, src, dst, src_stride, dst_stride, sl, dl, prime_map,
                            I_LOG_TILE, 0, 0, pmask);
        pass_i<false, false>(t, k2, tpl, npolys, limbs, s, dst, dst, dst_stride, dst_stride, dl, dl, prime_map,
                             I_LOG_TILE, 1, 0, pmask);
    } else {
        pass_i<true, false>(t, k2, tpl, npolys, limbs, s, src, dst, src_stride, dst_stride, sl, dl, prime_map,
                            I_LOG_TILE, 0, 0, pmask);
        pass_i<true, true>(t, k1, tpl, npolys, limbs, s, dst, dst, dst_stride, dst_stride, dl, dl, prime_map,
                           I_LOG_TILE, 1, 1, pmask);
    }
}

void launch_ntt(const DevTables &t, bool inverse, const u64 *src, u64 *dst, u64 src_stride,
                u64 dst_stride, u32 npolys, u32 limbs, const u32 *prime_map, hipStream_t s,
                u64 src_slot, u64 dst_slot) {
    if (npolys == 0 || limbs == 0) return;
    const u64 sl = src_slot ? src_slot : t.N, dl = dst_slot ? dst_slot : t.N;
    if (t.use_f64 && !(t.f64_off & 1)) {
        launch_ntt_f64(t, inverse, src, dst, src_stride, dst_stride, sl, dl, npolys, limbs, prime_map,
                       NttFuse(), s);
        return;
    }
    // mixed chain: limbs of primes below 2^50 on the FP64 transform, the rest here
    // (disjoint limbs, so the two launches do not touch each other's words)
    const u64 *pmask = nullptr;
    if (t.ntt_mixed && !(t.f64_off & 1)) {
        NttFuse fz;
        fz.pmask = t.pmask_small;
        launch_ntt_f64(t, inverse, src, dst, src_stride, dst_stride, sl, dl, npolys, limbs, prime_map, fz, s);
        pmask = t.pmask_big;
    }
    ntt_int(t, inverse, src, dst, src_stride, dst_stride, sl, dl, npolys, limbs, prime_map, s, pmask);
}

void launch_ntt_int(const DevTables &t, bool inverse, const u64 *src, u64 *dst, u64 src_stride,
                    u64 dst_stride, u32 npolys, u32 limbs, const u32 *prime_map, hipStream_t s,
                    const u64 *pmask) {
    if (npolys == 0 || limbs == 0) return;
    ntt_int(t, inverse, src, dst, src_stride, dst_stride, t.N, t.N, npolys, limbs, prime_map, s, pmask);
}

}  // namespace fhea
