// fhea_ntt_f64.hip -- negacyclic NTT / INTT on FP64 FMA (every prime < 2^50).
//
// Same transform as ntt_forward / intt_inverse (hardware_fpga_model/hls/src/
// ntt.cpp:4-80), bit for bit: CT with W = tw[m+i] (natural in, bit-reversed
// out), GS with W = itw[m+i] and the N^-1 scaling.  Arithmetic is exact
// (fhea_f64.hpp); values are signed residues between the stages and canonical
// [0, q) words at the transform's ends.
//
// Decomposition.  N <= 4096: one pass, the whole limb in one workgroup.  Larger
// N: two passes over a limb viewed as [2^k1 rows][2^k2 columns]; the forward
// transform runs its first k1 stages down the columns ("COLS" pass, 16..64
// adjacent columns per workgroup so that every global access is a run of
// consecutive words) and the last k2 stages along rows ("ROWS" pass, several
// rows per workgroup); the inverse runs ROWS then COLS.  A workgroup owns a
// 4096-word tile.  Each of its 256 threads keeps 16 tile elements in
// registers — the elements whose row index differs only in a 4-bit window —
// and runs up to four radix-2 stages on them ("round"); rounds exchange
// through LDS.  All round geometry is compile-time (template LOGR = log2 of
// the sub-transform length).  The twiddles of a round (15 (w, w/q) pairs per
// thread) are loaded before its arithmetic.  Tile loads and stores go straight
// between global memory and registers when the thread's window makes them runs
// of >= 8 consecutive words (every COLS round, ROWS rounds whose window starts
// at bit >= 3); otherwise they are staged through LDS.
//
// Value bounds (q < 2^50, see fhea_f64.hpp):
//   forward: a round starts at |x| <= q + 1 (canonical input, a centred value,
//            the RESCALE_IN product <= q/2 + |c|/4 of |c| <= q/2, or the MDRS_IN
//            difference of two products); stage s adds a product of magnitude
//            <= q/2 + 0.375 B_s (twiddles formed as products, see stage_twiddles),
//            so the four stages reach at most 7.01q < 8q; then every value is
//            centred (|x| <= q/2 + 2);
//   inverse: in the rounds whose twiddles are formed as products (the first
//            round along rows) the sum branch is centred every stage and the
//            difference (<= 4q) is multiplied back to <= 2q; in the rounds that load
//            canonical twiddles (a product then stays <= q/2 + 1) the sums are centred
//            at every second stage and at the round's last: from |x| <= 2q at the
//            round's start they stay below 8q < 2^53, so they add exactly.
//
// Fused ends (used by the rescale and ModDown pipelines to save HBM passes):
//   FUSE_RESCALE_IN   forward COLS prologue: x = centred(v[j]) * T_i (rescale.cpp:43-56)
//   FUSE_RESCALE_OUT  forward ROWS epilogue: out = a*q_l^-1 + x    (rescale.cpp:62-69)
//   FUSE_MODDOWN_OUT  forward ROWS epilogue: out = (acc - x)*P^-1 [+ add]
//                     (key_switch.cpp:253-273 with the composite's poly_add)
//   FUSE_MDRS_IN      forward COLS prologue: x = centred(v[j])*T_i - conv[j]*(P q_l)^-1
//   FUSE_MDRS_OUT     forward ROWS epilogue: out = acc*(P q_l)^-1 + d*q_l^-1 + x
//                     (the ModDown with its add and the rescale as one transform;
//                     every step is exact mod q_i, so the words are those of the
//                     two steps run one after the other)
#include "fhea_f64.hpp"
#include "fhea_kernels.hpp"
#include "fhea_ntt_tile.hpp"

namespace fhea {

static constexpr u32 F_TPB = 256;

// Build-time schedule knobs (defaults are the measured best on MI355X):
//   FHEA_NTT_PREFETCH  1: a tile's global loads are issued one tile ahead (registers)
//   FHEA_NTT_WAVES     >0: register budget for that many waves per SIMD
//   FHEA_NTT_STAGE_FENCE 1: scheduling barrier at every butterfly stage
#ifndef FHEA_NTT_PREFETCH
#define FHEA_NTT_PREFETCH 1
#endif
#ifndef FHEA_NTT_WAVES
#define FHEA_NTT_WAVES 0
#endif
#ifndef FHEA_NTT_STAGE_FENCE
#define FHEA_NTT_STAGE_FENCE 0
#endif
#if FHEA_NTT_WAVES > 0
#define FHEA_NTT_ATTR __attribute__((amdgpu_waves_per_eu(FHEA_NTT_WAVES)))
#else
#define FHEA_NTT_ATTR
#endif
static constexpr u32 F_LOG_TILE = 12;
static constexpr u32 F_TILE = 1u << F_LOG_TILE;

struct TileGeom {
    u32 logN, logS, logC, logT, b0, c0;
    const double *tw;
    double q, qinv;
    // global word index of tile element e
    __device__ __forceinline__ u32 gidx(u32 e, u32 logR) const { return tile_gidx(e, logR, logS, logC, b0, c0); }
};

// Twiddles of a stage.  The thread needs the 2^(3-wb) pairs tw[wbase | t]; wbase
// has those low bits clear.  Since tw[i] = psi^bitrev(i) and bit reversal adds over
// disjoint bit sets, tw[wbase | t] = tw[wbase] * tw[t] (mod q).  FACT: load tw[wbase]
// only and form the others as products with the constants tw[1..7] (same for every
// thread of a prime): one load per stage instead of up to eight.  A product is a
// signed residue, |w| < 0.75q, with w/q rounded twice (w * fl(1/q)), so a butterfly
// product obeys |w v mod q| <= q/2 + 0.375|v|; a forward round from |x| <= q + 1 then
// grows to 1.875q, 3.08q, 4.73q, 7.01q < 8q; in the inverse the products of the
// differences stay below 2q (B <= q/2 + 0.375 * 2B) and the differences below 4q.
template <bool FACT>
__device__ __forceinline__ void stage_twiddles(double (&w)[8], double (&wq)[8], const TileGeom &g, u32 wbase,
                                               int npairs) {
    const double2 *tw2 = reinterpret_cast<const double2 *>(g.tw);
    if constexpr (FACT) {
        const double2 p0 = tw2[wbase];
        w[0] = p0.x;
        wq[0] = p0.y;
#pragma unroll
        for (int t = 1; t < 8; ++t) {
            if (t < npairs) {
                const double2 c = tw2[t];
                w[t] = mulmod_f64(p0.x, c.x, c.y, g.q);
                wq[t] = w[t] * g.qinv;
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            if (t < npairs) {
                const double2 pr = tw2[wbase + (u32)t];  // wbase has its low bits clear
                w[t] = pr.x;
                wq[t] = pr.y;
            }
        }
    }
}

template <bool INV, int LOGR, int RD, bool COLS>
__device__ __forceinline__ void round_arith(double (&x)[16], const TileGeom &g, u32 j0) {
    using R = Round<INV, LOGR, RD>;
    // twiddles formed as products in every forward round and in the first inverse
    // round along rows (the rounds where each thread's pairs are its own); the
    // other inverse rounds load them (measured faster that way)
    constexpr bool FACT = !INV || (!COLS && RD == 0);
    // stage st of the round uses 2^(3-wb) distinct (w, w/q) pairs, wb = its window bit
#pragma unroll
    for (int st = 0; st < R::NS; ++st) {
#if FHEA_NTT_STAGE_FENCE
        // keep each stage's twiddle loads inside the stage (fewer live registers)
        __builtin_amdgcn_sched_barrier(0);
#endif
        const int b = INV ? R::BIT0 + st : R::BIT0 - st;
        const int wb = b - R::P;
        const u32 k = (u32)b + g.logS + 1;
        const u32 wbase = ((1u << g.logN) >> k) | (j0 >> k);
        double w[8], wq[8];
        stage_twiddles<FACT>(w, wq, g, wbase, 8 >> wb);
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            if (u & (1 << wb)) continue;
            const int u2 = u | (1 << wb);
            const double ww = w[u >> (wb + 1)], wwq = wq[u >> (wb + 1)];
            if (!INV) {
                const double rr = mulmod_f64(x[u2], ww, wwq, g.q);
                const double a = x[u];
                x[u] = a + rr;
                x[u2] = a - rr;
            } else {
                const double a = x[u], bb = x[u2];
                const bool centre = FACT || (st & 1) || st == R::NS - 1;
                x[u] = centre ? reduce_f64(a + bb, g.q, g.qinv) : a + bb;
                x[u2] = mulmod_f64(a - bb, ww, wwq, g.q);
            }
        }
    }
    if (!INV) {
#pragma unroll
        for (int u = 0; u < 16; ++u) x[u] = reduce_f64(x[u], g.q, g.qinv);
    }
}

// raw input word j of the pass (the RESCALE_IN prologue reads the INTT'd last limb)
__device__ __forceinline__ u64 load_raw(const char *sbase, u32 j, const NttFuse &fz, u32 poly) {
    if (fz.pro == FUSE_RESCALE_IN) return fz.v[(size_t)poly * fz.v_stride + j];
    return ((const u64 *)sbase)[j];
}

// raw word j of the pass -> signed residue (double)
__device__ __forceinline__ double raw_to_f64(u64 w, int in_canon, const NttFuse &fz, u32 slot, double q,
                                             double qinv, u32 j, u32 poly) {
    if (fz.pro == FUSE_RESCALE_IN) {
        const double c = (w > (fz.ql >> 1)) ? u64_to_f64(w) - (double)fz.ql : u64_to_f64(w);
        return mulmod_f64(c, fz.cst_in[2 * slot], fz.cst_in[2 * slot + 1], q);
    }
    if (fz.pro == FUSE_MDRS_IN) {
        // c*T_i - conv_i*(P q_l)^-1, c = centred lift of the last limb, centred
        const u64 vw = fz.v[(size_t)poly * fz.v_stride + j];
        const double c = (vw > (fz.ql >> 1)) ? u64_to_f64(vw) - (double)fz.ql : u64_to_f64(vw);
        const double d = mulmod_f64(c, fz.cst_in[2 * slot], fz.cst_in[2 * slot + 1], q) -
                         mulmod_f64(u64_to_f64(w), fz.cst_pq[2 * slot], fz.cst_pq[2 * slot + 1], q);
        return reduce_f64(d, q, qinv);
    }
    if (in_canon) return u64_to_f64(w);
    return __longlong_as_double((long long)w);
}

__device__ __forceinline__ void store_out(char *dbase, u32 j, double v, int out_canon, int scale,
                                          double ni, double niq, const NttFuse &fz, u32 poly, u32 slot,
                                          u32 logN, double q, double qinv, bool inv) {
    if (scale) v = mulmod_f64(v, ni, niq, q);
    if (inv) v = reduce_f64(v, q, qinv);
    const size_t off = ((size_t)slot << logN) + j;
    if (fz.epi == FUSE_RESCALE_OUT) {
        const RescaleOut A = fz.ro[poly];
        const double a = mulmod_f64(u64_to_f64(A.a[off]), fz.cst_out[2 * slot], fz.cst_out[2 * slot + 1], q);
        A.out[off] = f64_to_canon(reduce_f64(a + v, q, qinv), q);
        return;
    }
    if (fz.epi == FUSE_MDRS_OUT) {
        // acc*(P q_l)^-1 + d*q_l^-1 + x: |sum| <= 2q + 2
        const RescaleOut A = fz.ro[poly];
        const double x = mulmod_f64(u64_to_f64(fz.acc[(size_t)poly * fz.acc_stride + off]), fz.cst_pq[2 * slot],
                                    fz.cst_pq[2 * slot + 1], q);
        // a null: d already folded into acc (HMult with the tensor in the key switch)
        const double d = A.a ? mulmod_f64(u64_to_f64(A.a[off]), fz.cst_out[2 * slot], fz.cst_out[2 * slot + 1], q) : 0.0;
        A.out[off] = f64_to_canon(reduce_f64(x + d + v, q, qinv), q);
        return;
    }
    if (fz.epi == FUSE_MODDOWN_OUT) {
        const CombineArgs A = fz.cb[poly >> 1];
        const u32 half = poly & 1;
        const double x = u64_to_f64(fz.acc[(size_t)poly * fz.acc_stride + off]);
        // x - v: |x| < q, |v| <= q/2 + 1  ->  |x - v| < 2q
        double r = mulmod_f64(x - v, fz.cst_out[2 * slot], fz.cst_out[2 * slot + 1], q);
        const u64 *add = half ? A.add1 : A.add0;
        if (add) {
            const u64 ag = half ? 0 : A.add0_galois;
            r += u64_to_f64(add[ag ? ((size_t)slot << logN) + galois_src(j, ag, logN) : off]);
        }
        (half ? A.out1 : A.out0)[off] = f64_to_canon(reduce_f64(r, q, qinv), q);
        return;
    }
    if (out_canon) ((u64 *)dbase)[j] = f64_to_canon(v, q);
    else ((double *)dbase)[j] = v;
}

// The fused epilogues of a thread's 16 tile elements (e = tid + k*TPB, values in lds),
// in two groups of eight whose operand words are loaded first: one wait per group
// instead of a load, wait and store per element.  Same arithmetic as store_out.
template <u32 TPB>
__device__ __forceinline__ void store_fused16(const double *lds, u32 tid, const TileGeom &g, u32 logR,
                                              const NttFuse &fz, u32 poly, u32 slot, u32 logN, double q,
                                              double qinv) {
    const size_t sb = (size_t)slot << logN;
    const int epi = fz.epi;
    const RescaleOut R = (epi == FUSE_MODDOWN_OUT) ? RescaleOut{nullptr, nullptr} : fz.ro[poly];
    const CombineArgs C = (epi == FUSE_MODDOWN_OUT) ? fz.cb[poly >> 1] : CombineArgs{nullptr, nullptr, nullptr, nullptr};
    const u32 half = poly & 1;
    // first operand (acc, or a for the rescale), optional second (d, or the add)
    const u64 *p1 = (epi == FUSE_RESCALE_OUT) ? R.a : fz.acc + (size_t)poly * fz.acc_stride;
    const u64 *p2 = (epi == FUSE_MDRS_OUT) ? R.a : (epi == FUSE_MODDOWN_OUT) ? (half ? C.add1 : C.add0) : nullptr;
    u64 *po = (epi == FUSE_MODDOWN_OUT) ? (half ? C.out1 : C.out0) : R.out;
    const auto *a1 = gptr(p1 + sb);
    const auto *a2 = p2 ? gptr(p2 + sb) : nullptr;
    auto *out = gptr_w(po + sb);
    const double co = fz.cst_out[2 * slot], coq = fz.cst_out[2 * slot + 1];
    const double cp = (epi == FUSE_MDRS_OUT) ? fz.cst_pq[2 * slot] : 0.0;
    const double cpq = (epi == FUSE_MDRS_OUT) ? fz.cst_pq[2 * slot + 1] : 0.0;
    // the ModDown's add0 read through sigma_g (a rotation's sigma_g(c0))
    const u64 ag = (epi == FUSE_MODDOWN_OUT && !half) ? C.add0_galois : 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        u64 w1[8], w2[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) w1[k] = a1[g.gidx(tid + (u32)(8 * h + k) * TPB, logR)];
        if (a2) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const u32 j = g.gidx(tid + (u32)(8 * h + k) * TPB, logR);
                w2[k] = a2[ag ? galois_src(j, ag, logN) : j];
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const u32 e = tid + (u32)(8 * h + k) * TPB;
            const double v = lds[lds_pad(e)];
            double r;
            if (epi == FUSE_MDRS_OUT) {
                // acc*(P q_l)^-1 + d*q_l^-1 + x (d folded into acc: no second operand)
                r = mulmod_f64(u64_to_f64(w1[k]), cp, cpq, q) + v;
                if (a2) r += mulmod_f64(u64_to_f64(w2[k]), co, coq, q);
            } else if (epi == FUSE_MODDOWN_OUT) {
                // (acc - v)*P^-1 [+ add]; |acc - v| < 2q
                r = mulmod_f64(u64_to_f64(w1[k]) - v, co, coq, q);
                if (a2) r += u64_to_f64(w2[k]);
            } else {
                // a*q_l^-1 + x
                r = mulmod_f64(u64_to_f64(w1[k]), co, coq, q) + v;
            }
            out[g.gidx(e, logR)] = f64_to_canon(reduce_f64(r, q, qinv), q);
        }
    }
}

// The kernel is persistent (find_tile) and issues the next tile's global loads
// before the current tile's arithmetic, so loads overlap FP64 work.
// PRO (inverse first pass only): 1 the input word is the product a1 * b1 of fz.ta[poly]
// (FUSE_PROD_IN), both factors prefetched; 2 it is in[galois_src(j)] of fz.ga[poly]
// (FUSE_GATHER_IN, sigma_g applied while loading).
// FZ false: an instance for plain transforms (no fused prologue or epilogue compiled
// in, so it keeps fewer registers); the tile walk fields of fz still apply.
// Register budgets: plain instances 4 waves per SIMD, the product prologue 3 (168
// registers instead of the compiler's 185 at 2 waves: HMult +0.35%).
template <bool INV, int LOGR, bool COLS, u32 TPB = F_TPB, bool PF = FHEA_NTT_PREFETCH, int PRO = 0, bool FZ = true>
__global__ __launch_bounds__(TPB) FHEA_NTT_ATTR __attribute__((amdgpu_waves_per_eu(!FZ ? 4 : (PRO == 1 ? 3 : 1)))) void k_ntt64f(const void *src, void *dst, u64 src_stride,
                                                  u64 dst_stride, u64 src_slot, u64 dst_slot,
                                                  u32 limbs, const u32 *__restrict__ prime_map,
                                                  const double *__restrict__ twf,
                                                  const double *__restrict__ ninvf,
                                                  const ModParams *__restrict__ mp, u32 logN,
                                                  u32 logT, int in_canon, int out_canon, int scale,
                                                  NttFuse fz_in, u32 tiles_per_limb, u32 n_tiles,
                                                  u32 npolys) {
    NttFuse fz = fz_in;
    if constexpr (!FZ) {
        fz.pro = FUSE_NONE;
        fz.epi = FUSE_NONE;
    }
    __shared__ double lds[TPB * 16 + TPB];  // one tile of 16 words per thread, padded
    constexpr int NR = (LOGR + 3) / 4;
    using R0 = Round<INV, LOGR, 0>;
    using RL = Round<INV, LOGR, NR - 1>;
    constexpr bool DIRECT_IN = COLS || R0::P >= 3;
    const u32 tid0 = threadIdx.x;
    const u32 tid = tid0;
    const u32 T = 1u << logT;
    const bool active = tid < (T >> 4);
    TileGeom g;
    g.logN = logN;
    g.logT = logT;
    g.logC = COLS ? (logT - LOGR) : 0;
    g.logS = COLS ? (logN - LOGR) : 0;
    const u32 logNB = logT - LOGR - g.logC;
    const u32 ncgl = g.logS - g.logC;  // log2 of column groups per limb
    // column bits below the row window: a COLS tile (only two-pass transforms, 4096-word
    // tiles) has 2^(12 - LOGR) columns, a ROWS tile none; compile-time, so the LDS
    // offsets of a thread's 16 elements are immediates (lds_pad(eb | u << P) =
    // lds_pad(eb) + (u << P) + ((u << P) >> 4): disjoint bits)
    constexpr u32 LC = COLS ? F_LOG_TILE - LOGR : 0;
    constexpr u32 P0 = LC + R0::P;
    const u32 eb0 = (tid & ((1u << P0) - 1)) | ((tid >> P0) << (P0 + 4));
    const bool direct_out_possible = (COLS || RL::P >= 3);

    auto geom = [&](const TileId &id, TileGeom &gg) {
        gg.c0 = (id.tI & ((1u << ncgl) - 1)) << gg.logC;
        gg.b0 = (id.tI >> ncgl) << logNB;
    };
    constexpr bool PROD = PRO == 1;
    auto issue = [&](const TileId &id, u64 (&r)[16], u64 (&r2)[PROD ? 16 : 1]) {
        TileGeom gg = g;
        geom(id, gg);
        if constexpr (PRO == 2) {
            const GatherArgs G = fz.ga[id.poly];
            const auto *pin = gptr(G.in + (size_t)id.slot * src_slot);
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const u32 e = tid + (u32)k * TPB;
                if (e < T) r[k] = pin[galois_src(gg.gidx(e, LOGR), G.galois, logN)];
            }
            return;
        }
        if constexpr (PROD) {
            const TensorArgs TA = fz.ta[id.poly];
            const auto *pa = gptr(TA.a1 + (size_t)id.slot * src_slot), *pb = gptr(TA.b1 + (size_t)id.slot * src_slot);
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const u32 e = tid + (u32)k * TPB;
                if (e < T) {
                    const u32 j = gg.gidx(e, LOGR);
                    r[k] = pa[j];
                    r2[k] = pb[j];
                }
            }
            return;
        }
        const char *sbase = (const char *)src + 8 * ((size_t)id.poly * src_stride + (size_t)id.slot * src_slot);
        if (DIRECT_IN) {
            if (active) {
                const u32 jb = gg.gidx(eb0, LOGR);
#pragma unroll
                for (int u = 0; u < 16; ++u) r[u] = load_raw(sbase, jb + ((u32)u << (R0::P + gg.logS)), fz, id.poly);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const u32 e = tid + (u32)k * TPB;
                if (e < T) r[k] = load_raw(sbase, gg.gidx(e, LOGR), fz, id.poly);
            }
        }
    };

    const u32 nsl = fz.slots ? fz.nslots : limbs;
    auto next_tile = [&](u32 t) {
        return fz.unit_major ? find_tile_um(t, n_tiles, tiles_per_limb, npolys, nsl, prime_map, fz.slots, fz.pmask)
                             : find_tile(t, n_tiles, tiles_per_limb, npolys, prime_map, fz.slots, fz.pmask);
    };
    TileId cur = next_tile(blockIdx.x);
    static_assert(PRO == 0 || (INV && !COLS && !DIRECT_IN), "product / gather prologue: inverse first (row) pass");
    u64 raw[16], raw2[PROD ? 16 : 1];
    if (PF && cur.tile < n_tiles) issue(cur, raw, raw2);
    while (cur.tile < n_tiles) {
        // per-tile thread-derived indices (not kept live across the loop); the forward
        // row pass (fused epilogues) keeps them live: measured faster at 3 waves there
        const u32 tid = opaque_tid(tid0);
        const u32 eb0 = (tid & ((1u << P0) - 1)) | ((tid >> P0) << (P0 + 4));
        const TileId nxt = next_tile(cur.tile + gridDim.x);
        if (!PF) issue(cur, raw, raw2);
        geom(cur, g);
        g.q = mp[cur.pidx].qd;
        g.qinv = mp[cur.pidx].qinv;
        g.tw = twf + ((size_t)cur.pidx << (logN + 1));
        char *dbase = (char *)dst + 8 * ((size_t)cur.poly * dst_stride + (size_t)cur.slot * dst_slot);
        double x[16];
        // ---- round 0 input (this tile's words were loaded one iteration ago)
        if (DIRECT_IN) {
            if (active) {
                const u32 jb0 = g.gidx(eb0, LOGR);
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    x[u] = raw_to_f64(raw[u], in_canon, fz, cur.slot, g.q, g.qinv, jb0 + ((u32)u << (R0::P + g.logS)),
                                      cur.poly);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const u32 e = tid + (u32)k * TPB;
                if (e < T) {
                    if constexpr (PROD)
                        // d2 = a1 b1 of canonical words: |.| <= 0.875q
                        lds[lds_pad(e)] = mulmod_vv_f64(u64_to_f64(raw[k]), u64_to_f64(raw2[k]), g.q, g.qinv);
                    else
                        lds[lds_pad(e)] = raw_to_f64(raw[k], in_canon, fz, cur.slot, g.q, g.qinv, g.gidx(e, LOGR), cur.poly);
                }
            }
        }
        // prefetch the next tile
        if (PF && nxt.tile < n_tiles) issue(nxt, raw, raw2);
        if (!DIRECT_IN) {
            __syncthreads();
            if (active) {
#pragma unroll
                for (int u = 0; u < 16; ++u) x[u] = lds[lds_pad(eb0) + ((u32)u << P0) + (((u32)u << P0) >> 4)];
            }
        }
        if (active) round_arith<INV, LOGR, 0, COLS>(x, g, g.gidx(eb0, LOGR));
        // ---- later rounds (exchange through LDS)
        if constexpr (NR > 1) {
            {
                using Rp = Round<INV, LOGR, 0>;
                using R1 = Round<INV, LOGR, 1>;
                constexpr u32 Pp = LC + Rp::P, P1 = LC + R1::P;
                const u32 ebp = (tid & ((1u << Pp) - 1)) | ((tid >> Pp) << (Pp + 4));
                const u32 eb1 = (tid & ((1u << P1) - 1)) | ((tid >> P1) << (P1 + 4));
                __syncthreads();
                if (active) {
#pragma unroll
                    for (int u = 0; u < 16; ++u) lds[lds_pad(ebp) + ((u32)u << Pp) + (((u32)u << Pp) >> 4)] = x[u];
                }
                __syncthreads();
                if (active) {
#pragma unroll
                    for (int u = 0; u < 16; ++u) x[u] = lds[lds_pad(eb1) + ((u32)u << P1) + (((u32)u << P1) >> 4)];
                    round_arith<INV, LOGR, 1, COLS>(x, g, g.gidx(eb1, LOGR));
                }
            }
            if constexpr (NR > 2) {
                using Rp = Round<INV, LOGR, 1>;
                using R2 = Round<INV, LOGR, 2>;
                constexpr u32 Pp = LC + Rp::P, P2 = LC + R2::P;
                const u32 ebp = (tid & ((1u << Pp) - 1)) | ((tid >> Pp) << (Pp + 4));
                const u32 eb2 = (tid & ((1u << P2) - 1)) | ((tid >> P2) << (P2 + 4));
                __syncthreads();
                if (active) {
#pragma unroll
                    for (int u = 0; u < 16; ++u) lds[lds_pad(ebp) + ((u32)u << Pp) + (((u32)u << Pp) >> 4)] = x[u];
                }
                __syncthreads();
                if (active) {
#pragma unroll
                    for (int u = 0; u < 16; ++u) x[u] = lds[lds_pad(eb2) + ((u32)u << P2) + (((u32)u << P2) >> 4)];
                    round_arith<INV, LOGR, 2, COLS>(x, g, g.gidx(eb2, LOGR));
                }
            }
            if constexpr (NR > 3) {
                using Rp = Round<INV, LOGR, 2>;
                using R3 = Round<INV, LOGR, 3>;
                constexpr u32 Pp = LC + Rp::P, P3 = LC + R3::P;
                const u32 ebp = (tid & ((1u << Pp) - 1)) | ((tid >> Pp) << (Pp + 4));
                const u32 eb3 = (tid & ((1u << P3) - 1)) | ((tid >> P3) << (P3 + 4));
                __syncthreads();
                if (active) {
#pragma unroll
                    for (int u = 0; u < 16; ++u) lds[lds_pad(ebp) + ((u32)u << Pp) + (((u32)u << Pp) >> 4)] = x[u];
                }
                __syncthreads();
                if (active) {
#pragma unroll
                    for (int u = 0; u < 16; ++u) x[u] = lds[lds_pad(eb3) + ((u32)u << P3) + (((u32)u << P3) >> 4)];
                    round_arith<INV, LOGR, 3, COLS>(x, g, g.gidx(eb3, LOGR));
                }
            }
        }
        // ---- output of the last round
        {
            constexpr u32 P = LC + RL::P;
            const u32 eb = (tid & ((1u << P) - 1)) | ((tid >> P) << (P + 4));
            const double ni = scale ? ninvf[2 * cur.pidx] : 0.0, niq = scale ? ninvf[2 * cur.pidx + 1] : 0.0;
            if (direct_out_possible && fz.epi == FUSE_NONE) {
                if (active) {
                    const u32 jb = g.gidx(eb, LOGR);
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        store_out(dbase, jb + ((u32)u << (RL::P + g.logS)), x[u], out_canon, scale, ni, niq, fz,
                                  cur.poly, cur.slot, logN, g.q, g.qinv, INV);
                }
            } else {
                __syncthreads();
                if (active) {
#pragma unroll
                    for (int u = 0; u < 16; ++u) lds[lds_pad(eb) + ((u32)u << P) + (((u32)u << P) >> 4)] = x[u];
                }
                __syncthreads();
                if (!INV && !COLS && fz.epi != FUSE_NONE && T == 16 * TPB)
                    store_fused16<TPB>(lds, tid, g, LOGR, fz, cur.poly, cur.slot, logN, g.q, g.qinv);
                else
                    for (u32 e = tid; e < T; e += TPB)
                        store_out(dbase, g.gidx(e, LOGR), lds[lds_pad(e)], out_canon, scale, ni, niq, fz, cur.poly,
                                  cur.slot, logN, g.q, g.qinv, INV);
            }
        }
        __syncthreads();  // LDS reads of this tile are done before the next tile writes
        cur = nxt;
    }
}

template <bool INV, bool COLS, int PRO, bool FZ>
static void launch_pass_fz(int logR, dim3 grid, u32 npolys, u32 max_wg, u32 wg_cap, hipStream_t s, const void *src, void *dst, u64 ss, u64 ds,
                        u64 sl, u64 dl, u32 limbs, const u32 *pm, const double *twf, const double *ninvf,
                        const ModParams *mp, u32 logN, u32 logT, int ic, int oc, int sc, const NttFuse &fz) {
    const u32 tpl = grid.x;
    u32 ntiles = grid.x * grid.y;
    if (fz.unit_major) ntiles = (npolys * tpl + 7) / 8 * 8 * (grid.y / npolys);  // find_tile_um numbering
    // persistent grid: as many workgroups as can be resident at once (max_wg is
    // 4 per CU; a kernel instance that fits fewer per CU gets fewer)
#define FHEA_NTT_CASE(LR)                                                                            \
    case LR: {                                                                                       \
        static int per_cu = 0;                                                                       \
        if (per_cu == 0) {                                                                           \
            int n = 0;                                                                               \
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_ntt64f<INV, LR, COLS, F_TPB, FHEA_NTT_PREFETCH, PRO, FZ>, F_TPB, 0) != \
                    hipSuccess || n <= 0)                                                            \
                n = 4;                                                                               \
            per_cu = n < 4 ? n : 4;                                                                  \
        }                                                                                            \
        const u32 use_cu = (wg_cap && wg_cap < (u32)per_cu) ? wg_cap : (u32)per_cu;                   \
        const u32 cap = max_wg / 4 * use_cu;                                                          \
        const u32 nwg = ntiles < cap ? ntiles : cap;                                                 \
        hipLaunchKernelGGL((k_ntt64f<INV, LR, COLS, F_TPB, FHEA_NTT_PREFETCH, PRO, FZ>), dim3(nwg), dim3(F_TPB), 0, s, src, dst, ss, ds, sl,   \
                           dl, limbs, pm, twf, ninvf, mp, logN, logT, ic, oc, sc, fz, tpl, ntiles,   \
                           npolys);                                                                  \
        break;                                                                                       \
    }
    switch (logR) {
        FHEA_NTT_CASE(6)
        FHEA_NTT_CASE(7)
        FHEA_NTT_CASE(8)
        FHEA_NTT_CASE(9)
        FHEA_NTT_CASE(10)
        FHEA_NTT_CASE(11)
        FHEA_NTT_CASE(12)
    default:
        break;
    }
#undef FHEA_NTT_CASE
}

template <bool INV, bool COLS, int PRO = 0>
static void launch_pass(int logR, dim3 grid, u32 npolys, u32 max_wg, u32 wg_cap, hipStream_t s, const void *src, void *dst, u64 ss, u64 ds,
                        u64 sl, u64 dl, u32 limbs, const u32 *pm, const double *twf, const double *ninvf,
                        const ModParams *mp, u32 logN, u32 logT, int ic, int oc, int sc, const NttFuse &fz) {
    if (PRO == 0 && fz.pro == FUSE_NONE && fz.epi == FUSE_NONE)
        launch_pass_fz<INV, COLS, PRO, false>(logR, grid, npolys, max_wg, wg_cap, s, src, dst, ss, ds, sl, dl, limbs, pm,
                                              twf, ninvf, mp, logN, logT, ic, oc, sc, fz);
    else
        launch_pass_fz<INV, COLS, PRO, true>(logR, grid, npolys, max_wg, wg_cap, s, src, dst, ss, ds, sl, dl, limbs, pm,
                                             twf, ninvf, mp, logN, logT, ic, oc, sc, fz);
}

// One pass over a whole limb of N = 2^13 or 2^14: N/16 threads, the limb in LDS,
// each limb read and written once (no register prefetch: it would not fit).
template <bool INV, int LG>
static void launch_limb_pass(u32 ntiles, u32 npolys, u32 max_wg, u32 wg_cap, hipStream_t s, const void *src,
                             void *dst, u64 ss, u64 ds, u64 sl, u64 dl, u32 limbs, const u32 *pm,
                             const double *twf, const double *ninvf, const ModParams *mp, int ic, int oc, int sc,
                             const NttFuse &fz) {
    constexpr u32 TPB = (1u << LG) / 16;
    static int per_cu = 0;
    if (per_cu == 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_ntt64f<INV, LG, false, TPB, false>, TPB, 0) !=
                hipSuccess ||
            n <= 0)
            n = 1;
        per_cu = n < 4 ? n : 4;
    }
    const u32 use_cu = (wg_cap && wg_cap < (u32)per_cu) ? wg_cap : (u32)per_cu;
    const u32 cap = max_wg / 4 * use_cu;
    const u32 nwg = ntiles < cap ? ntiles : cap;
    hipLaunchKernelGGL((k_ntt64f<INV, LG, false, TPB, false>), dim3(nwg), dim3(TPB), 0, s, src, dst, ss, ds, sl, dl,
                       limbs, pm, twf, ninvf, mp, (u32)LG, (u32)LG, ic, oc, sc, fz, 1u, ntiles, npolys);
}

void launch_ntt_f64(const DevTables &t, bool inverse, const u64 *src, u64 *dst, u64 ss, u64 ds, u64 sl,
                    u64 dl, u32 npolys, u32 limbs, const u32 *pm, const NttFuse &fz, hipStream_t s) {
    const u32 lg = t.logN;
    const double *tw = inverse ? t.itwf : t.twf;
    NttFuse none;
    none.slots = fz.slots;
    none.nslots = fz.nslots;
    none.pmask = fz.pmask;
    const u32 ylimbs = fz.slots ? fz.nslots : limbs;
    if (lg <= F_LOG_TILE) {
        dim3 grid(1, npolys * ylimbs);
        if (inverse)
            launch_pass<true, false>(lg, grid, npolys, t.max_wg, t.wg_per_cu, s, src, dst, ss, ds, sl, dl, limbs, pm, tw, t.ninvf, t.mp, lg, lg,
                                     1, 1, 1, none);
        else
            launch_pass<false, false>(lg, grid, npolys, t.max_wg, t.wg_per_cu, s, src, dst, ss, ds, sl, dl, limbs, pm, tw, t.ninvf, t.mp, lg,
                                      lg, 1, 1, 0, fz);
        return;
    }
    const u32 k1s = inverse ? t.k1_inv : t.k1_fwd;
    if ((lg == 13 || lg == 14) && k1s == lg) {
        const u32 nt = npolys * ylimbs;
#define FHEA_LIMB_PASS(LG)                                                                                   \
    if (inverse)                                                                                              \
        launch_limb_pass<true, LG>(nt, npolys, t.max_wg, t.wg_per_cu, s, src, dst, ss, ds, sl, dl, limbs, pm, tw, \
                                   t.ninvf, t.mp, 1, 1, 1, none);                                             \
    else                                                                                                      \
        launch_limb_pass<false, LG>(nt, npolys, t.max_wg, t.wg_per_cu, s, src, dst, ss, ds, sl, dl, limbs, pm, tw, \
                                    t.ninvf, t.mp, 1, 1, 0, fz);
        if (lg == 13) {
            FHEA_LIMB_PASS(13)
        } else {
            FHEA_LIMB_PASS(14)
        }
#undef FHEA_LIMB_PASS
        return;
    }
    const u32 k1 = k1s ? k1s : lg / 2, k2 = lg - k1;
    dim3 grid(1u << (lg - F_LOG_TILE), npolys * ylimbs);
    const bool first = fz.only_pass != 2, second = fz.only_pass != 1;
    // the second pass reads the first's output at dst, or at src when it runs alone
    const void *s2 = first ? (const void *)dst : (const void *)src;
    const u64 ss2 = first ? ds : ss, sl2 = first ? dl : sl;
    if (!inverse) {
        NttFuse pro = fz, epi = fz;
        pro.epi = FUSE_NONE;
        epi.pro = FUSE_NONE;
        // the first pass of a transform whose prologue reads the rescale lift walks its
        // tiles unit-major, so every slot reads the lift's words from one XCD's L2
        // (ModDown + rescale first pass 336 -> 319 us per 16-item chunk); FHEA_F64_OFF
        // bit 256 keeps the slot-major walk
        pro.unit_major = !(t.f64_off & 256) && (fz.pro == FUSE_RESCALE_IN || fz.pro == FUSE_MDRS_IN);
        epi.unit_major = 0;
        if (first)
            launch_pass<false, true>(k1, grid, npolys, t.max_wg, t.wg_per_cu, s, src, dst, ss, ds, sl, dl, limbs, pm, tw,
                                     t.ninvf, t.mp, lg, F_LOG_TILE, 1, 0, 0, pro);
        if (second)
            launch_pass<false, false>(k2, grid, npolys, t.max_wg, t.wg_per_cu, s, s2, dst, ss2, ds, sl2, dl, limbs, pm, tw,
                                      t.ninvf, t.mp, lg, F_LOG_TILE, 0, 1, 0, epi);
    } else {
        // FUSE_PROD_IN: the first pass reads its input as the product a1 * b1 (fz.ta)
        // FUSE_GATHER_IN: ... as sigma_g of fz.ga's polynomials
        NttFuse prod = none;
        prod.pro = fz.pro;
        prod.ta = fz.ta;
        prod.ga = fz.ga;
        if (first && fz.pro == FUSE_PROD_IN)
            launch_pass<true, false, 1>(k2, grid, npolys, t.max_wg, t.wg_per_cu, s, src, dst, ss, ds, sl, dl, limbs, pm,
                                        tw, t.ninvf, t.mp, lg, F_LOG_TILE, 1, 0, 0, prod);
        else if (first && fz.pro == FUSE_GATHER_IN)
            launch_pass<true, false, 2>(k2, grid, npolys, t.max_wg, t.wg_per_cu, s, src, dst, ss, ds, sl, dl, limbs, pm,
                                        tw, t.ninvf, t.mp, lg, F_LOG_TILE, 1, 0, 0, prod);
        else if (first)
            launch_pass<true, false>(k2, grid, npolys, t.max_wg, t.wg_per_cu, s, src, dst, ss, ds, sl, dl, limbs, pm, tw,
                                     t.ninvf, t.mp, lg, F_LOG_TILE, 1, 0, 0, none);
        if (second)
            launch_pass<true, true>(k1, grid, npolys, t.max_wg, t.wg_per_cu, s, s2, dst, ss2, ds, sl2, dl, limbs, pm, tw,
                                    t.ninvf, t.mp, lg, F_LOG_TILE, 0, 1, 1, none);
    }
}

// ---------------------------------------------------------------------------
// Key-switch inner product fused into the transforms around it (see KsRowsArgs).
// A work item is one 4096-word tile (positions tI*4096 ..) of one QlP slot i of one
// batch item z; numbering: tile fastest (so that on a persistent grid of a
// multiple of 8 workgroups every XCD keeps the same tiles, and the key tile of
// (i, tI) is reused from its L2 by the batch items that follow), then item, then slot.
// Arithmetic as k_ks_inner_f64: canonical digit words times canonical key words
// by mulmod_vv_f64 (<= 0.875q each), at most KS_FUSE_MAX_PARTS terms, centred once.
// ---------------------------------------------------------------------------
static constexpr u32 KS_FUSE_MAX_PARTS = 8;
#ifndef FHEA_KS_ROWS_WAVES
#define FHEA_KS_ROWS_WAVES 0
#endif
#if FHEA_KS_ROWS_WAVES > 0
#define FHEA_KS_ROWS_ATTR __attribute__((amdgpu_waves_per_eu(FHEA_KS_ROWS_WAVES)))
#else
#define FHEA_KS_ROWS_ATTR
#endif

// Rounds 1.. of a tile transform after round 0 (x: round-0 layout in, last-round
// layout out), exchanging through lds.
template <bool INV, int LOGR, bool COLS, int RD>
__device__ __forceinline__ void next_round(double (&x)[16], double *lds, const TileGeom &g, u32 tid) {
    using Rp = Round<INV, LOGR, RD - 1>;
    using Rn = Round<INV, LOGR, RD>;
    constexpr u32 Pp = (COLS ? F_LOG_TILE - LOGR : 0) + Rp::P, Pn = (COLS ? F_LOG_TILE - LOGR : 0) + Rn::P;
    const u32 ebp = (tid & ((1u << Pp) - 1)) | ((tid >> Pp) << (Pp + 4));
    const u32 ebn = (tid & ((1u << Pn) - 1)) | ((tid >> Pn) << (Pn + 4));
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) lds[lds_pad(ebp) + ((u32)u << Pp) + (((u32)u << Pp) >> 4)] = x[u];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) x[u] = lds[lds_pad(ebn) + ((u32)u << Pn) + (((u32)u << Pn) >> 4)];
    round_arith<INV, LOGR, RD, COLS>(x, g, g.gidx(ebn, LOGR));
}

template <bool INV, int LOGR, bool COLS>
__device__ __forceinline__ void later_rounds(double (&x)[16], double *lds, const TileGeom &g, u32 tid) {
    constexpr int NR = (LOGR + 3) / 4;
    if constexpr (NR > 1) next_round<INV, LOGR, COLS, 1>(x, lds, g, tid);
    if constexpr (NR > 2) next_round<INV, LOGR, COLS, 2>(x, lds, g, tid);
    if constexpr (NR > 3) next_round<INV, LOGR, COLS, 3>(x, lds, g, tid);
}

// x (last-round layout of a row tile) into lds in position order: lds[lds_pad(e)] =
// element e of the tile (position tI*4096 + e).
template <bool INV, int LOGR>
__device__ __forceinline__ void rows_to_lds(const double (&x)[16], double *lds, u32 tid) {
    using RL = Round<INV, LOGR, (LOGR + 3) / 4 - 1>;
    constexpr u32 P = RL::P;
    const u32 eb = (tid & ((1u << P) - 1)) | ((tid >> P) << (P + 4));
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) lds[lds_pad(eb) + ((u32)u << P) + (((u32)u << P) >> 4)] = x[u];
    __syncthreads();
}

// Forward row pass (LOGRF stages) of the 4096-word tile at src + (tile base), the
// result left in lds in position order (element e at lds_pad(e)).  src holds
// first-pass intermediates (|x| <= q/2 + 2).
template <int LOGRF>
__device__ __forceinline__ void fwd_rows_to_lds(const __attribute__((address_space(1))) double *src, double *lds,
                                                const TileGeom &g, u32 tid, u32 base) {
    using RF0 = Round<false, LOGRF, 0>;
    double x[16];
    const u32 eb0 = (tid & ((1u << RF0::P) - 1)) | ((tid >> RF0::P) << (RF0::P + 4));
    if (RF0::P >= 3) {
        const u32 jb = g.gidx(eb0, LOGRF);
#pragma unroll
        for (int u = 0; u < 16; ++u) x[u] = src[jb + ((u32)u << RF0::P)];
    } else {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) lds[lds_pad(tid + (u32)k * F_TPB)] = src[base + tid + (u32)k * F_TPB];
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 16; ++u) x[u] = lds[lds_pad(eb0) + ((u32)u << RF0::P) + (((u32)u << RF0::P) >> 4)];
    }
    round_arith<false, LOGRF, 0, false>(x, g, g.gidx(eb0, LOGRF));
    later_rounds<false, LOGRF, false>(x, lds, g, tid);
    rows_to_lds<false, LOGRF>(x, lds, tid);
}

// OWN: how an item's own limbs are read: 0 as words of c, 1 formed from the HMult
// factors (ta: d2 = a1 b1, with P d0 / P d1 added), 2 through sigma_g (c read at
// galois_src); 3 as 1 and 4 as 2, with the last row pass of the ModDown + rescale
// (3) or of the ModDown (4) and its epilogue in place of the stores of the sums
// (KsRowsArgs::tt); 5 as 1, with the ModDown INTT's first (row) pass in place of the
// stores (KsRowsArgs::lift; P slots have no own limb, so rotations use it too); one
// instance each, so each keeps its own register budget.
template <int LOGRF, int LOGRI, bool PSLOT, int OWN>
__global__ __launch_bounds__(F_TPB) FHEA_KS_ROWS_ATTR __attribute__((amdgpu_waves_per_eu(PSLOT ? 1 : 3))) void k_ks_rows_f64(KsRowsArgs a, const double *__restrict__ twf,
                                                     const double *__restrict__ itwf,
                                                     const ModParams *__restrict__ mp, u32 logN, u32 slot0,
                                                     u32 n_work) {
    __shared__ double lds[F_TPB * 16 + F_TPB];
    const u32 tid0 = threadIdx.x;
    const u32 T = 1u << (logN - F_LOG_TILE);
    const u32 N = 1u << logN;
    const u32 lk = a.l + a.K;
    const u64 kstride = (u64)(a.key_limbs + a.K) << logN;
    for (u32 w = blockIdx.x; w < n_work; w += gridDim.x) {
        const u32 tid = opaque_tid(tid0);
        const u32 tI = w % T;
        const u32 zi = w / T;
        const u32 z = zi % a.nb;
        const u32 i = slot0 + zi / a.nb;
        const u32 pidx = (i < a.l) ? i : (a.L + i - a.l);
        const u32 kl = (i < a.l) ? i : (a.key_limbs + i - a.l);
        const double q = mp[pidx].qd, qinv = mp[pidx].qinv;
        const u32 base = tI << F_LOG_TILE;
        const KsIpArgs A = a.args[z];
        double s0[16], s1[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) s0[k] = s1[k] = 0.0;
        for (u32 p = 0; p < a.nparts; ++p) {
            const u32 tid = opaque_tid(tid0);
            const u32 start = p * a.alpha;
            const u32 sz = (p == a.nparts - 1) ? (a.l - start) : a.alpha;
            const auto *ka = gptr(A.key + (size_t)p * kstride + ((size_t)kl << logN) + base + tid);
            const auto *kb = ka + (size_t)a.dnum * kstride;
            if ((OWN == 1 || OWN == 3 || OWN == 5) && i >= start && i < start + sz) {
                // HMult with the tensor folded in: the own limb d2 = a1 b1, and
                // acc0 += P d0, acc1 += P d1 (d0 = a0 b0, d1 = a0 b1 + a1 b0), every product
                // of canonical words (mulmod_vv_f64, <= 0.875q); the sums are centred after
                // (<= q/2 + 2) so the later digits' terms keep them below 8q
                const TensorArgs TA = a.ta[z];
                const size_t o = ((size_t)i << logN) + base + tid;
                const auto *pa0 = gptr(TA.a0 + o), *pa1 = gptr(TA.a1 + o);
                const auto *pb0 = gptr(TA.b0 + o), *pb1 = gptr(TA.b1 + o);
                const double pm = a.pmodq[2 * i], pmq = a.pmodq[2 * i + 1];
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    // two elements' six loads in flight at a time (register budget)
                    if ((k & 1) == 0) __builtin_amdgcn_sched_barrier(0);
                    const size_t e = (size_t)k * F_TPB;
                    const double a0 = u64_to_f64(pa0[e]), a1 = u64_to_f64(pa1[e]);
                    const double b0 = u64_to_f64(pb0[e]), b1 = u64_to_f64(pb1[e]);
                    double d2 = mulmod_vv_f64(a1, b1, q, qinv);
                    d2 = d2 < 0.0 ? d2 + q : d2;  // the canonical word
                    const double d0 = mulmod_vv_f64(a0, b0, q, qinv);
                    const double d1 = mulmod_vv_f64(a0, b1, q, qinv) + mulmod_vv_f64(a1, b0, q, qinv);
                    s0[k] = reduce_f64(s0[k] + mulmod_vv_f64(d2, u64_to_f64(kb[e]), q, qinv) +
                                           mulmod_f64(d0, pm, pmq, q),
                                       q, qinv);
                    s1[k] = reduce_f64(s1[k] + mulmod_vv_f64(d2, u64_to_f64(ka[e]), q, qinv) +
                                           mulmod_f64(d1, pm, pmq, q),
                                       q, qinv);
                }
                continue;
            }
            if (i >= start && i < start + sz) {
                // the digit's own limb: the input itself (eval domain, canonical)
                // (through sigma_g when the item's input is the unpermuted polynomial)
                const auto *c = gptr(A.c + ((size_t)i << logN));
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    // groups of four elements: the loads of one group are in flight together
                    // (fewer live registers than all 48 at once)
                    if ((k & 3) == 0) __builtin_amdgcn_sched_barrier(0);
                    const u32 pos = base + tid + (u32)k * F_TPB;
                    const double d = u64_to_f64(c[(OWN == 2 || OWN == 4) ? galois_src(pos, A.galois, logN) : pos]);
                    s0[k] += mulmod_vv_f64(d, u64_to_f64(kb[(size_t)k * F_TPB]), q, qinv);
                    s1[k] += mulmod_vv_f64(d, u64_to_f64(ka[(size_t)k * F_TPB]), q, qinv);
                }
                continue;
            }
            // second forward pass of the converted limb, in registers
            const u32 pos = (i < start) ? i : (i - sz);
            const auto *src = gptr((const double *)(a.compact + (size_t)z * a.cz + ((size_t)(p * lk + pos) << logN)));
            TileGeom g;
            g.logN = logN;
            g.logT = F_LOG_TILE;
            g.logC = 0;
            g.logS = 0;
            g.c0 = 0;
            g.b0 = tI << (F_LOG_TILE - LOGRF);
            g.q = q;
            g.qinv = qinv;
            g.tw = twf + ((size_t)pidx << (logN + 1));
            fwd_rows_to_lds<LOGRF>(src, lds, g, tid, base);
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                if ((k & 3) == 0) __builtin_amdgcn_sched_barrier(0);
                const double v = lds[lds_pad(tid + (u32)k * F_TPB)];
                const double d = v < 0.0 ? v + q : v;  // |v| <= q/2 + 2: the canonical word
                s0[k] += mulmod_vv_f64(d, u64_to_f64(kb[(size_t)k * F_TPB]), q, qinv);
                s1[k] += mulmod_vv_f64(d, u64_to_f64(ka[(size_t)k * F_TPB]), q, qinv);
            }
        }
        if (OWN == 3 || OWN == 4) {
            // OWN 3, the ModDown + rescale: out_s = acc_s (P q_l)^-1 + x_s, x_s the row
            // pass of tt_s, the column-pass output of c T_i - conv_i (P q_l)^-1
            // (FUSE_MDRS_OUT); |acc (P q_l)^-1| <= q/2 + 1, |x| <= q/2 + 2.
            // OWN 4, the ModDown: out_s = (acc_s - x_s) P^-1 [+ add_s], x_s the row pass
            // of conv(P part) (FUSE_MODDOWN_OUT); |acc - x| <= q + 3
            const u32 tid = opaque_tid(tid0);
            TileGeom g;
            g.logN = logN;
            g.logT = F_LOG_TILE;
            g.logC = 0;
            g.logS = 0;
            g.c0 = 0;
            g.b0 = tI << (F_LOG_TILE - LOGRF);
            g.q = q;
            g.qinv = qinv;
            g.tw = twf + ((size_t)pidx << (logN + 1));
            // (item and slot through opaque_tid: their addresses are formed here, not
            // kept live across the digit loop)
            const u32 zo = opaque_tid(z), io = opaque_tid(i);
            const double cp = a.cst[2 * io], cpq = a.cst[2 * io + 1];
            const size_t so = ((size_t)io << logN) + base + tid;
            // one polynomial's row pass and epilogue (called for s = 0 and s = 1)
            auto finish = [&](const double (&acc)[16], u32 sp) {
                const auto *src = gptr(a.tt + (size_t)(2 * zo + sp) * a.tt_pstride + ((size_t)io << logN));
                fwd_rows_to_lds<LOGRF>(src, lds, g, tid, base);
                if (OWN == 3) {
                    auto *o = gptr_w(a.ro[2 * zo + sp].out + so);
#pragma unroll
                    for (int k = 0; k < 16; ++k) {
                        const double v = lds[lds_pad(tid + (u32)k * F_TPB)];
                        const double r = mulmod_f64(reduce_f64(acc[k], q, qinv), cp, cpq, q) + v;
                        o[(size_t)k * F_TPB] = f64_to_canon(reduce_f64(r, q, qinv), q);
                    }
                } else {
                    const CombineArgs C = a.cb[zo];
                    auto *o = gptr_w((sp ? C.out1 : C.out0) + so);
                    const u64 *add = sp ? C.add1 : C.add0;
                    const u64 ag = sp ? 0 : C.add0_galois;
                    const auto *ad = add ? gptr(add + ((size_t)io << logN)) : nullptr;
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        // the add's words of eight elements first (one wait), through
                        // sigma_g's source index when ag
                        u64 w[8];
                        if (ad && ag) {
#pragma unroll
                            for (int k = 0; k < 8; ++k)
                                w[k] = ad[galois_src(base + tid + (u32)(8 * h + k) * F_TPB, ag, logN)];
                        } else if (ad) {
#pragma unroll
                            for (int k = 0; k < 8; ++k) w[k] = ad[base + tid + (u32)(8 * h + k) * F_TPB];
                        }
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            const int e = 8 * h + k;
                            const double v = lds[lds_pad(tid + (u32)e * F_TPB)];
                            double r = mulmod_f64(reduce_f64(acc[e], q, qinv) - v, cp, cpq, q);
                            if (ad) r += u64_to_f64(w[k]);
                            o[(size_t)e * F_TPB] = f64_to_canon(reduce_f64(r, q, qinv), q);
                        }
                    }
                }
            };
            finish(s0, 0);
            finish(s1, 1);
        } else if (OWN == 5) {
            // the ModDown INTT's first pass on each sum (input: the canonical words):
            // P slot i to pb, the rescaled limb l-1 to lift
            using RI0 = Round<true, LOGRI, 0>;
            using RIL = Round<true, LOGRI, (LOGRI + 3) / 4 - 1>;
            const u32 tid = opaque_tid(tid0);
            TileGeom g;
            g.logN = logN;
            g.logT = F_LOG_TILE;
            g.logC = 0;
            g.logS = 0;
            g.c0 = 0;
            g.b0 = tI << (F_LOG_TILE - LOGRI);
            g.q = q;
            g.qinv = qinv;
            g.tw = itwf + ((size_t)pidx << (logN + 1));
            const u32 zo = opaque_tid(z), io = opaque_tid(i);
            auto finish = [&](const double (&acc)[16], u32 sp) {
                double *dp = (io >= a.l) ? (double *)a.pb + ((size_t)(2 * zo + sp) * a.K + (io - a.l)) * N
                                         : a.lift + (size_t)(2 * zo + sp) * N;
                auto *dst = gptr_w(dp);
                __syncthreads();
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const double v = reduce_f64(acc[k], q, qinv);
                    lds[lds_pad(tid + (u32)k * F_TPB)] = v < 0.0 ? v + q : v;
                }
                __syncthreads();
                double x[16];
                const u32 eb0 = (tid & ((1u << RI0::P) - 1)) | ((tid >> RI0::P) << (RI0::P + 4));
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    x[u] = lds[lds_pad(eb0) + ((u32)u << RI0::P) + (((u32)u << RI0::P) >> 4)];
                round_arith<true, LOGRI, 0, false>(x, g, g.gidx(eb0, LOGRI));
                later_rounds<true, LOGRI, false>(x, lds, g, tid);
                if (RIL::P >= 3) {
                    const u32 eb = (tid & ((1u << RIL::P) - 1)) | ((tid >> RIL::P) << (RIL::P + 4));
                    const u32 jb = g.gidx(eb, LOGRI);
#pragma unroll
                    for (int u = 0; u < 16; ++u) dst[jb + ((u32)u << RIL::P)] = reduce_f64(x[u], q, qinv);
                } else {
                    rows_to_lds<true, LOGRI>(x, lds, tid);
#pragma unroll
                    for (int k = 0; k < 16; ++k)
                        dst[base + tid + (u32)k * F_TPB] = reduce_f64(lds[lds_pad(tid + (u32)k * F_TPB)], q, qinv);
                }
            };
            finish(s0, 0);
            finish(s1, 1);
        } else if (!PSLOT) {
            auto *o0 = gptr_w(a.acc + (size_t)z * a.az + ((size_t)i << logN) + base + tid);
            auto *o1 = o0 + ((size_t)lk << logN);
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                o0[(size_t)k * F_TPB] = f64_to_canon(reduce_f64(s0[k], q, qinv), q);
                o1[(size_t)k * F_TPB] = f64_to_canon(reduce_f64(s1[k], q, qinv), q);
            }
        } else {
            // first pass of the ModDown INTT on each sum (input: the canonical words)
            using RI0 = Round<true, LOGRI, 0>;
            using RIL = Round<true, LOGRI, (LOGRI + 3) / 4 - 1>;
            TileGeom g;
            g.logN = logN;
            g.logT = F_LOG_TILE;
            g.logC = 0;
            g.logS = 0;
            g.c0 = 0;
            g.b0 = tI << (F_LOG_TILE - LOGRI);
            g.q = q;
            g.qinv = qinv;
            g.tw = itwf + ((size_t)pidx << (logN + 1));
            const u32 j = i - a.l;
            for (u32 sp = 0; sp < 2; ++sp) {
                auto *dst = gptr_w((double *)(a.pb + ((size_t)(2 * z + sp) * a.K + j) * N));
                __syncthreads();
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const double v = reduce_f64(sp ? s1[k] : s0[k], q, qinv);
                    lds[lds_pad(tid + (u32)k * F_TPB)] = v < 0.0 ? v + q : v;
                }
                __syncthreads();
                double x[16];
                const u32 eb0 = (tid & ((1u << RI0::P) - 1)) | ((tid >> RI0::P) << (RI0::P + 4));
#pragma unroll
                for (int u = 0; u < 16; ++u) x[u] = lds[lds_pad(eb0) + ((u32)u << RI0::P) + (((u32)u << RI0::P) >> 4)];
                round_arith<true, LOGRI, 0, false>(x, g, g.gidx(eb0, LOGRI));
                later_rounds<true, LOGRI, false>(x, lds, g, tid);
                if (RIL::P >= 3) {
                    const u32 eb = (tid & ((1u << RIL::P) - 1)) | ((tid >> RIL::P) << (RIL::P + 4));
                    const u32 jb = g.gidx(eb, LOGRI);
#pragma unroll
                    for (int u = 0; u < 16; ++u) dst[jb + ((u32)u << RIL::P)] = reduce_f64(x[u], q, qinv);
                } else {
                    rows_to_lds<true, LOGRI>(x, lds, tid);
#pragma unroll
                    for (int k = 0; k < 16; ++k)
                        dst[base + tid + (u32)k * F_TPB] = reduce_f64(lds[lds_pad(tid + (u32)k * F_TPB)], q, qinv);
                }
            }
        }
        __syncthreads();  // this item's LDS reads are done before the next item writes
    }
}

template <int LOGRF, int LOGRI, bool PSLOT, int OWN>
static void launch_ks_rows_own(const DevTables &t, const KsRowsArgs &a, u32 slot0, u32 nslots, hipStream_t s) {
    static int per_cu = 0;
    if (per_cu == 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_ks_rows_f64<LOGRF, LOGRI, PSLOT, OWN>, F_TPB, 0) !=
                hipSuccess ||
            n <= 0)
            n = 2;
        per_cu = n < 4 ? n : 4;
    }
    const u32 n_work = nslots * a.nb * (1u << (t.logN - F_LOG_TILE));
    const u32 cap = t.max_wg / 4 * (u32)per_cu;
    const u32 nwg = n_work < cap ? n_work : cap;
    hipLaunchKernelGGL((k_ks_rows_f64<LOGRF, LOGRI, PSLOT, OWN>), dim3(nwg), dim3(F_TPB), 0, s, a, t.twf, t.itwf, t.mp,
                       t.logN, slot0, n_work);
}

template <int LOGRF, int LOGRI, bool PSLOT>
static void launch_ks_rows_inst(const DevTables &t, const KsRowsArgs &a, u32 slot0, u32 nslots, hipStream_t s) {
    if constexpr (!PSLOT) {
        if (a.tt && a.ta) {
            launch_ks_rows_own<LOGRF, LOGRI, false, 3>(t, a, slot0, nslots, s);
            return;
        }
        if (a.tt && a.gather) {
            launch_ks_rows_own<LOGRF, LOGRI, false, 4>(t, a, slot0, nslots, s);
            return;
        }
    }
    if (a.ta)
        launch_ks_rows_own<LOGRF, LOGRI, PSLOT, 1>(t, a, slot0, nslots, s);
    else if (a.gather)
        launch_ks_rows_own<LOGRF, LOGRI, PSLOT, 2>(t, a, slot0, nslots, s);
    else
        launch_ks_rows_own<LOGRF, LOGRI, PSLOT, 0>(t, a, slot0, nslots, s);
}

// second-pass lengths (stages along rows) of the current forward / inverse splits
static void rows_logs(const DevTables &t, u32 &lf, u32 &li) {
    const u32 lg = t.logN;
    lf = lg - (t.k1_fwd ? t.k1_fwd : lg / 2);
    li = lg - (t.k1_inv ? t.k1_inv : lg / 2);
}

bool ks_rows_fusable(const DevTables &t) {
    if (!(t.use_f64 || t.q_mixed_fusable) || (t.f64_off & 1) || t.logN <= F_LOG_TILE) return false;
    if ((t.logN == 13 || t.logN == 14) && (t.k1_fwd == t.logN || t.k1_inv == t.logN)) return false;
    u32 lf, li;
    rows_logs(t, lf, li);
    return lf >= 6 && lf <= 9 && li >= 6 && li <= 9;
}

void launch_ks_rows_f64(const DevTables &t, const KsRowsArgs &a, hipStream_t s) {
    u32 lf, li;
    rows_logs(t, lf, li);
    // the Q-slot kernel does not use LOGRI: one instance per forward split
#define FHEA_KS_ROWS(LF, LI)                                                  \
    if (lf == LF && li == LI) {                                               \
        if (!(t.f64_off & 128)) {                                             \
            launch_ks_rows_inst<LF, 8, false>(t, a, 0, a.l + a.K, s);         \
            return;                                                           \
        }                                                                     \
        launch_ks_rows_inst<LF, 8, false>(t, a, 0, a.l, s);                   \
        launch_ks_rows_inst<LF, LI, true>(t, a, a.l, a.K, s);                 \
        return;                                                               \
    }
#define FHEA_KS_ROWS_F(LF) FHEA_KS_ROWS(LF, 6) FHEA_KS_ROWS(LF, 7) FHEA_KS_ROWS(LF, 8) FHEA_KS_ROWS(LF, 9)
    FHEA_KS_ROWS_F(6)
    FHEA_KS_ROWS_F(7)
    FHEA_KS_ROWS_F(8)
    FHEA_KS_ROWS_F(9)
#undef FHEA_KS_ROWS_F
#undef FHEA_KS_ROWS
}

void launch_ks_rows_f64_early(const DevTables &t, const KsRowsArgs &a, u32 slot0, u32 nslots, hipStream_t s) {
    u32 lf, li;
    rows_logs(t, lf, li);
#define FHEA_KS_EARLY(LF, LI)                                                          \
    if (lf == LF && li == LI) {                                                        \
        launch_ks_rows_own<LF, LI, false, 5>(t, a, slot0, nslots, s);                  \
        return;                                                                        \
    }
#define FHEA_KS_EARLY_F(LF) FHEA_KS_EARLY(LF, 6) FHEA_KS_EARLY(LF, 7) FHEA_KS_EARLY(LF, 8) FHEA_KS_EARLY(LF, 9)
    FHEA_KS_EARLY_F(6)
    FHEA_KS_EARLY_F(7)
    FHEA_KS_EARLY_F(8)
    FHEA_KS_EARLY_F(9)
#undef FHEA_KS_EARLY_F
#undef FHEA_KS_EARLY
}

void launch_ks_rows_f64_q(const DevTables &t, const KsRowsArgs &a, u32 slot0, u32 nslots, hipStream_t s) {
    u32 lf, li;
    rows_logs(t, lf, li);
    switch (lf) {
    case 6: launch_ks_rows_inst<6, 8, false>(t, a, slot0, nslots, s); break;
    case 7: launch_ks_rows_inst<7, 8, false>(t, a, slot0, nslots, s); break;
    case 8: launch_ks_rows_inst<8, 8, false>(t, a, slot0, nslots, s); break;
    case 9: launch_ks_rows_inst<9, 8, false>(t, a, slot0, nslots, s); break;
    default: break;
    }
}

}  // namespace fhea
