// fhea_ntt.hip -- batched negacyclic NTT / INTT for gfx950.
//
// Computes exactly ntt_forward / intt_inverse of the reference
// (hardware_fpga_model/hls/src/ntt.cpp:4-80): Cooley-Tukey with W = tw[m+i]
// (natural in, bit-reversed out) and Gentleman-Sande with W = itw[m+i]
// followed by the N^-1 scaling.  Every butterfly reduces fully into [0, q),
// so the output words are identical to the reference's.
//
// Decomposition: a limb of N words is viewed as [N/(R*S)][R][S]; one pass runs
// log2(R) consecutive stages on the sub-transforms along R (stride S).  N <= 4096
// runs in one pass (the whole limb in LDS); larger N in two passes
// (k1 = log2(N)/2 stages over strided columns, then the rest over contiguous
// rows; the inverse runs rows first).  Each workgroup stages a 4096-word tile
// (32 KiB of LDS) made of several sub-transforms: C adjacent columns (so every
// global access is a run of C consecutive words) or several adjacent rows.
#include "fhea_f64.hpp"
#include "fhea_kernels.hpp"

namespace fhea {

static constexpr u32 NTT_TPB = 256;
static constexpr u32 NTT_LOG_TILE = 12;

template <bool INV>
__global__ __launch_bounds__(NTT_TPB) void k_ntt_pass(
    const u64 *src, u64 *dst, u64 src_stride, u64 dst_stride,
    u64 src_slot, u64 dst_slot, u32 limbs, const u32 *__restrict__ prime_map, const u64 *__restrict__ twtab,
    const u64 *__restrict__ ninv, const ModParams *__restrict__ mp, u32 logN, u32 logR, u32 logS,
    u32 logC, u32 logT, int scale_out) {
    extern __shared__ u64 lds[];
    const u32 slot = blockIdx.y % limbs;
    const u32 poly = blockIdx.y / limbs;
    const u32 pidx = prime_map[slot];
    if (pidx == 0xFFFFFFFFu) return;
    const u64 q = mp[pidx].q;
    const u32 N = 1u << logN, R = 1u << logR, C = 1u << logC, T = 1u << logT;
    const u64 *tw = twtab + (size_t)pidx * 2 * N;
    const u64 *s = src + (size_t)poly * src_stride + (size_t)slot * src_slot;
    u64 *d = dst + (size_t)poly * dst_stride + (size_t)slot * dst_slot;
    const u32 logNB = logT - logR - logC;
    const u32 ncg = 1u << (logS - logC);
    const u32 c0 = (blockIdx.x & (ncg - 1)) << logC;
    const u32 b0 = (blockIdx.x >> (logS - logC)) << logNB;

    for (u32 e = threadIdx.x; e < T; e += NTT_TPB) {
        u32 c = e & (C - 1), r = (e >> logC) & (R - 1), ob = e >> (logC + logR);
        u32 j = ((b0 + ob) << (logR + logS)) + (r << logS) + c0 + c;
        lds[e] = s[j];
    }
    __syncthreads();

    for (u32 ls = 0; ls < logR; ++ls) {
        const u32 logh = INV ? ls : (logR - 1 - ls);
        const u32 h = 1u << logh;
        const u32 logt = logh + logS;
        for (u32 bi = threadIdx.x; bi < (T >> 1); bi += NTT_TPB) {
            u32 c = bi & (C - 1);
            u32 rest = bi >> logC;
            u32 p = rest & ((R >> 1) - 1);
            u32 ob = rest >> (logR - 1);
            u32 r1 = ((p >> logh) << (logh + 1)) | (p & (h - 1));
            u32 e1 = (((ob << logR) | r1) << logC) | c;
            u32 e2 = e1 + (h << logC);
            u32 j1 = ((b0 + ob) << (logR + logS)) + (r1 << logS) + c0 + c;
            u32 widx = (N >> (logt + 1)) + (j1 >> (logt + 1));
            u64 w = tw[2 * widx], wsh = tw[2 * widx + 1];
            u64 u = lds[e1], v = lds[e2];
            if (!INV) {
                u64 x = mul_shoup(v, w, wsh, q);
                lds[e1] = add_mod(u, x, q);
                lds[e2] = sub_mod(u, x, q);
            } else {
                lds[e1] = add_mod(u, v, q);
                lds[e2] = mul_shoup(sub_mod(u, v, q), w, wsh, q);
            }
        }
        __syncthreads();
    }

    const u64 ni = scale_out ? ninv[2 * pidx] : 0, nish = scale_out ? ninv[2 * pidx + 1] : 0;
    for (u32 e = threadIdx.x; e < T; e += NTT_TPB) {
        u32 c = e & (C - 1), r = (e >> logC) & (R - 1), ob = e >> (logC + logR);
        u32 j = ((b0 + ob) << (logR + logS)) + (r << logS) + c0 + c;
        u64 v = lds[e];
        if (scale_out) v = mul_shoup(v, ni, nish, q);
        d[j] = v;
    }
}

// ---------------------------------------------------------------------------
// FP64 path (all primes < 2^50): same tile decomposition, but each thread keeps
// 16 elements of a sub-transform in registers and runs up to four radix-2
// stages on them per LDS round trip (a "round").  The 16 elements differ in a
// 4-bit window of the row index r; the stages of a round act on window bits.
// Twiddles come as (w, w/q) double pairs.  Values are signed residues:
//   forward:  each round starts at |x| <= q (canonical input) or q/2+1, grows
//             by <= 1.5q per stage (mulmod_f64 bound) to <= 7q, then every
//             value is centred with reduce_f64;
//   inverse:  the sum branch is centred every stage, so |x| <= 1.5q throughout.
// The intermediate between the two passes is stored as doubles (centred); the
// transform's final output is the canonical residue in [0, q).
// ---------------------------------------------------------------------------
static constexpr u32 F_TILE_LOG = 12;
static constexpr u32 F_TILE = 1u << F_TILE_LOG;

__device__ __forceinline__ u32 lds_pad(u32 e) { return e + (e >> 4); }

template <bool INV>
__global__ __launch_bounds__(NTT_TPB) void k_nttf_pass(
    const void *src, void *dst, u64 src_stride, u64 dst_stride, u64 src_slot, u64 dst_slot,
    u32 limbs, const u32 *__restrict__ prime_map, const double *__restrict__ twf,
    const double *__restrict__ ninvf, const ModParams *__restrict__ mp, u32 logN, u32 logR,
    u32 logS, u32 logC, u32 logT, int in_canon, int out_canon, int scale_out) {
    __shared__ double lds[F_TILE + F_TILE / 16];
    const u32 slot = blockIdx.y % limbs;
    const u32 poly = blockIdx.y / limbs;
    const u32 pidx = prime_map[slot];
    if (pidx == 0xFFFFFFFFu) return;
    const double q = (double)mp[pidx].q;
    const double qinv = 1.0 / q;
    const u32 N = 1u << logN, R = 1u << logR, C = 1u << logC, T = 1u << logT;
    const double *tw = twf + (size_t)pidx * 2 * N;
    const char *sbase = (const char *)src + 8 * ((size_t)poly * src_stride + (size_t)slot * src_slot);
    char *dbase = (char *)dst + 8 * ((size_t)poly * dst_stride + (size_t)slot * dst_slot);
    const u32 logNB = logT - logR - logC;
    const u32 ncg = 1u << (logS - logC);
    const u32 c0 = (blockIdx.x & (ncg - 1)) << logC;
    const u32 b0 = (blockIdx.x >> (logS - logC)) << logNB;
    const u32 tid = threadIdx.x;

    // coalesced tile load
    for (u32 e = tid; e < T; e += NTT_TPB) {
        u32 c = e & (C - 1), r = (e >> logC) & (R - 1), ob = e >> (logC + logR);
        u32 j = ((b0 + ob) << (logR + logS)) + (r << logS) + c0 + c;
        double v;
        if (in_canon) v = u64_to_f64(((const u64 *)sbase)[j]);
        else v = ((const double *)sbase)[j];
        lds[lds_pad(e)] = v;
    }
    __syncthreads();

    const u32 nrounds = (logR + 3) >> 2;
    const bool active = tid < (T >> 4);
    for (u32 rd = 0; rd < nrounds; ++rd) {
        const u32 ns = (rd + 1 < nrounds) ? 4 : (logR - 4 * rd);  // stages in this round
        // window of 4 r-bits [p, p+4) containing the round's stage bits
        u32 p, bit0;
        if (!INV) {
            const u32 top = logR - 1 - 4 * rd;
            p = top >= 3 ? top - 3 : 0;
            bit0 = top;  // bit of the first stage; later stages go down
        } else {
            p = (4 * rd + 4 <= logR) ? 4 * rd : logR - 4;
            bit0 = 4 * rd;  // later stages go up
        }
        const u32 P = logC + p;
        if (active) {
            const u32 ebase = (tid & ((1u << P) - 1)) | ((tid >> P) << (P + 4));
            double x[16];
#pragma unroll
            for (u32 u = 0; u < 16; ++u) x[u] = lds[lds_pad(ebase | (u << P))];
            // global index of element u=0 of this thread
            const u32 c = ebase & (C - 1), r = (ebase >> logC) & (R - 1), ob = ebase >> (logC + logR);
            const u32 j0 = ((b0 + ob) << (logR + logS)) + (r << logS) + c0 + c;
            for (u32 st = 0; st < ns; ++st) {
                const u32 b = INV ? bit0 + st : bit0 - st;  // r-bit of this stage
                const u32 wb = b - p;                          // bit inside the window
                const u32 k = b + logS + 1;                    // log2(2t)
                const u32 wbase = (N >> k) | (j0 >> k);
#pragma unroll
                for (u32 u = 0; u < 16; ++u) {
                    if (u & (1u << wb)) continue;
                    const u32 u2 = u | (1u << wb);
                    const u32 widx = wbase | (u >> (wb + 1));
                    const double w = tw[2 * widx], wq = tw[2 * widx + 1];
                    if (!INV) {
                        const double rr = mulmod_f64(x[u2], w, wq, q);
                        const double a = x[u];
                        x[u] = a + rr;
                        x[u2] = a - rr;
                    } else {
                        const double a = x[u], bb = x[u2];
                        x[u] = reduce_f64(a + bb, q, qinv);
                        x[u2] = mulmod_f64(a - bb, w, wq, q);
                    }
                }
            }
            if (!INV) {
#pragma unroll
                for (u32 u = 0; u < 16; ++u) x[u] = reduce_f64(x[u], q, qinv);
            }
#pragma unroll
            for (u32 u = 0; u < 16; ++u) lds[lds_pad(ebase | (u << P))] = x[u];
        }
        __syncthreads();
    }

    const double ni = scale_out ? ninvf[2 * pidx] : 0.0, niq = scale_out ? ninvf[2 * pidx + 1] : 0.0;
    for (u32 e = tid; e < T; e += NTT_TPB) {
        u32 c = e & (C - 1), r = (e >> logC) & (R - 1), ob = e >> (logC + logR);
        u32 j = ((b0 + ob) << (logR + logS)) + (r << logS) + c0 + c;
        double v = lds[lds_pad(e)];
        if (scale_out) v = mulmod_f64(v, ni, niq, q);
        if (INV) v = reduce_f64(v, q, qinv);
        if (out_canon) ((u64 *)dbase)[j] = f64_to_canon(v, q);
        else ((double *)dbase)[j] = v;
    }
}

static void passf(const DevTables &t, bool inv, const void *src, void *dst, u64 ss, u64 ds, u64 sl,
                  u64 dl, u32 npolys, u32 limbs, const u32 *pm, u32 logR, u32 logS, u32 logC,
                  u32 logT, int in_canon, int out_canon, int scale, hipStream_t s) {
    dim3 grid(1u << (t.logN - logT), npolys * limbs);
    if (inv)
        hipLaunchKernelGGL(k_nttf_pass<true>, grid, dim3(NTT_TPB), 0, s, src, dst, ss, ds, sl, dl, limbs,
                           pm, t.itwf, t.ninvf, t.mp, t.logN, logR, logS, logC, logT, in_canon,
                           out_canon, scale);
    else
        hipLaunchKernelGGL(k_nttf_pass<false>, grid, dim3(NTT_TPB), 0, s, src, dst, ss, ds, sl, dl,
                           limbs, pm, t.twf, t.ninvf, t.mp, t.logN, logR, logS, logC, logT, in_canon,
                           out_canon, scale);
}

static void launch_ntt_f64(const DevTables &t, bool inverse, const u64 *src, u64 *dst, u64 ss,
                           u64 ds, u64 sl, u64 dl, u32 npolys, u32 limbs, const u32 *pm,
                           hipStream_t s) {
    const u32 lg = t.logN;
    if (lg <= F_TILE_LOG) {
        passf(t, inverse, src, dst, ss, ds, sl, dl, npolys, limbs, pm, lg, 0, 0, lg, 1, 1,
              inverse ? 1 : 0, s);
        return;
    }
    const u32 k1 = lg / 2, k2 = lg - k1, lt = F_TILE_LOG;
    if (!inverse) {
        passf(t, false, src, dst, ss, ds, sl, dl, npolys, limbs, pm, k1, k2, lt - k1, lt, 1, 0, 0, s);
        passf(t, false, dst, dst, ds, ds, dl, dl, npolys, limbs, pm, k2, 0, 0, lt, 0, 1, 0, s);
    } else {
        passf(t, true, src, dst, ss, ds, sl, dl, npolys, limbs, pm, k2, 0, 0, lt, 1, 0, 0, s);
        passf(t, true, dst, dst, ds, ds, dl, dl, npolys, limbs, pm, k1, k2, lt - k1, lt, 0, 1, 1, s);
    }
}

static void pass(const DevTables &t, bool inv, const u64 *src, u64 *dst, u64 ss, u64 ds, u64 sl,
                 u64 dl, u32 npolys, u32 limbs, const u32 *pm, u32 logR, u32 logS, u32 logC,
                 u32 logT, int scale, hipStream_t s) {
    dim3 grid(1u << (t.logN - logT), npolys * limbs);
    size_t shm = (size_t)8 << logT;
    if (inv)
        hipLaunchKernelGGL(k_ntt_pass<true>, grid, dim3(NTT_TPB), shm, s, src, dst, ss, ds, sl, dl, limbs,
                           pm, t.itw, t.ninv, t.mp, t.logN, logR, logS, logC, logT, scale);
    else
        hipLaunchKernelGGL(k_ntt_pass<false>, grid, dim3(NTT_TPB), shm, s, src, dst, ss, ds, sl, dl,
                           limbs, pm, t.tw, t.ninv, t.mp, t.logN, logR, logS, logC, logT, scale);
}

void launch_ntt(const DevTables &t, bool inverse, const u64 *src, u64 *dst, u64 src_stride,
                u64 dst_stride, u32 npolys, u32 limbs, const u32 *prime_map, hipStream_t s,
                u64 src_slot, u64 dst_slot) {
    if (npolys == 0 || limbs == 0) return;
    const u32 lg = t.logN;
    const u64 sl = src_slot ? src_slot : t.N, dl = dst_slot ? dst_slot : t.N;
    if (t.use_f64) {
        launch_ntt_f64(t, inverse, src, dst, src_stride, dst_stride, sl, dl, npolys, limbs, prime_map, s);
        return;
    }
    if (lg <= NTT_LOG_TILE) {
        pass(t, inverse, src, dst, src_stride, dst_stride, sl, dl, npolys, limbs, prime_map, lg, 0, 0,
             lg, inverse ? 1 : 0, s);
        return;
    }
    const u32 k1 = lg / 2, k2 = lg - k1, lt = NTT_LOG_TILE;
    if (!inverse) {
        // stages 0..k1-1 over columns (stride 2^k2), then k1..lg-1 over rows
        pass(t, false, src, dst, src_stride, dst_stride, sl, dl, npolys, limbs, prime_map, k1, k2,
             lt - k1, lt, 0, s);
        pass(t, false, dst, dst, dst_stride, dst_stride, dl, dl, npolys, limbs, prime_map, k2, 0, 0, lt,
             0, s);
    } else {
        pass(t, true, src, dst, src_stride, dst_stride, sl, dl, npolys, limbs, prime_map, k2, 0, 0, lt,
             0, s);
        pass(t, true, dst, dst, dst_stride, dst_stride, dl, dl, npolys, limbs, prime_map, k1, k2,
             lt - k1, lt, 1, s);
    }
}

}  // namespace fhea
