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
    if (t.use_f64 && !(t.f64_off & 1)) {
        launch_ntt_f64(t, inverse, src, dst, src_stride, dst_stride, sl, dl, npolys, limbs, prime_map,
                       NttFuse(), s);
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
