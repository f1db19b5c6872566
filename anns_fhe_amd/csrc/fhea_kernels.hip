// fhea_kernels.hip -- elementwise, permutation, basis-conversion, key-switch
// inner-product, ModDown and rescale kernels for gfx950.
//
// All of these are streaming kernels over (limb x coefficient): one thread per
// 64-bit word, consecutive threads on consecutive coefficients, so every global
// access is coalesced (the automorphism gather aside).  Per-limb constants are
// wave-uniform and come through the scalar cache.
#include "fhea_f64.hpp"
#include "fhea_kernels.hpp"
#include "fhea_ntt_tile.hpp"

namespace fhea {

static constexpr u32 TPB = 256;

// ---------------------------------------------------------------------------
// poly_add / poly_sub / poly_mul (poly_arith.cpp:3-42), limb i mod prime i.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_binop(int op, const u64 *__restrict__ a,
                                               const u64 *__restrict__ b, u64 *__restrict__ o,
                                               const ModParams *__restrict__ mp, u32 logN) {
    const u32 limb = blockIdx.y;
    const size_t i = ((size_t)limb << logN) + (size_t)blockIdx.x * TPB + threadIdx.x;
    const ModParams m = mp[limb];
    u64 x = a[i], y = b[i], r;
    if (op == 0) r = add_mod(x, y, m.q);
    else if (op == 1) r = sub_mod(x, y, m.q);
    else r = mulmod(x, y, m);
    o[i] = r;
}

void launch_binop(const DevTables &t, int op, const u64 *a, const u64 *b, u64 *o, u32 limbs,
                  hipStream_t s) {
    dim3 grid(t.N / TPB, limbs);
    hipLaunchKernelGGL(k_binop, grid, dim3(TPB), 0, s, op, a, b, o, t.mp, t.logN);
}

// ---------------------------------------------------------------------------
// Tensor product (fhe_host_driver.cpp:280-286, four OP 5 + one OP 3 fused).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_tensor(const TensorArgs *__restrict__ args,
                                                const ModParams *__restrict__ mp, u32 logN) {
    const u32 limb = blockIdx.y;
    const TensorArgs A = args[blockIdx.z];
    const size_t i = ((size_t)limb << logN) + (size_t)blockIdx.x * TPB + threadIdx.x;
    const ModParams m = mp[limb];
    u64 a0 = A.a0[i], a1 = A.a1[i], b0 = A.b0[i], b1 = A.b1[i];
    A.d0[i] = mulmod(a0, b0, m);
    A.d2[i] = mulmod(a1, b1, m);
    A.d1[i] = reduce128((u128)a0 * b1 + (u128)a1 * b0, m);
}

// FP64 form (all primes < 2^50): products by mulmod_vv_f64, |a*b mod q| <= 0.875q.
__global__ __launch_bounds__(TPB) void k_tensor_f64(const TensorArgs *__restrict__ args,
                                                    const ModParams *__restrict__ mp, u32 logN) {
    const u32 limb = blockIdx.y;
    const TensorArgs A = args[blockIdx.z];
    const size_t i = ((size_t)limb << logN) + (size_t)blockIdx.x * TPB + threadIdx.x;
    const double q = mp[limb].qd, qinv = mp[limb].qinv;
    const double a1 = u64_to_f64(A.a1[i]), b1 = u64_to_f64(A.b1[i]);
    A.d2[i] = f64_to_canon(mulmod_vv_f64(a1, b1, q, qinv), q);
    if (!A.d0) return;  // d2 only (the key switch forms d0, d1 itself)
    const double a0 = u64_to_f64(A.a0[i]), b0 = u64_to_f64(A.b0[i]);
    A.d0[i] = f64_to_canon(mulmod_vv_f64(a0, b0, q, qinv), q);
    const double x = mulmod_vv_f64(a0, b1, q, qinv) + mulmod_vv_f64(a1, b0, q, qinv);
    A.d1[i] = f64_to_canon(reduce_f64(x, q, qinv), q);
}

void launch_tensor(const DevTables &t, const TensorArgs *args_dev, u32 nbatch, u32 limbs,
                   hipStream_t s) {
    dim3 grid(t.N / TPB, limbs, nbatch);
    if (t.use_f64 && !(t.f64_off & 8))
        hipLaunchKernelGGL(k_tensor_f64, grid, dim3(TPB), 0, s, args_dev, t.mp, t.logN);
    else
        hipLaunchKernelGGL(k_tensor, grid, dim3(TPB), 0, s, args_dev, t.mp, t.logN);
}

// ---------------------------------------------------------------------------
// Automorphism with an explicit packed map (automorphism.cpp:3-12, OP 6).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_automorph_map(const u64 *__restrict__ in,
                                                       u64 *__restrict__ out,
                                                       const u64 *__restrict__ pmap, u32 logN) {
    const u32 limb = blockIdx.y;
    const u32 i = blockIdx.x * TPB + threadIdx.x;
    u64 p = pmap[i >> 1];
    u32 idx = (i & 1) ? (u32)(p >> 32) : (u32)p;
    out[((size_t)limb << logN) + i] = in[((size_t)limb << logN) + idx];
}

void launch_automorph_map(const DevTables &t, const u64 *in, u64 *out, const u64 *packed_map,
                          u32 limbs, hipStream_t s) {
    dim3 grid(t.N / TPB, limbs);
    hipLaunchKernelGGL(k_automorph_map, grid, dim3(TPB), 0, s, in, out, packed_map, t.logN);
}

// ---------------------------------------------------------------------------
// sigma_g on c0 and c1 with the closed-form map
//   map[i] = br(((g (2 br(i) + 1)) mod 2N - 1) / 2)
// (the evaluation-domain image of a(X) -> a(X^g) in bit-reversed slot order).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(TPB) void k_automorph(const AutoArgs *__restrict__ args, u32 limbs,
                                                   u32 logN) {
    const AutoArgs A = args[blockIdx.z];
    const u32 y = blockIdx.y;  // [0, 2*limbs): c0 limbs then c1 limbs
    const u32 limb = y % limbs;
    const u32 i = blockIdx.x * TPB + threadIdx.x;
    const u32 src = galois_src(i, A.galois, logN);
    const u64 *in = (y < limbs) ? A.in0 : A.in1;
    u64 *out = (y < limbs) ? A.out0 : A.out1;
    out[((size_t)limb << logN) + i] = in[((size_t)limb << logN) + src];
}

void launch_automorph(const DevTables &t, const AutoArgs *args_dev, u32 nbatch, u32 limbs,
                      hipStream_t s, u32 npolys) {
    dim3 grid(t.N / TPB, npolys * limbs, nbatch);
    hipLaunchKernelGGL(k_automorph, grid, dim3(TPB), 0, s, args_dev, limbs, t.logN);
}

// ---------------------------------------------------------------------------
// Fast basis conversion (key_switch.cpp:7-60):
//   y_i   = x_i * qhat_i^-1 mod q_i
//   out_j = sum_i y_i * (qhat_i mod p_j) mod p_j
// The reference reduces y_i mod p_j before each product; the residue of the
// sum mod p_j is the same, so the products are accumulated in 128 bits and
// reduced once per output.
// ---------------------------------------------------------------------------
static constexpr u32 FBC_MAX_IN = 32;

__global__ __launch_bounds__(TPB) void k_fbc(const ConvDesc *__restrict__ descs,
                                             const u32 *__restrict__ in_first,
                                             const ModParams *__restrict__ mp,
                                             const u64 *__restrict__ in, u64 in_zstride,
                                             u64 *__restrict__ out, u64 out_pstride,
                                             u64 out_zstride, u32 logN) {
    const u32 p = blockIdx.y, z = blockIdx.z;
    const ConvDesc D = descs[p];
    const u32 r = blockIdx.x * TPB + threadIdx.x;
    const u64 *x = in + (size_t)z * in_zstride + ((size_t)in_first[p] << logN) + r;
    u64 *o = out + (size_t)z * out_zstride + (size_t)p * out_pstride + r;
    u64 y[FBC_MAX_IN];
#pragma unroll 4
    for (u32 i = 0; i < D.n_in; ++i) {
        const u64 q = mp[D.in_prime[i]].q;
        y[i] = mul_shoup(x[(size_t)i << logN], D.in_cst[2 * i], D.in_cst[2 * i + 1], q);
    }
    for (u32 j = 0; j < D.n_out; ++j) {
        u128 acc = 0;
        const u64 *h = D.hat + j;
#pragma unroll 4
        for (u32 i = 0; i < D.n_in; ++i) acc += (u128)y[i] * h[(size_t)i * D.n_out];
        o[(size_t)j << logN] = reduce128(acc, mp[D.out_prime[j]]);
    }
}

// Same arithmetic with the inputs held in registers (NIN >= n_in, fully unrolled)
// and the descriptor's constants in LDS: qhat_i mod p_j for every (i, j) and each
// output prime's (q, floor(2^128/q)) for the 128-bit reduction.
static constexpr u32 FBC_I_MAX_PAIRS = 2048;  // n_in * n_out staged (16 KiB)
static constexpr u32 FBC_I_MAX_OUT = 256;

template <u32 NIN>
__global__ __launch_bounds__(TPB) void k_fbc_i(const ConvDesc *__restrict__ descs,
                                               const u32 *__restrict__ in_first,
                                               const ModParams *__restrict__ mp,
                                               const u64 *__restrict__ in, u64 in_zstride,
                                               u64 *__restrict__ out, u64 out_pstride,
                                               u64 out_zstride, u32 logN) {
    __shared__ u64 s_hat[FBC_I_MAX_PAIRS];  // [i][j]
    __shared__ u64 s_q[FBC_I_MAX_OUT], s_rlo[FBC_I_MAX_OUT], s_rhi[FBC_I_MAX_OUT];
    const u32 p = blockIdx.y, z = blockIdx.z;
    const ConvDesc D = descs[p];
    const u32 n_in = D.n_in, n_out = D.n_out;
    for (u32 k = threadIdx.x; k < n_in * n_out; k += TPB) s_hat[k] = D.hat[k];
    for (u32 j = threadIdx.x; j < n_out; j += TPB) {
        const ModParams &m = mp[D.out_prime[j]];
        s_q[j] = m.q;
        s_rlo[j] = m.r_lo;
        s_rhi[j] = m.r_hi;
    }
    const u32 r = blockIdx.x * TPB + threadIdx.x;
    const u64 *x = in + (size_t)z * in_zstride + ((size_t)in_first[p] << logN) + r;
    u64 *o = out + (size_t)z * out_zstride + (size_t)p * out_pstride + r;
    u64 y[NIN];
#pragma unroll
    for (u32 i = 0; i < NIN; ++i) {
        y[i] = 0;
        if (i < n_in) y[i] = mul_shoup(x[(size_t)i << logN], D.in_cst[2 * i], D.in_cst[2 * i + 1], mp[D.in_prime[i]].q);
    }
    __syncthreads();
    for (u32 j = 0; j < n_out; ++j) {
        u128 acc = 0;
#pragma unroll
        for (u32 i = 0; i < NIN; ++i)
            if (i < n_in) acc += (u128)y[i] * s_hat[i * n_out + j];
        ModParams m;
        m.q = s_q[j];
        m.r_lo = s_rlo[j];
        m.r_hi = s_rhi[j];
        o[(size_t)j << logN] = reduce128(acc, m);
    }
}

// FP64 form: y_i = [x_i * qhat_i^-1]_{q_i} in [0, q_i) (x_i, qhat_i^-1 < q_i, so
// the centred product lies in (-q_i, q_i) and one conditional add makes it
// canonical; y_i < q_i < 4 p_j because max(q)/min(q) < 4), each term
// y_i * (qhat_i mod p_j) by mulmod_f64 (|term| <= p_j/2 + y_i/4 < 1.5 p_j), the running sum
// centred every four terms (<= 6.5 p_j).
static constexpr u32 FBC_F64_MAX_IN = 8;
static constexpr u32 FBC_F64_MAX_PAIRS = 1024;  // n_in * n_out staged in LDS (16 KiB)

__global__ __launch_bounds__(TPB) void k_fbc_f64(const ConvDesc *__restrict__ descs,
                                                 const u32 *__restrict__ in_first,
                                                 const ModParams *__restrict__ mp,
                                                 const u64 *__restrict__ in, u64 in_zstride,
                                                 u64 *__restrict__ out, u64 out_pstride,
                                                 u64 out_zstride, u32 logN) {
    __shared__ double2 s_hat[FBC_F64_MAX_PAIRS];  // (qhat_i mod p_j, same / p_j), [i][j]
    __shared__ double2 s_p[FBC_F64_MAX_PAIRS / FBC_F64_MAX_IN];  // (p_j, 1/p_j)
    const u32 p = blockIdx.y, z = blockIdx.z;
    const ConvDesc D = descs[p];
    const u32 n_in = D.n_in, n_out = D.n_out;
    for (u32 k = threadIdx.x; k < n_in * n_out; k += TPB)
        s_hat[k] = make_double2(D.hatf[2 * k], D.hatf[2 * k + 1]);
    for (u32 j = threadIdx.x; j < n_out; j += TPB) {
        const ModParams &m = mp[D.out_prime[j]];
        s_p[j] = make_double2(m.qd, m.qinv);
    }
    const u32 r = blockIdx.x * TPB + threadIdx.x;
    const u64 *x = in + (size_t)z * in_zstride + ((size_t)in_first[p] << logN) + r;
    u64 *o = out + (size_t)z * out_zstride + (size_t)p * out_pstride + r;
    double y[FBC_F64_MAX_IN];
#pragma unroll
    for (u32 i = 0; i < FBC_F64_MAX_IN; ++i) {
        if (i < n_in) {
            // the conversion needs the integer [x_i qhat_i^-1]_{q_i} in [0, q_i), not just
            // its residue: the centred product (|.| < q_i) is moved into [0, q_i)
            const double q = mp[D.in_prime[i]].qd;
            const double t = mulmod_f64(u64_to_f64(x[(size_t)i << logN]), D.in_cstf[2 * i], D.in_cstf[2 * i + 1], q);
            y[i] = t < 0.0 ? t + q : t;
        } else {
            y[i] = 0.0;
        }
    }
    __syncthreads();
    for (u32 j = 0; j < n_out; ++j) {
        const double2 pp = s_p[j];
        double acc = 0.0;
#pragma unroll
        for (u32 i = 0; i < FBC_F64_MAX_IN; ++i) {
            if (i < n_in) {
                const double2 h = s_hat[i * n_out + j];
                acc += mulmod_f64(y[i], h.x, h.y, pp.x);
                if ((i & 3) == 3) acc = reduce_f64(acc, pp.x, pp.y);
            }
        }
        o[(size_t)j << logN] = f64_to_canon(reduce_f64(acc, pp.x, pp.y), pp.x);
    }
}

void launch_fbc(const DevTables &t, const ConvDesc *descs_dev, const u32 *in_first_dev, u32 ndesc,
                u32 max_nin, u32 max_nout, const u64 *in, u64 in_zstride, u64 *out, u64 out_pstride,
                u64 out_zstride, u32 nbatch, hipStream_t s) {
    dim3 grid(t.N / TPB, ndesc, nbatch);
    if (t.use_f64 && !(t.f64_off & 2) && max_nin <= FBC_F64_MAX_IN && max_nin * max_nout <= FBC_F64_MAX_PAIRS &&
        max_nout <= FBC_F64_MAX_PAIRS / FBC_F64_MAX_IN)
        hipLaunchKernelGGL(k_fbc_f64, grid, dim3(TPB), 0, s, descs_dev, in_first_dev, t.mp, in,
                           in_zstride, out, out_pstride, out_zstride, t.logN);
    else if (max_nin * max_nout <= FBC_I_MAX_PAIRS && max_nout <= FBC_I_MAX_OUT && max_nin <= 32) {
#define FHEA_FBC_I(NI)                                                                                      \
    hipLaunchKernelGGL(k_fbc_i<NI>, grid, dim3(TPB), 0, s, descs_dev, in_first_dev, t.mp, in, in_zstride, out, \
                       out_pstride, out_zstride, t.logN)
        if (max_nin <= 8)
            FHEA_FBC_I(8);
        else if (max_nin <= 16)
            FHEA_FBC_I(16);
        else
            FHEA_FBC_I(32);
#undef FHEA_FBC_I
    } else
        hipLaunchKernelGGL(k_fbc, grid, dim3(TPB), 0, s, descs_dev, in_first_dev, t.mp, in, in_zstride,
                           out, out_pstride, out_zstride, t.logN);
}

// ---------------------------------------------------------------------------
// Key-switch inner product (key_switch.cpp:332-381).  For QlP slot i of the
// extended basis: acc0_i = sum_p digit_p[i] * evk_b[p][i],
// acc1_i = sum_p digit_p[i] * evk_a[p][i], where digit_p[i] is the input limb
// itself when i lies in digit p and the converted limb otherwise.  Products
// accumulate in 128 bits (< nparts * q^2) and are reduced once.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_ks_inner(const KsIpArgs *__restrict__ args,
                                                  const u64 *__restrict__ compact, u64 cz,
                                                  u64 *__restrict__ acc, u64 az,
                                                  const ModParams *__restrict__ mp, u32 logN, u32 l,
                                                  u32 K, u32 L, u32 alpha, u32 nparts,
                                                  u32 key_limbs, u32 dnum, u32 slot0) {
    const u32 i = slot0 + blockIdx.y;  // QlP slot
    const u32 z = blockIdx.z;
    const u32 r = blockIdx.x * TPB + threadIdx.x;
    const KsIpArgs A = args[z];
    const u32 pidx = (i < l) ? i : (L + i - l);
    const u32 kl = (i < l) ? i : (key_limbs + i - l);
    const u64 kstride = (u64)(key_limbs + K) << logN;
    const u64 *ka = A.key + ((u64)kl << logN) + r;
    const u64 *kb = ka + (u64)dnum * kstride;
    const u64 *cz_base = compact + (size_t)z * cz + r;
    u128 s0 = 0, s1 = 0;
    for (u32 p = 0; p < nparts; ++p) {
        const u32 start = p * alpha;
        const u32 sz = (p == nparts - 1) ? (l - start) : alpha;
        u64 d;
        if (i >= start && i < start + sz) {
            d = A.c[((size_t)i << logN) + r];
        } else {
            const u32 pos = (i < start) ? i : (i - sz);
            d = cz_base[((size_t)p * (l + K) + pos) << logN];
        }
        const u64 b = kb[(size_t)p * kstride], a = ka[(size_t)p * kstride];
        s0 += (u128)d * b;
        s1 += (u128)d * a;
    }
    const ModParams m = mp[pidx];
    u64 *o = acc + (size_t)z * az + ((size_t)i << logN) + r;
    o[0] = reduce128(s0, m);
    o[(size_t)(l + K) << logN] = reduce128(s1, m);
}

// FP64 form with a stationary key: block (x, i, g) handles key slot i for the
// `gs` consecutive items of group g, which share one key (a relinearisation
// batch: one group; per-item rotation keys: gs = 1).  The key words are loaded
// once per group.  Terms are mulmod_vv_f64 products (<= 0.875q each), at most
// KS_F64_MAX_PARTS of them per sum (<= 7.2q), centred once.
static constexpr u32 KS_F64_MAX_PARTS = 8;

__global__ __launch_bounds__(TPB) void k_ks_inner_f64(const KsIpArgs *__restrict__ args,
                                                      const u64 *__restrict__ compact, u64 cz,
                                                      u64 *__restrict__ acc, u64 az,
                                                      const ModParams *__restrict__ mp, u32 logN,
                                                      u32 l, u32 K, u32 L, u32 alpha, u32 nparts,
                                                      u32 key_limbs, u32 dnum, u32 gs, u32 nb) {
    const u32 i = blockIdx.y;
    const u32 z0 = blockIdx.z * gs, z1 = min(z0 + gs, nb);
    const u32 r = blockIdx.x * TPB + threadIdx.x;
    const u32 pidx = (i < l) ? i : (L + i - l);
    const u32 kl = (i < l) ? i : (key_limbs + i - l);
    const double q = mp[pidx].qd, qinv = mp[pidx].qinv;
    const u64 kstride = (u64)(key_limbs + K) << logN;
    const u64 *ka = args[z0].key + ((u64)kl << logN) + r;
    const u64 *kb = ka + (u64)dnum * kstride;
    double va[KS_F64_MAX_PARTS], vb[KS_F64_MAX_PARTS];
#pragma unroll
    for (u32 p = 0; p < KS_F64_MAX_PARTS; ++p) {
        va[p] = p < nparts ? u64_to_f64(ka[(size_t)p * kstride]) : 0.0;
        vb[p] = p < nparts ? u64_to_f64(kb[(size_t)p * kstride]) : 0.0;
    }
    for (u32 z = z0; z < z1; ++z) {
        const u64 *c = args[z].c;
        const u64 *cb = compact + (size_t)z * cz + r;
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (u32 p = 0; p < KS_F64_MAX_PARTS; ++p) {
            if (p < nparts) {
                const u32 start = p * alpha;
                const u32 sz = (p == nparts - 1) ? (l - start) : alpha;
                u64 dw;
                if (i >= start && i < start + sz) dw = c[((size_t)i << logN) + r];
                else dw = cb[((size_t)p * (l + K) + ((i < start) ? i : (i - sz))) << logN];
                const double d = u64_to_f64(dw);
                s0 += mulmod_vv_f64(d, vb[p], q, qinv);
                s1 += mulmod_vv_f64(d, va[p], q, qinv);
            }
        }
        u64 *o = acc + (size_t)z * az + ((size_t)i << logN) + r;
        o[0] = f64_to_canon(reduce_f64(s0, q, qinv), q);
        o[(size_t)(l + K) << logN] = f64_to_canon(reduce_f64(s1, q, qinv), q);
    }
}

void launch_ks_inner(const DevTables &t, const KsIpArgs *args_dev, const u64 *compact,
                     u64 compact_zstride, u64 *acc, u64 acc_zstride, u32 nbatch, u32 l, u32 alpha,
                     u32 nparts, u32 key_limbs, u32 dnum, bool shared_key, hipStream_t s, u32 slot0,
                     u32 nslots) {
    if (nslots) {
        // a slot range on the integer kernel (mixed chains: the slots of primes >= 2^50)
        dim3 grid(t.N / TPB, nslots, nbatch);
        hipLaunchKernelGGL(k_ks_inner, grid, dim3(TPB), 0, s, args_dev, compact, compact_zstride, acc,
                           acc_zstride, t.mp, t.logN, l, t.K, t.L, alpha, nparts, key_limbs, dnum, slot0);
        return;
    }
    if (t.use_f64 && !(t.f64_off & 4) && nparts <= KS_F64_MAX_PARTS) {
        const u32 gs = shared_key ? nbatch : 1;
        dim3 grid(t.N / TPB, l + t.K, (nbatch + gs - 1) / gs);
        hipLaunchKernelGGL(k_ks_inner_f64, grid, dim3(TPB), 0, s, args_dev, compact, compact_zstride,
                           acc, acc_zstride, t.mp, t.logN, l, t.K, t.L, alpha, nparts, key_limbs, dnum,
                           gs, nbatch);
        return;
    }
    dim3 grid(t.N / TPB, l + t.K, nbatch);
    hipLaunchKernelGGL(k_ks_inner, grid, dim3(TPB), 0, s, args_dev, compact, compact_zstride, acc,
                       acc_zstride, t.mp, t.logN, l, t.K, t.L, alpha, nparts, key_limbs, dnum, 0u);
}

// ---------------------------------------------------------------------------
// ModDown combine (key_switch.cpp:253-273), optionally fused with the
// composite's following poly_add (fhe_host_driver.cpp:289-290).
// acc layout per batch item: [2][l+K][N] (pstride between the two polys).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_moddown_combine(
    const CombineArgs *__restrict__ args, const u64 *__restrict__ acc, u64 az, u64 ap,
    const u64 *__restrict__ tt, u64 tz, u64 tp, const u64 *__restrict__ pinv,
    const ModParams *__restrict__ mp, u32 logN, u32 l) {
    const u32 y = blockIdx.y;  // [0, 2l)
    const u32 poly = y / l, i = y % l, z = blockIdx.z;
    const u32 r = blockIdx.x * TPB + threadIdx.x;
    const CombineArgs A = args[z];
    const u64 q = mp[i].q;
    const size_t off = ((size_t)i << logN) + r;
    u64 x = acc[(size_t)z * az + poly * ap + off];
    u64 v = tt[(size_t)z * tz + poly * tp + off];
    u64 res = mul_shoup(sub_mod(x, v, q), pinv[2 * i], pinv[2 * i + 1], q);
    const u64 *add = poly ? A.add1 : A.add0;
    if (add) res = add_mod(add[off], res, q);
    (poly ? A.out1 : A.out0)[off] = res;
}

__global__ __launch_bounds__(TPB) void k_moddown_combine_f64(
    const CombineArgs *__restrict__ args, const u64 *__restrict__ acc, u64 az, u64 ap,
    const u64 *__restrict__ tt, u64 tz, u64 tp, const double *__restrict__ pinvf,
    const ModParams *__restrict__ mp, u32 logN, u32 l) {
    const u32 y = blockIdx.y;
    const u32 poly = y / l, i = y % l, z = blockIdx.z;
    const u32 r = blockIdx.x * TPB + threadIdx.x;
    const CombineArgs A = args[z];
    const double q = mp[i].qd, qinv = mp[i].qinv;
    const size_t off = ((size_t)i << logN) + r;
    const double x = u64_to_f64(acc[(size_t)z * az + poly * ap + off]);
    const double v = u64_to_f64(tt[(size_t)z * tz + poly * tp + off]);
    double res = mulmod_f64(x - v, pinvf[2 * i], pinvf[2 * i + 1], q);  // |x-v| < q
    const u64 *add = poly ? A.add1 : A.add0;
    if (add) res += u64_to_f64(add[off]);
    (poly ? A.out1 : A.out0)[off] = f64_to_canon(reduce_f64(res, q, qinv), q);
}

void launch_moddown_combine(const DevTables &t, const CombineArgs *args_dev, const u64 *acc,
                            u64 acc_zstride, u64 acc_pstride, const u64 *tt, u64 tt_zstride,
                            u64 tt_pstride, const u64 *pinv, const double *pinvf, u32 nbatch, u32 l,
                            hipStream_t s) {
    dim3 grid(t.N / TPB, 2 * l, nbatch);
    if (t.use_f64 && !(t.f64_off & 32))
        hipLaunchKernelGGL(k_moddown_combine_f64, grid, dim3(TPB), 0, s, args_dev, acc, acc_zstride,
                           acc_pstride, tt, tt_zstride, tt_pstride, pinvf, t.mp, t.logN, l);
    else
        hipLaunchKernelGGL(k_moddown_combine, grid, dim3(TPB), 0, s, args_dev, acc, acc_zstride,
                           acc_pstride, tt, tt_zstride, tt_pstride, pinv, t.mp, t.logN, l);
}

// ---------------------------------------------------------------------------
// Rescale (rescale.cpp:43-70).  With v the coefficient-domain last limb and
// T_i = QlQlInvModqlDivqlModq[i] (== -q_l^-1 mod q_i):
//   prep:    tmp_i = centred(v) * T_i mod q_i
//          = v * T_i - [v > floor(q_l/2)] * (q_l * T_i mod q_i)
//   combine: out_i = a_i * q_l^-1 + NTT_i(tmp_i)  mod q_i
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_rescale_prep(const u64 *__restrict__ v, u64 vz,
                                                      u64 *__restrict__ tmp, u64 tz,
                                                      const u64 *__restrict__ qlql,
                                                      const u64 *__restrict__ sub,
                                                      const ModParams *__restrict__ mp, u32 logN,
                                                      u32 l) {
    const u32 i = blockIdx.y, z = blockIdx.z;
    const u32 r = blockIdx.x * TPB + threadIdx.x;
    const u64 q = mp[i].q, ql = mp[l - 1].q;
    const u64 x = v[(size_t)z * vz + r];
    u64 t = mul_shoup(x, qlql[2 * i], qlql[2 * i + 1], q);
    if (x > (ql >> 1)) t = sub_mod(t, sub[i], q);
    tmp[(size_t)z * tz + ((size_t)i << logN) + r] = t;
}

// FP64 form: c = centred(v) with |c| <= q_l/2 < q_i, tmp = c * T_i by mulmod_f64.
__global__ __launch_bounds__(TPB) void k_rescale_prep_f64(const u64 *__restrict__ v, u64 vz,
                                                          u64 *__restrict__ tmp, u64 tz,
                                                          const double *__restrict__ qlqlf,
                                                          const ModParams *__restrict__ mp,
                                                          u32 logN, u32 l) {
    const u32 i = blockIdx.y, z = blockIdx.z;
    const u32 r = blockIdx.x * TPB + threadIdx.x;
    const u64 qlw = mp[l - 1].q;
    const double q = mp[i].qd, qinv = mp[i].qinv;
    const u64 x = v[(size_t)z * vz + r];
    const double c = (x > (qlw >> 1)) ? u64_to_f64(x) - (double)qlw : u64_to_f64(x);
    const double t = mulmod_f64(c, qlqlf[2 * i], qlqlf[2 * i + 1], q);
    tmp[(size_t)z * tz + ((size_t)i << logN) + r] = f64_to_canon(reduce_f64(t, q, qinv), q);
}

void launch_rescale_prep(const DevTables &t, const u64 *v, u64 v_stride, u64 *tmp, u64 tmp_stride,
                         const u64 *qlql, const u64 *sub, const double *qlqlf, u32 npolys, u32 l,
                         hipStream_t s, u32 nlimbs) {
    dim3 grid(t.N / TPB, nlimbs ? nlimbs : l - 1, npolys);
    if (t.use_f64 && !(t.f64_off & 32))
        hipLaunchKernelGGL(k_rescale_prep_f64, grid, dim3(TPB), 0, s, v, v_stride, tmp, tmp_stride,
                           qlqlf, t.mp, t.logN, l);
    else
        hipLaunchKernelGGL(k_rescale_prep, grid, dim3(TPB), 0, s, v, v_stride, tmp, tmp_stride, qlql,
                           sub, t.mp, t.logN, l);
}

__global__ __launch_bounds__(TPB) void k_rescale_combine(const RescaleOut *__restrict__ args,
                                                         const u64 *__restrict__ tmp, u64 tz,
                                                         const u64 *__restrict__ qlinv,
                                                         const ModParams *__restrict__ mp,
                                                         u32 logN) {
    const u32 i = blockIdx.y, z = blockIdx.z;
    const u32 r = blockIdx.x * TPB + threadIdx.x;
    const RescaleOut A = args[z];
    const u64 q = mp[i].q;
    const size_t off = ((size_t)i << logN) + r;
    u64 a = mul_shoup(A.a[off], qlinv[2 * i], qlinv[2 * i + 1], q);
    A.out[off] = add_mod(a, tmp[(size_t)z * tz + off], q);
}

__global__ __launch_bounds__(TPB) void k_rescale_combine_f64(const RescaleOut *__restrict__ args,
                                                             const u64 *__restrict__ tmp, u64 tz,
                                                             const double *__restrict__ qlinvf,
                                                             const ModParams *__restrict__ mp,
                                                             u32 logN) {
    const u32 i = blockIdx.y, z = blockIdx.z;
    const u32 r = blockIdx.x * TPB + threadIdx.x;
    const RescaleOut A = args[z];
    const double q = mp[i].qd, qinv = mp[i].qinv;
    const size_t off = ((size_t)i << logN) + r;
    const double a = mulmod_f64(u64_to_f64(A.a[off]), qlinvf[2 * i], qlinvf[2 * i + 1], q);
    const double x = a + u64_to_f64(tmp[(size_t)z * tz + off]);
    A.out[off] = f64_to_canon(reduce_f64(x, q, qinv), q);
}

void launch_rescale_combine(const DevTables &t, const RescaleOut *args_dev, const u64 *tmp,
                            u64 tmp_stride, const u64 *qlinv, const double *qlinvf, u32 npolys, u32 l,
                            hipStream_t s, u32 nlimbs) {
    dim3 grid(t.N / TPB, nlimbs ? nlimbs : l - 1, npolys);
    if (t.use_f64 && !(t.f64_off & 32))
        hipLaunchKernelGGL(k_rescale_combine_f64, grid, dim3(TPB), 0, s, args_dev, tmp, tmp_stride,
                           qlinvf, t.mp, t.logN);
    else
        hipLaunchKernelGGL(k_rescale_combine, grid, dim3(TPB), 0, s, args_dev, tmp, tmp_stride, qlinv,
                           t.mp, t.logN);
}

// ---------------------------------------------------------------------------
// Hoisted rotations (no reference counterpart: the reference permutes, then key
// switches each rotation, fhe_host_driver.cpp:341-358).  The ModUp digits of one
// ciphertext are computed once; rotation z reads them through its Galois
// permutation (sigma_g acts on every limb of the extended basis by the same
// evaluation-domain permutation) and multiplies with its own key.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_ks_inner_rot(const KsRotArgs *__restrict__ args,
                                                      u64 *__restrict__ acc,
                                                      u64 az, const ModParams *__restrict__ mp, u32 logN, u32 l,
                                                      u32 K, u32 L, u32 alpha, u32 nparts, u32 key_limbs,
                                                      u32 dnum) {
    const u32 i = blockIdx.y, z = blockIdx.z;
    const u32 r = blockIdx.x * TPB + threadIdx.x;
    const KsRotArgs A = args[z];
    const u32 src = galois_src(r, A.galois, logN);
    const u32 pidx = (i < l) ? i : (L + i - l);
    const u32 kl = (i < l) ? i : (key_limbs + i - l);
    const u64 kstride = (u64)(key_limbs + K) << logN;
    const u64 *ka = A.key + ((u64)kl << logN) + r;
    const u64 *kb = ka + (u64)dnum * kstride;
    u128 s0 = 0, s1 = 0;
    for (u32 p = 0; p < nparts; ++p) {
        const u32 start = p * alpha;
        const u32 sz = (p == nparts - 1) ? (l - start) : alpha;
        u64 d;
        if (i >= start && i < start + sz)
            d = A.c[((size_t)i << logN) + src];
        else
            d = A.compact[(((size_t)p * (l + K) + ((i < start) ? i : (i - sz))) << logN) + src];
        s0 += (u128)d * kb[(size_t)p * kstride];
        s1 += (u128)d * ka[(size_t)p * kstride];
    }
    const ModParams m = mp[pidx];
    u64 *o = acc + (size_t)z * az + ((size_t)i << logN) + r;
    o[0] = reduce128(s0, m);
    o[(size_t)(l + K) << logN] = reduce128(s1, m);
}

__global__ __launch_bounds__(TPB) void k_ks_inner_rot_f64(const KsRotArgs *__restrict__ args,
                                                          u64 *__restrict__ acc, u64 az,
                                                          const ModParams *__restrict__ mp, u32 logN, u32 l,
                                                          u32 K, u32 L, u32 alpha, u32 nparts, u32 key_limbs,
                                                          u32 dnum) {
    const u32 i = blockIdx.y, z = blockIdx.z;
    const u32 r = blockIdx.x * TPB + threadIdx.x;
    const KsRotArgs A = args[z];
    const u32 src = galois_src(r, A.galois, logN);
    const u32 pidx = (i < l) ? i : (L + i - l);
    const u32 kl = (i < l) ? i : (key_limbs + i - l);
    const double q = mp[pidx].qd, qinv = mp[pidx].qinv;
    const u64 kstride = (u64)(key_limbs + K) << logN;
    const u64 *ka = A.key + ((u64)kl << logN) + r;
    const u64 *kb = ka + (u64)dnum * kstride;
    double s0 = 0.0, s1 = 0.0;  // <= KS_F64_MAX_PARTS terms of <= 0.875q
    for (u32 p = 0; p < nparts; ++p) {
        const u32 start = p * alpha;
        const u32 sz = (p == nparts - 1) ? (l - start) : alpha;
        u64 dw;
        if (i >= start && i < start + sz)
            dw = A.c[((size_t)i << logN) + src];
        else
            dw = A.compact[(((size_t)p * (l + K) + ((i < start) ? i : (i - sz))) << logN) + src];
        const double d = u64_to_f64(dw);
        s0 += mulmod_vv_f64(d, u64_to_f64(kb[(size_t)p * kstride]), q, qinv);
        s1 += mulmod_vv_f64(d, u64_to_f64(ka[(size_t)p * kstride]), q, qinv);
    }
    u64 *o = acc + (size_t)z * az + ((size_t)i << logN) + r;
    o[0] = f64_to_canon(reduce_f64(s0, q, qinv), q);
    o[(size_t)(l + K) << logN] = f64_to_canon(reduce_f64(s1, q, qinv), q);
}

void launch_ks_inner_rot(const DevTables &t, const KsRotArgs *args_dev, u64 *acc, u64 acc_zstride, u32 nbatch,
                         u32 l, u32 alpha, u32 nparts, u32 key_limbs, u32 dnum, hipStream_t s) {
    dim3 grid(t.N / TPB, l + t.K, nbatch);
    if (t.use_f64 && !(t.f64_off & 4) && nparts <= KS_F64_MAX_PARTS)
        hipLaunchKernelGGL(k_ks_inner_rot_f64, grid, dim3(TPB), 0, s, args_dev, acc, acc_zstride, t.mp, t.logN, l,
                           t.K, t.L, alpha, nparts, key_limbs, dnum);
    else
        hipLaunchKernelGGL(k_ks_inner_rot, grid, dim3(TPB), 0, s, args_dev, acc, acc_zstride, t.mp, t.logN, l, t.K,
                           t.L, alpha, nparts, key_limbs, dnum);
}

// ---------------------------------------------------------------------------
// Batched ciphertext addition (OP 3 on both polynomials, fhe_host_driver.cpp:248-256).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_add_batch(const AddArgs *__restrict__ args,
                                                   const ModParams *__restrict__ mp, u32 logN) {
    const u32 limb = blockIdx.y;
    const AddArgs A = args[blockIdx.z];
    const size_t i = ((size_t)limb << logN) + (size_t)blockIdx.x * TPB + threadIdx.x;
    const u64 q = mp[limb].q;
    A.d0[i] = add_mod(A.a0[i], A.b0[i], q);
    A.d1[i] = add_mod(A.a1[i], A.b1[i], q);
}

void launch_add_batch(const DevTables &t, const AddArgs *args_dev, u32 nbatch, u32 limbs, hipStream_t s) {
    dim3 grid(t.N / TPB, limbs, nbatch);
    hipLaunchKernelGGL(k_add_batch, grid, dim3(TPB), 0, s, args_dev, t.mp, t.logN);
}

// ---------------------------------------------------------------------------
// Diagonal-method inner products: out_{q,g} = bias_q + ptb_g + sum_r ct_{q,r} (.) pt_{g,r}.
// The reference evaluates each candidate batch with its own rotate-and-sum
// (interactive_query_manager.cpp:190-274); with each query's rotations computed
// once, every (query, candidate block) pair is one plaintext multiply-accumulate
// over r.  A group is a set of queries that all need the same blocks (the queries
// probing one cluster, say); a unit is a rectangle of a group's queries x blocks.  FP64 path: a 512-thread workgroup covers 64 coefficient positions of a
// unit of up to 128 pairs; per r it loads each of the unit's 2*uq + ug words once
// per position into LDS (double-buffered, the next r's loads in flight during this
// r's products) and each wave accumulates 16 pairs, so a word read from HBM feeds
// 2*uq*ug/(2*uq + ug) products (6.9 for 9 queries x 14 blocks).  Integer path
// (primes of 50 bits and more): one query x 8 blocks per thread.  Every unit of a call is one
// slice of the grid, so one launch fills the GPU.
// ---------------------------------------------------------------------------

// an absent operand (LIN_NONE) reads word idx of the arena and yields 0, so that
// the loads of a prologue issue back to back without branches
__device__ __forceinline__ u64 lin_word(const u64 *P, u64 off, size_t idx) {
    const u64 v = P[(off == LIN_NONE ? 0 : off) + idx];
    return off == LIN_NONE ? 0 : v;
}

// A unit's offset table is uniform and fixed for the launch: read through the
// constant address space it comes in by scalar loads, which neither occupy vector
// registers nor make the vector loads behind them wait.
typedef __attribute__((address_space(4))) const u64 CU64;

struct LinView {
    const CU64 *ct0, *ct1, *b0, *b1, *pt, *ptb, *o0, *o1;
    __device__ LinView(const LinUnit &U, u32 R) {
        ct0 = (const CU64 *)U.offs;
        ct1 = ct0 + (size_t)U.nq * R;
        b0 = ct1 + (size_t)U.nq * R;
        b1 = b0 + U.nq;
        pt = b1 + U.nq;
        ptb = pt + (size_t)U.ng * R;
        o0 = ptb + U.ng;
        o1 = o0 + (size_t)U.nq * U.ng;
    }
};

// Integer form: 128-bit lazy sums, reduced every 32 products (< 2^127).
template <int QG, int GG>
__global__ __launch_bounds__(TPB) void k_linear_sum(u64 *__restrict__ P, const LinUnit *__restrict__ units, u32 R,
                                                    const ModParams *__restrict__ mp, u32 logN) {
    const LinUnit U = units[blockIdx.z];
    const LinView V(U, R);
    const u32 limb = blockIdx.y;
    const size_t idx = ((size_t)limb << logN) + (size_t)blockIdx.x * TPB + threadIdx.x;
    const ModParams m = mp[limb];
    u128 a0[QG][GG], a1[QG][GG];
#pragma unroll
    for (int i = 0; i < QG; ++i) {
        const bool qv = U.q0 + i < U.nq;
        const u64 b0 = qv ? lin_word(P, V.b0[U.q0 + i], idx) : 0, b1 = qv ? lin_word(P, V.b1[U.q0 + i], idx) : 0;
#pragma unroll
        for (int k = 0; k < GG; ++k) {
            a0[i][k] = (u128)b0 + (U.g0 + k < U.ng ? lin_word(P, V.ptb[U.g0 + k], idx) : 0);
            a1[i][k] = b1;
        }
    }
    for (u32 r = 0; r < R; ++r) {
        u64 x0[QG], x1[QG], p[GG];
#pragma unroll
        for (int i = 0; i < QG; ++i) {
            const bool qv = U.q0 + i < U.nq;
            x0[i] = qv ? P[V.ct0[(size_t)(U.q0 + i) * R + r] + idx] : 0;
            x1[i] = qv ? P[V.ct1[(size_t)(U.q0 + i) * R + r] + idx] : 0;
        }
#pragma unroll
        for (int k = 0; k < GG; ++k) p[k] = U.g0 + k < U.ng ? P[V.pt[(size_t)(U.g0 + k) * R + r] + idx] : 0;
#pragma unroll
        for (int i = 0; i < QG; ++i)
#pragma unroll
            for (int k = 0; k < GG; ++k) {
                a0[i][k] += (u128)x0[i] * p[k];
                a1[i][k] += (u128)x1[i] * p[k];
            }
        if ((r & 31) == 31) {
#pragma unroll
            for (int i = 0; i < QG; ++i)
#pragma unroll
                for (int k = 0; k < GG; ++k) {
                    a0[i][k] = reduce128(a0[i][k], m);
                    a1[i][k] = reduce128(a1[i][k], m);
                }
        }
    }
#pragma unroll
    for (int i = 0; i < QG; ++i)
#pragma unroll
        for (int k = 0; k < GG; ++k)
            if (U.q0 + i < U.nq && U.g0 + k < U.ng) {
                const size_t o = (size_t)(U.q0 + i) * U.ng + U.g0 + k;
                P[V.o0[o] + idx] = reduce128(a0[i][k], m);
                P[V.o1[o] + idx] = reduce128(a1[i][k], m);
            }
}

// FP64 form (primes < 2^50): mulmod_vv_f64 terms (<= 0.875q), sums centred every
// 8 products (<= q/2 + 2 + 7q).  Each wave owns a tile of 2 queries x 8 blocks of
// the unit; the unit is qa tiles high and 8/qa tiles wide (qa = 8, 4 or 2: 16 x 8,
// 8 x 16 or 4 x 32 pairs), chosen per group so that the tiles are well filled.
static constexpr u32 LIN_TPB = 512, LIN_POS = 64, LIN_NW = LIN_TPB / LIN_POS;

__global__ __launch_bounds__(LIN_TPB) void k_linear_sum_lds_f64(u64 *__restrict__ P,
                                                               const LinUnit *__restrict__ units, u32 R,
                                                               const ModParams *__restrict__ mp, u32 logN) {
    // row LIN_NS takes the words of the streams a unit does not have
    __shared__ double buf[2][LIN_NS + 1][LIN_POS];
    const LinUnit U = units[blockIdx.z];
    const LinView V(U, R);
    // the wave index is uniform across the wave: stream and tile bookkeeping in SGPRs
    const u32 limb = blockIdx.y, lane = threadIdx.x % LIN_POS;
    const u32 w = __builtin_amdgcn_readfirstlane(threadIdx.x / LIN_POS);
    const size_t idx = ((size_t)limb << logN) + (size_t)blockIdx.x * LIN_POS + lane;
    const double q = mp[limb].qd, qinv = mp[limb].qinv;
    const u32 uq = U.uq, ug = U.ug, ns = 2 * uq + ug;
    // streams: 2i (ct0 of query i), 2i+1 (ct1 of query i), 2*uq + k (block k);
    // thread (w, lane) moves streams w, w + 8, ... of its position
    constexpr u32 PER = (LIN_NS + LIN_NW - 1) / LIN_NW;
    // Every thread loads and deposits all PER slots whether or not the unit has that
    // stream (a missing one reads block 0's words into the spare row): the loads of
    // a step are straight-line code.
    const CU64 *sbase[PER];
    u32 srow[PER];
#pragma unroll
    for (u32 k = 0; k < PER; ++k) {
        const u32 j = w + k * LIN_NW;
        srow[k] = j < ns ? j : LIN_NS;
        if (j < 2 * uq)
            sbase[k] = ((j & 1) ? V.ct1 : V.ct0) + (size_t)(U.q0 + j / 2) * R;
        else
            sbase[k] = V.pt + (size_t)(U.g0 + (j < ns ? j - 2 * uq : 0)) * R;
    }
    // this wave's tile: queries tq, tq+1 x blocks tg .. tg+7 of the unit
    const u32 tq = 2 * (w % U.qa), tg = 8 * (w / U.qa);
    const bool v0 = tq < uq, v1 = tq + 1 < uq;
    const u32 nb = tg < ug ? min(8u, ug - tg) : 0;
    // biases: indices clamped into the unit's tables so that all loads issue back to
    // back; the accumulators of pairs outside the unit are never stored
    double a0[2][8], a1[2][8];
    {
        double pb[8], b0[2], b1[2];
#pragma unroll
        for (u32 k = 0; k < 8; ++k) pb[k] = u64_to_f64(lin_word(P, V.ptb[min(U.g0 + tg + k, U.ng - 1)], idx));
#pragma unroll
        for (u32 i = 0; i < 2; ++i) {
            const u32 qq = min(U.q0 + tq + i, U.nq - 1);
            b0[i] = u64_to_f64(lin_word(P, V.b0[qq], idx));
            b1[i] = u64_to_f64(lin_word(P, V.b1[qq], idx));
        }
#pragma unroll
        for (u32 i = 0; i < 2; ++i)
#pragma unroll
            for (u32 k = 0; k < 8; ++k) {
                a0[i][k] = reduce_f64(b0[i] + pb[k], q, qinv);
                a1[i][k] = b1[i];
            }
    }
    // two register slots: the loads of step r+2 are issued at step r, the words of
    // step r+1 (issued a step earlier) go into LDS after step r's products
    u64 la[PER], lb[PER];
#define LIN_ISSUE(ld, rr)                                                      \
    _Pragma("unroll") for (u32 k = 0; k < PER; ++k) ld[k] = P[sbase[k][rr] + idx];
#define LIN_DEPOSIT(ld, b)                                                     \
    _Pragma("unroll") for (u32 k = 0; k < PER; ++k) buf[b][srow[k]][lane] = u64_to_f64(ld[k]);
#define LIN_PRODUCTS(b, rr)                                                    \
    _Pragma("unroll") for (u32 i = 0; i < 2; ++i) {                            \
        if (!(i ? v1 : v0)) continue;                                          \
        const double x0 = buf[b][2 * (tq + i)][lane], x1 = buf[b][2 * (tq + i) + 1][lane]; \
        _Pragma("unroll") for (u32 k = 0; k < 8; ++k) {                        \
            if (k < nb) {                                                      \
                const double p = buf[b][2 * uq + tg + k][lane];                \
                a0[i][k] += mulmod_vv_f64(x0, p, q, qinv);                     \
                a1[i][k] += mulmod_vv_f64(x1, p, q, qinv);                     \
            }                                                                  \
        }                                                                      \
    }                                                                          \
    if (((rr) & 7) == 7) {                                                     \
        _Pragma("unroll") for (u32 i = 0; i < 2; ++i)                          \
        _Pragma("unroll") for (u32 k = 0; k < 8; ++k) {                        \
            a0[i][k] = reduce_f64(a0[i][k], q, qinv);                          \
            a1[i][k] = reduce_f64(a1[i][k], q, qinv);                          \
        }                                                                      \
    }
// G: the step may lie past R (the tail loop); in the main loop it never does
#define LIN_STEP(G, X, Y, rr, b)                                               \
    if (!(G) || (rr) + 2 < R) {                                                \
        LIN_ISSUE(X, (rr) + 2)                                                 \
    }                                                                          \
    if (!(G) || (rr) < R) {                                                    \
        LIN_PRODUCTS(b, rr)                                                    \
    }                                                                          \
    if (!(G) || (rr) + 1 < R) {                                                \
        LIN_DEPOSIT(Y, (b) ^ 1)                                                \
    }                                                                          \
    __syncthreads();
    LIN_ISSUE(la, 0u)
    LIN_ISSUE(lb, R > 1 ? 1u : 0u)
    LIN_DEPOSIT(la, 0)
    __syncthreads();
    u32 r = 0;
    for (; r + 4 <= R; r += 2) {
        LIN_STEP(false, la, lb, r, 0)
        LIN_STEP(false, lb, la, r + 1, 1)
    }
    for (; r < R; r += 2) {
        LIN_STEP(true, la, lb, r, 0)
        LIN_STEP(true, lb, la, r + 1, 1)
    }
#undef LIN_STEP
#undef LIN_PRODUCTS
#undef LIN_ISSUE
#undef LIN_DEPOSIT
#pragma unroll
    for (u32 i = 0; i < 2; ++i) {
        if (!(i ? v1 : v0)) continue;
#pragma unroll
        for (u32 k = 0; k < 8; ++k)
            if (k < nb) {
                const size_t o = (size_t)(U.q0 + tq + i) * U.ng + U.g0 + tg + k;
                P[V.o0[o] + idx] = f64_to_canon(reduce_f64(a0[i][k], q, qinv), q);
                P[V.o1[o] + idx] = f64_to_canon(reduce_f64(a1[i][k], q, qinv), q);
            }
    }
}

// FP64 form for small groups: each wave owns one tile (uq <= 2 queries x ug <= 8
// blocks of a group, 64 positions) and reads its words itself; a workgroup holds
// LIN_TW independent tiles (no LDS, no barriers), so groups probed by one or two
// queries do not each hold a whole LDS unit.  Loads of step r+1 are issued before
// the products of step r; tile rows and blocks past uq / ug repeat the last valid
// stream (cache hits) so that every step's loads are straight-line code.
static constexpr u32 LIN_TW = 8;

__global__ __launch_bounds__(LIN_TW * LIN_POS) void k_linear_sum_tiles_f64(u64 *__restrict__ P,
                                                                         const LinUnit *__restrict__ tiles,
                                                                         u32 n_tiles, u32 R,
                                                                         const ModParams *__restrict__ mp,
                                                                         u32 logN) {
    const u32 t = blockIdx.z * LIN_TW + __builtin_amdgcn_readfirstlane(threadIdx.x / LIN_POS);
    if (t >= n_tiles) return;
    const LinUnit U = tiles[t];
    const LinView V(U, R);
    const u32 limb = blockIdx.y, lane = threadIdx.x % LIN_POS;
    const size_t idx = ((size_t)limb << logN) + (size_t)blockIdx.x * LIN_POS + lane;
    const double q = mp[limb].qd, qinv = mp[limb].qinv;
    const u32 uq = U.uq, ug = U.ug;
    const CU64 *c0[2], *c1[2], *pr[8];
#pragma unroll
    for (u32 i = 0; i < 2; ++i) {
        const size_t row = (size_t)(U.q0 + min(i, uq - 1)) * R;
        c0[i] = V.ct0 + row;
        c1[i] = V.ct1 + row;
    }
#pragma unroll
    for (u32 k = 0; k < 8; ++k) pr[k] = V.pt + (size_t)(U.g0 + min(k, ug - 1)) * R;
    double a0[2][8], a1[2][8];
    {
        double pb[8], b0[2], b1[2];
#pragma unroll
        for (u32 k = 0; k < 8; ++k) pb[k] = u64_to_f64(lin_word(P, V.ptb[U.g0 + min(k, ug - 1)], idx));
#pragma unroll
        for (u32 i = 0; i < 2; ++i) {
            const u32 qq = U.q0 + min(i, uq - 1);
            b0[i] = u64_to_f64(lin_word(P, V.b0[qq], idx));
            b1[i] = u64_to_f64(lin_word(P, V.b1[qq], idx));
        }
#pragma unroll
        for (u32 i = 0; i < 2; ++i)
#pragma unroll
            for (u32 k = 0; k < 8; ++k) {
                a0[i][k] = reduce_f64(b0[i] + pb[k], q, qinv);
                a1[i][k] = b1[i];
            }
    }
    u64 nx0[2], nx1[2], np[8];
#define LIN_TLOAD(rr)                                                          \
    _Pragma("unroll") for (u32 i = 0; i < 2; ++i) {                            \
        nx0[i] = P[c0[i][rr] + idx];                                           \
        nx1[i] = P[c1[i][rr] + idx];                                           \
    }                                                                          \
    _Pragma("unroll") for (u32 k = 0; k < 8; ++k) np[k] = P[pr[k][rr] + idx];
    LIN_TLOAD(0u)
    for (u32 r = 0; r < R; ++r) {
        double x0[2], x1[2], p[8];
#pragma unroll
        for (u32 i = 0; i < 2; ++i) {
            x0[i] = u64_to_f64(nx0[i]);
            x1[i] = u64_to_f64(nx1[i]);
        }
#pragma unroll
        for (u32 k = 0; k < 8; ++k) p[k] = u64_to_f64(np[k]);
        LIN_TLOAD(min(r + 1, R - 1))
#pragma unroll
        for (u32 i = 0; i < 2; ++i)
#pragma unroll
            for (u32 k = 0; k < 8; ++k) {
                a0[i][k] += mulmod_vv_f64(x0[i], p[k], q, qinv);
                a1[i][k] += mulmod_vv_f64(x1[i], p[k], q, qinv);
            }
        if ((r & 7) == 7) {
#pragma unroll
            for (u32 i = 0; i < 2; ++i)
#pragma unroll
                for (u32 k = 0; k < 8; ++k) {
                    a0[i][k] = reduce_f64(a0[i][k], q, qinv);
                    a1[i][k] = reduce_f64(a1[i][k], q, qinv);
                }
        }
    }
#undef LIN_TLOAD
#pragma unroll
    for (u32 i = 0; i < 2; ++i)
#pragma unroll
        for (u32 k = 0; k < 8; ++k)
            if (i < uq && k < ug) {
                const size_t o = (size_t)(U.q0 + i) * U.ng + U.g0 + k;
                P[V.o0[o] + idx] = f64_to_canon(reduce_f64(a0[i][k], q, qinv), q);
                P[V.o1[o] + idx] = f64_to_canon(reduce_f64(a1[i][k], q, qinv), q);
            }
}

bool linear_sum_f64(const DevTables &t) { return t.use_f64 && !(t.f64_off & 8); }

void launch_linear_sum(const DevTables &t, u64 *P, const LinUnit *units_dev, u32 n_units, u32 R, u32 limbs,
                       int shape, hipStream_t s) {
    if (!n_units) return;
    if (shape == LIN_SHAPE_LDS) {
        dim3 grid(t.N / LIN_POS, limbs, n_units);
        hipLaunchKernelGGL(k_linear_sum_lds_f64, grid, dim3(LIN_TPB), 0, s, P, units_dev, R, t.mp, t.logN);
    } else if (shape == LIN_SHAPE_TILES) {
        dim3 grid(t.N / LIN_POS, limbs, (n_units + LIN_TW - 1) / LIN_TW);
        hipLaunchKernelGGL(k_linear_sum_tiles_f64, grid, dim3(LIN_TW * LIN_POS), 0, s, P, units_dev, n_units, R,
                           t.mp, t.logN);
    } else {
        dim3 grid(t.N / TPB, limbs, n_units);
        hipLaunchKernelGGL((k_linear_sum<1, 8>), grid, dim3(TPB), 0, s, P, units_dev, R, t.mp, t.logN);
    }
}

}  // namespace fhea
