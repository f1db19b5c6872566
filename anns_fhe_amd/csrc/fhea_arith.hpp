// fhea_arith.hpp -- 64-bit RNS modular arithmetic for the HIP engine (host and device).
//
// Replaces the reference's L0 (hardware_fpga_model/hls/src/mod_arith.cpp:3-25).
// Every function returns the unique residue in [0, q), so results are
// identical to the reference's Barrett mod_mul whatever reduction is used
// here.  Primes are limited to q < 2^61 (the reference's largest is 60 bits).
//
// On gfx950 the 64x64->128 products lower to v_mad_u64_u32 / v_mul_hi_u32 /
// v_mul_lo_u32 sequences (there is no 64-bit multiplier; each of those issues at
// about the FP64 FMA rate), so the reductions below are chosen for the fewest
// 32-bit multiplies:
//   * mul_shoup   a*w mod q for a precomputed constant w (twiddles, CRT
//                 constants): 1 mulhi64 + 2 mullo64.
//   * mulmod      a*b mod q for a, b < q: pre-shifted Barrett with
//                 mu = floor(2^(2b)/q), b = bitlen(q).
//   * reduce128   any 128-bit x mod q with ratio = floor(2^128/q) (used for
//                 lazily accumulated sums of products).
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

typedef uint64_t u64;
typedef uint32_t u32;
typedef unsigned __int128 u128;

#define FHEA_HD __host__ __device__ __forceinline__

FHEA_HD u64 mulhi64(u64 a, u64 b) { return (u64)(((u128)a * b) >> 64); }

// Conditional corrections are branch-free (q < 2^61, so the sign bit of a
// wrapped difference marks "went negative"): data-dependent branches cost more
// than the arithmetic in the host-side transforms.
FHEA_HD u64 neg_mask(u64 x) { return (u64)((int64_t)x >> 63); }

// x in [0, 2q) -> x mod q
FHEA_HD u64 csub_q(u64 x, u64 q) {
    const u64 t = x - q;
    return t + (q & neg_mask(t));
}

FHEA_HD u64 add_mod(u64 a, u64 b, u64 q) { return csub_q(a + b, q); }

FHEA_HD u64 sub_mod(u64 a, u64 b, u64 q) {
    const u64 d = a - b;
    return d + (q & neg_mask(d));
}

// a * w mod q, wsh = floor(w * 2^64 / q), w < q, any a < 2^64 (Harvey/Shoup).
FHEA_HD u64 mul_shoup(u64 a, u64 w, u64 wsh, u64 q) {
    u64 qh = mulhi64(a, wsh);
    return csub_q(a * w - qh * q, q);
}

// Same, result left in [0, 2q) (lazy form inside transforms).
FHEA_HD u64 mul_shoup_lazy(u64 a, u64 w, u64 wsh, u64 q) {
    u64 qh = mulhi64(a, wsh);
    return a * w - qh * q;
}

// Per-prime constants for the variable x variable product.
struct ModParams {
    u64 q;
    u64 mu;      // floor(2^(2*bq) / q)  (< 2^(bq+1))
    u64 r_lo;    // floor(2^128 / q) low word
    u64 r_hi;    // floor(2^128 / q) high word
    u32 bq;      // bit length of q
    u32 pad;
    double qd;   // q as a double (exact for q < 2^53)
    double qinv; // 1/q rounded to double (FP64 kernels)
};

// x mod q for x < 2^(2*bq) (in particular x = a*b with a, b < q).
FHEA_HD u64 barrett_reduce(u128 x, const ModParams &m) {
    u64 t = (u64)(x >> (m.bq - 1));
    u128 pr = (u128)t * m.mu;
    u64 qhat = (u64)(pr >> (m.bq + 1));
    u64 r = (u64)x - qhat * m.q;
    if (r >= m.q) r -= m.q;
    if (r >= m.q) r -= m.q;
    return r;
}

FHEA_HD u64 mulmod(u64 a, u64 b, const ModParams &m) { return barrett_reduce((u128)a * b, m); }

// Any 128-bit x mod q (q < 2^61): qhat = floor(x * floor(2^128/q) / 2^128)
// is within 1 of floor(x/q), so one correction suffices.
FHEA_HD u64 reduce128(u128 x, const ModParams &m) {
    u64 x0 = (u64)x, x1 = (u64)(x >> 64);
    u64 carry = mulhi64(x0, m.r_lo);
    u128 t = (u128)x0 * m.r_hi;
    u128 s1 = (u128)(u64)t + carry;
    u64 tmp1 = (u64)s1;
    u64 tmp3 = (u64)(t >> 64) + (u64)(s1 >> 64);
    u128 t2 = (u128)x1 * m.r_lo;
    u128 s2 = (u128)tmp1 + (u64)t2;
    u64 c2 = (u64)(t2 >> 64) + (u64)(s2 >> 64);
    u64 qhat = x1 * m.r_hi + tmp3 + c2;
    u64 r = x0 - qhat * m.q;
    if (r >= m.q) r -= m.q;
    if (r >= m.q) r -= m.q;
    return r;
}

// Host-side constant generation.
inline u64 shoup_companion(u64 w, u64 q) { return (u64)(((u128)w << 64) / q); }

inline ModParams make_mod_params(u64 q) {
    ModParams m;
    m.q = q;
    u32 b = 64 - (u32)__builtin_clzll(q);
    m.bq = b;
    m.mu = (u64)(((u128)1 << (2 * b)) / q);
    // floor(2^128 / q) = floor((2^128 - 1) / q) because q is odd and > 1
    u128 r = (~(u128)0) / q;
    m.r_lo = (u64)r;
    m.r_hi = (u64)(r >> 64);
    m.pad = 0;
    m.qd = (double)q;
    m.qinv = 1.0 / (double)q;
    return m;
}
