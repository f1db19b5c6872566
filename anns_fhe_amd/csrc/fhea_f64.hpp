// fhea_f64.hpp -- exact modular arithmetic on FP64 FMA for primes q < 2^50.
//
// On gfx950 FP64 FMA and the 32-bit integer multiplies issue at about the same
// rate (measured per lane: v_fma_f64 34.5, v_mul_lo_u32 37.3, v_mul_hi_u32 37.4,
// v_mad_u64_u32 33.1 T/s), but a 64x64-bit modular product needs about ten
// multiplies plus the carries between their halves, against six FP64
// instructions here.  For q < 2^50 every residue is an exact double, and the
// product v*w splits exactly into hi = fl(v*w) and lo = fma(v, w, -hi)
// (error-free transformation: |v*w| < 2^106, both integers).  With a quotient
// estimate k = rint(v * (w/q)), the value v*w - k*q is an integer of magnitude
// below 2^51 and is computed exactly as fma(-k, q, hi) + lo.  Residues are kept
// signed ("centred"), |x| < 8q < 2^53, and brought back to [0, q) only when
// they leave a transform, so the result words equal the integer path's.
//
// Bounds used by the callers (q < 2^50):
//   mulmod_f64(v, w, wq): |v| <= 8q, 0 <= w < q  ->  result == v*w (mod q) exactly and
//                         |result| <= q (1/2 + |v| 2^-52), i.e. <= q/2 + |v|/4:
//                         k is within 1/2 + |v w/q| 2^-52 of v w / q (one rounding of
//                         w/q, one of the product).  So |v| <= q -> 0.75q,
//                         |v| <= 4q -> 1.5q, |v| <= 8q -> 2.5q.
//   mulmod_vv_f64(a, b):  canonical a, b -> |result| <= 0.875q.
//   reduce_f64(x):        |x| <= 8q  ->  |result| <= q/2 + 2.
// A forward radix-16 round from |x| <= q grows at most to 1.75q, 2.69q, 3.86q, 5.33q
// over its four stages (< 8q < 2^53); tests/test_f64_arith.py checks these in exact
// binary64 emulation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

__device__ __forceinline__ double mulmod_f64(double v, double w, double wq, double q) {
    const double k = __builtin_rint(v * wq);
    const double hi = v * w;
    const double lo = __builtin_fma(v, w, -hi);
    return __builtin_fma(-k, q, hi) + lo;
}

// a*b mod q for two canonical residues a, b in [0, q): |result| <= 0.875q
// (k is within 3 * 2^-53 * ab/q < 3/8 of ab/q).
__device__ __forceinline__ double mulmod_vv_f64(double a, double b, double q, double qinv) {
    const double hi = a * b;
    const double lo = __builtin_fma(a, b, -hi);
    const double k = __builtin_rint(hi * qinv);
    return __builtin_fma(-k, q, hi) + lo;
}

__device__ __forceinline__ double reduce_f64(double x, double q, double qinv) {
    return __builtin_fma(-__builtin_rint(x * qinv), q, x);
}

// A pointer read from device memory (an argument array) is a generic pointer, and
// loads through it become flat loads with a 64-bit address per element; these
// casts state that it addresses global memory (loads with a scalar base + offsets).
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gptr(const T *p) {
    return (const __attribute__((address_space(1))) T *)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gptr_w(T *p) {
    return (__attribute__((address_space(1))) T *)p;
}

// x in [0, 2^52) integer <-> double, exact (exponent-bias trick).
__device__ __forceinline__ double u64_to_f64(uint64_t x) {
    return __longlong_as_double((long long)(x | 0x4330000000000000ull)) - 4503599627370496.0;
}

// |x| <= q/2 + 2 (after reduce_f64)  ->  canonical residue in [0, q).
__device__ __forceinline__ uint64_t f64_to_canon(double x, double q) {
    double y = x < 0.0 ? x + q : x;
    return (uint64_t)__double_as_longlong(y + 4503599627370496.0) & 0x000FFFFFFFFFFFFFull;
}
