// fhea_probe.hip -- VALU issue-rate probe (no reference counterpart).
//
// The 64-bit integer kernels (a prime >= 2^50) are bound by the vector ALU, and
// mostly by its 32-bit multiplies: a 64x64-bit modular product lowers to
// v_mul_lo_u32 / v_mul_hi_u32 / v_mad_u64_u32 sequences.  This probe measures the
// issue rate of each of those instructions (and of FP64 FMA for comparison) on
// the device, as the ceiling for the integer kernels' op rate.  Every thread runs
// eight independent dependent chains of one instruction kind over `iters` steps
// (enough chains and waves to hide the latency); the result words are stored so
// that nothing is dead.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fhea.h"

namespace {

typedef uint64_t u64;
typedef uint32_t u32;

constexpr u32 P_TPB = 256;
constexpr u32 P_CHAINS = 8;

template <int KIND>
__global__ __launch_bounds__(P_TPB) void k_valu_probe(u64 *out, u32 iters, u32 m0) {
    const u32 tid = blockIdx.x * P_TPB + threadIdx.x;
    u64 res = 0;
    if constexpr (KIND == 0 || KIND == 1) {
        u32 a[P_CHAINS], m[P_CHAINS];
#pragma unroll
        for (u32 k = 0; k < P_CHAINS; ++k) {
            a[k] = tid * (2 * k + 1) + m0;
            m[k] = m0 + 2 * k + 1;
        }
#pragma unroll 4
        for (u32 it = 0; it < iters; ++it) {
#pragma unroll
            // squaring for v_mul_lo_u32: a chain a *= m would be folded into a *= m^4 by
            // the compiler (multiplication mod 2^32 is associative)
            for (u32 k = 0; k < P_CHAINS; ++k) a[k] = KIND == 0 ? a[k] * a[k] : __umulhi(a[k], m[k]);
        }
#pragma unroll
        for (u32 k = 0; k < P_CHAINS; ++k) res ^= a[k];
    } else if constexpr (KIND == 2) {
        u64 acc[P_CHAINS];
        u32 m[P_CHAINS];
#pragma unroll
        for (u32 k = 0; k < P_CHAINS; ++k) {
            acc[k] = (u64)tid * (2 * k + 1) + m0;
            m[k] = m0 + 2 * k + 1;
        }
#pragma unroll 4
        for (u32 it = 0; it < iters; ++it) {
#pragma unroll
            for (u32 k = 0; k < P_CHAINS; ++k) acc[k] = (u64)(u32)acc[k] * m[k] + acc[k];  // v_mad_u64_u32
        }
#pragma unroll
        for (u32 k = 0; k < P_CHAINS; ++k) res ^= acc[k];
    } else {
        double d[P_CHAINS], f[P_CHAINS];
#pragma unroll
        for (u32 k = 0; k < P_CHAINS; ++k) {
            d[k] = (double)(tid + k);
            f[k] = 1.0 + 1.0 / (double)(m0 + 2 * k + 1);
        }
#pragma unroll 4
        for (u32 it = 0; it < iters; ++it) {
#pragma unroll
            for (u32 k = 0; k < P_CHAINS; ++k) d[k] = __builtin_fma(d[k], f[k], -0.5);
        }
#pragma unroll
        for (u32 k = 0; k < P_CHAINS; ++k) res ^= (u64)__double_as_longlong(d[k]);
    }
    out[tid] = res;
}

template <int KIND>
hipError_t launch_probe(u64 *out, u32 blocks, u32 iters, hipStream_t s) {
    hipLaunchKernelGGL(k_valu_probe<KIND>, dim3(blocks), dim3(P_TPB), 0, s, out, iters, 12345u);
    return hipGetLastError();
}

hipError_t launch_kind(u32 kind, u64 *out, u32 blocks, u32 iters, hipStream_t s) {
    switch (kind) {
    case 0: return launch_probe<0>(out, blocks, iters, s);
    case 1: return launch_probe<1>(out, blocks, iters, s);
    case 2: return launch_probe<2>(out, blocks, iters, s);
    default: return launch_probe<3>(out, blocks, iters, s);
    }
}

}  // namespace

extern "C" int fhea_probe_valu_rate(uint32_t device, uint32_t kind, uint32_t iters, double *lane_ops_per_s) {
    if (!lane_ops_per_s || kind > 3 || iters == 0) return FHEA_ERR_ARG;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return FHEA_ERR_DEVICE;
    if ((int)device != cur && hipSetDevice((int)device) != hipSuccess) return FHEA_ERR_DEVICE;
    int st = FHEA_OK;
    int cus = 0;
    u64 *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, (int)device) != hipSuccess || cus <= 0)
        st = FHEA_ERR_DEVICE;
    // 8 workgroups of 256 threads per CU: 8 waves per SIMD
    const u32 blocks = 8u * (u32)(cus > 0 ? cus : 1);
    if (st == FHEA_OK && hipMalloc(&out, (size_t)blocks * P_TPB * sizeof(u64)) != hipSuccess) st = FHEA_ERR_NO_MEM;
    if (st == FHEA_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        st = FHEA_ERR_DEVICE;
    if (st == FHEA_OK) {
        float ms = 0.0f;
        // one untimed launch (code object load, clocks up), then the timed one
        if (launch_kind(kind, out, blocks, iters, nullptr) != hipSuccess || hipEventRecord(e0, nullptr) != hipSuccess ||
            launch_kind(kind, out, blocks, iters, nullptr) != hipSuccess || hipEventRecord(e1, nullptr) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.0f)
            st = FHEA_ERR_DEVICE;
        else
            *lane_ops_per_s = (double)blocks * P_TPB * iters * P_CHAINS / ((double)ms * 1e-3);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out) (void)hipFree(out);
    if ((int)device != cur) (void)hipSetDevice(cur);
    return st;
}
