// fhea_internal.hpp -- engine-internal structures shared by the host code and
// the kernel launchers.  Not part of the C-ABI (include/fhea.h is).
#pragma once
#include <stdint.h>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fhea.h"
#include "fhea_arith.hpp"

namespace fhea {

// Per-prime device tables.  Index 0..L-1 = Q primes, L..L+K-1 = P primes.
struct PrimeTables {
    const ModParams *mp = nullptr;  // [T]
    const u64 *tw = nullptr;        // [T*N] psi^br(j)                (Shoup pairs interleaved:
    const u64 *itw = nullptr;       // [T*N] psi^-br(j)                 see NttTw below)
    const u64 *ninv = nullptr;      // [T*2] (N^-1, shoup)
};

// Twiddle layout on the device: for prime i, 2*N words: tw[2j] = w_j,
// tw[2j+1] = floor(w_j 2^64 / q) so one 16-byte load fetches a Shoup pair.

// Basis conversion descriptor (fast_bases_conv, key_switch.cpp:7-60).
struct ConvDesc {
    u32 n_in = 0, n_out = 0;
    const u32 *in_prime = nullptr;   // [n_in]  prime index of each input limb
    const u64 *in_cst = nullptr;     // [2*n_in] (qhat_i^-1 mod q_i, shoup)
    const u32 *out_prime = nullptr;  // [n_out] prime index of each output limb
    const u32 *out_pos = nullptr;    // [n_out] limb slot in the destination
    const u64 *hat = nullptr;        // [n_in*n_out] qhat_i mod p_j, row-major by input
    // FP64 forms (only when every prime is below 2^50): (value, value / modulus)
    const double *in_cstf = nullptr; // [2*n_in]
    const double *hatf = nullptr;    // [2*n_in*n_out]
    const double *in_ncstf = nullptr; // [2*n_in] (N^-1 qhat_i^-1 mod q_i, / q_i): INTT scaling and conversion in one product
};

// Per-level tables (sizeQl = l).
struct LevelTables {
    u32 l = 0, nparts = 0;
    std::vector<ConvDesc> modup;           // [nparts]
    std::vector<u32> part_start, part_size;
    const u32 *qlp_prime = nullptr;        // [l+K] prime index of each QlP slot
    const u32 *q_prime = nullptr;          // [l] identity map 0..l-1
    const u32 *p_prime = nullptr;          // [K] L..L+K-1
    const u32 *modup_ntt_prime = nullptr;  // [nparts*(l+K)] prime index per digit slot
    const u32 *modup_slots = nullptr;      // the valid (converted) digit slots, in order
    u32 n_modup_slots = 0;
    ConvDesc moddown;                      // P -> Q_l
    const u64 *pinv = nullptr;             // [2*l] (P^-1 mod q_i, shoup)
    const u64 *rs_qlinv = nullptr;         // [2*(l-1)] (q_l^-1 mod q_i, shoup)
    const u64 *rs_qlql = nullptr;          // [2*(l-1)] (-q_l^-1 mod q_i, shoup)
    // FP64 forms (value, value / q_i)
    const double *pinvf = nullptr, *rs_qlinvf = nullptr, *rs_qlqlf = nullptr;
    const double *rs_pqf = nullptr;        // [2*(l-1)] ((P q_l)^-1 mod q_i, / q_i)
    const double *rs_liftf = nullptr;      // [6] (N^-1, P^-1 N^-1, P^-1) mod q_{l-1}, each (value, / q_{l-1})
};

// Twiddles of the CPU FP64 transforms as doubles (w and w/q, forward and inverse),
// built per prime on first use.
struct HostF64Twiddles {
    struct Prime {
        std::once_flag once;
        std::vector<double> w, wq, iw, iwq;  // [N] each, the order of HostTables::tw / itw
    };
    std::unique_ptr<Prime[]> p;
};

struct HostTables {
    u32 N = 0, logN = 0, L = 0, K = 0, dnum = 0, alpha = 0;
    std::vector<u64> primes, psi, ninv;
    std::vector<u64> psi_given;  // explicit roots from the parameter block, if any
    std::vector<u64> tw, itw;  // [T*N] plain (for export / CPU client)
    std::vector<u64> tw_sh, itw_sh;  // [T*N] Shoup companions of tw / itw (CPU transforms)
    // CPU transforms: 0 = FP64 (exact, AVX2+FMA) where q < 2^50 and the CPU has it,
    // else 64-bit integer; 1 = always 64-bit integer (same words either way)
    u32 host_ntt_mode = 0;
    std::shared_ptr<HostF64Twiddles> f64tw;  // set with the tables; copies share it
};

struct RotKey {
    int32_t step;
    u64 galois;
    u64 evk_off;
    u32 key_limbs;
};

// Host-side table generation (fhea_context.cpp).
bool is_prime_u64(u64 n);
u32 gen_primes(u32 two_n, u32 bits, u32 count, const std::vector<u64> &skip, std::vector<u64> &out);
u64 min_primitive_root(u64 q, u32 two_n);
u64 inv_mod(u64 a, u64 q);
u64 pow_mod(u64 a, u64 e, u64 q);
u32 bit_reverse(u32 x, u32 bits);
void build_host_tables(HostTables &h);
// Validates a parameter block (N in [min_n, 2^17], L, K, dnum, primes) and fills h.
// Returns an FHEA_* status.
int host_tables_from_params(const fhea_params *p, u32 min_n, HostTables &h);

}  // namespace fhea
