// fhea_context.cpp -- parameter and per-prime table generation (host, one-time).
//
// The reference reads these tables from files dumped by OpenFHE
// (cpp_host_model/src/fhe_host_driver.cpp:18-102, produced by
// integration_tools/generate_interactive_test_vectors.cpp:246-282).  OpenFHE
// is not available here, so the engine generates them from documented rules:
//   * primes q == 1 (mod 2N), q < 2^bits, taken in descending order;
//   * psi = the smallest primitive 2N-th root of unity mod q (or a given one);
//   * tw[j] = psi^br(j), itw[j] = psi^-br(j), br over log2(N) bits
//     (generate_interactive_test_vectors.cpp:263-275);
//   * n_inv = N^-1 mod q.
#include <cstring>

#include "fhea_internal.hpp"

namespace fhea {

u64 pow_mod(u64 a, u64 e, u64 q) {
    u128 r = 1 % q, b = a % q;
    while (e) {
        if (e & 1) r = r * b % q;
        b = b * b % q;
        e >>= 1;
    }
    return (u64)r;
}

u64 inv_mod(u64 a, u64 q) {
    // q prime: a^(q-2)
    if (a % q == 0) return 0;
    return pow_mod(a, q - 2, q);
}

bool is_prime_u64(u64 n) {
    if (n < 2) return false;
    const u64 small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (u64 p : small) {
        if (n == p) return true;
        if (n % p == 0) return false;
    }
    u64 d = n - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; ++s; }
    for (u64 a : small) {
        u64 x = pow_mod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool composite = true;
        for (int r = 1; r < s; ++r) {
            x = (u64)((u128)x * x % n);
            if (x == n - 1) { composite = false; break; }
        }
        if (composite) return false;
    }
    return true;
}

u32 gen_primes(u32 two_n, u32 bits, u32 count, const std::vector<u64> &skip, std::vector<u64> &out) {
    u64 top = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    u64 k = (top - 1) / two_n;
    u32 found = 0;
    while (found < count && k > 0) {
        u64 q = k * two_n + 1;
        --k;
        bool dup = false;
        for (u64 s : skip) dup |= (s == q);
        for (u64 s : out) dup |= (s == q);
        if (dup) continue;
        if (is_prime_u64(q)) { out.push_back(q); ++found; }
    }
    return found;
}

u64 min_primitive_root(u64 q, u32 two_n) {
    u64 g = 0;
    for (u64 x = 2; x < q; ++x) {
        u64 c = pow_mod(x, (q - 1) / two_n, q);
        if (pow_mod(c, two_n / 2, q) == q - 1) { g = c; break; }
    }
    if (!g) return 0;
    // every primitive 2N-th root is g^odd
    u64 g2 = (u64)((u128)g * g % q), cur = g, best = g;
    for (u32 i = 1; i < two_n / 2; ++i) {
        cur = (u64)((u128)cur * g2 % q);
        if (cur < best) best = cur;
    }
    return best;
}

u32 bit_reverse(u32 x, u32 bits) {
    u32 r = 0;
    for (u32 i = 0; i < bits; ++i) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

void build_host_tables(HostTables &h) {
    const u32 T = h.L + h.K, N = h.N;
    h.psi.resize(T);
    h.ninv.resize(T);
    h.tw.assign((size_t)T * N, 0);
    h.itw.assign((size_t)T * N, 0);
    h.tw_sh.assign((size_t)T * N, 0);
    h.itw_sh.assign((size_t)T * N, 0);
    h.f64tw = std::make_shared<HostF64Twiddles>();
    h.f64tw->p.reset(new HostF64Twiddles::Prime[T]);
    for (u32 i = 0; i < T; ++i) {
        u64 q = h.primes[i];
        u64 psi = h.psi_given.empty() ? min_primitive_root(q, 2 * N) : h.psi_given[i];
        h.psi[i] = psi;
        h.ninv[i] = inv_mod(N, q);
        u64 psi_inv = inv_mod(psi, q);
        // powers psi^e for e in [0, N): fill by repeated multiplication
        u64 cur = 1, curi = 1;
        std::vector<u64> p1(N), p2(N);
        for (u32 e = 0; e < N; ++e) {
            p1[e] = cur;
            p2[e] = curi;
            cur = (u64)((u128)cur * psi % q);
            curi = (u64)((u128)curi * psi_inv % q);
        }
        for (u32 j = 0; j < N; ++j) {
            u32 br = bit_reverse(j, h.logN);
            h.tw[(size_t)i * N + j] = p1[br];
            h.itw[(size_t)i * N + j] = p2[br];
            h.tw_sh[(size_t)i * N + j] = shoup_companion(p1[br], q);
            h.itw_sh[(size_t)i * N + j] = shoup_companion(p2[br], q);
        }
    }
}

int host_tables_from_params(const fhea_params *p, u32 min_n, HostTables &h) {
    if (!p) return FHEA_ERR_ARG;
    const u32 N = p->N;
    if (N < min_n || N > (1u << 17) || (N & (N - 1))) return FHEA_ERR_ARG;
    if (p->L < 1 || p->K < 1 || p->dnum < 1 || p->dnum > p->L || p->L + p->K > 255) return FHEA_ERR_ARG;
    const u32 alpha = (p->L + p->dnum - 1) / p->dnum;
    if (alpha > 32 || p->K > 32) return FHEA_ERR_ARG;
    const u32 bits = p->q_bits ? p->q_bits : 50;
    if (bits < 20 || bits > 61) return FHEA_ERR_ARG;
    h = HostTables();
    h.N = N;
    while ((1u << h.logN) < N) ++h.logN;
    h.L = p->L;
    h.K = p->K;
    h.dnum = p->dnum;
    h.alpha = alpha;
    std::vector<u64> gen;
    if (!p->q_primes || !p->p_primes) {
        if (gen_primes(2 * N, bits, p->L + p->K, {}, gen) != p->L + p->K) return FHEA_ERR_ARG;
    }
    for (u32 i = 0; i < p->L; ++i) h.primes.push_back(p->q_primes ? p->q_primes[i] : gen[i]);
    for (u32 j = 0; j < p->K; ++j) h.primes.push_back(p->p_primes ? p->p_primes[j] : gen[p->L + j]);
    for (u64 q : h.primes)
        if (q < 3 || q >= (1ull << 61) || (q - 1) % (2ull * N) != 0 || !is_prime_u64(q)) return FHEA_ERR_ARG;
    for (size_t i = 0; i < h.primes.size(); ++i)
        for (size_t j = i + 1; j < h.primes.size(); ++j)
            if (h.primes[i] == h.primes[j]) return FHEA_ERR_ARG;
    if (p->psi) {
        // a primitive 2N-th root: psi^N == -1 (then its order is exactly 2N)
        for (u32 i = 0; i < p->L + p->K; ++i) {
            const u64 q = h.primes[i], w = p->psi[i];
            if (w == 0 || w >= q || pow_mod(w, N, q) != q - 1) return FHEA_ERR_ARG;
            h.psi_given.push_back(w);
        }
    }
    build_host_tables(h);
    return FHEA_OK;
}

}  // namespace fhea
