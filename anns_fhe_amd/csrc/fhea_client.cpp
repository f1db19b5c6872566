// fhea_client.cpp -- CPU client side: secret and evaluation-key generation.
//
// The reference takes its keys from OpenFHE (KeyGen / EvalMultKeyGen /
// EvalRotateKeyGen, integration_tools/generate_interactive_test_vectors.cpp:223-224)
// and loads them with initialize_context_from_directory
// (cpp_host_model/src/fhe_host_driver.cpp:105-172).  This file generates keys of
// the same hybrid key-switching form and writes them straight into the key
// arena.  Digit p of a key switching s' -> s is, over the Q limbs of the key
// level and all P limbs (evaluation domain):
//     a_p uniform,   b_p = -a_p s + e_p + [limb i in digit p] (P mod q_i) s'
// with e_p Gaussian (sigma 3.2).  This is client-side, one-time work; the
// evaluation path never runs on the CPU.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "fhea_engine.hpp"

using namespace fhea;

namespace {

struct HostNtt {
    const HostTables &h;
    explicit HostNtt(const HostTables &ht) : h(ht) {}
    // forward negacyclic NTT of one limb (ntt.cpp:4-43), Shoup products
    void forward(u64 *a, u32 pidx) const {
        const u32 N = h.N;
        const u64 q = h.primes[pidx];
        const u64 *tw = h.tw.data() + (size_t)pidx * N;
        u32 t = N >> 1;
        for (u32 m = 1; m < N; m <<= 1, t >>= 1) {
            for (u32 i = 0; i < m; ++i) {
                const u64 W = tw[m + i], Wsh = shoup_companion(W, q);
                u64 *x = a + 2 * i * t;
                for (u32 j = 0; j < t; ++j) {
                    u64 u = x[j], v = mul_shoup(x[j + t], W, Wsh, q);
                    x[j] = add_mod(u, v, q);
                    x[j + t] = sub_mod(u, v, q);
                }
            }
        }
    }
};

// Signed coefficients -> evaluation-domain limbs for the given prime indices.
std::vector<u64> to_eval(const HostTables &h, const int64_t *coef, const std::vector<u32> &pidx) {
    const u32 N = h.N;
    std::vector<u64> out((size_t)pidx.size() * N);
    HostNtt ntt(h);
    for (size_t r = 0; r < pidx.size(); ++r) {
        const u64 q = h.primes[pidx[r]];
        u64 *row = out.data() + r * N;
        for (u32 j = 0; j < N; ++j) {
            int64_t v = coef[j];
            row[j] = v >= 0 ? (u64)v % q : q - ((u64)(-v) % q);
            if (row[j] == q) row[j] = 0;
        }
        ntt.forward(row, pidx[r]);
    }
    return out;
}

std::vector<u32> key_primes(const HostTables &h, u32 key_limbs) {
    std::vector<u32> p;
    for (u32 i = 0; i < key_limbs; ++i) p.push_back(i);
    for (u32 j = 0; j < h.K; ++j) p.push_back(h.L + j);
    return p;
}

u32 galois_src(u32 i, u64 g, u32 logN) {
    u32 br = bit_reverse(i, logN);
    u64 e = (g * (2ull * br + 1)) % (2ull << logN);
    return bit_reverse((u32)((e - 1) / 2), logN);
}

// Writes the key for s_new -> s (both evaluation domain on key_primes) at evk.
int write_switch_key(fhea_ctx *c, const std::vector<u64> &s_e, const std::vector<u64> &sn_e,
                     u32 key_limbs, uint64_t seed, uint64_t evk) {
    const HostTables &h = c->h;
    const u32 N = h.N, K = h.K, dnum = h.dnum, A = h.alpha;
    const std::vector<u32> kp = key_primes(h, key_limbs);
    const u32 R = (u32)kp.size();
    const u64 words = 2ull * dnum * R * N;
    if (evk & 7) return FHEA_ERR_ALIGN;
    if (evk > c->key_words || words > c->key_words - evk) return FHEA_ERR_RANGE;
    std::vector<u64> key(words);
    u64 *ka = key.data(), *kb = key.data() + (size_t)dnum * R * N;
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> gauss(0.0, 3.2);
    std::vector<int64_t> e(N);
    // P mod q_i
    std::vector<u64> pmod(R);
    for (u32 r = 0; r < R; ++r) {
        u64 q = h.primes[kp[r]], pm = 1 % q;
        for (u32 j = 0; j < K; ++j) pm = (u64)((u128)pm * (h.primes[h.L + j] % q) % q);
        pmod[r] = pm;
    }
    for (u32 p = 0; p < dnum; ++p) {
        for (u32 j = 0; j < N; ++j) e[j] = (int64_t)std::llround(gauss(rng));
        std::vector<u64> e_e = to_eval(h, e.data(), kp);
        for (u32 r = 0; r < R; ++r) {
            const u64 q = h.primes[kp[r]];
            std::uniform_int_distribution<u64> uni(0, q - 1);
            const bool in_digit = kp[r] < key_limbs && kp[r] >= p * A && kp[r] < (p + 1) * A;
            u64 *ar = ka + ((size_t)p * R + r) * N, *br = kb + ((size_t)p * R + r) * N;
            const u64 *sr = s_e.data() + (size_t)r * N, *snr = sn_e.data() + (size_t)r * N;
            const u64 *er = e_e.data() + (size_t)r * N;
            for (u32 j = 0; j < N; ++j) {
                u64 a = uni(rng);
                ar[j] = a;
                u64 as = (u64)((u128)a * sr[j] % q);
                u64 b = add_mod(sub_mod(0, as, q), er[j], q);
                if (in_digit) b = add_mod(b, (u64)((u128)pmod[r] * snr[j] % q), q);
                br[j] = b;
            }
        }
    }
    return fhea_write_key(c, evk, key.data(), words);
}

std::vector<int64_t> s_to_i64(const int8_t *s, u32 N) {
    std::vector<int64_t> v(N);
    for (u32 j = 0; j < N; ++j) v[j] = s[j];
    return v;
}

}  // namespace

extern "C" {

int fhea_keygen_secret(const fhea_ctx *c, uint64_t seed, int8_t *s_out) {
    if (!c || !s_out) return FHEA_ERR_STATE;
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int> d(-1, 1);
    for (u32 j = 0; j < c->h.N; ++j) s_out[j] = (int8_t)d(rng);
    return FHEA_OK;
}

uint64_t fhea_key_size_words(const fhea_ctx *c, uint32_t key_limbs) {
    if (!c) return 0;
    return 2ull * c->h.dnum * (key_limbs + c->h.K) * c->h.N;
}

uint64_t fhea_galois_for_step(const fhea_ctx *c, int32_t step) {
    if (!c) return 0;
    const int64_t half = c->h.N / 2;
    int64_t e = ((int64_t)step % half + half) % half;
    return pow_mod(5, (u64)e, 2ull * c->h.N);
}

int fhea_keygen_switch_key(fhea_ctx *c, const int8_t *s, const int64_t *s_new, uint32_t key_limbs,
                           uint64_t seed, uint64_t evk) {
    if (!c || !c->poly) return FHEA_ERR_STATE;
    if (!s || !s_new || key_limbs == 0 || key_limbs > c->h.L) return FHEA_ERR_ARG;
    const std::vector<u32> kp = key_primes(c->h, key_limbs);
    std::vector<int64_t> si = s_to_i64(s, c->h.N);
    return write_switch_key(c, to_eval(c->h, si.data(), kp), to_eval(c->h, s_new, kp), key_limbs, seed,
                            evk);
}

int fhea_keygen_relin_key(fhea_ctx *c, const int8_t *s, uint32_t key_limbs, uint64_t seed, uint64_t evk) {
    if (!c || !c->poly) return FHEA_ERR_STATE;
    if (!s || key_limbs == 0 || key_limbs > c->h.L) return FHEA_ERR_ARG;
    const u32 N = c->h.N;
    const std::vector<u32> kp = key_primes(c->h, key_limbs);
    std::vector<int64_t> si = s_to_i64(s, N);
    std::vector<u64> s_e = to_eval(c->h, si.data(), kp), s2(s_e.size());
    for (size_t r = 0; r < kp.size(); ++r) {
        const u64 q = c->h.primes[kp[r]];
        for (u32 j = 0; j < N; ++j) s2[r * N + j] = (u64)((u128)s_e[r * N + j] * s_e[r * N + j] % q);
    }
    return write_switch_key(c, s_e, s2, key_limbs, seed, evk);
}

int fhea_keygen_rotation_key(fhea_ctx *c, const int8_t *s, uint64_t galois, uint32_t key_limbs,
                             uint64_t seed, uint64_t evk) {
    if (!c || !c->poly) return FHEA_ERR_STATE;
    if (!s || key_limbs == 0 || key_limbs > c->h.L) return FHEA_ERR_ARG;
    if ((galois & 1) == 0 || galois >= 2ull * c->h.N) return FHEA_ERR_ARG;
    const u32 N = c->h.N;
    const std::vector<u32> kp = key_primes(c->h, key_limbs);
    std::vector<int64_t> si = s_to_i64(s, N);
    std::vector<u64> s_e = to_eval(c->h, si.data(), kp), sg(s_e.size());
    // sigma_g(s) in the evaluation domain: out[i] = in[map_g(i)]
    std::vector<u32> map(N);
    for (u32 i = 0; i < N; ++i) map[i] = galois_src(i, galois, c->h.logN);
    for (size_t r = 0; r < kp.size(); ++r)
        for (u32 i = 0; i < N; ++i) sg[r * N + i] = s_e[r * N + map[i]];
    return write_switch_key(c, s_e, sg, key_limbs, seed, evk);
}

}  // extern "C"
