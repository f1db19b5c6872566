// fhea_client.cpp -- CPU client side of CKKS: secret / public / evaluation keys,
// encoding, encryption, decryption, decoding.
//
// The reference's client is OpenFHE (KeyGen, EvalMultKeyGen, EvalRotateKeyGen,
// MakeCKKSPackedPlaintext, Encrypt, Decrypt, GetRealPackedValue;
// integration_tools/generate_interactive_test_vectors.cpp:161-224,
// software_golden_model/fhe_core/fhe_context_manager.cpp:81-128,457-512).  OpenFHE
// is not available, so this file is a client of its own with documented
// conventions; the evaluation path never runs here.
//
// Keys.  Digit p of a key switching s' -> s is, over the Q limbs of the key level
// and all P limbs (evaluation domain):
//     a_p uniform,   b_p = -a_p s + e_p + [limb i in digit p] (P mod q_i) s'
// with e_p Gaussian (sigma 3.2).  Public key: (b, a) = (-a s + e, a) over the Q limbs.
//
// Encoding (canonical embedding).  zeta = exp(i pi / N).  Slot j holds
// m(zeta^(5^j mod 2N)) / scale, j < N/2; the conjugate slots are implied, so m
// has real (integer) coefficients.  Since zeta^(2t+1) runs over the odd powers,
//     m(zeta^(2t+1)) = sum_k (m_k zeta^k) omega^(t k),   omega = zeta^2,
// is a length-N DFT of the twisted coefficients; slot j is DFT index
// t_j = (5^j mod 2N - 1) / 2 and its conjugate t'_j = (2N - 5^j mod 2N - 1) / 2.
// Encoding runs the inverse DFT from those N values and rounds.  A vector of n
// values (n a power of two dividing N/2) is replicated with period n, so a
// rotation by r (the Galois element 5^r) rotates it cyclically.
//
// Decryption m = c0 + c1 s (evaluation domain); decoding lifts each coefficient
// from its residues to the centred integer by Garner's mixed-radix CRT, then
// runs the forward DFT.
#include <cmath>
#include <complex>
#include <cstring>
#include <algorithm>
#include <random>
#include <thread>
#include <vector>

#include "fhea_engine.hpp"

using namespace fhea;

struct fhea_client {
    HostTables h;
    std::vector<std::complex<double>> roots;  // exp(2 pi i k / N), k < N
    std::vector<std::complex<double>> zeta;   // exp(i pi k / N), k < N
    std::vector<u32> slot_t, slot_tc;          // DFT index of slot j and of its conjugate, j < N/2
    void init_embedding();
};

namespace {

typedef std::complex<double> cplx;
const double kPi = 3.14159265358979323846;

struct HostNtt {
    const HostTables &h;
    explicit HostNtt(const HostTables &ht) : h(ht) {}
    // forward negacyclic NTT of one limb (ntt.cpp:4-43), Shoup products
    void forward(u64 *a, u32 pidx) const {
        const u32 N = h.N;
        const u64 q = h.primes[pidx];
        const u64 *tw = h.tw.data() + (size_t)pidx * N, *tsh = h.tw_sh.data() + (size_t)pidx * N;
        u32 t = N >> 1;
        for (u32 m = 1; m < N; m <<= 1, t >>= 1) {
            for (u32 i = 0; i < m; ++i) {
                const u64 W = tw[m + i], Wsh = tsh[m + i];
                u64 *x = a + 2 * i * t;
                for (u32 j = 0; j < t; ++j) {
                    u64 u = x[j], v = mul_shoup(x[j + t], W, Wsh, q);
                    x[j] = add_mod(u, v, q);
                    x[j + t] = sub_mod(u, v, q);
                }
            }
        }
    }
    // inverse (ntt.cpp:46-80): Gentleman-Sande with itw, then N^-1
    void inverse(u64 *a, u32 pidx) const {
        const u32 N = h.N;
        const u64 q = h.primes[pidx];
        const u64 *itw = h.itw.data() + (size_t)pidx * N, *ish = h.itw_sh.data() + (size_t)pidx * N;
        u32 t = 1;
        for (u32 m = N >> 1; m >= 1; m >>= 1, t <<= 1) {
            for (u32 i = 0; i < m; ++i) {
                const u64 W = itw[m + i], Wsh = ish[m + i];
                u64 *x = a + 2 * i * t;
                for (u32 j = 0; j < t; ++j) {
                    u64 u = x[j], v = x[j + t];
                    x[j] = add_mod(u, v, q);
                    x[j + t] = mul_shoup(sub_mod(u, v, q), W, Wsh, q);
                }
            }
        }
        const u64 ni = h.ninv[pidx], nish = shoup_companion(ni, q);
        for (u32 j = 0; j < N; ++j) a[j] = mul_shoup(a[j], ni, nish, q);
    }
};

u64 to_residue(int64_t v, u64 q) {
    u64 r = v >= 0 ? (u64)v % q : q - ((u64)(-(v + 1)) % q) - 1;
    return r;
}

// Signed coefficients -> evaluation-domain limbs for the given prime indices.
std::vector<u64> to_eval(const HostTables &h, const int64_t *coef, const std::vector<u32> &pidx) {
    const u32 N = h.N;
    std::vector<u64> out((size_t)pidx.size() * N);
    HostNtt ntt(h);
    for (size_t r = 0; r < pidx.size(); ++r) {
        const u64 q = h.primes[pidx[r]];
        u64 *row = out.data() + r * N;
        for (u32 j = 0; j < N; ++j) row[j] = to_residue(coef[j], q);
        ntt.forward(row, pidx[r]);
    }
    return out;
}

std::vector<u32> q_primes(u32 limbs) {
    std::vector<u32> p(limbs);
    for (u32 i = 0; i < limbs; ++i) p[i] = i;
    return p;
}

std::vector<u32> key_primes(const HostTables &h, u32 key_limbs) {
    std::vector<u32> p = q_primes(key_limbs);
    for (u32 j = 0; j < h.K; ++j) p.push_back(h.L + j);
    return p;
}

u32 galois_src(u32 i, u64 g, u32 logN) {
    u32 br = bit_reverse(i, logN);
    u64 e = (g * (2ull * br + 1)) % (2ull << logN);
    return bit_reverse((u32)((e - 1) / 2), logN);
}

std::vector<int64_t> s_to_i64(const int8_t *s, u32 N) {
    std::vector<int64_t> v(N);
    for (u32 j = 0; j < N; ++j) v[j] = s[j];
    return v;
}

void gaussian(std::mt19937_64 &rng, std::vector<int64_t> &e) {
    std::normal_distribution<double> g(0.0, 3.2);
    for (auto &x : e) x = (int64_t)std::llround(g(rng));
}

// Key for s_new -> s (both evaluation domain on key_primes), host layout
// [a_0..a_{dnum-1} | b_0..b_{dnum-1}], each [key_limbs + K][N].
std::vector<u64> make_switch_key(const HostTables &h, const std::vector<u64> &s_e, const std::vector<u64> &sn_e,
                                 u32 key_limbs, uint64_t seed) {
    const u32 N = h.N, K = h.K, dnum = h.dnum, A = h.alpha;
    const std::vector<u32> kp = key_primes(h, key_limbs);
    const u32 R = (u32)kp.size();
    std::vector<u64> key(2ull * dnum * R * N);
    u64 *ka = key.data(), *kb = key.data() + (size_t)dnum * R * N;
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> gauss(0.0, 3.2);
    std::vector<int64_t> e(N);
    std::vector<u64> pmod(R);  // P mod q_i
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
            const ModParams mq = make_mod_params(q);
            std::uniform_int_distribution<u64> uni(0, q - 1);
            const bool in_digit = kp[r] < key_limbs && kp[r] >= p * A && kp[r] < (p + 1) * A;
            u64 *ar = ka + ((size_t)p * R + r) * N, *br = kb + ((size_t)p * R + r) * N;
            const u64 *sr = s_e.data() + (size_t)r * N, *snr = sn_e.data() + (size_t)r * N;
            const u64 *er = e_e.data() + (size_t)r * N;
            for (u32 j = 0; j < N; ++j) {
                u64 a = uni(rng);
                ar[j] = a;
                u64 as = mulmod(a, sr[j], mq);
                u64 b = add_mod(sub_mod(0, as, q), er[j], q);
                if (in_digit) b = add_mod(b, mulmod(pmod[r], snr[j], mq), q);
                br[j] = b;
            }
        }
    }
    return key;
}

int key_for(const HostTables &h, int kind, const int8_t *s, const int64_t *s_new, u64 galois, u32 key_limbs,
            uint64_t seed, std::vector<u64> &out) {
    if (!s || key_limbs == 0 || key_limbs > h.L) return FHEA_ERR_ARG;
    const u32 N = h.N;
    const std::vector<u32> kp = key_primes(h, key_limbs);
    std::vector<int64_t> si = s_to_i64(s, N);
    std::vector<u64> s_e = to_eval(h, si.data(), kp), sn;
    if (kind == 0) {  // explicit s_new
        if (!s_new) return FHEA_ERR_ARG;
        sn = to_eval(h, s_new, kp);
    } else if (kind == 1) {  // relinearisation: s^2
        sn.resize(s_e.size());
        for (size_t r = 0; r < kp.size(); ++r) {
            const ModParams mq = make_mod_params(h.primes[kp[r]]);
            for (u32 j = 0; j < N; ++j) sn[r * N + j] = mulmod(s_e[r * N + j], s_e[r * N + j], mq);
        }
    } else {  // rotation: sigma_g(s), out[i] = in[map_g(i)] in the evaluation domain
        if ((galois & 1) == 0 || galois >= 2ull * N) return FHEA_ERR_ARG;
        sn.resize(s_e.size());
        std::vector<u32> map(N);
        for (u32 i = 0; i < N; ++i) map[i] = galois_src(i, galois, h.logN);
        for (size_t r = 0; r < kp.size(); ++r)
            for (u32 i = 0; i < N; ++i) sn[r * N + i] = s_e[r * N + map[i]];
    }
    out = make_switch_key(h, s_e, sn, key_limbs, seed);
    return FHEA_OK;
}

int write_key_to_arena(fhea_ctx *c, const std::vector<u64> &key, uint64_t evk) {
    if (evk & 7) return FHEA_ERR_ALIGN;
    if (evk > c->key_words || key.size() > c->key_words - evk) return FHEA_ERR_RANGE;
    return fhea_write_key(c, evk, key.data(), key.size());
}

int secret(u32 N, uint64_t seed, int8_t *s_out) {
    if (!s_out) return FHEA_ERR_ARG;
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int> d(-1, 1);
    for (u32 j = 0; j < N; ++j) s_out[j] = (int8_t)d(rng);
    return FHEA_OK;
}

// ---- canonical embedding

// In-place iterative radix-2 DFT of size n = roots.size(),
// X_t = sum_k x_k exp(sign 2 pi i t k / n).
void dft(std::vector<cplx> &a, int sign, const std::vector<cplx> &roots) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t half = len >> 1, step = n / len;
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < half; ++k) {
                const cplx w = sign > 0 ? roots[k * step] : std::conj(roots[k * step]);
                cplx u = a[i + k], v = a[i + k + half] * w;
                a[i + k] = u + v;
                a[i + k + half] = u - v;
            }
    }
}

// DFT index of slot j and of its conjugate
void slot_index(u32 N, u32 j, u32 &t, u32 &tc) {
    const u64 e = pow_mod(5, j, 2ull * N);
    t = (u32)((e - 1) / 2);
    tc = (u32)((2ull * N - e - 1) / 2);
}

bool valid_slots(u32 N, u32 n) { return n >= 1 && n <= N / 2 && (n & (n - 1)) == 0; }

// Integer-valued double -> residue mod q, exact for any magnitude: |v| = mant * 2^e
// with a 53-bit integer mantissa.
u64 double_to_residue(double v, u64 q) {
    const double a = std::fabs(v);
    u64 r;
    if (a < 4.0e18) {
        r = (u64)(int64_t)a % q;
    } else {
        int e = 0;
        const double m = std::frexp(a, &e);  // a = m 2^e, m in [0.5, 1)
        const u64 mant = (u64)std::ldexp(m, 53);
        r = (u64)((u128)(mant % q) * pow_mod(2, (u64)(e - 53), q) % q);
    }
    return (v < 0 && r) ? q - r : r;
}

// Encodes n values (replicated to N/2 slots) into `limbs` evaluation-domain limbs.
int encode_one(const fhea_client *cl, const double *re, const double *im, u32 n, u32 limbs, double scale,
               u64 *pt) {
    const HostTables &h = cl->h;
    const u32 N = h.N;
    std::vector<cplx> Y(N);
    for (u32 j = 0; j < N / 2; ++j) {
        const cplx z(re[j % n] * scale, im ? im[j % n] * scale : 0.0);
        Y[cl->slot_t[j]] = z;
        Y[cl->slot_tc[j]] = std::conj(z);
    }
    dft(Y, -1, cl->roots);
    std::vector<double> m(N);
    for (u32 k = 0; k < N; ++k) {
        const double v = std::nearbyint((Y[k] * std::conj(cl->zeta[k])).real() / (double)N);
        if (!std::isfinite(v) || std::fabs(v) > 0x1p1000) return FHEA_ERR_ARG;
        m[k] = v;
    }
    HostNtt ntt(h);
    for (u32 i = 0; i < limbs; ++i) {
        const u64 q = h.primes[i];
        u64 *row = pt + (size_t)i * N;
        for (u32 k = 0; k < N; ++k) row[k] = double_to_residue(m[k], q);
        ntt.forward(row, i);
    }
    return FHEA_OK;
}

// centred integer value of residues x[0..l) (Garner), as a double
struct Crt {
    std::vector<u64> q;
    std::vector<ModParams> mp;
    std::vector<u64> inv;  // inv[i*l + k] = q_k^-1 mod q_i
    u32 l;
    Crt(const HostTables &h, u32 limbs) : l(limbs) {
        q.assign(h.primes.begin(), h.primes.begin() + limbs);
        for (u64 x : q) mp.push_back(make_mod_params(x));
        inv.assign((size_t)l * l, 0);
        for (u32 i = 0; i < l; ++i)
            for (u32 k = 0; k < i; ++k) inv[(size_t)i * l + k] = inv_mod(q[k] % q[i], q[i]);
    }
    double value(const u64 *x, size_t stride) const {
        u64 v[256];
        for (u32 i = 0; i < l; ++i) {
            u64 t = x[i * stride];
            for (u32 k = 0; k < i; ++k) {
                const u64 d = sub_mod(t, v[k] % q[i], q[i]);
                t = mulmod(d, inv[(size_t)i * l + k], mp[i]);
            }
            v[i] = t;
        }
        const bool neg = v[l - 1] > (q[l - 1] - 1) / 2;
        double acc = 0.0;
        for (int i = (int)l - 1; i >= 0; --i) {
            const double d = neg ? -(double)(q[i] - 1 - v[i]) : (double)v[i];
            acc = (i == (int)l - 1) ? d : acc * (double)q[i] + d;
        }
        return neg ? acc - 1.0 : acc;
    }
};

}  // namespace

void fhea_client::init_embedding() {
    const u32 N = h.N;
    roots.resize(N);
    zeta.resize(N);
    for (u32 k = 0; k < N; ++k) {
        const double a = 2.0 * kPi * (double)k / (double)N, b = kPi * (double)k / (double)N;
        roots[k] = cplx(std::cos(a), std::sin(a));
        zeta[k] = cplx(std::cos(b), std::sin(b));
    }
    slot_t.resize(N / 2);
    slot_tc.resize(N / 2);
    for (u32 j = 0; j < N / 2; ++j) slot_index(N, j, slot_t[j], slot_tc[j]);
}

extern "C" {

// ---- device-context forms (keys go straight into the key arena)

int fhea_keygen_secret(const fhea_ctx *c, uint64_t seed, int8_t *s_out) {
    if (!c) return FHEA_ERR_STATE;
    return secret(c->h.N, seed, s_out);
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
    std::vector<u64> key;
    int st = key_for(c->h, 0, s, s_new, 0, key_limbs, seed, key);
    return st ? st : write_key_to_arena(c, key, evk);
}

int fhea_keygen_relin_key(fhea_ctx *c, const int8_t *s, uint32_t key_limbs, uint64_t seed, uint64_t evk) {
    if (!c || !c->poly) return FHEA_ERR_STATE;
    std::vector<u64> key;
    int st = key_for(c->h, 1, s, nullptr, 0, key_limbs, seed, key);
    return st ? st : write_key_to_arena(c, key, evk);
}

int fhea_keygen_rotation_key(fhea_ctx *c, const int8_t *s, uint64_t galois, uint32_t key_limbs,
                             uint64_t seed, uint64_t evk) {
    if (!c || !c->poly) return FHEA_ERR_STATE;
    std::vector<u64> key;
    int st = key_for(c->h, 2, s, nullptr, galois, key_limbs, seed, key);
    return st ? st : write_key_to_arena(c, key, evk);
}

// ---- host-only client

int fhea_client_create(const fhea_params *p, fhea_client **out) {
    if (!out) return FHEA_ERR_ARG;
    *out = nullptr;
    fhea_client *cl = new fhea_client();
    const int st = host_tables_from_params(p, 16, cl->h);
    if (st != FHEA_OK) {
        delete cl;
        return st;
    }
    cl->init_embedding();
    *out = cl;
    return FHEA_OK;
}

int fhea_client_from_ctx(const fhea_ctx *c, fhea_client **out) {
    if (!c || !out) return FHEA_ERR_STATE;
    fhea_client *cl = new fhea_client();
    cl->h = c->h;
    cl->init_embedding();
    *out = cl;
    return FHEA_OK;
}

int fhea_client_destroy(fhea_client *cl) {
    delete cl;
    return FHEA_OK;
}

int fhea_client_info(const fhea_client *cl, uint32_t *info5) {
    if (!cl) return FHEA_ERR_STATE;
    if (!info5) return FHEA_ERR_ARG;
    info5[0] = cl->h.N;
    info5[1] = cl->h.L;
    info5[2] = cl->h.K;
    info5[3] = cl->h.dnum;
    info5[4] = cl->h.alpha;
    return FHEA_OK;
}

int fhea_client_primes(const fhea_client *cl, uint64_t *out) {
    if (!cl) return FHEA_ERR_STATE;
    if (!out) return FHEA_ERR_ARG;
    std::memcpy(out, cl->h.primes.data(), cl->h.primes.size() * 8);
    return FHEA_OK;
}

int fhea_client_table(const fhea_client *cl, int which, uint64_t *out, uint64_t count) {
    if (!cl) return FHEA_ERR_STATE;
    if (!out) return FHEA_ERR_ARG;
    const std::vector<u64> *v = which == 0 ? &cl->h.psi : which == 1 ? &cl->h.ninv
                              : which == 2 ? &cl->h.tw : which == 3 ? &cl->h.itw : nullptr;
    if (!v) return FHEA_ERR_ARG;
    if (count != v->size()) return FHEA_ERR_RANGE;
    std::memcpy(out, v->data(), count * 8);
    return FHEA_OK;
}

int fhea_client_keygen_secret(const fhea_client *cl, uint64_t seed, int8_t *s_out) {
    if (!cl) return FHEA_ERR_STATE;
    return secret(cl->h.N, seed, s_out);
}

int fhea_client_switch_key(const fhea_client *cl, const int8_t *s, const int64_t *s_new, uint32_t key_limbs,
                           uint64_t seed, uint64_t *key_out) {
    if (!cl) return FHEA_ERR_STATE;
    if (!key_out) return FHEA_ERR_ARG;
    std::vector<u64> key;
    int st = key_for(cl->h, 0, s, s_new, 0, key_limbs, seed, key);
    if (!st) std::memcpy(key_out, key.data(), key.size() * 8);
    return st;
}

int fhea_client_relin_key(const fhea_client *cl, const int8_t *s, uint32_t key_limbs, uint64_t seed,
                          uint64_t *key_out) {
    if (!cl) return FHEA_ERR_STATE;
    if (!key_out) return FHEA_ERR_ARG;
    std::vector<u64> key;
    int st = key_for(cl->h, 1, s, nullptr, 0, key_limbs, seed, key);
    if (!st) std::memcpy(key_out, key.data(), key.size() * 8);
    return st;
}

int fhea_client_rotation_key(const fhea_client *cl, const int8_t *s, uint64_t galois, uint32_t key_limbs,
                             uint64_t seed, uint64_t *key_out) {
    if (!cl) return FHEA_ERR_STATE;
    if (!key_out) return FHEA_ERR_ARG;
    std::vector<u64> key;
    int st = key_for(cl->h, 2, s, nullptr, galois, key_limbs, seed, key);
    if (!st) std::memcpy(key_out, key.data(), key.size() * 8);
    return st;
}

int fhea_encode(const fhea_client *cl, const double *re, const double *im, uint32_t n, uint32_t limbs,
                double scale, uint64_t *pt) {
    if (!cl) return FHEA_ERR_STATE;
    const HostTables &h = cl->h;
    if (!re || !pt || !valid_slots(h.N, n) || limbs == 0 || limbs > h.L || !(scale > 0.0)) return FHEA_ERR_ARG;
    return encode_one(cl, re, im, n, limbs, scale, pt);
}

int fhea_encode_batch(const fhea_client *cl, const double *re, const double *im, uint32_t count, uint32_t n,
                      uint32_t limbs, double scale, uint64_t *pt, uint32_t threads) {
    if (!cl) return FHEA_ERR_STATE;
    const HostTables &h = cl->h;
    if (!re || !pt || !valid_slots(h.N, n) || limbs == 0 || limbs > h.L || !(scale > 0.0)) return FHEA_ERR_ARG;
    if (count == 0) return FHEA_OK;
    u32 T = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
    T = std::min<u32>(T, count);
    std::vector<int> st(T, FHEA_OK);
    std::vector<std::thread> pool;
    const size_t W = (size_t)limbs * h.N;
    for (u32 t = 0; t < T; ++t)
        pool.emplace_back([&, t]() {
            for (u32 k = t; k < count; k += T) {
                const int r = encode_one(cl, re + (size_t)k * n, im ? im + (size_t)k * n : nullptr, n, limbs, scale,
                                         pt + k * W);
                if (r != FHEA_OK) {
                    st[t] = r;
                    return;
                }
            }
        });
    for (auto &th : pool) th.join();
    for (int r : st)
        if (r != FHEA_OK) return r;
    return FHEA_OK;
}

int fhea_decode(const fhea_client *cl, const uint64_t *pt, uint32_t limbs, double scale, uint32_t n,
                double *re, double *im) {
    if (!cl) return FHEA_ERR_STATE;
    const HostTables &h = cl->h;
    const u32 N = h.N;
    if (!pt || !re || !valid_slots(N, n) || limbs == 0 || limbs > h.L || !(scale > 0.0)) return FHEA_ERR_ARG;
    std::vector<u64> c(pt, pt + (size_t)limbs * N);
    HostNtt ntt(h);
    for (u32 i = 0; i < limbs; ++i) ntt.inverse(c.data() + (size_t)i * N, i);
    Crt crt(h, limbs);
    std::vector<cplx> Y(N);
    for (u32 k = 0; k < N; ++k) Y[k] = crt.value(c.data() + k, N) * cl->zeta[k];
    dft(Y, +1, cl->roots);
    for (u32 j = 0; j < n; ++j) {
        const u32 t = cl->slot_t[j];
        re[j] = Y[t].real() / scale;
        if (im) im[j] = Y[t].imag() / scale;
    }
    return FHEA_OK;
}

int fhea_keygen_public(const fhea_client *cl, const int8_t *s, uint64_t seed, uint64_t *pk) {
    if (!cl) return FHEA_ERR_STATE;
    const HostTables &h = cl->h;
    const u32 N = h.N, L = h.L;
    if (!s || !pk) return FHEA_ERR_ARG;
    std::vector<int64_t> si = s_to_i64(s, N), e(N);
    const std::vector<u32> qp = q_primes(L);
    std::vector<u64> s_e = to_eval(h, si.data(), qp);
    std::mt19937_64 rng(seed);
    gaussian(rng, e);
    std::vector<u64> e_e = to_eval(h, e.data(), qp);
    u64 *b = pk, *a = pk + (size_t)L * N;
    for (u32 i = 0; i < L; ++i) {
        const u64 q = h.primes[i];
        const ModParams mq = make_mod_params(q);
        std::uniform_int_distribution<u64> uni(0, q - 1);
        for (u32 j = 0; j < N; ++j) {
            const size_t o = (size_t)i * N + j;
            a[o] = uni(rng);
            b[o] = add_mod(sub_mod(0, mulmod(a[o], s_e[o], mq), q), e_e[o], q);
        }
    }
    return FHEA_OK;
}

int fhea_encrypt(const fhea_client *cl, const uint64_t *pk, const uint64_t *pt, uint32_t limbs, uint64_t seed,
                 uint64_t *ct) {
    if (!cl) return FHEA_ERR_STATE;
    const HostTables &h = cl->h;
    const u32 N = h.N, L = h.L;
    if (!pk || !pt || !ct || limbs == 0 || limbs > L) return FHEA_ERR_ARG;
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int> tern(-1, 1);
    std::vector<int64_t> v(N), e0(N), e1(N);
    for (auto &x : v) x = tern(rng);
    gaussian(rng, e0);
    gaussian(rng, e1);
    const std::vector<u32> qp = q_primes(limbs);
    std::vector<u64> v_e = to_eval(h, v.data(), qp), e0_e = to_eval(h, e0.data(), qp),
                     e1_e = to_eval(h, e1.data(), qp);
    const u64 *b = pk, *a = pk + (size_t)L * N;
    u64 *c0 = ct, *c1 = ct + (size_t)limbs * N;
    for (u32 i = 0; i < limbs; ++i) {
        const u64 q = h.primes[i];
        const ModParams mq = make_mod_params(q);
        for (u32 j = 0; j < N; ++j) {
            const size_t o = (size_t)i * N + j;
            c0[o] = add_mod(add_mod(mulmod(v_e[o], b[o], mq), e0_e[o], q), pt[o], q);
            c1[o] = add_mod(mulmod(v_e[o], a[o], mq), e1_e[o], q);
        }
    }
    return FHEA_OK;
}

int fhea_decrypt(const fhea_client *cl, const int8_t *s, const uint64_t *ct, uint32_t limbs, uint64_t *pt) {
    if (!cl) return FHEA_ERR_STATE;
    const HostTables &h = cl->h;
    const u32 N = h.N;
    if (!s || !ct || !pt || limbs == 0 || limbs > h.L) return FHEA_ERR_ARG;
    std::vector<int64_t> si = s_to_i64(s, N);
    std::vector<u64> s_e = to_eval(h, si.data(), q_primes(limbs));
    const u64 *c0 = ct, *c1 = ct + (size_t)limbs * N;
    for (u32 i = 0; i < limbs; ++i) {
        const u64 q = h.primes[i];
        const ModParams mq = make_mod_params(q);
        for (u32 j = 0; j < N; ++j) {
            const size_t o = (size_t)i * N + j;
            pt[o] = add_mod(c0[o], mulmod(c1[o], s_e[o], mq), q);
        }
    }
    return FHEA_OK;
}

}  // extern "C"
