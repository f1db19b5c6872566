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
// with e_p rounded Gaussian (sigma 3.2).  Public key: (b, a) = (-a s + e, a) over the Q limbs.
//
// Encoding (canonical embedding).  zeta = exp(i pi / N).  Slot j holds
// m(zeta^(5^j mod 2N)) / scale, j < N/2; the conjugate slots are implied, so m
// has real (integer) coefficients.  Every 5^j is 1 mod 4, and the powers of 5
// run over all N/2 such residues mod 2N, so the slots are m at zeta^(4u+1),
// u < M = N/2.  With zeta^((4u+1)M) = i,
//     m(zeta^(4u+1)) = sum_{k<M} w_k omega^(u k),
//     w_k = (m_k + i m_{k+M}) zeta^k,   omega = zeta^4 = exp(2 pi i / M),
// is one length-M DFT; slot j is DFT index u_j = (5^j mod 2N - 1) / 4.  Encoding
// runs the inverse DFT from the M slot values, untwists w_k and rounds its real
// and imaginary parts to m_k and m_{k+M}.  A vector of n values (n a power of
// two dividing N/2) is replicated with period n, so a rotation by r (the Galois
// element 5^r) rotates it cyclically.
//
// Decryption m = c0 + c1 s (evaluation domain); decoding lifts each coefficient
// from its residues to the centred integer by Garner's mixed-radix CRT, then
// runs the forward DFT.
#include <cmath>
#include <cstring>
#include <algorithm>
#include <random>
#include <thread>
#include <vector>

#include <cerrno>
#include <fcntl.h>
#include <sys/random.h>
#include <unistd.h>

#include "fhea_engine.hpp"

using namespace fhea;

struct fhea_client {
    HostTables h;
    // canonical embedding as one DFT of length M = N/2 (split real / imaginary parts)
    std::vector<double> tw_re, tw_im;      // stage of length 2h: exp(2 pi i k / 2h) at [h + k], k < h
    std::vector<double> zeta_re, zeta_im;  // exp(i pi k / N), k < M
    std::vector<u32> slot_u;               // DFT index of slot j, j < M
    std::vector<u32> brev;                 // bit reversal on log2(M) bits
    void init_embedding();
};

namespace {

const double kPi = 3.14159265358979323846;

// FP64 form of the CPU transforms for q < 2^50 (the device's exact products:
// for |v| <= 8q, v w - rint(v w/q) q is computed exactly as fma(-k, q, v w) plus the
// low part fma(v, w, -v w), and is at most q/2 + |v|/4 in magnitude).  Values are
// centred every stage (|x| <= q/2 + 2, and |x| <= q on the inverse's products).
// Four lanes per AVX2 instruction, the stages with butterfly distance 1 and 2
// looping over the groups so that they vectorize too; the words equal the integer
// transform's.  Every function here keeps products and sums as written (no
// contraction into other FMAs).
__attribute__((target("avx2,fma"))) inline double hmulmod(double v, double w, double wq, double q) {
#pragma clang fp contract(off)
    const double k = std::rint(v * wq);
    const double hi = v * w;
    const double lo = std::fma(v, w, -hi);
    return std::fma(-k, q, hi) + lo;
}
__attribute__((target("avx2,fma"))) inline double hreduce(double x, double q, double qinv) {
#pragma clang fp contract(off)
    return std::fma(-std::rint(x * qinv), q, x);
}
// [0, 2^52) integers <-> doubles, exact (exponent-bias trick)
inline double u52_to_f64(u64 x) {
    double d;
    const u64 b = x | 0x4330000000000000ull;
    std::memcpy(&d, &b, 8);
    return d - 4503599627370496.0;
}
inline u64 f64_to_u52(double y) {
    const double z = y + 4503599627370496.0;
    u64 b;
    std::memcpy(&b, &z, 8);
    return b & 0x000FFFFFFFFFFFFFull;
}
__attribute__((target("avx2,fma"))) inline void fwd_bfly(double &u, double &v, double w, double wq, double q,
                                                        double qinv) {
#pragma clang fp contract(off)
    const double r = hmulmod(v, w, wq, q), a = u;
    u = hreduce(a + r, q, qinv);
    v = hreduce(a - r, q, qinv);
}
__attribute__((target("avx2,fma"))) inline void fwd_bfly_lazy(double &u, double &v, double w, double wq, double q) {
#pragma clang fp contract(off)
    const double r = hmulmod(v, w, wq, q), a = u;
    u = a + r;
    v = a - r;
}
__attribute__((target("avx2,fma"))) inline void inv_bfly(double &u, double &v, double w, double wq, double q,
                                                        double qinv) {
#pragma clang fp contract(off)
    const double a = u, b = v;
    u = hreduce(a + b, q, qinv);
    v = hmulmod(a - b, w, wq, q);
}
__attribute__((target("avx2,fma"))) void ntt_forward_f64_host(u64 *a, u32 N, u64 q, const double *W,
                                                             const double *WQ) {
#pragma clang fp contract(off)
    std::vector<double> xv(N);
    double *x = xv.data();
    const double qd = (double)q, qinv = 1.0 / qd;
    for (u32 j = 0; j < N; ++j) x[j] = u52_to_f64(a[j]);
    // centred every fourth stage and at the last: from |x| < q three unreduced stages
    // reach 3.86q (< 8q, the products' bound), from |x| <= q/2 + 2 they reach 2.9q
    u32 t = N >> 1, st = 0;
    for (u32 m = 1; m < N; m <<= 1, t >>= 1, ++st) {
        const bool red = (st & 3) == 3 || 2 * m == N;
#define FHEA_FWD_BFLY(U, V, W_, WQ_)                      \
    if (red)                                              \
        fwd_bfly(U, V, W_, WQ_, qd, qinv);                \
    else                                                  \
        fwd_bfly_lazy(U, V, W_, WQ_, qd);
        if (t >= 4) {
            for (u32 i = 0; i < m; ++i) {
                const double w = W[m + i], wq = WQ[m + i];
                double *xa = x + 2 * i * t, *xb = xa + t;
                if (red)
                    for (u32 j = 0; j < t; ++j) fwd_bfly(xa[j], xb[j], w, wq, qd, qinv);
                else
                    for (u32 j = 0; j < t; ++j) fwd_bfly_lazy(xa[j], xb[j], w, wq, qd);
            }
        } else if (t == 2) {
            for (u32 i = 0; i < m; ++i) {
                double *xa = x + 4 * i;
                FHEA_FWD_BFLY(xa[0], xa[2], W[m + i], WQ[m + i])
                FHEA_FWD_BFLY(xa[1], xa[3], W[m + i], WQ[m + i])
            }
        } else {
            for (u32 i = 0; i < m; ++i) {
                FHEA_FWD_BFLY(x[2 * i], x[2 * i + 1], W[m + i], WQ[m + i])
            }
        }
#undef FHEA_FWD_BFLY
    }
    for (u32 j = 0; j < N; ++j) a[j] = f64_to_u52(x[j] < 0.0 ? x[j] + qd : x[j]);
}
__attribute__((target("avx2,fma"))) void ntt_inverse_f64_host(u64 *a, u32 N, u64 q, const double *W,
                                                             const double *WQ, double ni, double niq) {
#pragma clang fp contract(off)
    std::vector<double> xv(N);
    double *x = xv.data();
    const double qd = (double)q, qinv = 1.0 / qd;
    for (u32 j = 0; j < N; ++j) x[j] = u52_to_f64(a[j]);
    u32 t = 1;
    for (u32 m = N >> 1; m >= 1; m >>= 1, t <<= 1) {
        if (t == 1) {
            for (u32 i = 0; i < m; ++i) inv_bfly(x[2 * i], x[2 * i + 1], W[m + i], WQ[m + i], qd, qinv);
        } else if (t == 2) {
            for (u32 i = 0; i < m; ++i) {
                double *xa = x + 4 * i;
                inv_bfly(xa[0], xa[2], W[m + i], WQ[m + i], qd, qinv);
                inv_bfly(xa[1], xa[3], W[m + i], WQ[m + i], qd, qinv);
            }
        } else {
            for (u32 i = 0; i < m; ++i) {
                const double w = W[m + i], wq = WQ[m + i];
                double *xa = x + 2 * i * t, *xb = xa + t;
                for (u32 j = 0; j < t; ++j) inv_bfly(xa[j], xb[j], w, wq, qd, qinv);
            }
        }
    }
    for (u32 j = 0; j < N; ++j) {
        const double y = hreduce(hmulmod(x[j], ni, niq, qd), qd, qinv);
        a[j] = f64_to_u52(y < 0.0 ? y + qd : y);
    }
}

// c = c0 + c1 s mod q, canonical inputs and output (|c1 s - k q| <= 0.875q exactly,
// then the sum is centred and made canonical)
__attribute__((target("avx2,fma"))) void dec_mul_f64(u64 *c, const u64 *c0, const u64 *c1, const u64 *s, u32 N,
                                                    u64 q) {
#pragma clang fp contract(off)
    const double qd = (double)q, qinv = 1.0 / qd;
    for (u32 j = 0; j < N; ++j) {
        const double a = u52_to_f64(c1[j]), b = u52_to_f64(s[j]);
        const double hi = a * b;
        const double lo = std::fma(a, b, -hi);
        const double k = std::rint(hi * qinv);
        const double y = hreduce(std::fma(-k, qd, hi) + lo + u52_to_f64(c0[j]), qd, qinv);
        c[j] = f64_to_u52(y < 0.0 ? y + qd : y);
    }
}

// c0 = v b + e0 + m, c1 = v a + e1 (mod q < 2^50), every operand canonical: the
// product as in dec_mul_f64 (|r| <= 0.875q), the sums below 3q, one reduction.
__attribute__((target("avx2,fma"))) void enc_mul_f64(u64 *c0, u64 *c1, const u64 *v, const u64 *b, const u64 *a,
                                                    const u64 *e0, const u64 *e1, const u64 *m, u32 N, u64 q) {
#pragma clang fp contract(off)
    const double qd = (double)q, qinv = 1.0 / qd;
    for (u32 j = 0; j < N; ++j) {
        const double x = u52_to_f64(v[j]), pb = u52_to_f64(b[j]), pa = u52_to_f64(a[j]);
        const double hb = x * pb, lb = std::fma(x, pb, -hb), kb = std::rint(hb * qinv);
        const double ha = x * pa, la = std::fma(x, pa, -ha), ka = std::rint(ha * qinv);
        const double y0 = hreduce(std::fma(-kb, qd, hb) + lb + u52_to_f64(e0[j]) + u52_to_f64(m[j]), qd, qinv);
        const double y1 = hreduce(std::fma(-ka, qd, ha) + la + u52_to_f64(e1[j]), qd, qinv);
        c0[j] = f64_to_u52(y0 < 0.0 ? y0 + qd : y0);
        c1[j] = f64_to_u52(y1 < 0.0 ? y1 + qd : y1);
    }
}

// the FP64 twiddles of prime pidx (built on first use)
const HostF64Twiddles::Prime &f64_twiddles(const HostTables &h, u32 pidx) {
    HostF64Twiddles::Prime &P = h.f64tw->p[pidx];
    std::call_once(P.once, [&] {
        const u32 N = h.N;
        const double q = (double)h.primes[pidx];
        P.w.resize(N);
        P.wq.resize(N);
        P.iw.resize(N);
        P.iwq.resize(N);
        for (u32 k = 0; k < N; ++k) {
            P.w[k] = (double)h.tw[(size_t)pidx * N + k];
            P.wq[k] = P.w[k] / q;
            P.iw[k] = (double)h.itw[(size_t)pidx * N + k];
            P.iwq[k] = P.iw[k] / q;
        }
    });
    return P;
}

bool host_f64_ok() {
    static const bool ok = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
    return ok;
}

struct HostNtt {
    const HostTables &h;
    explicit HostNtt(const HostTables &ht) : h(ht) {}
    // forward negacyclic NTT of one limb (ntt.cpp:4-43), Shoup products, lazy
    // butterflies: values stay in [0, 4q) (q < 2^62), canonical at the end
    bool f64(u64 q) const { return h.host_ntt_mode == 0 && h.f64tw && q < (1ull << 50) && host_f64_ok(); }
    void forward(u64 *a, u32 pidx) const {
        const u32 N = h.N;
        const u64 q = h.primes[pidx], q2 = 2 * q;
        const u64 *tw = h.tw.data() + (size_t)pidx * N, *tsh = h.tw_sh.data() + (size_t)pidx * N;
        if (f64(q)) {
            const HostF64Twiddles::Prime &P = f64_twiddles(h, pidx);
            return ntt_forward_f64_host(a, N, q, P.w.data(), P.wq.data());
        }
        u32 t = N >> 1;
        for (u32 m = 1; m < N; m <<= 1, t >>= 1) {
            for (u32 i = 0; i < m; ++i) {
                const u64 W = tw[m + i], Wsh = tsh[m + i];
                u64 *x = a + 2 * i * t;
                for (u32 j = 0; j < t; ++j) {
                    u64 u = x[j];
                    if (u >= q2) u -= q2;
                    const u64 v = mul_shoup_lazy(x[j + t], W, Wsh, q);  // [0, 2q)
                    x[j] = u + v;
                    x[j + t] = u - v + q2;
                }
            }
        }
        for (u32 j = 0; j < N; ++j) {
            u64 v = a[j];
            if (v >= q2) v -= q2;
            if (v >= q) v -= q;
            a[j] = v;
        }
    }
    // inverse (ntt.cpp:46-80): Gentleman-Sande with itw, then N^-1; values stay
    // in [0, 2q) between the stages
    void inverse(u64 *a, u32 pidx) const {
        const u32 N = h.N;
        const u64 q = h.primes[pidx], q2 = 2 * q;
        const u64 *itw = h.itw.data() + (size_t)pidx * N, *ish = h.itw_sh.data() + (size_t)pidx * N;
        if (f64(q)) {
            const HostF64Twiddles::Prime &P = f64_twiddles(h, pidx);
            const double ni = (double)h.ninv[pidx];
            return ntt_inverse_f64_host(a, N, q, P.iw.data(), P.iwq.data(), ni, ni / (double)q);
        }
        u32 t = 1;
        for (u32 m = N >> 1; m >= 1; m >>= 1, t <<= 1) {
            for (u32 i = 0; i < m; ++i) {
                const u64 W = itw[m + i], Wsh = ish[m + i];
                u64 *x = a + 2 * i * t;
                for (u32 j = 0; j < t; ++j) {
                    const u64 u = x[j], v = x[j + t];
                    u64 w = u + v;
                    if (w >= q2) w -= q2;
                    x[j] = w;
                    x[j + t] = mul_shoup_lazy(u - v + q2, W, Wsh, q);
                }
            }
        }
        const u64 ni = h.ninv[pidx], nish = shoup_companion(ni, q);
        for (u32 j = 0; j < N; ++j) a[j] = mul_shoup(a[j], ni, nish, q);
    }
};

// x mod q for any 64-bit x, q < 2^62: with r = floor((2^64 - 1) / q) the quotient
// estimate floor(x r / 2^64) is at most 2 below floor(x / q).
struct Reducer {
    u64 q, r;
    explicit Reducer(u64 qq) : q(qq), r(~0ull / qq) {}
    u64 operator()(u64 x) const {
        u64 t = x - mulhi64(x, r) * q;
        if (t >= q) t -= q;
        if (t >= q) t -= q;
        return t;
    }
};

u64 to_residue(int64_t v, const Reducer &red) {
    return v >= 0 ? red((u64)v) : red.q - red((u64)(-(v + 1))) - 1;
}

// Signed coefficients -> evaluation-domain limbs for the given prime indices.
std::vector<u64> to_eval(const HostTables &h, const int64_t *coef, const std::vector<u32> &pidx) {
    const u32 N = h.N;
    std::vector<u64> out((size_t)pidx.size() * N);
    HostNtt ntt(h);
    // keys and noise are small: |c| < q gives the residue by one conditional add
    u64 mag = 0;
    for (u32 j = 0; j < N; ++j) mag = std::max<u64>(mag, coef[j] < 0 ? (u64)(-(coef[j] + 1)) + 1 : (u64)coef[j]);
    for (size_t r = 0; r < pidx.size(); ++r) {
        const u64 q = h.primes[pidx[r]];
        u64 *row = out.data() + r * N;
        if (mag < q) {
            for (u32 j = 0; j < N; ++j) row[j] = (u64)(coef[j] + (coef[j] < 0 ? (int64_t)q : 0));
        } else {
            const Reducer red(q);
            for (u32 j = 0; j < N; ++j) row[j] = to_residue(coef[j], red);
        }
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

// Rounded Gaussian noise: x = round(y), y ~ N(0, 3.2^2), sampled by inversion of
// the distribution of |x| (64-bit thresholds t[k] = 2^64 P(|x| <= k), P(|x| <= k) =
// 1 - erfc((k + 1/2) / (3.2 sqrt 2))) and a sign bit; |x| > 40 has probability
// below 2^-64 and is not produced.
// ---------------------------------------------------------------------------
// Randomness.  A seed of FHEA_SEED_RANDOM (the default of the Python client) draws
// from a ChaCha20 keystream (RFC 8439: 20 rounds, 256-bit key from getrandom(2),
// zero nonce, a fresh key per call), the cryptographically secure source that key
// generation and encryption need.  Any other seed selects std::mt19937_64 seeded
// with it: reproducible, for tests and for checking against the oracle, and NOT
// secure (its state follows from its outputs, which the public key exposes).
// ---------------------------------------------------------------------------
inline u32 rotl32(u32 x, int r) { return (x << r) | (x >> (32 - r)); }

inline void chacha_quarter(u32 &a, u32 &b, u32 &c, u32 &d) {
    a += b; d ^= a; d = rotl32(d, 16);
    c += d; b ^= c; b = rotl32(b, 12);
    a += b; d ^= a; d = rotl32(d, 8);
    c += d; b ^= c; b = rotl32(b, 7);
}

// RFC 8439 section 2.3: the 64-byte block for (key, counter, nonce).
void chacha20_block(const u32 key[8], u32 counter, const u32 nonce[3], u32 out[16]) {
    u32 st[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key[0], key[1], key[2], key[3],
                  key[4],      key[5],      key[6],      key[7],      counter, nonce[0], nonce[1], nonce[2]};
    u32 x[16];
    std::memcpy(x, st, sizeof x);
    for (int i = 0; i < 10; ++i) {
        chacha_quarter(x[0], x[4], x[8], x[12]);
        chacha_quarter(x[1], x[5], x[9], x[13]);
        chacha_quarter(x[2], x[6], x[10], x[14]);
        chacha_quarter(x[3], x[7], x[11], x[15]);
        chacha_quarter(x[0], x[5], x[10], x[15]);
        chacha_quarter(x[1], x[6], x[11], x[12]);
        chacha_quarter(x[2], x[7], x[8], x[13]);
        chacha_quarter(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + st[i];
}

bool os_random(void *buf, size_t n) {
    uint8_t *p = (uint8_t *)buf;
    while (n) {
        const ssize_t r = getrandom(p, n, 0);
        if (r < 0) {
            if (errno == EINTR) continue;
            // no getrandom: the kernel's pool through /dev/urandom
            const int fd = open("/dev/urandom", O_RDONLY | O_CLOEXEC);
            if (fd < 0) return false;
            while (n) {
                const ssize_t k = read(fd, p, n);
                if (k <= 0) {
                    close(fd);
                    return false;
                }
                p += k;
                n -= (size_t)k;
            }
            close(fd);
            return true;
        }
        p += r;
        n -= (size_t)r;
    }
    return true;
}

// UniformRandomBitGenerator over the ChaCha20 keystream, 64 bits per draw.
class ChaChaRng {
public:
    using result_type = uint64_t;
    static constexpr result_type min() { return 0; }
    static constexpr result_type max() { return ~0ull; }
    // the key is drawn in the body, after every member has been initialised
    ChaChaRng() { ok_ = os_random(key_, sizeof key_); }
    bool ok() const { return ok_; }
    result_type operator()() {
        if (pos_ == 16) {
            const u32 nonce[3] = {0, 0, 0};
            chacha20_block(key_, counter_++, nonce, buf_);
            pos_ = 0;
        }
        const u64 v = (u64)buf_[pos_] | ((u64)buf_[pos_ + 1] << 32);
        pos_ += 2;
        return v;
    }

private:
    bool ok_ = false;
    u32 key_[8] = {};
    u32 counter_ = 0;
    u32 buf_[16];
    int pos_ = 16;
};

// The generator a seed selects (see above); both satisfy UniformRandomBitGenerator.
class Rng {
public:
    using result_type = uint64_t;
    static constexpr result_type min() { return 0; }
    static constexpr result_type max() { return ~0ull; }
    explicit Rng(uint64_t seed) : secure_(seed == FHEA_SEED_RANDOM), mt_(seed) {}
    // false when the secure source could not be keyed (no OS randomness)
    bool ok() const { return !secure_ || cc_.ok(); }
    result_type operator()() { return secure_ ? cc_() : mt_(); }

private:
    bool secure_;
    std::mt19937_64 mt_;
    ChaChaRng cc_;
};

struct GaussTable {
    static constexpr u32 K = 40;
    u64 t[K + 1];
    GaussTable() {
        const long double two64 = 18446744073709551616.0L, d = 3.2L * std::sqrt(2.0L);
        for (u32 k = 0; k <= K; ++k) {
            const long double tail = std::erfc(((long double)k + 0.5L) / d) * two64;  // 2^64 P(|x| > k)
            t[k] = tail < 1.0L ? ~0ull : (u64)(two64 - tail);
        }
    }
};

void gaussian(Rng &rng, std::vector<int64_t> &e) {
    static const GaussTable g;
    u64 signs = 0;
    u32 left = 0;
    for (auto &x : e) {
        const u64 u = rng();
        u32 k = 0;
        while (k < GaussTable::K && u >= g.t[k]) ++k;
        if (left == 0) {
            signs = rng();
            left = 64;
        }
        x = (signs & 1) ? -(int64_t)k : (int64_t)k;
        signs >>= 1;
        --left;
    }
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
    Rng rng(seed);
    if (!rng.ok()) return {};
    std::vector<int64_t> e(N);
    std::vector<u64> pmod(R);  // P mod q_i
    for (u32 r = 0; r < R; ++r) {
        u64 q = h.primes[kp[r]], pm = 1 % q;
        for (u32 j = 0; j < K; ++j) pm = (u64)((u128)pm * (h.primes[h.L + j] % q) % q);
        pmod[r] = pm;
    }
    for (u32 p = 0; p < dnum; ++p) {
        gaussian(rng, e);
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
    return out.empty() ? FHEA_ERR_STATE : FHEA_OK;
}

int write_key_to_arena(fhea_ctx *c, const std::vector<u64> &key, uint64_t evk) {
    if (evk & 7) return FHEA_ERR_ALIGN;
    if (evk > c->key_words || key.size() > c->key_words - evk) return FHEA_ERR_RANGE;
    return fhea_write_key(c, evk, key.data(), key.size());
}

int secret(u32 N, uint64_t seed, int8_t *s_out) {
    if (!s_out) return FHEA_ERR_ARG;
    Rng rng(seed);
    if (!rng.ok()) return FHEA_ERR_STATE;
    std::uniform_int_distribution<int> d(-1, 1);
    for (u32 j = 0; j < N; ++j) s_out[j] = (int8_t)d(rng);
    return FHEA_OK;
}

// ---- canonical embedding

// In-place iterative radix-2 DFT of length M on split parts (re, im):
// X_t = sum_k x_k exp(sign 2 pi i t k / M); the twiddles of the stage of length 2h
// are contiguous at tw[h + k] (k < h), and the stages of length 2 and 4 loop over
// the groups, so that every stage's inner loop runs over consecutive words.
// Compiled twice (AVX2+FMA and the baseline target), picked at run time.
__attribute__((always_inline)) inline void dft_split_body(double *re, double *im, u32 M, int sign,
                                                          const double *tw_re, const double *tw_im,
                                                          const u32 *brev) {
    for (u32 i = 1; i < M; ++i) {
        const u32 j = brev[i];
        if (i < j) {
            std::swap(re[i], re[j]);
            std::swap(im[i], im[j]);
        }
    }
    const double sg = sign > 0 ? 1.0 : -1.0;
    if (M >= 2)  // length 2: w = 1
        for (u32 i = 0; i < M; i += 2) {
            const double ar = re[i], ai = im[i], br = re[i + 1], bi = im[i + 1];
            re[i] = ar + br;
            im[i] = ai + bi;
            re[i + 1] = ar - br;
            im[i + 1] = ai - bi;
        }
    if (M >= 4)  // length 4: w = 1, sign i
        for (u32 i = 0; i < M; i += 4) {
            double ar = re[i], ai = im[i], br = re[i + 2], bi = im[i + 2];
            re[i] = ar + br;
            im[i] = ai + bi;
            re[i + 2] = ar - br;
            im[i + 2] = ai - bi;
            ar = re[i + 1];
            ai = im[i + 1];
            const double vr = -sg * im[i + 3], vi = sg * re[i + 3];  // (br + i bi) * (sign i)
            re[i + 1] = ar + vr;
            im[i + 1] = ai + vi;
            re[i + 3] = ar - vr;
            im[i + 3] = ai - vi;
        }
    for (u32 half = 4; half < M; half <<= 1) {
        const double *wr = tw_re + half, *wi = tw_im + half;
        for (u32 i = 0; i < M; i += 2 * half) {
            double *ar = re + i, *ai = im + i, *br = ar + half, *bi = ai + half;
            for (u32 k = 0; k < half; ++k) {
                const double cr = wr[k], ci = sg * wi[k];
                const double vr = br[k] * cr - bi[k] * ci, vi = br[k] * ci + bi[k] * cr;
                br[k] = ar[k] - vr;
                bi[k] = ai[k] - vi;
                ar[k] += vr;
                ai[k] += vi;
            }
        }
    }
}

void dft_split_base(double *re, double *im, u32 M, int sign, const double *tw_re, const double *tw_im,
                    const u32 *brev) {
    dft_split_body(re, im, M, sign, tw_re, tw_im, brev);
}
__attribute__((target("avx2,fma"))) void dft_split_avx2(double *re, double *im, u32 M, int sign,
                                                       const double *tw_re, const double *tw_im, const u32 *brev) {
    dft_split_body(re, im, M, sign, tw_re, tw_im, brev);
}
void dft_split(const fhea_client *cl, double *re, double *im, int sign) {
    const u32 M = cl->h.N / 2;
    if (host_f64_ok())
        dft_split_avx2(re, im, M, sign, cl->tw_re.data(), cl->tw_im.data(), cl->brev.data());
    else
        dft_split_base(re, im, M, sign, cl->tw_re.data(), cl->tw_im.data(), cl->brev.data());
}

bool valid_slots(u32 N, u32 n) { return n >= 1 && n <= N / 2 && (n & (n - 1)) == 0; }

// Integer-valued double -> residue mod q, exact for any magnitude: |v| = mant * 2^e
// with a 53-bit integer mantissa.
u64 double_to_residue(double v, const Reducer &red) {
    const u64 q = red.q;
    const double a = std::fabs(v);
    u64 r;
    if (a < 4.0e18) {
        r = red((u64)a);
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
    const u32 N = h.N, M = N / 2;
    std::vector<double> xr(M), xi(M);
    for (u32 j = 0; j < M; ++j) {
        const u32 u = cl->slot_u[j];
        xr[u] = re[j % n] * scale;
        xi[u] = im ? im[j % n] * scale : 0.0;
    }
    dft_split(cl, xr.data(), xi.data(), -1);
    // w_k zeta^-k / M = m_k + i m_{k+M}
    std::vector<double> m(N);
    const double inv = 1.0 / (double)M;
    for (u32 k = 0; k < M; ++k) {
        const double zr = cl->zeta_re[k], zi = cl->zeta_im[k];
        const double a = std::nearbyint((xr[k] * zr + xi[k] * zi) * inv);
        const double b = std::nearbyint((xi[k] * zr - xr[k] * zi) * inv);
        if (!std::isfinite(a) || !std::isfinite(b) || std::fabs(a) > 0x1p1000 || std::fabs(b) > 0x1p1000)
            return FHEA_ERR_ARG;
        m[k] = a;
        m[k + M] = b;
    }
    HostNtt ntt(h);
    for (u32 i = 0; i < limbs; ++i) {
        const Reducer red(h.primes[i]);
        u64 *row = pt + (size_t)i * N;
        for (u32 k = 0; k < N; ++k) row[k] = double_to_residue(m[k], red);
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

// First n slots of a plaintext whose limbs c (limbs*N words) are in the
// coefficient domain: lift (Garner), twist, one length-N/2 DFT, slot order.
void decode_coeffs(const fhea_client *cl, const Crt &crt, const u64 *c, double scale, u32 n, double *re,
                   double *im) {
    const u32 N = cl->h.N, M = N / 2;
    std::vector<double> xr(M), xi(M);
    if (crt.l == 1) {  // one limb: the centred residue itself
        const u64 q = crt.q[0], hq = (q - 1) / 2;
        const double qd = (double)q;
        for (u32 k = 0; k < M; ++k) {
            const double a = (double)c[k] - (c[k] > hq ? qd : 0.0);
            const double b = (double)c[k + M] - (c[k + M] > hq ? qd : 0.0);
            const double zr = cl->zeta_re[k], zi = cl->zeta_im[k];
            xr[k] = a * zr - b * zi;
            xi[k] = a * zi + b * zr;
        }
    } else {
        for (u32 k = 0; k < M; ++k) {
            const double a = crt.value(c + k, N), b = crt.value(c + k + M, N);
            const double zr = cl->zeta_re[k], zi = cl->zeta_im[k];
            xr[k] = a * zr - b * zi;
            xi[k] = a * zi + b * zr;
        }
    }
    dft_split(cl, xr.data(), xi.data(), +1);
    for (u32 j = 0; j < n; ++j) {
        const u32 u = cl->slot_u[j];
        re[j] = xr[u] / scale;
        if (im) im[j] = xi[u] / scale;
    }
}

}  // namespace

void fhea_client::init_embedding() {
    const u32 N = h.N, M = N / 2;
    tw_re.assign(M, 0.0);
    tw_im.assign(M, 0.0);
    for (u32 half = 1; half < M; half <<= 1)
        for (u32 k = 0; k < half; ++k) {
            const double a = kPi * (double)k / (double)half;  // 2 pi k / (2 half)
            tw_re[half + k] = std::cos(a);
            tw_im[half + k] = std::sin(a);
        }
    zeta_re.resize(M);
    zeta_im.resize(M);
    for (u32 k = 0; k < M; ++k) {
        const double b = kPi * (double)k / (double)N;
        zeta_re[k] = std::cos(b);
        zeta_im[k] = std::sin(b);
    }
    brev.resize(M);
    u32 lm = 0;
    while ((1u << lm) < M) ++lm;
    for (u32 i = 0; i < M; ++i) brev[i] = bit_reverse(i, lm);
    slot_u.resize(M);
    for (u32 j = 0; j < M; ++j) slot_u[j] = (u32)((pow_mod(5, j, 2ull * N) - 1) / 4);
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

int fhea_client_set_host_ntt(fhea_client *cl, int mode) {
    if (!cl) return FHEA_ERR_STATE;
    if (mode != 0 && mode != 1) return FHEA_ERR_ARG;
    cl->h.host_ntt_mode = (u32)mode;
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
    decode_coeffs(cl, Crt(h, limbs), c.data(), scale, n, re, im);
    return FHEA_OK;
}

int fhea_chacha20_block(const uint32_t *key, uint32_t counter, const uint32_t *nonce, uint32_t *out) {
    if (!key || !nonce || !out) return FHEA_ERR_ARG;
    chacha20_block(key, counter, nonce, out);
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
    Rng rng(seed);
    if (!rng.ok()) return FHEA_ERR_STATE;
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
    Rng rng(seed);
    if (!rng.ok()) return FHEA_ERR_STATE;
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
    const HostNtt ntt(h);
    for (u32 i = 0; i < limbs; ++i) {
        const u64 q = h.primes[i];
        const size_t at = (size_t)i * N;
        if (ntt.f64(q)) {
            enc_mul_f64(c0 + at, c1 + at, v_e.data() + at, b + at, a + at, e0_e.data() + at, e1_e.data() + at, pt + at,
                        N, q);
            continue;
        }
        const ModParams mq = make_mod_params(q);
        for (u32 j = 0; j < N; ++j) {
            const size_t o = (size_t)i * N + j;
            c0[o] = add_mod(add_mod(mulmod(v_e[o], b[o], mq), e0_e[o], q), pt[o], q);
            c1[o] = add_mod(mulmod(v_e[o], a[o], mq), e1_e[o], q);
        }
    }
    return FHEA_OK;
}

int fhea_secret_eval(const fhea_client *cl, const int8_t *s, uint32_t limbs, uint64_t *s_eval) {
    if (!cl) return FHEA_ERR_STATE;
    if (!s || !s_eval || limbs == 0 || limbs > cl->h.L) return FHEA_ERR_ARG;
    const std::vector<u64> e = to_eval(cl->h, s_to_i64(s, cl->h.N).data(), q_primes(limbs));
    std::memcpy(s_eval, e.data(), e.size() * 8);
    return FHEA_OK;
}

int fhea_decrypt_decode_batch(const fhea_client *cl, const uint64_t *s_eval, const uint64_t *cts, uint32_t count,
                              uint32_t limbs, double scale, uint32_t n, double *re, double *im, uint32_t threads) {
    if (!cl) return FHEA_ERR_STATE;
    const HostTables &h = cl->h;
    const u32 N = h.N;
    if (!s_eval || !cts || !re || !valid_slots(N, n) || limbs == 0 || limbs > h.L || !(scale > 0.0))
        return FHEA_ERR_ARG;
    if (count == 0) return FHEA_OK;
    u32 T = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
    T = std::min<u32>(T, count);
    const Crt crt(h, limbs);
    std::vector<ModParams> mq(limbs);
    for (u32 i = 0; i < limbs; ++i) mq[i] = make_mod_params(h.primes[i]);
    const size_t W = (size_t)limbs * N;
    auto work = [&](u32 t) {
        std::vector<u64> c(W);
        HostNtt ntt(h);
        for (u32 k = t; k < count; k += T) {
            const u64 *c0 = cts + 2 * W * k, *c1 = c0 + W;
            for (u32 i = 0; i < limbs; ++i) {
                const u64 q = h.primes[i];
                const size_t o = (size_t)i * N;
                if (ntt.f64(q))
                    dec_mul_f64(c.data() + o, c0 + o, c1 + o, s_eval + o, N, q);
                else
                    for (u32 j = 0; j < N; ++j)
                        c[o + j] = add_mod(c0[o + j], mulmod(c1[o + j], s_eval[o + j], mq[i]), q);
                ntt.inverse(c.data() + o, i);
            }
            decode_coeffs(cl, crt, c.data(), scale, n, re + (size_t)k * n, im ? im + (size_t)k * n : nullptr);
        }
    };
    if (T == 1) {
        work(0);
        return FHEA_OK;
    }
    std::vector<std::thread> pool;
    for (u32 t = 0; t < T; ++t) pool.emplace_back(work, t);
    for (auto &th : pool) th.join();
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
