/*
 * fhe_oracle.c -- CPU restatement of the reference CKKS evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY (see fhe_oracle.h).  Nothing in the product
 * (anns_fhe_amd/) links or calls this file.
 *
 * Pinning (each function is checked in tests/test_oracle_*.py by a check that
 * does not retype its formula):
 *   orc_mod_add/sub/mul      exhaustive edge cases + random pairs vs Python
 *                            big-int (a op b) % q, for 30..61-bit primes
 *   orc_barrett_constants    Python floor(2^k/q), k = 2*bitlen(q)+3
 *   orc_gen_primes           Python primality (Miller-Rabin, independent code)
 *                            and q == 1 mod 2N; descending order
 *   orc_min_primitive_root   psi^N == -1 and minimality by brute force (small N)
 *   orc_twiddles             Python pow(psi, bitrev(j), q)
 *   orc_ntt_forward          O(N^2) evaluation A[i] = sum_j a_j psi^((2br(i)+1)j)
 *                            at N <= 256; negacyclic convolution theorem at
 *                            larger N (checked against a Python schoolbook product)
 *   orc_intt_inverse         INTT(NTT(x)) == x; closed-form inverse at small N
 *   orc_automorphism /       coefficient-domain identity a(X) -> a(X^g) (with the
 *   orc_galois_map           X^N = -1 sign flip) pushed through the NTT
 *   orc_*_tables             Python big-int CRT definitions (OpenFHE's)
 *   orc_rescale              CRT brute force: reconstruct x in Z_Q, check
 *                            y == (x - c)/q_l with c the centred last residue
 *   orc_fast_bases_conv      CRT closed form sum_i [x_i qhat_i^-1]_qi qhat_i mod p
 *   orc_approx_mod_down      CRT closed form with exact big-int P division bound
 *   orc_key_switch           decryption invariant with keys made in the test
 *                            (|Dec(KS(d)) - d s'| small) and the exact noiseless
 *                            case (s = 0, e = 0: KS(d) = (d s', .) bit for bit)
 *   composites               composition of the above + decryption invariants
 */
#include "fhe_oracle.h"
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

#define ST_OK 0
#define ST_BAD_OPCODE (-1)
#define ST_RANGE (-2)
#define ST_ALIGN (-3)
#define ST_ARG (-4)
#define ST_NO_KEY (-5)
#define ST_NO_MEM (-6)

#define OP_NTT 1u
#define OP_INTT 2u
#define OP_POLY_ADD 3u
#define OP_POLY_SUB 4u
#define OP_POLY_MUL_NTT 5u
#define OP_AUTOMORPHISM 6u
#define OP_KEY_SWITCH 7u
#define OP_RESCALE 8u

/* ===================================================================== */
/* L0: modular arithmetic                                                */
/* ===================================================================== */

/* mod_arith.cpp:3-6 */
uint64_t orc_mod_add(uint64_t a, uint64_t b, uint64_t q) {
    uint64_t s = a + b;
    return (s >= q) ? (s - q) : s;
}

/* mod_arith.cpp:8-12 */
uint64_t orc_mod_sub(uint64_t a, uint64_t b, uint64_t q) {
    return (a >= b) ? (a - b) : (a + q - b);
}

/* Low 64 bits of floor(p * m / 2^k) with p, m < 2^128, computed at full
 * 256-bit precision (mod_arith.cpp:17-19: ap_uint<192> product then >> k). */
static uint64_t barrett_q1(u128 p, u128 m, uint32_t k) {
    uint64_t p0 = (uint64_t)p, p1 = (uint64_t)(p >> 64);
    uint64_t m0 = (uint64_t)m, m1 = (uint64_t)(m >> 64);
    u128 t00 = (u128)p0 * m0, t01 = (u128)p0 * m1, t10 = (u128)p1 * m0, t11 = (u128)p1 * m1;
    uint64_t w[4];
    w[0] = (uint64_t)t00;
    u128 mid = (t00 >> 64) + (uint64_t)t01 + (uint64_t)t10;
    w[1] = (uint64_t)mid;
    u128 hi = (mid >> 64) + (t01 >> 64) + (t10 >> 64) + (uint64_t)t11;
    w[2] = (uint64_t)hi;
    w[3] = (uint64_t)(hi >> 64) + (uint64_t)(t11 >> 64);
    uint32_t idx = k / 64, sh = k % 64;
    uint64_t r = w[idx] >> sh;
    if (sh && idx + 1 < 4) r |= w[idx + 1] << (64 - sh);
    return r;
}

/* mod_arith.cpp:15-25: Barrett with m = floor(2^k / q), k = 2*msb(q)+3;
 * r = p - q1*q computed mod 2^64, then at most two conditional subtractions. */
uint64_t orc_mod_mul(uint64_t a, uint64_t b, uint64_t q, uint64_t m_lo, uint64_t m_hi,
                     uint32_t k) {
    u128 p = (u128)a * b;
    u128 m = ((u128)m_hi << 64) | m_lo;
    uint64_t q1 = barrett_q1(p, m, k);
    uint64_t q2 = q1 * q;
    uint64_t r = (uint64_t)p - q2;
    if (r >= q) r -= q;
    if (r >= q) r -= q;
    return r;
}

/* generate_interactive_test_vectors.cpp:248-253 */
void orc_barrett_constants(uint64_t q, uint64_t *m_lo, uint64_t *m_hi, uint32_t *k) {
    uint32_t msb = 64 - (uint32_t)__builtin_clzll(q);
    uint32_t kk = 2 * msb + 3;
    u128 m = (((u128)1) << kk) / (u128)q;
    *m_lo = (uint64_t)m;
    *m_hi = (uint64_t)(m >> 64);
    *k = kk;
}

/* Extended Euclid on signed 128-bit (generate_interactive_test_vectors.cpp:29-44
 * computes the same inverse with int64, which overflows above 62 bits). */
uint64_t orc_inv_mod(uint64_t a, uint64_t q) {
    __int128 t = 0, nt = 1, r = q, nr = a % q;
    while (nr != 0) {
        __int128 qq = r / nr, tmp;
        tmp = t - qq * nt; t = nt; nt = tmp;
        tmp = r - qq * nr; r = nr; nr = tmp;
    }
    if (r != 1) return 0;
    if (t < 0) t += q;
    return (uint64_t)t;
}

uint64_t orc_pow_mod(uint64_t a, uint64_t e, uint64_t q) {
    u128 r = 1 % q, b = a % q;
    while (e) {
        if (e & 1) r = (r * b) % q;
        b = (b * b) % q;
        e >>= 1;
    }
    return (uint64_t)r;
}

/* Deterministic Miller-Rabin for n < 2^64. */
int orc_is_prime(uint64_t n) {
    static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; i++) {
        if (n == bases[i]) return 1;
        if (n % bases[i] == 0) return 0;
    }
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) { d >>= 1; s++; }
    for (int i = 0; i < 12; i++) {
        uint64_t x = orc_pow_mod(bases[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int r = 1; r < s; r++) {
            x = (uint64_t)(((u128)x * x) % n);
            if (x == n - 1) { comp = 0; break; }
        }
        if (comp) return 0;
    }
    return 1;
}

/* NTT-friendly primes q == 1 (mod two_n), q < 2^bits, descending from 2^bits,
 * skipping any listed in skip.  (The reference takes OpenFHE's primes from
 * files; the selection rule here is this build's documented convention.) */
uint32_t orc_gen_primes(uint32_t two_n, uint32_t bits, uint32_t count, const uint64_t *skip,
                        uint32_t n_skip, uint64_t *out) {
    uint64_t top = (bits >= 64) ? ~0ull : ((1ull << bits) - 1);
    uint64_t kk = (top - 1) / two_n;
    uint32_t found = 0;
    while (found < count && kk > 0) {
        uint64_t q = kk * two_n + 1;
        kk--;
        int skipped = 0;
        for (uint32_t i = 0; i < n_skip; i++)
            if (skip[i] == q) skipped = 1;
        for (uint32_t i = 0; i < found; i++)
            if (out[i] == q) skipped = 1;
        if (skipped) continue;
        if (orc_is_prime(q)) out[found++] = q;
    }
    return found;
}

/* Smallest primitive two_n-th root of unity mod q (two_n a power of two). */
uint64_t orc_min_primitive_root(uint64_t q, uint32_t two_n) {
    uint64_t g = 0;
    for (uint64_t x = 2; x < q; x++) {
        uint64_t c = orc_pow_mod(x, (q - 1) / two_n, q);
        if (orc_pow_mod(c, two_n / 2, q) == q - 1) { g = c; break; }
    }
    if (!g) return 0;
    uint64_t g2 = (uint64_t)(((u128)g * g) % q);
    uint64_t cur = g, best = g;
    for (uint32_t i = 1; i < two_n / 2; i++) {
        cur = (uint64_t)(((u128)cur * g2) % q);
        if (cur < best) best = cur;
    }
    return best;
}

/* generate_interactive_test_vectors.cpp:20-27 */
uint32_t orc_bit_reverse(uint32_t x, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

static uint32_t log2u(uint32_t n) { uint32_t l = 0; while ((1u << l) < n) l++; return l; }

/* generate_interactive_test_vectors.cpp:263-275: tw[j] = psi^br(j),
 * itw[j] = psi^-br(j), br over log2(N) bits. */
void orc_twiddles(uint64_t q, uint64_t psi, uint32_t N, uint64_t *tw, uint64_t *itw) {
    uint32_t lg = log2u(N);
    uint64_t psi_inv = orc_inv_mod(psi, q);
    uint64_t *pw = (uint64_t *)malloc((size_t)N * 8), *pwi = (uint64_t *)malloc((size_t)N * 8);
    uint64_t c = 1, ci = 1;
    for (uint32_t e = 0; e < N; e++) {
        pw[e] = c; pwi[e] = ci;
        c = (uint64_t)(((u128)c * psi) % q);
        ci = (uint64_t)(((u128)ci * psi_inv) % q);
    }
    for (uint32_t j = 0; j < N; j++) {
        uint32_t br = orc_bit_reverse(j, lg);
        tw[j] = pw[br];
        itw[j] = pwi[br];
    }
    free(pw); free(pwi);
}

/* ===================================================================== */
/* L1: polynomial kernels                                                 */
/* ===================================================================== */

/* ntt.cpp:4-43 (Cooley-Tukey, natural in -> bit-reversed out), runtime N. */
void orc_ntt_forward(uint64_t *a, uint32_t N, const uint64_t *tw, uint64_t q, uint64_t m_lo,
                     uint64_t m_hi, uint32_t k) {
    uint32_t t = N >> 1;
    for (uint32_t m = 1; m < N; m <<= 1) {
        for (uint32_t i = 0; i < m; i++) {
            uint32_t j1 = 2 * i * t;
            uint64_t W = tw[m + i];
            for (uint32_t jj = 0; jj < t; jj++) {
                uint32_t j = j1 + jj;
                uint64_t lo = a[j], hi = a[j + t];
                uint64_t w = orc_mod_mul(hi, W, q, m_lo, m_hi, k);
                uint64_t s = lo + w;
                if (s >= q) s -= q;
                uint64_t d = (lo < w) ? (lo + q - w) : (lo - w);
                a[j] = s;
                a[j + t] = d;
            }
        }
        t >>= 1;
    }
}

/* ntt.cpp:46-80 (Gentleman-Sande, then scale by N^-1), runtime N. */
void orc_intt_inverse(uint64_t *a, uint32_t N, const uint64_t *itw, uint64_t n_inv, uint64_t q,
                      uint64_t m_lo, uint64_t m_hi, uint32_t k) {
    uint32_t t = 1;
    for (uint32_t m = N >> 1; m >= 1; m >>= 1) {
        for (uint32_t i = 0; i < m; i++) {
            uint32_t j1 = 2 * i * t;
            uint64_t W = itw[m + i];
            for (uint32_t jj = 0; jj < t; jj++) {
                uint32_t j = j1 + jj;
                uint64_t lo = a[j], hi = a[j + t];
                a[j] = orc_mod_add(lo, hi, q);
                uint64_t d = orc_mod_sub(lo, hi, q);
                a[j + t] = orc_mod_mul(d, W, q, m_lo, m_hi, k);
            }
        }
        t <<= 1;
    }
    for (uint32_t i = 0; i < N; i++) a[i] = orc_mod_mul(a[i], n_inv, q, m_lo, m_hi, k);
}

/* automorphism.cpp:3-12: out[i] = a[map[i]], map packed two u32 per u64,
 * low half = even index. */
void orc_automorphism(const uint64_t *a, uint64_t *out, const uint64_t *packed_map, uint32_t N) {
    for (uint32_t i = 0; i < N; i++) {
        uint64_t p = packed_map[i / 2];
        uint32_t idx = (i % 2 == 0) ? (uint32_t)(p & 0xFFFFFFFFu) : (uint32_t)(p >> 32);
        out[i] = a[idx];
    }
}

/* Closed form of the evaluation-domain permutation for a(X) -> a(X^g) in the
 * bit-reversed slot order of orc_ntt_forward (the reference loads OpenFHE's
 * AutomorphismTransform map from a file, generate_interactive_test_vectors.cpp:307-322). */
void orc_galois_map(uint32_t N, uint64_t g, uint32_t *map_u32) {
    uint32_t lg = log2u(N);
    uint64_t two_n = 2ull * N;
    for (uint32_t i = 0; i < N; i++) {
        uint64_t e = 2ull * orc_bit_reverse(i, lg) + 1;
        uint64_t e2 = (g % two_n) * e % two_n;
        map_u32[i] = orc_bit_reverse((uint32_t)((e2 - 1) / 2), lg);
    }
}

/* ===================================================================== */
/* Context                                                                */
/* ===================================================================== */

struct orc_ctx {
    uint32_t N, logN, L, K, dnum, alpha;
    uint64_t *primes; /* L+K: Q then P */
    uint64_t *psi, *n_inv, *m_lo, *m_hi;
    uint32_t *k;
    uint64_t *tw, *itw; /* (L+K) * N */
};

orc_ctx *orc_ctx_create(uint32_t N, uint32_t L, uint32_t K, uint32_t dnum,
                        const uint64_t *q_primes, const uint64_t *p_primes) {
    if (N < 4 || (N & (N - 1)) || L == 0 || dnum == 0 || dnum > L) return NULL;
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    uint32_t T = L + K;
    c->N = N; c->logN = log2u(N); c->L = L; c->K = K; c->dnum = dnum;
    c->alpha = (L + dnum - 1) / dnum;
    c->primes = (uint64_t *)malloc(T * 8);
    c->psi = (uint64_t *)malloc(T * 8);
    c->n_inv = (uint64_t *)malloc(T * 8);
    c->m_lo = (uint64_t *)malloc(T * 8);
    c->m_hi = (uint64_t *)malloc(T * 8);
    c->k = (uint32_t *)malloc(T * 4);
    c->tw = (uint64_t *)malloc((size_t)T * N * 8);
    c->itw = (uint64_t *)malloc((size_t)T * N * 8);
    for (uint32_t i = 0; i < T; i++) {
        uint64_t q = (i < L) ? q_primes[i] : p_primes[i - L];
        c->primes[i] = q;
        orc_barrett_constants(q, &c->m_lo[i], &c->m_hi[i], &c->k[i]);
        c->n_inv[i] = orc_inv_mod(N, q);
        c->psi[i] = orc_min_primitive_root(q, 2 * N);
        orc_twiddles(q, c->psi[i], N, c->tw + (size_t)i * N, c->itw + (size_t)i * N);
    }
    return c;
}

/* Tables from given primitive 2N-th roots (Q then P) instead of the minimal ones:
 * the reference loads its twiddles from files (fhe_host_driver.cpp:56-66), whose
 * roots are OpenFHE's choice; psi[i] must satisfy psi^N == -1 mod q_i (returns -1
 * otherwise, tables unchanged). */
int orc_ctx_set_psi(orc_ctx *c, const uint64_t *psi) {
    uint32_t T = c->L + c->K;
    for (uint32_t i = 0; i < T; i++)
        if (orc_pow_mod(psi[i] % c->primes[i], c->N, c->primes[i]) != c->primes[i] - 1) return -1;
    for (uint32_t i = 0; i < T; i++) {
        c->psi[i] = psi[i] % c->primes[i];
        orc_twiddles(c->primes[i], c->psi[i], c->N, c->tw + (size_t)i * c->N, c->itw + (size_t)i * c->N);
    }
    return 0;
}

void orc_ctx_destroy(orc_ctx *c) {
    if (!c) return;
    free(c->primes); free(c->psi); free(c->n_inv); free(c->m_lo); free(c->m_hi);
    free(c->k); free(c->tw); free(c->itw); free(c);
}

uint32_t orc_ctx_alpha(const orc_ctx *c) { return c->alpha; }

void orc_ctx_tables(const orc_ctx *c, uint64_t *psi, uint64_t *n_inv, uint64_t *m_lo,
                    uint64_t *m_hi, uint32_t *k, uint64_t *tw, uint64_t *itw) {
    uint32_t T = c->L + c->K;
    if (psi) memcpy(psi, c->psi, T * 8);
    if (n_inv) memcpy(n_inv, c->n_inv, T * 8);
    if (m_lo) memcpy(m_lo, c->m_lo, T * 8);
    if (m_hi) memcpy(m_hi, c->m_hi, T * 8);
    if (k) memcpy(k, c->k, T * 4);
    if (tw) memcpy(tw, c->tw, (size_t)T * c->N * 8);
    if (itw) memcpy(itw, c->itw, (size_t)T * c->N * 8);
}

static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

/* Number of digits at a level and the size of digit `part`
 * (key_switch.cpp:229-231; the reference fixes numPartQ=3). */
static uint32_t num_parts(const orc_ctx *c, uint32_t sizeQl) {
    return (sizeQl + c->alpha - 1) / c->alpha;
}
static uint32_t part_size(const orc_ctx *c, uint32_t sizeQl, uint32_t part) {
    uint32_t np = num_parts(c, sizeQl);
    return (part == np - 1) ? (sizeQl - c->alpha * part) : c->alpha;
}

/* PartQlHatInvModq / PartQlHatModp at level sizeQl for digit `part`
 * (test_vector_generator.cpp:408-465; OpenFHE's definition: qhat_i = Q_part/q_i
 * over the digit's primes at this level; output basis = the other Q primes of
 * the level in ascending order, then P, as key_switch.cpp:256-268).  Returns n_out. */
uint32_t orc_modup_tables(const orc_ctx *c, uint32_t sizeQl, uint32_t part,
                          uint64_t *qHatInvModq, uint64_t *qHatModp) {
    uint32_t start = c->alpha * part, sz = part_size(c, sizeQl, part);
    uint32_t n_out = sizeQl - sz + c->K;
    uint64_t outp[256];
    uint32_t o = 0;
    for (uint32_t i = 0; i < sizeQl; i++)
        if (i < start || i >= start + sz) outp[o++] = c->primes[i];
    for (uint32_t j = 0; j < c->K; j++) outp[o++] = c->primes[c->L + j];
    for (uint32_t i = 0; i < sz; i++) {
        uint64_t qi = c->primes[start + i];
        uint64_t prod = 1 % qi;
        for (uint32_t kk = 0; kk < sz; kk++)
            if (kk != i) prod = mulmod(prod, c->primes[start + kk] % qi, qi);
        qHatInvModq[i] = orc_inv_mod(prod, qi);
        for (uint32_t j = 0; j < n_out; j++) {
            uint64_t pj = outp[j];
            uint64_t pr = 1 % pj;
            for (uint32_t kk = 0; kk < sz; kk++)
                if (kk != i) pr = mulmod(pr, c->primes[start + kk] % pj, pj);
            qHatModp[i * n_out + j] = pr;
        }
    }
    return n_out;
}

/* PInvModq (sizeQl), PHatInvModp (K), PHatModq (K x sizeQl, row-major by P
 * index) (test_vector_generator.cpp:467-494). */
void orc_moddown_tables(const orc_ctx *c, uint32_t sizeQl, uint64_t *PInvModq,
                        uint64_t *PHatInvModp, uint64_t *PHatModq) {
    const uint64_t *P = c->primes + c->L;
    for (uint32_t i = 0; i < sizeQl; i++) {
        uint64_t qi = c->primes[i], pr = 1 % qi;
        for (uint32_t j = 0; j < c->K; j++) pr = mulmod(pr, P[j] % qi, qi);
        PInvModq[i] = orc_inv_mod(pr, qi);
    }
    for (uint32_t j = 0; j < c->K; j++) {
        uint64_t pj = P[j], pr = 1 % pj;
        for (uint32_t kk = 0; kk < c->K; kk++)
            if (kk != j) pr = mulmod(pr, P[kk] % pj, pj);
        PHatInvModp[j] = orc_inv_mod(pr, pj);
        for (uint32_t i = 0; i < sizeQl; i++) {
            uint64_t qi = c->primes[i], h = 1 % qi;
            for (uint32_t kk = 0; kk < c->K; kk++)
                if (kk != j) h = mulmod(h, P[kk] % qi, qi);
            PHatModq[j * sizeQl + i] = h;
        }
    }
}

/* Rescale tables for dropping q_l, l = sizeQl-1 (test_vector_generator.cpp:304-325).
 * qlInvModq[i] = q_l^-1 mod q_i.  OpenFHE's QlQlInvModqlDivqlModq[i] =
 * floor(Q'u/q_l) mod q_i with Q' = Q_l/q_l, u = Q'^-1 mod q_l, which is
 * congruent to -q_l^-1 mod q_i (the test checks this against big integers). */
void orc_rescale_tables(const orc_ctx *c, uint32_t sizeQl, uint64_t *QlQlInvModqlDivqlModq,
                        uint64_t *qlInvModq) {
    uint64_t ql = c->primes[sizeQl - 1];
    for (uint32_t i = 0; i + 1 < sizeQl; i++) {
        uint64_t qi = c->primes[i];
        uint64_t inv = orc_inv_mod(ql % qi, qi);
        qlInvModq[i] = inv;
        QlQlInvModqlDivqlModq[i] = inv ? qi - inv : 0;
    }
}

/* ===================================================================== */
/* L2: CKKS-RNS ops                                                       */
/* ===================================================================== */

/* rescale.cpp:3-72, runtime N; tables passed as in the reference. */
static void rescale_core(uint32_t N, const uint64_t *a_in, uint64_t *a_out, uint32_t num_limbs,
                         const uint64_t *q, const uint64_t *m_lo, const uint64_t *m_hi,
                         const uint32_t *kk, const uint64_t *inv_twiddles_last,
                         const uint64_t *twiddles, uint64_t n_inv_last,
                         const uint64_t *QlQl, const uint64_t *qlInv) {
    uint32_t last = num_limbs - 1;
    uint64_t *tmp = (uint64_t *)malloc((size_t)N * 8);
    uint64_t *tmp_i = (uint64_t *)malloc((size_t)N * 8);
    memcpy(tmp, a_in + (size_t)last * N, (size_t)N * 8);
    orc_intt_inverse(tmp, N, inv_twiddles_last, n_inv_last, q[last], m_lo[last], m_hi[last], kk[last]);
    for (uint32_t i = 0; i < last; i++) {
        uint64_t qi = q[i];
        uint64_t old_q = q[last], half = old_q >> 1;
        uint64_t old_q_mod = orc_mod_mul(old_q, 1, qi, m_lo[i], m_hi[i], kk[i]);
        for (uint32_t j = 0; j < N; j++) {
            uint64_t v = tmp[j];
            uint64_t vm = orc_mod_mul(v, 1, qi, m_lo[i], m_hi[i], kk[i]);
            uint64_t val = (v > half) ? orc_mod_sub(vm, old_q_mod, qi) : vm;
            tmp_i[j] = orc_mod_mul(val, QlQl[i], qi, m_lo[i], m_hi[i], kk[i]);
        }
        orc_ntt_forward(tmp_i, N, twiddles + (size_t)i * N, qi, m_lo[i], m_hi[i], kk[i]);
        for (uint32_t j = 0; j < N; j++) {
            uint64_t s = orc_mod_mul(a_in[(size_t)i * N + j], qlInv[i], qi, m_lo[i], m_hi[i], kk[i]);
            a_out[(size_t)i * N + j] = orc_mod_add(s, tmp_i[j], qi);
        }
    }
    free(tmp); free(tmp_i);
}

void orc_rescale(const orc_ctx *c, const uint64_t *a_in, uint64_t *a_out, uint32_t num_limbs) {
    uint64_t QlQl[256], qlInv[256];
    orc_rescale_tables(c, num_limbs, QlQl, qlInv);
    uint32_t last = num_limbs - 1;
    rescale_core(c->N, a_in, a_out, num_limbs, c->primes, c->m_lo, c->m_hi, c->k,
                 c->itw + (size_t)last * c->N, c->tw, c->n_inv[last], QlQl, qlInv);
}

/* key_switch.cpp:7-60 (ApproxSwitchCRTBasis without overflow correction,
 * coefficient domain), runtime N. */
void orc_fast_bases_conv(uint32_t N, const uint64_t *in_poly, uint64_t *out_poly, uint32_t n_in,
                         uint32_t n_out, const uint64_t *in_primes, const uint64_t *out_primes,
                         const uint64_t *qHatInvModq, const uint64_t *qHatModp,
                         const uint64_t *bm_in_lo, const uint64_t *bm_in_hi, const uint32_t *bk_in,
                         const uint64_t *bm_out_lo, const uint64_t *bm_out_hi,
                         const uint32_t *bk_out) {
    uint64_t *buf = (uint64_t *)malloc((size_t)n_in * N * 8);
    for (uint32_t i = 0; i < n_in; i++)
        for (uint32_t r = 0; r < N; r++)
            buf[(size_t)i * N + r] = orc_mod_mul(in_poly[(size_t)i * N + r], qHatInvModq[i],
                                                 in_primes[i], bm_in_lo[i], bm_in_hi[i], bk_in[i]);
    for (uint32_t j = 0; j < n_out; j++) {
        uint64_t pj = out_primes[j];
        for (uint32_t r = 0; r < N; r++) {
            uint64_t sum = 0;
            for (uint32_t i = 0; i < n_in; i++) {
                uint64_t x = orc_mod_mul(buf[(size_t)i * N + r], 1, pj, bm_out_lo[j], bm_out_hi[j], bk_out[j]);
                uint64_t pm = orc_mod_mul(x, qHatModp[i * n_out + j], pj, bm_out_lo[j], bm_out_hi[j], bk_out[j]);
                sum = orc_mod_add(sum, pm, pj);
            }
            out_poly[(size_t)j * N + r] = sum;
        }
    }
    free(buf);
}

/* key_switch.cpp:62-163: out_i = (x_Q,i - NTT_i(Conv_{P->Q}(INTT(x_P))))_i * P^-1. */
void orc_approx_mod_down(const orc_ctx *c, const uint64_t *in_poly, uint64_t *out_poly,
                         uint32_t sizeQl) {
    uint32_t N = c->N, K = c->K, L = c->L;
    uint64_t PInvModq[256], PHatInvModp[64], *PHatModq = (uint64_t *)malloc((size_t)K * sizeQl * 8);
    orc_moddown_tables(c, sizeQl, PInvModq, PHatInvModp, PHatModq);
    uint64_t *bufP = (uint64_t *)malloc((size_t)K * N * 8);
    uint64_t *sw = (uint64_t *)malloc((size_t)sizeQl * N * 8);
    memcpy(bufP, in_poly + (size_t)sizeQl * N, (size_t)K * N * 8);
    for (uint32_t j = 0; j < K; j++)
        orc_intt_inverse(bufP + (size_t)j * N, N, c->itw + (size_t)(L + j) * N, c->n_inv[L + j],
                         c->primes[L + j], c->m_lo[L + j], c->m_hi[L + j], c->k[L + j]);
    orc_fast_bases_conv(N, bufP, sw, K, sizeQl, c->primes + L, c->primes, PHatInvModp, PHatModq,
                        c->m_lo + L, c->m_hi + L, c->k + L, c->m_lo, c->m_hi, c->k);
    for (uint32_t i = 0; i < sizeQl; i++)
        orc_ntt_forward(sw + (size_t)i * N, N, c->tw + (size_t)i * N, c->primes[i], c->m_lo[i],
                        c->m_hi[i], c->k[i]);
    for (uint32_t i = 0; i < sizeQl; i++) {
        uint64_t qi = c->primes[i];
        for (uint32_t r = 0; r < N; r++) {
            uint64_t vq = in_poly[(size_t)i * N + r], vp = sw[(size_t)i * N + r];
            uint64_t d = (vq >= vp) ? (vq - vp) : (vq + qi - vp);
            out_poly[(size_t)i * N + r] = orc_mod_mul(d, PInvModq[i], qi, c->m_lo[i], c->m_hi[i], c->k[i]);
        }
    }
    free(PHatModq); free(bufP); free(sw);
}

/* key_switch.cpp:167-439 (EvalFastKeySwitchCore; the caller adds c0). */
/* ModUp of every digit (key_switch.cpp:229-349): digits[p] holds the QlP limbs of
 * digit p in the evaluation domain (own limbs copied, the others converted). */
static void modup_digits(const orc_ctx *c, const uint64_t *c_in, uint32_t sizeQl, uint64_t *digits) {
    uint32_t N = c->N, K = c->K, L = c->L;
    uint32_t QlP = sizeQl + K;
    uint64_t *parts = (uint64_t *)malloc((size_t)c->alpha * N * 8);
    uint64_t *compl_ = (uint64_t *)malloc((size_t)QlP * N * 8);
    uint32_t np = num_parts(c, sizeQl);
    for (uint32_t part = 0; part < np; part++) {
        uint64_t *digit = digits + (size_t)part * QlP * N;
        uint32_t sz = part_size(c, sizeQl, part), start = c->alpha * part;
        memcpy(parts, c_in + (size_t)start * N, (size_t)sz * N * 8);
        for (uint32_t i = 0; i < sz; i++) {
            uint32_t idx = start + i;
            orc_intt_inverse(parts + (size_t)i * N, N, c->itw + (size_t)idx * N, c->n_inv[idx],
                             c->primes[idx], c->m_lo[idx], c->m_hi[idx], c->k[idx]);
        }
        uint32_t n_out = sizeQl - sz + K;
        uint64_t outp[256], olo[256], ohi[256];
        uint32_t ok[256], oidx[256], o = 0;
        for (uint32_t i = 0; i < sizeQl; i++)
            if (i < start || i >= start + sz) oidx[o++] = i;
        for (uint32_t j = 0; j < K; j++) oidx[o++] = L + j;
        for (uint32_t j = 0; j < n_out; j++) {
            outp[j] = c->primes[oidx[j]]; olo[j] = c->m_lo[oidx[j]];
            ohi[j] = c->m_hi[oidx[j]]; ok[j] = c->k[oidx[j]];
        }
        uint64_t qHatInv[256];
        uint64_t *qHatModp = (uint64_t *)malloc((size_t)sz * n_out * 8);
        orc_modup_tables(c, sizeQl, part, qHatInv, qHatModp);
        orc_fast_bases_conv(N, parts, compl_, sz, n_out, c->primes + start, outp, qHatInv, qHatModp,
                            c->m_lo + start, c->m_hi + start, c->k + start, olo, ohi, ok);
        free(qHatModp);
        for (uint32_t j = 0; j < n_out; j++) {
            uint32_t pi = oidx[j];
            orc_ntt_forward(compl_ + (size_t)j * N, N, c->tw + (size_t)pi * N, c->primes[pi],
                            c->m_lo[pi], c->m_hi[pi], c->k[pi]);
        }
        /* digit assembly (key_switch.cpp:332-349) */
        o = 0;
        for (uint32_t i = 0; i < sizeQl; i++) {
            if (i >= start && i < start + sz)
                memcpy(digit + (size_t)i * N, c_in + (size_t)i * N, (size_t)N * 8);
            else
                memcpy(digit + (size_t)i * N, compl_ + (size_t)(o++) * N, (size_t)N * 8);
        }
        for (uint32_t j = 0; j < K; j++)
            memcpy(digit + (size_t)(sizeQl + j) * N, compl_ + (size_t)(o++) * N, (size_t)N * 8);
    }
    free(parts); free(compl_);
}

/* evk inner product (key_switch.cpp:352-381) of the digits, each read through the
 * evaluation-domain permutation `map` (NULL = identity), then ModDown of both sums. */
static void inner_moddown(const orc_ctx *c, const uint64_t *digits, const uint32_t *map,
                          const uint64_t *key_mem, uint64_t evk_off, uint32_t key_limbs,
                          uint64_t *c_out0, uint64_t *c_out1, uint32_t sizeQl) {
    uint32_t N = c->N, K = c->K, L = c->L;
    uint32_t QlP = sizeQl + K;
    uint64_t stride = (uint64_t)(key_limbs + K) * N;
    uint64_t evk_a = evk_off, evk_b = evk_off + (uint64_t)c->dnum * stride;
    uint64_t *ext0 = (uint64_t *)calloc((size_t)QlP * N, 8);
    uint64_t *ext1 = (uint64_t *)calloc((size_t)QlP * N, 8);
    uint32_t np = num_parts(c, sizeQl);
    for (uint32_t part = 0; part < np; part++) {
        const uint64_t *digit = digits + (size_t)part * QlP * N;
        for (uint32_t i = 0; i < QlP; i++) {
            uint32_t pi = (i < sizeQl) ? i : (L + i - sizeQl);
            uint32_t kl = (i < sizeQl) ? i : (key_limbs + i - sizeQl);
            uint64_t q = c->primes[pi];
            const uint64_t *ka = key_mem + evk_a + part * stride + (uint64_t)kl * N;
            const uint64_t *kb = key_mem + evk_b + part * stride + (uint64_t)kl * N;
            for (uint32_t r = 0; r < N; r++) {
                uint64_t d = digit[(size_t)i * N + (map ? map[r] : r)];
                uint64_t cb = orc_mod_mul(d, kb[r], q, c->m_lo[pi], c->m_hi[pi], c->k[pi]);
                uint64_t ca = orc_mod_mul(d, ka[r], q, c->m_lo[pi], c->m_hi[pi], c->k[pi]);
                ext0[(size_t)i * N + r] = orc_mod_add(ext0[(size_t)i * N + r], cb, q);
                ext1[(size_t)i * N + r] = orc_mod_add(ext1[(size_t)i * N + r], ca, q);
            }
        }
    }
    orc_approx_mod_down(c, ext0, c_out0, sizeQl);
    orc_approx_mod_down(c, ext1, c_out1, sizeQl);
    free(ext0); free(ext1);
}

void orc_key_switch(const orc_ctx *c, const uint64_t *c_in, const uint64_t *key_mem,
                    uint64_t evk_off, uint32_t key_limbs, uint64_t *c_out0, uint64_t *c_out1,
                    uint32_t sizeQl) {
    uint32_t QlP = sizeQl + c->K;
    uint64_t *digits = (uint64_t *)malloc((size_t)num_parts(c, sizeQl) * QlP * c->N * 8);
    modup_digits(c, c_in, sizeQl, digits);
    inner_moddown(c, digits, NULL, key_mem, evk_off, key_limbs, c_out0, c_out1, sizeQl);
    free(digits);
}

/* ===================================================================== */
/* L3: op dispatcher (fhe_accel_top.cpp:220-327)                          */
/* ===================================================================== */

static int in_range(uint64_t off, uint64_t words, uint64_t total) {
    return off <= total && words <= total - off;
}

int orc_accel_op(const orc_ctx *c, uint64_t *pm, uint64_t pw, const uint64_t *km, uint64_t kw,
                 uint32_t op, uint64_t src_a, uint64_t src_b, uint64_t dst, uint64_t evk,
                 uint32_t limbs, uint32_t key_limbs) {
    uint32_t N = c->N;
    if (op < OP_NTT || op > OP_RESCALE) return ST_BAD_OPCODE;
    if ((src_a | src_b | dst | evk) & 7) return ST_ALIGN;
    if (limbs == 0 || limbs > c->L) return ST_ARG;
    uint64_t W = (uint64_t)limbs * N;
    if (!in_range(src_a, W, pw)) return ST_RANGE;
    uint64_t *A = (uint64_t *)malloc(W * 8), *B = NULL, *O = NULL;
    memcpy(A, pm + src_a, W * 8);
    int st = ST_OK;
    switch (op) {
    case OP_NTT:
    case OP_INTT:
        if (!in_range(dst, W, pw)) { st = ST_RANGE; break; }
        for (uint32_t l = 0; l < limbs; l++) {
            if (op == OP_NTT)
                orc_ntt_forward(A + (size_t)l * N, N, c->tw + (size_t)l * N, c->primes[l], c->m_lo[l], c->m_hi[l], c->k[l]);
            else
                orc_intt_inverse(A + (size_t)l * N, N, c->itw + (size_t)l * N, c->n_inv[l], c->primes[l], c->m_lo[l], c->m_hi[l], c->k[l]);
        }
        memcpy(pm + dst, A, W * 8);
        break;
    case OP_POLY_ADD:
    case OP_POLY_SUB:
    case OP_POLY_MUL_NTT:
        if (!in_range(src_b, W, pw) || !in_range(dst, W, pw)) { st = ST_RANGE; break; }
        B = (uint64_t *)malloc(W * 8);
        memcpy(B, pm + src_b, W * 8);
        for (uint32_t l = 0; l < limbs; l++) {
            uint64_t q = c->primes[l];
            for (uint32_t r = 0; r < N; r++) {
                size_t x = (size_t)l * N + r;
                if (op == OP_POLY_ADD) A[x] = orc_mod_add(A[x], B[x], q);          /* poly_arith.cpp:3-14 */
                else if (op == OP_POLY_SUB) A[x] = orc_mod_sub(A[x], B[x], q);     /* poly_arith.cpp:16-27 */
                else A[x] = orc_mod_mul(A[x], B[x], q, c->m_lo[l], c->m_hi[l], c->k[l]); /* :29-42 */
            }
        }
        memcpy(pm + dst, A, W * 8);
        break;
    case OP_AUTOMORPHISM:
        if (!in_range(src_b, N / 2, pw) || !in_range(dst, W, pw)) { st = ST_RANGE; break; }
        O = (uint64_t *)malloc(W * 8);
        for (uint32_t l = 0; l < limbs; l++)
            orc_automorphism(A + (size_t)l * N, O + (size_t)l * N, pm + src_b, N);
        memcpy(pm + dst, O, W * 8);
        break;
    case OP_KEY_SWITCH: {
        uint32_t kl = key_limbs ? key_limbs : limbs;
        uint64_t kwords = 2ull * c->dnum * (kl + c->K) * N;
        if (kl < limbs) { st = ST_ARG; break; }
        if (!in_range(dst, 2 * W, pw) || !in_range(evk, kwords, kw)) { st = ST_RANGE; break; }
        O = (uint64_t *)malloc(2 * W * 8);
        orc_key_switch(c, A, km, evk, kl, O, O + W, limbs);
        memcpy(pm + dst, O, 2 * W * 8);
        break;
    }
    case OP_RESCALE: {
        if (limbs < 2) { st = ST_ARG; break; }
        uint64_t W1 = (uint64_t)(limbs - 1) * N;
        if (!in_range(src_b, W, pw) || !in_range(dst, 2 * W1, pw)) { st = ST_RANGE; break; }
        B = (uint64_t *)malloc(W * 8);
        memcpy(B, pm + src_b, W * 8);
        O = (uint64_t *)malloc(2 * W1 * 8);
        orc_rescale(c, A, O, limbs);
        orc_rescale(c, B, O + W1, limbs);
        memcpy(pm + dst, O, 2 * W1 * 8);
        break;
    }
    }
    free(A); free(B); free(O);
    return st;
}

/* ===================================================================== */
/* L5: composites (fhe_host_driver.cpp:248-404)                           */
/* ===================================================================== */

static void poly_binop(const orc_ctx *c, int op, const uint64_t *a, const uint64_t *b, uint64_t *o,
                       uint32_t limbs) {
    uint32_t N = c->N;
    for (uint32_t l = 0; l < limbs; l++) {
        uint64_t q = c->primes[l];
        for (uint32_t r = 0; r < N; r++) {
            size_t x = (size_t)l * N + r;
            if (op == 0) o[x] = orc_mod_add(a[x], b[x], q);
            else if (op == 1) o[x] = orc_mod_sub(a[x], b[x], q);
            else o[x] = orc_mod_mul(a[x], b[x], q, c->m_lo[l], c->m_hi[l], c->k[l]);
        }
    }
}

#define CHECK_RANGE(off, words) do { if (((off) & 7) != 0) return ST_ALIGN; \
    if (!in_range((off), (words), pw)) return ST_RANGE; } while (0)

/* fhe_host_driver.cpp:248-256 */
int orc_eval_add(const orc_ctx *c, uint64_t *pm, uint64_t pw, uint64_t a0, uint64_t a1, uint64_t b0,
                 uint64_t b1, uint64_t d0, uint64_t d1, uint32_t limbs) {
    if (limbs == 0 || limbs > c->L) return ST_ARG;
    uint64_t W = (uint64_t)limbs * c->N;
    CHECK_RANGE(a0, W); CHECK_RANGE(a1, W); CHECK_RANGE(b0, W); CHECK_RANGE(b1, W);
    CHECK_RANGE(d0, W); CHECK_RANGE(d1, W);
    uint64_t *t0 = (uint64_t *)malloc(W * 8), *t1 = (uint64_t *)malloc(W * 8);
    poly_binop(c, 0, pm + a0, pm + b0, t0, limbs);
    poly_binop(c, 0, pm + a1, pm + b1, t1, limbs);
    memcpy(pm + d0, t0, W * 8);
    memcpy(pm + d1, t1, W * 8);
    free(t0); free(t1);
    return ST_OK;
}

/* fhe_host_driver.cpp:258-270 */
int orc_eval_sub_plain(const orc_ctx *c, uint64_t *pm, uint64_t pw, uint64_t a0, uint64_t a1,
                       uint64_t pt, uint64_t d0, uint64_t d1, uint32_t limbs) {
    if (limbs == 0 || limbs > c->L) return ST_ARG;
    uint64_t W = (uint64_t)limbs * c->N;
    CHECK_RANGE(a0, W); CHECK_RANGE(a1, W); CHECK_RANGE(pt, W); CHECK_RANGE(d0, W); CHECK_RANGE(d1, W);
    uint64_t *t0 = (uint64_t *)malloc(W * 8), *t1 = (uint64_t *)malloc(W * 8);
    poly_binop(c, 1, pm + a0, pm + pt, t0, limbs);
    memcpy(t1, pm + a1, W * 8);
    memcpy(pm + d0, t0, W * 8);
    memcpy(pm + d1, t1, W * 8);
    free(t0); free(t1);
    return ST_OK;
}

/* fhe_host_driver.cpp:272-306 */
int orc_eval_mult_relin_rescale(const orc_ctx *c, uint64_t *pm, uint64_t pw, const uint64_t *km,
                                uint64_t kw, uint64_t relin_evk, uint32_t key_limbs, uint64_t a0,
                                uint64_t a1, uint64_t b0, uint64_t b1, uint64_t d0, uint64_t d1,
                                uint32_t L) {
    if (L < 2 || L > c->L) return ST_ARG;
    uint32_t N = c->N;
    uint32_t kl = key_limbs ? key_limbs : L;
    if (kl < L) return ST_ARG;
    uint64_t W = (uint64_t)L * N, W1 = (uint64_t)(L - 1) * N;
    CHECK_RANGE(a0, W); CHECK_RANGE(a1, W); CHECK_RANGE(b0, W); CHECK_RANGE(b1, W);
    CHECK_RANGE(d0, W1); CHECK_RANGE(d1, W1);
    if ((relin_evk & 7) || !in_range(relin_evk, 2ull * c->dnum * (kl + c->K) * N, kw)) return ST_RANGE;
    uint64_t *D0 = (uint64_t *)malloc(W * 8), *D1 = (uint64_t *)malloc(W * 8);
    uint64_t *D2 = (uint64_t *)malloc(W * 8), *T = (uint64_t *)malloc(W * 8);
    uint64_t *KS = (uint64_t *)malloc(2 * W * 8), *O = (uint64_t *)malloc(2 * W1 * 8);
    poly_binop(c, 2, pm + a0, pm + b0, D0, L);
    poly_binop(c, 2, pm + a1, pm + b1, D2, L);
    poly_binop(c, 2, pm + a0, pm + b1, D1, L);
    poly_binop(c, 2, pm + a1, pm + b0, T, L);
    poly_binop(c, 0, D1, T, D1, L);
    orc_key_switch(c, D2, km, relin_evk, kl, KS, KS + W, L);
    poly_binop(c, 0, D0, KS, D0, L);
    poly_binop(c, 0, D1, KS + W, D1, L);
    orc_rescale(c, D0, O, L);
    orc_rescale(c, D1, O + W1, L);
    memcpy(pm + d0, O, W1 * 8);
    memcpy(pm + d1, O + W1, W1 * 8);
    free(D0); free(D1); free(D2); free(T); free(KS); free(O);
    return ST_OK;
}

/* fhe_host_driver.cpp:308-329 */
int orc_eval_mult_plain_rescale(const orc_ctx *c, uint64_t *pm, uint64_t pw, uint64_t a0,
                                uint64_t a1, uint64_t pt, uint64_t d0, uint64_t d1, uint32_t L) {
    if (L < 2 || L > c->L) return ST_ARG;
    uint64_t W = (uint64_t)L * c->N, W1 = (uint64_t)(L - 1) * c->N;
    CHECK_RANGE(a0, W); CHECK_RANGE(a1, W); CHECK_RANGE(pt, W); CHECK_RANGE(d0, W1); CHECK_RANGE(d1, W1);
    uint64_t *D0 = (uint64_t *)malloc(W * 8), *D1 = (uint64_t *)malloc(W * 8);
    uint64_t *O = (uint64_t *)malloc(2 * W1 * 8);
    poly_binop(c, 2, pm + a0, pm + pt, D0, L);
    poly_binop(c, 2, pm + a1, pm + pt, D1, L);
    orc_rescale(c, D0, O, L);
    orc_rescale(c, D1, O + W1, L);
    memcpy(pm + d0, O, W1 * 8);
    memcpy(pm + d1, O + W1, W1 * 8);
    free(D0); free(D1); free(O);
    return ST_OK;
}

/* One rotation step on host buffers (fhe_host_driver.cpp:331-359):
 * sigma_g(c0), sigma_g(c1), KS(sigma_g(c1)), c0' = sigma_g(c0) + ks0, c1' = ks1. */
static void rotate_step(const orc_ctx *c, const uint64_t *km, uint64_t g, uint64_t evk,
                        uint32_t key_limbs, const uint64_t *in0, const uint64_t *in1,
                        uint64_t *out0, uint64_t *out1, uint32_t limbs) {
    uint32_t N = c->N;
    uint64_t W = (uint64_t)limbs * N;
    uint32_t *map = (uint32_t *)malloc((size_t)N * 4);
    orc_galois_map(N, g, map);
    uint64_t *packed = (uint64_t *)malloc((size_t)N / 2 * 8);
    for (uint32_t i = 0; i < N / 2; i++) packed[i] = ((uint64_t)map[2 * i + 1] << 32) | map[2 * i];
    uint64_t *s0 = (uint64_t *)malloc(W * 8), *s1 = (uint64_t *)malloc(W * 8);
    uint64_t *ks = (uint64_t *)malloc(2 * W * 8);
    for (uint32_t l = 0; l < limbs; l++) {
        orc_automorphism(in0 + (size_t)l * N, s0 + (size_t)l * N, packed, N);
        orc_automorphism(in1 + (size_t)l * N, s1 + (size_t)l * N, packed, N);
    }
    orc_key_switch(c, s1, km, evk, key_limbs, ks, ks + W, limbs);
    poly_binop(c, 0, s0, ks, out0, limbs);
    memcpy(out1, ks + W, W * 8);
    free(map); free(packed); free(s0); free(s1); free(ks);
}

/* fhe_host_driver.cpp:361-404: signed power-of-two decomposition of total_rot,
 * largest step first; a missing step key is an error (:336-339). */
int orc_eval_rotate(const orc_ctx *c, uint64_t *pm, uint64_t pw, const uint64_t *km, uint64_t kw,
                    uint32_t n_keys, const int32_t *steps, const uint64_t *galois,
                    const uint64_t *evk_offs, const uint32_t *key_limbs, uint64_t in0,
                    uint64_t in1, uint64_t d0, uint64_t d1, int32_t total_rot, uint32_t limbs) {
    if (limbs == 0 || limbs > c->L) return ST_ARG;
    uint64_t W = (uint64_t)limbs * c->N;
    CHECK_RANGE(in0, W); CHECK_RANGE(in1, W); CHECK_RANGE(d0, W); CHECK_RANGE(d1, W);
    uint64_t *c0 = (uint64_t *)malloc(W * 8), *c1 = (uint64_t *)malloc(W * 8);
    uint64_t *n0 = (uint64_t *)malloc(W * 8), *n1 = (uint64_t *)malloc(W * 8);
    memcpy(c0, pm + in0, W * 8);
    memcpy(c1, pm + in1, W * 8);
    int abs_rot = total_rot < 0 ? -total_rot : total_rot;
    int sign = total_rot > 0 ? 1 : -1;
    int st = ST_OK;
    while (abs_rot > 0) {
        int step_size = 1;
        while ((step_size << 1) <= abs_rot) step_size <<= 1;
        int step = sign * step_size;
        int found = -1;
        for (uint32_t i = 0; i < n_keys; i++)
            if (steps[i] == step) found = (int)i;
        if (found < 0) { st = ST_NO_KEY; break; }
        uint32_t kl = key_limbs[found] ? key_limbs[found] : limbs;
        if (kl < limbs || (evk_offs[found] & 7) ||
            !in_range(evk_offs[found], 2ull * c->dnum * (kl + c->K) * c->N, kw)) { st = ST_RANGE; break; }
        rotate_step(c, km, galois[found], evk_offs[found], kl, c0, c1, n0, n1, limbs);
        uint64_t *t = c0; c0 = n0; n0 = t;
        t = c1; c1 = n1; n1 = t;
        abs_rot -= step_size;
    }
    if (st == ST_OK) {
        memcpy(pm + d0, c0, W * 8);
        memcpy(pm + d1, c1, W * 8);
    }
    free(c0); free(c1); free(n0); free(n1);
    return st;
}

int orc_eval_automorphism_ks(const orc_ctx *c, uint64_t *pm, uint64_t pw, const uint64_t *km,
                             uint64_t kw, uint64_t galois, uint64_t evk, uint32_t key_limbs,
                             uint64_t in0, uint64_t in1, uint64_t d0, uint64_t d1, uint32_t limbs) {
    if (limbs == 0 || limbs > c->L) return ST_ARG;
    uint64_t W = (uint64_t)limbs * c->N;
    CHECK_RANGE(in0, W); CHECK_RANGE(in1, W); CHECK_RANGE(d0, W); CHECK_RANGE(d1, W);
    uint32_t kl = key_limbs ? key_limbs : limbs;
    if (kl < limbs) return ST_ARG;
    if ((evk & 7) || !in_range(evk, 2ull * c->dnum * (kl + c->K) * c->N, kw)) return ST_RANGE;
    uint64_t *o0 = (uint64_t *)malloc(W * 8), *o1 = (uint64_t *)malloc(W * 8);
    rotate_step(c, km, galois, evk, kl, pm + in0, pm + in1, o0, o1, limbs);
    memcpy(pm + d0, o0, W * 8);
    memcpy(pm + d1, o1, W * 8);
    free(o0); free(o1);
    return ST_OK;
}

int orc_eval_mult_relin(const orc_ctx *c, uint64_t *pm, uint64_t pw, const uint64_t *km, uint64_t kw,
                        uint64_t relin_evk, uint32_t key_limbs, uint64_t a0, uint64_t a1, uint64_t b0,
                        uint64_t b1, uint64_t d0, uint64_t d1, uint32_t L) {
    if (L < 1 || L > c->L) return ST_ARG;
    uint32_t N = c->N;
    uint32_t kl = key_limbs ? key_limbs : L;
    if (kl < L) return ST_ARG;
    uint64_t W = (uint64_t)L * N;
    CHECK_RANGE(a0, W); CHECK_RANGE(a1, W); CHECK_RANGE(b0, W); CHECK_RANGE(b1, W);
    CHECK_RANGE(d0, W); CHECK_RANGE(d1, W);
    if ((relin_evk & 7) || !in_range(relin_evk, 2ull * c->dnum * (kl + c->K) * N, kw)) return ST_RANGE;
    uint64_t *D0 = (uint64_t *)malloc(W * 8), *D1 = (uint64_t *)malloc(W * 8);
    uint64_t *D2 = (uint64_t *)malloc(W * 8), *T = (uint64_t *)malloc(W * 8);
    uint64_t *KS = (uint64_t *)malloc(2 * W * 8);
    poly_binop(c, 2, pm + a0, pm + b0, D0, L);
    poly_binop(c, 2, pm + a1, pm + b1, D2, L);
    poly_binop(c, 2, pm + a0, pm + b1, D1, L);
    poly_binop(c, 2, pm + a1, pm + b0, T, L);
    poly_binop(c, 0, D1, T, D1, L);
    orc_key_switch(c, D2, km, relin_evk, kl, KS, KS + W, L);
    poly_binop(c, 0, D0, KS, pm + d0, L);
    poly_binop(c, 0, D1, KS + W, pm + d1, L);
    free(D0); free(D1); free(D2); free(T); free(KS);
    return ST_OK;
}

int orc_linear_sum(const orc_ctx *c, uint64_t *pm, uint64_t pw, const uint64_t *ct0, const uint64_t *ct1,
                   uint32_t R, const uint64_t *pt, const uint64_t *ptb, uint32_t G, uint64_t bias0,
                   uint64_t bias1, const uint64_t *out0, const uint64_t *out1, uint32_t L) {
    const uint64_t NONE = ~(uint64_t)0;
    if (L < 1 || L > c->L || R == 0 || G == 0) return ST_ARG;
    uint64_t W = (uint64_t)L * c->N;
    if (bias0 != NONE) { CHECK_RANGE(bias0, W); CHECK_RANGE(bias1, W); }
    for (uint32_t r = 0; r < R; ++r) { CHECK_RANGE(ct0[r], W); CHECK_RANGE(ct1[r], W); }
    for (uint32_t g = 0; g < G; ++g) {
        CHECK_RANGE(out0[g], W); CHECK_RANGE(out1[g], W);
        if (ptb && ptb[g] != NONE) CHECK_RANGE(ptb[g], W);
        for (uint32_t r = 0; r < R; ++r) CHECK_RANGE(pt[(uint64_t)g * R + r], W);
    }
    uint64_t *A0 = (uint64_t *)malloc(W * 8), *A1 = (uint64_t *)malloc(W * 8), *T = (uint64_t *)malloc(W * 8);
    for (uint32_t g = 0; g < G; ++g) {
        memset(A0, 0, W * 8);
        memset(A1, 0, W * 8);
        if (bias0 != NONE) {
            memcpy(A0, pm + bias0, W * 8);
            memcpy(A1, pm + bias1, W * 8);
        }
        if (ptb && ptb[g] != NONE) poly_binop(c, 0, A0, pm + ptb[g], A0, L);
        for (uint32_t r = 0; r < R; ++r) {
            uint64_t p = pt[(uint64_t)g * R + r];
            poly_binop(c, 2, pm + ct0[r], pm + p, T, L);
            poly_binop(c, 0, A0, T, A0, L);
            poly_binop(c, 2, pm + ct1[r], pm + p, T, L);
            poly_binop(c, 0, A1, T, A1, L);
        }
        memcpy(pm + out0[g], A0, W * 8);
        memcpy(pm + out1[g], A1, W * 8);
    }
    free(A0); free(A1); free(T);
    return ST_OK;
}

int orc_eval_rotations_hoisted(const orc_ctx *c, uint64_t *pm, uint64_t pw, const uint64_t *km,
                               uint64_t kw, uint64_t in0, uint64_t in1, uint32_t n,
                               const uint64_t *galois, const uint64_t *evk, uint32_t key_limbs,
                               const uint64_t *d0, const uint64_t *d1, uint32_t L) {
    uint32_t N = c->N;
    if (L < 1 || L > c->L || n == 0) return ST_ARG;
    uint32_t kl = key_limbs ? key_limbs : L;
    if (kl < L || kl > c->L) return ST_ARG;
    uint64_t W = (uint64_t)L * N;
    CHECK_RANGE(in0, W); CHECK_RANGE(in1, W);
    for (uint32_t z = 0; z < n; z++) {
        CHECK_RANGE(d0[z], W); CHECK_RANGE(d1[z], W);
        if ((galois[z] & 1) == 0 || galois[z] >= 2ull * N) return ST_ARG;
        if ((evk[z] & 7) || !in_range(evk[z], 2ull * c->dnum * (kl + c->K) * N, kw)) return ST_RANGE;
    }
    uint32_t QlP = L + c->K;
    uint64_t *c0 = (uint64_t *)malloc(W * 8), *c1 = (uint64_t *)malloc(W * 8);
    memcpy(c0, pm + in0, W * 8);
    memcpy(c1, pm + in1, W * 8);
    uint64_t *digits = (uint64_t *)malloc((size_t)num_parts(c, L) * QlP * N * 8);
    modup_digits(c, c1, L, digits);
    uint32_t *map = (uint32_t *)malloc((size_t)N * 4);
    uint64_t *ks0 = (uint64_t *)malloc(W * 8), *ks1 = (uint64_t *)malloc(W * 8);
    for (uint32_t z = 0; z < n; z++) {
        orc_galois_map(N, galois[z], map);
        inner_moddown(c, digits, map, km, evk[z], kl, ks0, ks1, L);
        for (uint32_t i = 0; i < L; i++) {
            uint64_t q = c->primes[i];
            for (uint32_t r = 0; r < N; r++) {
                uint64_t a = c0[(size_t)i * N + map[r]];
                pm[d0[z] + (size_t)i * N + r] = orc_mod_add(a, ks0[(size_t)i * N + r], q);
                pm[d1[z] + (size_t)i * N + r] = ks1[(size_t)i * N + r];
            }
        }
    }
    free(c0); free(c1); free(digits); free(map); free(ks0); free(ks1);
    return ST_OK;
}
