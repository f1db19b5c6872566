/*
 * fhe_oracle.h -- CPU restatement of the reference's CKKS evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * engine in anns_fhe_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (or the
 * timed CPU baseline), never as part of the product path.
 *
 * Every function restates one function of the reference HLS C-model
 * (hardware_fpga_model/hls/src/ *.cpp files) or of its host driver
 * (cpp_host_model/src/fhe_host_driver.cpp), generalised from the compiled-in
 * N=16384 / MAX_LIMBS=15 / P<=4 / 3-digit bounds to runtime N, L, K, dnum.
 * Each definition in fhe_oracle.c cites the reference file:line it follows.
 *
 * Pinning status (what independent check fixes each function) is recorded in
 * the header comment of fhe_oracle.c; all functions are pinned by closed forms,
 * big-integer brute force or decryption invariants in tests/test_oracle_*.py.
 *
 * Arithmetic: all residues are uint64 in [0, q); intermediates use
 * unsigned __int128.  Word offsets into the emulated memories are uint64.
 */
#ifndef FHE_ORACLE_H
#define FHE_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- L0 modular arithmetic (mod_arith.cpp:3-25) ---------------------- */
uint64_t orc_mod_add(uint64_t a, uint64_t b, uint64_t q);
uint64_t orc_mod_sub(uint64_t a, uint64_t b, uint64_t q);
uint64_t orc_mod_mul(uint64_t a, uint64_t b, uint64_t q,
                     uint64_t m_lo, uint64_t m_hi, uint32_t k);
void orc_barrett_constants(uint64_t q, uint64_t *m_lo, uint64_t *m_hi, uint32_t *k);
uint64_t orc_inv_mod(uint64_t a, uint64_t q);
uint64_t orc_pow_mod(uint64_t a, uint64_t e, uint64_t q);
int orc_is_prime(uint64_t n);
uint32_t orc_gen_primes(uint32_t two_n, uint32_t bits, uint32_t count,
                        const uint64_t *skip, uint32_t n_skip, uint64_t *out);
uint64_t orc_min_primitive_root(uint64_t q, uint32_t two_n);
uint32_t orc_bit_reverse(uint32_t x, uint32_t bits);
void orc_twiddles(uint64_t q, uint64_t psi, uint32_t N, uint64_t *tw, uint64_t *itw);

/* ---- L1 polynomial kernels ------------------------------------------- */
void orc_ntt_forward(uint64_t *a, uint32_t N, const uint64_t *tw, uint64_t q,
                     uint64_t m_lo, uint64_t m_hi, uint32_t k);
void orc_intt_inverse(uint64_t *a, uint32_t N, const uint64_t *itw, uint64_t n_inv,
                      uint64_t q, uint64_t m_lo, uint64_t m_hi, uint32_t k);
void orc_automorphism(const uint64_t *a, uint64_t *out, const uint64_t *packed_map,
                      uint32_t N);
void orc_galois_map(uint32_t N, uint64_t g, uint32_t *map_u32);

/* ---- context (parameter block + per-prime tables) -------------------- */
typedef struct orc_ctx orc_ctx;
/* q_primes: L, p_primes: K.  Twiddles use the minimal primitive 2N-th root. */
orc_ctx *orc_ctx_create(uint32_t N, uint32_t L, uint32_t K, uint32_t dnum,
                        const uint64_t *q_primes, const uint64_t *p_primes);
int orc_ctx_set_psi(orc_ctx *c, const uint64_t *psi);
void orc_ctx_destroy(orc_ctx *c);
uint32_t orc_ctx_alpha(const orc_ctx *c);
/* Copies of the per-prime tables, index 0..L-1 = Q, L..L+K-1 = P. */
void orc_ctx_tables(const orc_ctx *c, uint64_t *psi, uint64_t *n_inv,
                    uint64_t *m_lo, uint64_t *m_hi, uint32_t *k,
                    uint64_t *tw, uint64_t *itw);
/* Per-level CRT tables (OpenFHE definitions, test_vector_generator.cpp:408-498). */
uint32_t orc_modup_tables(const orc_ctx *c, uint32_t sizeQl, uint32_t part,
                          uint64_t *qHatInvModq, uint64_t *qHatModp);
void orc_moddown_tables(const orc_ctx *c, uint32_t sizeQl, uint64_t *PInvModq,
                        uint64_t *PHatInvModp, uint64_t *PHatModq);
void orc_rescale_tables(const orc_ctx *c, uint32_t sizeQl,
                        uint64_t *QlQlInvModqlDivqlModq, uint64_t *qlInvModq);

/* ---- L2 CKKS-RNS ops -------------------------------------------------- */
void orc_rescale(const orc_ctx *c, const uint64_t *a_in, uint64_t *a_out, uint32_t num_limbs);
void orc_fast_bases_conv(uint32_t N, const uint64_t *in_poly, uint64_t *out_poly,
                         uint32_t n_in, uint32_t n_out,
                         const uint64_t *in_primes, const uint64_t *out_primes,
                         const uint64_t *qHatInvModq, const uint64_t *qHatModp,
                         const uint64_t *bm_in_lo, const uint64_t *bm_in_hi, const uint32_t *bk_in,
                         const uint64_t *bm_out_lo, const uint64_t *bm_out_hi, const uint32_t *bk_out);
void orc_approx_mod_down(const orc_ctx *c, const uint64_t *in_poly, uint64_t *out_poly,
                         uint32_t sizeQl);
/* Hybrid key switch of one polynomial c_in (sizeQl limbs, eval domain).
 * evk layout (key words): a at evk_off as [dnum][key_limbs+K][N], then b at
 * evk_off + dnum*(key_limbs+K)*N.  Q limb i of digit p is key limb i, P limb j is
 * key limb key_limbs+j.  With key_limbs == sizeQl this is the reference layout. */
void orc_key_switch(const orc_ctx *c, const uint64_t *c_in, const uint64_t *key_mem,
                    uint64_t evk_off, uint32_t key_limbs, uint64_t *c_out0,
                    uint64_t *c_out1, uint32_t sizeQl);

/* ---- L3 op dispatcher over word-addressed memories ------------------- */
/* Status codes mirror the C-ABI (include/fhea.h). */
int orc_accel_op(const orc_ctx *c, uint64_t *poly_mem, uint64_t poly_words,
                 const uint64_t *key_mem, uint64_t key_words, uint32_t op_code,
                 uint64_t src_a, uint64_t src_b, uint64_t dst, uint64_t evk_off,
                 uint32_t active_limbs, uint32_t key_limbs);

/* ---- L5 composites (fhe_host_driver.cpp:248-404) --------------------- */
int orc_eval_add(const orc_ctx *c, uint64_t *poly_mem, uint64_t poly_words,
                 uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1,
                 uint64_t d0, uint64_t d1, uint32_t limbs);
int orc_eval_sub_plain(const orc_ctx *c, uint64_t *poly_mem, uint64_t poly_words,
                       uint64_t a0, uint64_t a1, uint64_t pt, uint64_t d0, uint64_t d1,
                       uint32_t limbs);
int orc_eval_mult_relin_rescale(const orc_ctx *c, uint64_t *poly_mem, uint64_t poly_words,
                                const uint64_t *key_mem, uint64_t key_words,
                                uint64_t relin_evk, uint32_t key_limbs,
                                uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1,
                                uint64_t d0, uint64_t d1, uint32_t in_limbs);
int orc_eval_mult_plain_rescale(const orc_ctx *c, uint64_t *poly_mem, uint64_t poly_words,
                                uint64_t a0, uint64_t a1, uint64_t pt,
                                uint64_t d0, uint64_t d1, uint32_t in_limbs);
/* Rotation-key directory: n_keys entries of (step, galois element, evk offset,
 * key limbs).  The Galois map is the closed form of orc_galois_map. */
int orc_eval_rotate(const orc_ctx *c, uint64_t *poly_mem, uint64_t poly_words,
                    const uint64_t *key_mem, uint64_t key_words, uint32_t n_keys,
                    const int32_t *steps, const uint64_t *galois, const uint64_t *evk_offs,
                    const uint32_t *key_limbs, uint64_t in0, uint64_t in1,
                    uint64_t d0, uint64_t d1, int32_t total_rot, uint32_t limbs);
/* One automorphism + key switch with an explicit Galois element (the batched
 * single-g rotation of the key-switching benchmark). */
int orc_eval_automorphism_ks(const orc_ctx *c, uint64_t *poly_mem, uint64_t poly_words,
                             const uint64_t *key_mem, uint64_t key_words, uint64_t galois,
                             uint64_t evk_off, uint32_t key_limbs, uint64_t in0,
                             uint64_t in1, uint64_t d0, uint64_t d1, uint32_t limbs);

/* The multiply + relinearise half of eval_mult_relin_rescale
 * (fhe_host_driver.cpp:280-294): d = (a0 b0 + ks0(a1 b1), a0 b1 + a1 b0 + ks1(a1 b1)),
 * at `limbs` limbs, no rescale. */
int orc_eval_mult_relin(const orc_ctx *c, uint64_t *poly_mem, uint64_t poly_words,
                        const uint64_t *key_mem, uint64_t key_words, uint64_t relin_evk,
                        uint32_t key_limbs, uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1,
                        uint64_t d0, uint64_t d1, uint32_t limbs);
/* out_g = bias + ptb[g] + sum_r ct_r (.) pt[g*R + r] by OP 5 / OP 3 per term
 * (poly_arith.cpp:3-42); FHEA_NONE (all ones) marks an absent bias / ptb entry. */
int orc_linear_sum(const orc_ctx *c, uint64_t *poly_mem, uint64_t poly_words, const uint64_t *ct0,
                   const uint64_t *ct1, uint32_t R, const uint64_t *pt, const uint64_t *ptb,
                   uint32_t G, uint64_t bias0, uint64_t bias1, const uint64_t *out0,
                   const uint64_t *out1, uint32_t limbs);

/* Rotations of one ciphertext (in0, in1) by n Galois elements sharing one ModUp:
 * the digits of c1 are computed once and read through each sigma_g's
 * evaluation-domain map; out_z = (sigma_g(c0) + ks0, ks1).  The engine's hoisted
 * rotation (no reference counterpart), composed of pinned parts: modup_digits and
 * inner_moddown are the two halves of orc_key_switch. */
int orc_eval_rotations_hoisted(const orc_ctx *c, uint64_t *poly_mem, uint64_t poly_words,
                               const uint64_t *key_mem, uint64_t key_words, uint64_t in0,
                               uint64_t in1, uint32_t n, const uint64_t *galois,
                               const uint64_t *evk, uint32_t key_limbs, const uint64_t *d0,
                               const uint64_t *d1, uint32_t limbs);

#ifdef __cplusplus
}
#endif
#endif
