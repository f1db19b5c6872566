/*
 * fhea.h -- C-ABI of the MI355X CKKS evaluation engine (libfhea.so).
 *
 * Drop-in boundary for the reference's accelerator interface:
 *   - the op-code kernel  fhe_accel_top(...)
 *       hardware_fpga_model/hls/src/fhe_accel_top.h:8-55, op codes hls_params.h:18-26
 *   - the emulated device memories  SimMemoryBus
 *       cpp_host_model/src/sim_memory_bus.h:10-34
 *   - the composite calls of  FHEHostDriver
 *       cpp_host_model/src/fhe_host_driver.h:74-111
 *
 * Conventions kept from the reference:
 *   - polynomials are limb-major: limb i of an L-limb polynomial occupies words
 *     [i*N, (i+1)*N) from its offset (docs/data_format.md:51-73);
 *   - all offsets are 64-bit WORD offsets into the poly or key arena and must be
 *     multiples of 8 (one 512-bit beat, fhe_accel_top.cpp:39);
 *   - op-code output placement: OP 7 writes ks0 at dst and ks1 at dst+L*N,
 *     OP 8 writes c0' at dst and c1' at dst+(L-1)*N (fhe_accel_top.cpp:282-326);
 *   - evaluation keys (OP 7): a-part at evk as [dnum][key_limbs+K][N], b-part at
 *     evk + dnum*(key_limbs+K)*N (fhe_accel_top.cpp:305).  Q limb i uses key limb
 *     i and P limb j uses key limb key_limbs+j; key_limbs == L (or 0) is exactly
 *     the reference layout, key_limbs > L lets one full-level key serve every level.
 *
 * Differences (deliberate, see the call-by-call comments):
 *   - N, L, K and dnum are runtime parameters (the reference compiles in N=16384,
 *     MAX_LIMBS=15, K<=4, 3 digits; hls_params.h:6-8, key_switch.cpp:271-272);
 *   - the context tables are generated at fhea_ctx_create (replaces OP 99 and the
 *     table loading of fhe_host_driver.cpp:18-175) for EVERY level, not one;
 *   - every call returns an int status (FHEA_*), never throws, and rejects
 *     unknown op codes, unaligned or out-of-range offsets and missing keys
 *     (the reference ignores, truncates or throws; fhe_accel_top.cpp:220-327,
 *     fhe_host_driver.cpp:336-339, sim_memory_bus.cpp:13-15).
 *
 * Threading: a context is externally synchronised (one host thread at a time);
 * all work is issued on the context's HIP stream in call order.
 */
#ifndef FHEA_H
#define FHEA_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define FHEA_OK 0
#define FHEA_ERR_BAD_OPCODE (-1) /* unknown op code                              */
#define FHEA_ERR_RANGE (-2)      /* offset/size outside an arena                 */
#define FHEA_ERR_ALIGN (-3)      /* offset not a multiple of 8 words             */
#define FHEA_ERR_ARG (-4)        /* bad limb count / parameter                   */
#define FHEA_ERR_NO_KEY (-5)     /* missing relinearisation / rotation key       */
#define FHEA_ERR_NO_MEM (-6)     /* device allocation failed / arena exhausted   */
#define FHEA_ERR_DEVICE (-7)     /* HIP runtime error                            */
#define FHEA_ERR_STATE (-8)      /* null context / not initialised               */
#define FHEA_ERR_SCRATCH (-9)    /* work exceeds the no-allocation scratch area  */

/* ---- op codes (hls_params.h:18-26) -------------------------------------- */
#define FHEA_OP_NTT 1
#define FHEA_OP_INTT 2
#define FHEA_OP_POLY_ADD 3
#define FHEA_OP_POLY_SUB 4
#define FHEA_OP_POLY_MUL_NTT 5
#define FHEA_OP_AUTOMORPHISM 6
#define FHEA_OP_KEY_SWITCH 7
#define FHEA_OP_RESCALE 8
#define FHEA_OP_LOAD_CONTEXT 99

typedef struct fhea_ctx fhea_ctx;

/* Parameter block: replaces FHEContextConfig (fhe_host_driver.h:17-53). */
typedef struct {
    uint32_t N;                /* ring dimension, power of two, 2^10..2^17        */
    uint32_t L;                /* number of Q primes (top level)                  */
    uint32_t K;                /* number of special P primes                      */
    uint32_t dnum;             /* key-switching digits at level L (alpha=ceil(L/dnum)) */
    uint32_t q_bits;           /* bit size for generated primes (<= 61), 0 = 50   */
    uint32_t device;           /* HIP device ordinal                              */
    const uint64_t *q_primes;  /* optional explicit Q primes (L), NULL = generate */
    const uint64_t *p_primes;  /* optional explicit P primes (K), NULL = generate */
    uint64_t poly_words;       /* poly arena size in 64-bit words (poly_gmem)      */
    uint64_t key_words;        /* key arena size in 64-bit words (key_gmem)        */
    const uint64_t *psi;       /* optional primitive 2N-th root of unity per prime
                                  (L+K, Q then P); NULL = the smallest one.  Lets
                                  tables dumped by another library (whose twiddle
                                  file holds psi at index N/2) be reproduced.    */
} fhea_params;

/* ---- context (replaces OP 99 + initialize_context_from_directory,
 *      fhe_host_driver.cpp:18-175,206-208; fhe_accel_top.cpp:194-218) ------- */
int fhea_ctx_create(const fhea_params *params, fhea_ctx **out);
int fhea_ctx_destroy(fhea_ctx *ctx);
/* N, L, K, dnum, alpha */
int fhea_ctx_info(const fhea_ctx *ctx, uint32_t *info5);
/* L+K primes, Q first (rns_primes_Q.bin / rns_primes_P.bin). */
int fhea_ctx_primes(const fhea_ctx *ctx, uint64_t *out);
/* Exported tables for verification: which = 0 psi[T], 1 n_inv[T], 2 tw[T*N],
 * 3 itw[T*N] (twiddles.bin / inv_twiddles.bin layout). */
int fhea_ctx_table(const fhea_ctx *ctx, int which, uint64_t *out, uint64_t count);
/* Use an external HIP stream (hipStream_t) for all subsequent work; NULL = own. */
int fhea_ctx_set_stream(fhea_ctx *ctx, void *hip_stream);
void *fhea_ctx_stream(const fhea_ctx *ctx);
const char *fhea_status_string(int status);

/* ---- arenas (SimMemoryBus::write_poly/read_poly/write_key/read_key,
 *      sim_memory_bus.cpp:12-64) ------------------------------------------- */
int fhea_write_poly(fhea_ctx *ctx, uint64_t word_off, const uint64_t *host, uint64_t count);
int fhea_read_poly(fhea_ctx *ctx, uint64_t word_off, uint64_t *host, uint64_t count);
/* n arena ranges [off[k], off[k] + count[k]) copied back to back into h (one
 * synchronisation; page-locked h from fhea_host_alloc makes the copies DMA). */
int fhea_read_polys(fhea_ctx *ctx, uint32_t n, const uint64_t *off, const uint64_t *count, uint64_t *h);
/* The reverse: h (the ranges back to back) into n arena ranges, one synchronisation
 * (SimMemoryBus::write_poly per range, sim_memory_bus.cpp:12-30). */
int fhea_write_polys(fhea_ctx *ctx, uint32_t n, const uint64_t *off, const uint64_t *count,
                     const uint64_t *h);
/* The same gathered read, returned at once: the copies run on a copy stream after
 * the work submitted so far, beside the kernels submitted next; *ticket names the
 * read for fhea_read_wait (a ticket 16 or more reads old is complete already: the
 * call returns FHEA_OK at once; a ticket not yet issued is FHEA_ERR_ARG).  Work submitted before
 * the wait must not write the ranges; h must stay valid until then (page-locked h
 * lets the copies overlap kernels).  No reference counterpart (reads overlapped with
 * later evaluation, for the batched query stages). */
int fhea_read_polys_async(fhea_ctx *ctx, uint32_t n, const uint64_t *off, const uint64_t *count,
                          uint64_t *h, uint32_t *ticket);
int fhea_read_wait(fhea_ctx *ctx, uint32_t ticket);
int fhea_write_key(fhea_ctx *ctx, uint64_t word_off, const uint64_t *host, uint64_t count);
int fhea_read_key(fhea_ctx *ctx, uint64_t word_off, uint64_t *host, uint64_t count);
int fhea_copy_poly(fhea_ctx *ctx, uint64_t dst_off, uint64_t src_off, uint64_t count);
/* Device base pointers of the arenas (poly_gmem_ptr / key_gmem_ptr, sim_memory_bus.h:16-17). */
/* Free and total device memory of HIP device `device` in bytes (hipMemGetInfo), for
 * sizing arenas and admission budgets from what the GPU actually has.  No reference
 * counterpart (its memories are fixed-size host vectors, sim_memory_bus.h:13). */
int fhea_mem_info(uint32_t device, uint64_t *free_bytes, uint64_t *total_bytes);
/* VALU issue rate of one instruction kind on HIP device `device`, in lane operations
 * per second: kind 0 v_mul_lo_u32, 1 v_mul_hi_u32, 2 v_mad_u64_u32 (the 32-bit
 * multiplies a 64-bit modular product lowers to), 3 v_fma_f64.  Eight workgroups of
 * 256 threads per CU, eight independent chains of `iters` dependent steps per thread.
 * A diagnostic for the integer kernels' op-rate ceiling; no reference counterpart. */
int fhea_probe_valu_rate(uint32_t device, uint32_t kind, uint32_t iters, double *lane_ops_per_s);
/* Page-locked host memory (hipHostMalloc) for arena reads and writes at full DMA
 * speed; NULL on failure.  Free with fhea_host_free. */
void *fhea_host_alloc(uint64_t bytes);
int fhea_host_free(void *p);
uint64_t *fhea_poly_arena(fhea_ctx *ctx);
uint64_t *fhea_key_arena(fhea_ctx *ctx);
uint64_t fhea_poly_words(const fhea_ctx *ctx);
uint64_t fhea_key_words(const fhea_ctx *ctx);

/* ---- L3: one op code (fhe_accel_top(...) / FHEHostDriver::dispatch,
 *      fhe_accel_top.cpp:220-327, fhe_host_driver.cpp:177-242) --------------
 * src_a, src_b, dst: poly-arena word offsets; evk: key-arena word offset;
 * active_limbs: L of the operands; key_limbs: stored Q limbs of the key (OP 7),
 * 0 = active_limbs.  OP 6 reads the packed Galois map (N/2 words, two u32 per
 * word, low half = even index) at src_b.  OP 99 is accepted as a no-op (tables
 * are built at context creation).  In-place operation (dst == src) is allowed. */
int fhea_op(fhea_ctx *ctx, uint32_t op_code, uint64_t src_a, uint64_t src_b, uint64_t dst,
            uint64_t evk, uint32_t active_limbs, uint32_t key_limbs);

/* ---- L5: composites (FHEHostDriver, fhe_host_driver.cpp:248-404) --------- */
int fhea_eval_add(fhea_ctx *ctx, uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1,
                  uint64_t d0, uint64_t d1, uint32_t limbs);
int fhea_eval_sub_plain(fhea_ctx *ctx, uint64_t a0, uint64_t a1, uint64_t pt,
                        uint64_t d0, uint64_t d1, uint32_t limbs);
/* relin key location (FHEContextConfig::relin_key_offset, fhe_host_driver.h:52). */
int fhea_set_relin_key(fhea_ctx *ctx, uint64_t evk, uint32_t key_limbs);
/* d = Rescale(Relin(a * b)), in_limbs -> in_limbs-1 (fhe_host_driver.cpp:272-306). */
int fhea_eval_mult_relin_rescale(fhea_ctx *ctx, uint64_t a0, uint64_t a1, uint64_t b0,
                                 uint64_t b1, uint64_t d0, uint64_t d1, uint32_t in_limbs);
/* No reference counterpart.  The same words as fhea_eval_mult_relin_rescale, for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): no hipMalloc/hipFree or other
 * allocating call is made; the pipeline's scratch is a static device array of
 * FHEA_NOALLOC_SCRATCH_WORDS words (64 MiB) that every GPU the library runs on holds
 * until the process ends.  Work needing more than that, i.e.
 * N * (3*l + max(l + nparts*(l+K) + 2*(l+K) + 2*K + 2*l, 2*l)) words for one
 * ciphertext pair at level l (nparts = ceil(l/alpha)), returns FHEA_ERR_SCRATCH
 * (N=2^14 fits up to l=24 with K<=8; N=2^16, L=24, K=8 does not).  Every call on one
 * GPU shares that GPU's array, which keeps the last call's contents: calls that
 * overlap on one GPU (contexts with different streams, or several host threads)
 * race and must be serialised by the caller.  Not allowed while capturing with
 * fhea_graph_begin (FHEA_ERR_STATE). */
#define FHEA_NOALLOC_SCRATCH_WORDS (8u << 20)
int fhea_eval_mult_relin_rescale_noalloc(fhea_ctx *ctx, uint64_t a0, uint64_t a1, uint64_t b0,
                                         uint64_t b1, uint64_t d0, uint64_t d1, uint32_t in_limbs);
/* d = Rescale(a * pt) (fhe_host_driver.cpp:308-329). */
int fhea_eval_mult_plain_rescale(fhea_ctx *ctx, uint64_t a0, uint64_t a1, uint64_t pt,
                                 uint64_t d0, uint64_t d1, uint32_t in_limbs);
/* Rotation-key directory entry (RotationKeyMeta, fhe_host_driver.h:10-15):
 * step -> Galois element + key location. */
int fhea_add_rotation_key(fhea_ctx *ctx, int32_t step, uint64_t galois, uint64_t evk,
                          uint32_t key_limbs);
/* Rotation by total_rot via its signed power-of-two decomposition, largest step
 * first (fhe_host_driver.cpp:361-404); FHEA_ERR_NO_KEY if a step key is missing. */
int fhea_eval_rotate(fhea_ctx *ctx, uint64_t in0, uint64_t in1, uint64_t d0, uint64_t d1,
                     int32_t total_rot, uint32_t limbs);
/* One step of eval_rotate_single_step (fhe_host_driver.cpp:331-359) with an
 * explicit Galois element and key: d = (sigma_g(c0) + ks0, ks1), ks = KS(sigma_g(c1)). */
int fhea_eval_automorphism_ks(fhea_ctx *ctx, uint64_t galois, uint64_t evk, uint32_t key_limbs,
                              uint64_t in0, uint64_t in1, uint64_t d0, uint64_t d1,
                              uint32_t limbs);

/* ---- batched submission (no reference counterpart: the reference issues one
 *      ciphertext at a time).  Items of the same kind and level are fused into
 *      one set of launches over the whole batch. ---------------------------- */
#define FHEA_WORK_MULT_RELIN_RESCALE 1 /* a0,a1,b0,b1 -> d0,d1                    */
#define FHEA_WORK_AUTOMORPHISM_KS 2    /* a0,a1 -> d0,d1 with galois, evk         */
#define FHEA_WORK_NTT 3                /* a0 -> d0 (limbs), forward               */
#define FHEA_WORK_INTT 4               /* a0 -> d0 (limbs), inverse               */
#define FHEA_WORK_MULT_RELIN 5         /* a0,a1,b0,b1 -> d0,d1, no rescale        */
#define FHEA_WORK_ADD 6                /* (a0,a1) + (b0,b1) -> (d0,d1)            */
#define FHEA_WORK_ROTATE_HOISTED 7     /* like 2; consecutive items with the same
                                          input share one ModUp (sigma_g applied to
                                          the digits; outputs must not overlap the
                                          input).  No reference counterpart.       */
typedef struct {
    uint32_t kind;
    uint32_t limbs;
    uint64_t a0, a1, b0, b1;
    uint64_t d0, d1;
    uint64_t galois;
    uint64_t evk;
    uint32_t key_limbs;
    uint32_t reserved;
} fhea_work;
/* Consecutive items of the same kind, level and key level form one group, run as
 * one pipeline; groups run in order.  Within a group of kinds 1, 2, 5 and 6 an item
 * may write over its own inputs, but no output may overlap another output or an
 * input of another item of the group (FHEA_ERR_ARG): the group's items run in
 * chunks and, with FHEA_TUNE_STREAMS = n > 1, on n streams at once. */
int fhea_submit(fhea_ctx *ctx, const fhea_work *items, uint32_t n_items);
/* Batched forward (inverse != 0: inverse) NTT of npolys contiguous polynomials
 * of `limbs` limbs each (poly z at src + z*limbs*N), limb i mod prime i
 * (OP 1 / OP 2 over a whole batch in one set of launches). */
int fhea_ntt_batch(fhea_ctx *ctx, uint64_t src, uint64_t dst, uint32_t npolys, uint32_t limbs,
                   int inverse);
/* Wait for all work issued on the context's stream. */
int fhea_sync(fhea_ctx *ctx);

/* Rescale (OP 8 semantics) of n ciphertexts stored back to back: ciphertext z at
 * src + z*2*limbs*N (c0 then c1) -> dst + z*2*(limbs-1)*N.  Ranges must not overlap. */
int fhea_rescale_batch(fhea_ctx *ctx, uint64_t src, uint64_t dst, uint32_t n, uint32_t limbs);
/* The same with ciphertext z written to dst[z] (2*(limbs-1)*N words each); the
 * outputs must not overlap each other or the inputs. */
int fhea_rescale_scatter(fhea_ctx *ctx, uint64_t src, const uint64_t *dst, uint32_t n, uint32_t limbs);

/* HIP graph capture.  Engine work issued on the context between fhea_graph_begin
 * and fhea_graph_end (submits, ops, composites, NTT batches, linear sums, device
 * copies) is recorded instead of run; fhea_graph_launch replays it on the context's
 * stream with the same word offsets, so new data written to those offsets is
 * processed.  While capturing, calls that synchronise with the host (reads, writes,
 * fhea_sync, timers, set_stream) return FHEA_ERR_STATE, and work that needs a larger
 * workspace than is allocated returns FHEA_ERR_NO_MEM: run the same work once
 * before capturing it.  No reference counterpart (the FPGA kernel is called op by op). */
int fhea_graph_begin(fhea_ctx *ctx);
int fhea_graph_end(fhea_ctx *ctx, uint32_t *graph_id);
int fhea_graph_launch(fhea_ctx *ctx, uint32_t graph_id);
int fhea_graph_destroy(fhea_ctx *ctx, uint32_t graph_id);

/* Schedule tuning (results are identical for every setting; only speed changes).
 * Used by the autotuner (anns_fhe_amd/autotune.py). */
#define FHEA_TUNE_NTT_SPLIT_FWD 1 /* stages in the first pass of a two-pass forward NTT
                                     (both passes must have 6..12 stages), 0 = logN/2;
                                     logN itself (N = 2^13, 2^14): one pass per limb */
#define FHEA_TUNE_NTT_SPLIT_INV 2 /* the same for the inverse (its second pass)      */
#define FHEA_TUNE_NTT_WG_PER_CU 3 /* resident NTT workgroups per CU, 1..4, 0 = by occupancy */
#define FHEA_TUNE_CHUNK 4         /* batch items per key-switch pipeline chunk, 1..64 */
#define FHEA_TUNE_STREAMS 5       /* 1..4: a batched HMult runs its items in that many
                                     parts on as many streams, overlapping on the device */
int fhea_ctx_set_tuning(fhea_ctx *ctx, int knob, uint32_t value);
int fhea_ctx_get_tuning(const fhea_ctx *ctx, int knob, uint32_t *value);

/* Sentinel for an absent optional offset. */
#define FHEA_NONE (~(uint64_t)0)

/* Plaintext-matrix x encrypted-vector inner products by the diagonal method:
 *   out_g = bias + ptb[g] + sum_{r<R} ct_r (.) pt[g*R + r]     (g < G, on `limbs` limbs)
 * ct_r = (ct0[r], ct1[r]) (typically the rotations of one query by r), pt and ptb
 * plaintext offsets (ptb added to c0 only; ptb may be NULL or hold FHEA_NONE),
 * bias = (bias0, bias1) a ciphertext or FHEA_NONE twice.  Outputs (out0[g], out1[g])
 * must not overlap any input or each other (FHEA_ERR_ARG).  No reference
 * counterpart: it replaces the per-batch rotate-and-sum loop of
 * cpp_host_model/src/interactive_query_manager.cpp:190-274 with one pass per
 * candidate block. */
int fhea_eval_linear_sum(fhea_ctx *ctx, const uint64_t *ct0, const uint64_t *ct1, uint32_t R,
                         const uint64_t *pt, const uint64_t *ptb, uint32_t G, uint64_t bias0,
                         uint64_t bias1, const uint64_t *out0, const uint64_t *out1, uint32_t limbs);

/* The same for n_jobs jobs in one launch (for example every query of a batch
 * against its candidate blocks).  Job j: ct0/ct1[j*R + r], G[j] blocks whose
 * plaintexts, ptb and outputs follow those of job j-1 in pt (G[j]*R each), ptb,
 * out0, out1; bias0/bias1[j] (both arrays NULL, or per job a ciphertext or
 * FHEA_NONE twice).  No output may overlap any input of any job or another
 * output (FHEA_ERR_ARG).  No reference counterpart (see fhea_eval_linear_sum). */
int fhea_eval_linear_sum_batch(fhea_ctx *ctx, uint32_t n_jobs, uint32_t R, const uint64_t *ct0,
                               const uint64_t *ct1, const uint32_t *G, const uint64_t *pt,
                               const uint64_t *ptb, const uint64_t *bias0, const uint64_t *bias1,
                               const uint64_t *out0, const uint64_t *out1, uint32_t limbs);
/* Groups of queries that share their plaintext blocks (for example the queries
 * probing one cluster): group k has nq[k] ciphertext sets (ct0/ct1: nq[k]*R
 * offsets, bias0/bias1: nq[k]) and ng[k] blocks (pt: ng[k]*R, ptb: ng[k]), and
 * computes every (query, block) pair, outputs out0/out1[q*ng[k] + g]; the groups'
 * arrays follow one another.  The FP64 path evaluates up to 16 queries x 8 blocks
 * per workgroup, so each ciphertext and plaintext word read serves several pairs.
 * Overlap rules as for fhea_eval_linear_sum_batch.  No reference counterpart. */
int fhea_eval_linear_sum_groups(fhea_ctx *ctx, uint32_t n_groups, uint32_t R, const uint32_t *nq,
                                const uint32_t *ng, const uint64_t *ct0, const uint64_t *ct1,
                                const uint64_t *bias0, const uint64_t *bias1, const uint64_t *pt,
                                const uint64_t *ptb, const uint64_t *out0, const uint64_t *out1,
                                uint32_t limbs);

/* ---- timing on the context's stream (HIP events) ----------------------- */
int fhea_timer_start(fhea_ctx *ctx);
/* milliseconds since fhea_timer_start, synchronising the stream */
int fhea_timer_stop(fhea_ctx *ctx, float *ms);

/* ---- randomness of key generation and encryption -----------------------
 * Every seed argument below selects the source: FHEA_SEED_RANDOM draws from a
 * cryptographically secure generator (ChaCha20, RFC 8439, keyed from the OS with
 * getrandom(2); the reference relies on OpenFHE's PRNG for this); any other value
 * seeds std::mt19937_64, which is reproducible (tests, checks against the oracle)
 * and NOT secure: its state follows from the outputs a public key exposes. */
#define FHEA_SEED_RANDOM (~(uint64_t)0)
/* The keystream block of that generator: out = ChaCha20(key[8], counter, nonce[3])
 * (RFC 8439 section 2.3), exported so that its known-answer vectors can be checked. */
int fhea_chacha20_block(const uint32_t *key, uint32_t counter, const uint32_t *nonce, uint32_t *out);

/* ---- CPU client (key generation), the role of OpenFHE KeyGen /
 *      EvalMultKeyGen / EvalRotateKeyGen in
 *      integration_tools/generate_interactive_test_vectors.cpp:223-224 -------
 * s: N ternary coefficients in {-1,0,1} (int8).  Keys are written into the key
 * arena at evk in the layout described above, for level key_limbs. */
int fhea_keygen_secret(const fhea_ctx *ctx, uint64_t seed, int8_t *s_out);
/* hybrid key switching s_new -> s; s_new given by int64 coefficients. */
int fhea_keygen_switch_key(fhea_ctx *ctx, const int8_t *s, const int64_t *s_new,
                           uint32_t key_limbs, uint64_t seed, uint64_t evk);
int fhea_keygen_relin_key(fhea_ctx *ctx, const int8_t *s, uint32_t key_limbs, uint64_t seed,
                          uint64_t evk);
int fhea_keygen_rotation_key(fhea_ctx *ctx, const int8_t *s, uint64_t galois, uint32_t key_limbs,
                             uint64_t seed, uint64_t evk);
/* Key size in words for one key at level key_limbs. */
uint64_t fhea_key_size_words(const fhea_ctx *ctx, uint32_t key_limbs);
/* Galois element for a slot rotation by step: 5^step mod 2N (step may be negative). */
uint64_t fhea_galois_for_step(const fhea_ctx *ctx, int32_t step);

/* ---- Host-only CKKS client: the role OpenFHE plays for the reference's client
 *      (MakeCKKSPackedPlaintext / Encrypt / Decrypt / GetRealPackedValue and the
 *      key generators, software_golden_model/fhe_core/fhe_context_manager.cpp:81-128,
 *      457-512; integration_tools/generate_interactive_test_vectors.cpp:161-224).
 *      No GPU is touched; the same parameter block gives the same primes as
 *      fhea_ctx_create.  Polynomials are limb-major, evaluation domain.
 *      Slot j of a plaintext is m(zeta^(5^j mod 2N)) / scale, zeta = exp(i pi / N);
 *      n values (n a power of two, n <= N/2) are replicated with period n.       */
typedef struct fhea_client fhea_client;
int fhea_client_create(const fhea_params *params, fhea_client **out); /* N >= 16 allowed */
int fhea_client_from_ctx(const fhea_ctx *ctx, fhea_client **out);
int fhea_client_destroy(fhea_client *cl);
int fhea_client_info(const fhea_client *cl, uint32_t *info5);  /* N, L, K, dnum, alpha */
int fhea_client_primes(const fhea_client *cl, uint64_t *out);   /* L+K primes        */
/* which: 0 psi (L+K), 1 N^-1 (L+K), 2 twiddles, 3 inverse twiddles ((L+K)*N); as fhea_ctx_table */
int fhea_client_table(const fhea_client *cl, int which, uint64_t *out, uint64_t count);
/* CPU transforms of this client: 0 = FP64 (exact products, AVX2+FMA) for primes below
 * 2^50 when the CPU has it, else 64-bit integer (default); 1 = always 64-bit integer.
 * Both give the same words (no reference counterpart: OpenFHE picks its own). */
int fhea_client_set_host_ntt(fhea_client *cl, int mode);
int fhea_client_keygen_secret(const fhea_client *cl, uint64_t seed, int8_t *s_out);
/* evaluation keys into a host buffer of fhea_key_size_words words, same layout as the arena */
int fhea_client_switch_key(const fhea_client *cl, const int8_t *s, const int64_t *s_new,
                           uint32_t key_limbs, uint64_t seed, uint64_t *key_out);
int fhea_client_relin_key(const fhea_client *cl, const int8_t *s, uint32_t key_limbs, uint64_t seed,
                          uint64_t *key_out);
int fhea_client_rotation_key(const fhea_client *cl, const int8_t *s, uint64_t galois,
                             uint32_t key_limbs, uint64_t seed, uint64_t *key_out);
/* re/im: n values (im may be NULL); pt: limbs*N words.  Coefficients are rounded
 * to integers and reduced exactly (any scale up to 2^1000). */
int fhea_encode(const fhea_client *cl, const double *re, const double *im, uint32_t n,
                uint32_t limbs, double scale, uint64_t *pt);
/* count vectors of n values (re[k*n..], im likewise or NULL) -> count plaintexts of
 * limbs*N words, on `threads` host threads (0 = all cores) */
int fhea_encode_batch(const fhea_client *cl, const double *re, const double *im, uint32_t count,
                      uint32_t n, uint32_t limbs, double scale, uint64_t *pt, uint32_t threads);
/* first n slots of pt (limbs*N words); im may be NULL */
int fhea_decode(const fhea_client *cl, const uint64_t *pt, uint32_t limbs, double scale,
                uint32_t n, double *re, double *im);
/* pk: 2*L*N words, b = -a s + e then a (all L Q limbs) */
int fhea_keygen_public(const fhea_client *cl, const int8_t *s, uint64_t seed, uint64_t *pk);
/* ct = (v b + e0 + pt, v a + e1) on the first `limbs` Q limbs: 2*limbs*N words */
int fhea_encrypt(const fhea_client *cl, const uint64_t *pk, const uint64_t *pt, uint32_t limbs,
                 uint64_t seed, uint64_t *ct);
/* pt = c0 + c1 s (limbs*N words) */
int fhea_decrypt(const fhea_client *cl, const int8_t *s, const uint64_t *ct, uint32_t limbs,
                 uint64_t *pt);
/* the secret in the evaluation domain on the first `limbs` Q primes (limbs*N words),
 * for fhea_decrypt_decode_batch (OpenFHE Decrypt keeps the same precomputed form) */
int fhea_secret_eval(const fhea_client *cl, const int8_t *s, uint32_t limbs, uint64_t *s_eval);
/* Decrypt + decode of `count` ciphertexts of `limbs` limbs (ct k = c0 then c1 at
 * cts + 2*k*limbs*N) with s_eval from fhea_secret_eval (>= limbs limbs):
 * re[k*n + j] (im likewise, or NULL) = slot j / scale, the values fhea_decrypt then
 * fhea_decode give, on `threads` host threads (0 = all cores)
 * (OpenFHE Decrypt + GetRealPackedValue, fhe_context_manager.cpp:457-512). */
int fhea_decrypt_decode_batch(const fhea_client *cl, const uint64_t *s_eval, const uint64_t *cts,
                              uint32_t count, uint32_t limbs, double scale, uint32_t n, double *re,
                              double *im, uint32_t threads);

#ifdef __cplusplus
}
#endif
#endif
