// fhea_kernels.hpp -- launchers for the gfx950 kernels (fhea_ntt.hip, fhea_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "fhea_internal.hpp"

namespace fhea {

// Device-resident tables every launcher needs.
struct DevTables {
    u32 N = 0, logN = 0, L = 0, K = 0;
    const ModParams *mp = nullptr;  // [T]
    const u64 *tw = nullptr;        // [T][N][2]  (psi^br(j), shoup)
    const u64 *itw = nullptr;       // [T][N][2]  (psi^-br(j), shoup)
    const u64 *ninv = nullptr;      // [T][2]
    const u32 *ident = nullptr;     // [T] identity prime map 0..T-1
    // FP64 path, used when every prime is below 2^50; f64_off (FHEA_F64_OFF
    // bitmask) turns single components back to the integer kernels:
    // 1 transforms, 2 basis conversion, 4 inner product, 8 tensor,
    // 16 fused transform ends, 32 rescale / ModDown combines, 64 the inner
    // product fused into the transforms around it; 128 is a switch ON: with the
    // fused inner product, the P slots also run the ModDown INTT's first pass there;
    // 256: slot-major tile walk for the first pass of rescale-lift prologues;
    // 512: INTT column pass, basis conversion and NTT column pass as three launches
    // instead of the fused conversion kernel (fhea_conv_cols.hip); 1024: the HMult
    // tensor in full instead of folded into the key switch (tensor_folded); 4096:
    // rotations through the automorphism kernel instead of sigma_g applied while loading;
    // 2048: the HMult's Q-slot sums stored and the ModDown + rescale row pass as its own
    // launch instead of fused into the key-switch rows kernel (KsRowsArgs::tt); 8192:
    // the P-slot (and HMult limb l-1) sums stored and the ModDown INTT's row pass as its
    // own launch instead of fused into the rows kernel (KsRowsArgs::lift)
    bool use_f64 = false;
    u32 f64_off = 0;
    // kernel variant of the fused conversion column kernel (FHEA_CC_VARIANT)
    u32 cc_var = 0;
    u32 max_wg = 1024;  // persistent-kernel grid: resident workgroups (CUs x 4)
    // tuning (fhea_ctx_set_tuning): stages of the first pass of a two-pass FP64
    // forward / inverse NTT (0 = logN/2), persistent workgroups per CU (0 = occupancy)
    u32 k1_fwd = 0, k1_inv = 0, wg_per_cu = 0;
    const double *twf = nullptr;    // [T][N][2]  (psi^br(j), psi^br(j)/q)
    const double *itwf = nullptr;   // [T][N][2]
    const double *ninvf = nullptr;  // [T][2]
    const double *pmodqf = nullptr; // [L][2] (P mod q_i, / q_i): the HMult tensor folded into the key switch
    // mixed chain (some primes >= 2^50, others below): the integer kernels run
    // everything but the transforms of limbs whose prime is below 2^50, which go
    // to the FP64 transform (its tables are built for such chains too);
    // pmask_small / pmask_big: bit p of [p / 64] set for prime p below / not below 2^50
    bool ntt_mixed = false;
    const u64 *pmask_small = nullptr, *pmask_big = nullptr;
    // the first q_big_prefix Q primes are >= 2^50; q_mixed_fusable: every later Q
    // prime is below 2^50, so the ModDown / rescale transforms of the limbs past
    // the prefix take the FP64 fused ends and only the prefix the integer kernels
    u32 q_big_prefix = 0;
    bool q_mixed_fusable = false;
};

struct RescaleOut;
struct CombineArgs;
struct KsIpArgs;
// Tensor product of two ciphertexts (fhe_host_driver.cpp:280-286):
// d0 = a0 b0, d1 = a0 b1 + a1 b0, d2 = a1 b1 (d0 null: d2 only).
struct TensorArgs {
    const u64 *a0, *a1, *b0, *b1;
    u64 *d0, *d1, *d2;
};
// A polynomial read through sigma_g: word j of the result is in[galois_src(j)]
// (automorphism.cpp:3-12 with the closed-form map).
struct GatherArgs {
    const u64 *in;
    u64 galois;
};
enum : int {
    FUSE_NONE = 0,
    FUSE_RESCALE_IN = 1,
    FUSE_RESCALE_OUT = 2,
    FUSE_MODDOWN_OUT = 3,
    FUSE_MDRS_IN = 4,   // ModDown + rescale in one transform: prologue
    FUSE_MDRS_OUT = 5,  // ... and epilogue
    FUSE_PROD_IN = 6,   // inverse first pass: input word = a1 * b1 (the HMult tensor's d2, from ta[poly])
    FUSE_GATHER_IN = 7, // inverse first pass: input word j = in[galois_src(j)] (sigma_g applied on the fly, ga[poly])
};
// Fused prologue (first forward pass) / epilogue (last forward pass) of the
// FP64 transform; poly = batch index of the NTT launch.
struct NttFuse {
    int pro = FUSE_NONE, epi = FUSE_NONE;
    const u64 *v = nullptr;            // RESCALE_IN, MDRS_IN: INTT'd last limb, poly at v + poly*v_stride
    u64 v_stride = 0;
    u64 ql = 0;                        // RESCALE_IN, MDRS_IN: the dropped prime
    const double *cst_in = nullptr;    // RESCALE_IN, MDRS_IN: (T_i, T_i/q_i) per limb slot
    const double *cst_out = nullptr;   // RESCALE_OUT, MDRS_OUT: q_l^-1 / MODDOWN_OUT: P^-1, per slot
    const double *cst_pq = nullptr;    // MDRS_IN, MDRS_OUT: (P q_l)^-1 per slot
    const RescaleOut *ro = nullptr;    // RESCALE_OUT, MDRS_OUT: per poly (a, out)
    const CombineArgs *cb = nullptr;   // MODDOWN_OUT: per item (poly>>1), half = poly&1
    const u64 *acc = nullptr;          // MODDOWN_OUT, MDRS_OUT: poly at acc + poly*acc_stride
    u64 acc_stride = 0;
    // optional list of the limb slots to transform (instead of all `limbs` slots)
    const u32 *slots = nullptr;
    u32 nslots = 0;
    // optional prime bitmask: only limbs whose prime index p has bit p set
    // (DevTables::pmask_small on a mixed chain)
    const u64 *pmask = nullptr;
    // first pass: walk the tiles unit-major (find_tile_um) instead of slot-major
    int unit_major = 0;
    // FUSE_PROD_IN: per poly the factors (a1, b1 of TensorArgs; slot s at + s*src_slot)
    const TensorArgs *ta = nullptr;
    // FUSE_GATHER_IN: per poly the source polynomial and the Galois element
    const GatherArgs *ga = nullptr;
    // two-pass transforms (N > 4096): 0 both passes; 1 only the first (its output,
    // FP64 intermediates, left at dst); 2 only the second (reading such intermediates
    // at src)
    int only_pass = 0;
};

// Key switching's inner product fused into the transforms around it (FP64 path,
// two-pass N): after the ModUp's first forward pass (only_pass = 1 on the digits),
// work item (tile, item z, slot i) runs the second forward pass of every converted
// digit limb of slot i in registers, adds the digit's own limb where it has one,
// multiplies by the key and sums over the digits (key_switch.cpp:309-381), storing
// acc0/acc1 as canonical words for every QlP slot.  With FHEA_F64_OFF bit 128 the P
// slots (i >= l) instead run the first inverse pass of the ModDown INTT
// (key_switch.cpp:385-400) on their sums and store its FP64 intermediates in pb, so
// only the second inverse pass (only_pass = 2) remains.
struct KsRowsArgs {
    const u64 *compact;  // ModUp digits after the first forward pass: item z at + z*cz,
    u64 cz;              // digit p, converted position pos at + (p*(l+K) + pos)*N
    const KsIpArgs *args;  // per item: c (own limbs, eval domain), key base
    u64 *acc;            // acc0 of slot i at acc + z*az + i*N, acc1 at + (l+K)*N
    u64 az;
    u64 *pb;             // bit 128, P slots: intermediates of poly 2z+s, limb j = i-l at pb + (2z+s)*K*N + j*N
    u32 l, K, L, alpha, nparts, key_limbs, dnum, nb;
    // HMult with the tensor folded in (ta != null): item z's own limbs are d2 = a1 b1
    // from ta[z], and the Q slots also receive P*d0 (acc0) and P*d1 (acc1), d0 = a0 b0,
    // d1 = a0 b1 + a1 b0: the ModDown (acc - NTT(conv)) P^-1 then yields ModDown(acc) + d
    const TensorArgs *ta = nullptr;
    const double *pmodq = nullptr;  // DevTables::pmodqf
    u32 gather = 0;  // the items' own limbs are read through sigma_g (KsIpArgs::galois)
    // the ModDown's last transform fused in (tt != null, Q slots only): the sums
    // acc_s are not stored; the kernel runs the forward row pass x of the ModDown
    // conversion's column-pass output of item z, polynomial s, slot i at
    // tt + (2z+s)*tt_pstride + i*N and stores, canonical,
    //   with ta (ModDown + rescale): acc_s (P q_l)^-1 + x to ro[2z+s].out + i*N,
    //     cst = ((P q_l)^-1 mod q_i, / q_i);
    //   with gather (ModDown): (acc_s - x) P^-1 [+ add_s] to cb[z].out_s + i*N,
    //     add0 read through sigma_g when cb[z].add0_galois, cst = (P^-1 mod q_i, / q_i)
    const double *tt = nullptr;
    u64 tt_pstride = 0;
    const RescaleOut *ro = nullptr;
    const CombineArgs *cb = nullptr;
    const double *cst = nullptr;
    // the ModDown's INTT first (row) pass fused in (launch_ks_rows_f64_early; the P
    // slots, and with ta slot l-1): the sums are not stored; the kernel runs the
    // INTT's row pass on each and stores it, P slot i of polynomial 2z+s to
    // pb + ((2z+s)*K + i-l)*N (as doubles), the rescaled limb l-1 to lift + (2z+s)*N;
    // early = 1 tells the ModDown that these row passes are done
    double *lift = nullptr;
    u32 early = 0;
};
// true when launch_ks_rows_f64 has kernels for the context's current splits
bool ks_rows_fusable(const DevTables &t);
void launch_ks_rows_f64(const DevTables &t, const KsRowsArgs &a, hipStream_t s);
// the same over Q slots slot0 .. slot0+nslots-1 only (mixed chains: the Q slots
// whose primes are below 2^50)
void launch_ks_rows_f64_q(const DevTables &t, const KsRowsArgs &a, u32 slot0, u32 nslots, hipStream_t s);
// slots slot0 .. slot0+nslots-1 (the P slots, and slot l-1 with ta) with the ModDown
// INTT's row pass fused in (KsRowsArgs::lift)
void launch_ks_rows_f64_early(const DevTables &t, const KsRowsArgs &a, u32 slot0, u32 nslots, hipStream_t s);

// Basis conversion fused between the column passes of the transforms around it
// (fhea_conv_cols.hip; FP64 path, two-pass N, forward and inverse column passes
// of equal length).  Unit (polynomial, digit d): the n_in input limbs at
// in + poly*in_pstride + (in_first[d] + i)*N hold the INTT's first-pass (rows)
// intermediates; output j of descs[d] (for j in j0 .. j0+nj-1) receives the
// forward transform's first-pass (columns) intermediate of the converted limb at
// out + poly*out_pstride + d*out_dstride + j*N, ready for a second pass
// (only_pass = 2, or the fused key-switch rows kernel).
// With lift_in set (ModDown followed by the rescale by q_{l-1}, ndig = 1): the
// rows-pass intermediates of acc_{l-1} and d_{l-1} of poly sit at
// lift_in + poly*lift_stride (+ N for d); the kernel forms the rescaled limb's
// ModDown result v = INTT(d) + P^-1 (INTT(acc) - conv_{lift_j}) (mod q_{l-1},
// lift_prime) in registers, and output j (Q slot j) transforms
// c*T_j - conv_j*(P q_l)^-1 with c the centred lift of v (FUSE_MDRS_IN's prologue).
static constexpr u32 CC_MAX_IN = 8;  // inputs per digit held in registers
static constexpr u32 CC_MAX_OUT = 64; // outputs per digit (constants staged in LDS)
struct ConvColsArgs {
    const double *in;
    u64 in_pstride;
    const u32 *in_first;
    const ConvDesc *descs;
    double *out;
    u64 out_pstride, out_dstride;
    u32 j0, nj;
    u32 npolys, ndig;
    const double *lift_in = nullptr;
    u64 lift_stride = 0;
    u32 lift_prime = 0, lift_j = 0;
    u32 lift_d = 1;  // 0: lift_in holds acc'_{l-1} only (d folded into acc'), v = P^-1 (INTT(acc') - conv)
    const double *lift_cst = nullptr;  // (N^-1, P^-1 N^-1, P^-1) mod q_{l-1}, each (value, value/q)
    const double *cst_in = nullptr;    // (T_j, T_j/q_j)
    const double *cst_pq = nullptr;    // ((P q_l)^-1, /q_j)
};
// true when launch_conv_cols_f64 has a kernel for the context's splits and
// conversions of up to max_in inputs
bool conv_cols_fusable(const DevTables &t, u32 max_in);
void launch_conv_cols_f64(const DevTables &t, const ConvColsArgs &a, hipStream_t s);

// Batched negacyclic NTT / INTT (ntt.cpp:4-80).  Polynomial z of the batch
// has limb slot s at src + z*src_stride + s*src_slot (src_slot 0 = N); slot s
// uses prime prime_map[s] (0xFFFFFFFF = skip).  Output order is the
// reference's bit-reversed evaluation order.  In place is allowed (src == dst).
void launch_ntt(const DevTables &t, bool inverse, const u64 *src, u64 *dst, u64 src_stride,
                u64 dst_stride, u32 npolys, u32 limbs, const u32 *prime_map, hipStream_t s,
                u64 src_slot = 0, u64 dst_slot = 0);
// The same on the 64-bit integer kernel only, whatever the chain (on a mixed chain:
// the limbs whose prime is >= 2^50; callers pass only such slots).
void launch_ntt_int(const DevTables &t, bool inverse, const u64 *src, u64 *dst, u64 src_stride,
                    u64 dst_stride, u32 npolys, u32 limbs, const u32 *prime_map, hipStream_t s,
                    const u64 *pmask = nullptr);

// FP64 transform (t.use_f64) with optional fused ends (forward only).
void launch_ntt_f64(const DevTables &t, bool inverse, const u64 *src, u64 *dst, u64 src_stride,
                    u64 dst_stride, u64 src_slot, u64 dst_slot, u32 npolys, u32 limbs,
                    const u32 *prime_map, const NttFuse &fz, hipStream_t s);

// Elementwise ops on `limbs` limbs, limb i mod prime i (poly_arith.cpp:3-42).
// op: 0 add, 1 sub, 2 mul.
void launch_binop(const DevTables &t, int op, const u64 *a, const u64 *b, u64 *o, u32 limbs,
                  hipStream_t s);
// Tensor products of nbatch ciphertext pairs (TensorArgs above).
void launch_tensor(const DevTables &t, const TensorArgs *args_dev, u32 nbatch, u32 limbs,
                   hipStream_t s);
// out[i] = in[map[i]] with map packed two u32 per word (automorphism.cpp:3-12).
void launch_automorph_map(const DevTables &t, const u64 *in, u64 *out, const u64 *packed_map,
                          u32 limbs, hipStream_t s);
// out = sigma_g(in) for c0 and c1 of nbatch ciphertexts (closed-form map);
// npolys = 1: c0 only (in1 / out1 unused).
struct AutoArgs {
    const u64 *in0, *in1;
    u64 *out0, *out1;
    u64 galois;
};
void launch_automorph(const DevTables &t, const AutoArgs *args_dev, u32 nbatch, u32 limbs,
                      hipStream_t s, u32 npolys = 2);

// Basis conversion (key_switch.cpp:7-60) for a batch: block (x, p, z) converts
// descriptor descs[p] reading in + z*in_zstride + (in_first_p)*N and writing
// out + z*out_zstride + p*out_pstride + j*N.
// max_nin / max_nout: the largest n_in / n_out among the descriptors.
void launch_fbc(const DevTables &t, const ConvDesc *descs_dev, const u32 *in_first_dev,
                u32 ndesc, u32 max_nin, u32 max_nout, const u64 *in, u64 in_zstride, u64 *out,
                u64 out_pstride, u64 out_zstride, u32 nbatch, hipStream_t s);

// Key-switch inner product (key_switch.cpp:352-381) for nbatch inputs at level l.
struct KsIpArgs {
    const u64 *c;      // input polynomial (eval), l limbs
    const u64 *key;    // evk base (a then b)
    u64 galois = 0;    // fused rows kernel: c read through sigma_g (0: as is)
};
// shared_key: every item uses args_dev[0].key (read once per batch).
// nslots > 0: only QlP slots slot0 .. slot0+nslots-1, on the integer kernel.
void launch_ks_inner(const DevTables &t, const KsIpArgs *args_dev, const u64 *compact,
                     u64 compact_zstride, u64 *acc, u64 acc_zstride, u32 nbatch, u32 l, u32 alpha,
                     u32 nparts, u32 key_limbs, u32 dnum, bool shared_key, hipStream_t s, u32 slot0 = 0,
                     u32 nslots = 0);

// ModDown combine (key_switch.cpp:253-273): out_i = (x_i - t_i) * P^-1 [+ add_i].
struct CombineArgs {
    u64 *out0, *out1;
    const u64 *add0, *add1;  // may be null
    u64 add0_galois = 0;     // FP64 fused epilogue: add0 read through sigma_g (0: as is)
};
void launch_moddown_combine(const DevTables &t, const CombineArgs *args_dev, const u64 *acc,
                            u64 acc_zstride, u64 acc_pstride, const u64 *tt, u64 tt_zstride,
                            u64 tt_pstride, const u64 *pinv, const double *pinvf, u32 nbatch,
                            u32 l, hipStream_t s);

// Rescale (rescale.cpp:3-72) helpers.  v: INTT of the last limb, [z][2][N].
// Level l, output limbs 0 .. nlimbs-1 (0 = all l-1).
void launch_rescale_prep(const DevTables &t, const u64 *v, u64 v_stride, u64 *tmp,
                         u64 tmp_stride, const u64 *qlql, const u64 *sub, const double *qlqlf,
                         u32 npolys, u32 l, hipStream_t s, u32 nlimbs = 0);
struct RescaleOut {
    const u64 *a;  // input polynomial (l limbs)
    u64 *out;      // output polynomial (l-1 limbs)
};
void launch_rescale_combine(const DevTables &t, const RescaleOut *args_dev, const u64 *tmp,
                            u64 tmp_stride, const u64 *qlinv, const double *qlinvf, u32 npolys,
                            u32 l, hipStream_t s, u32 nlimbs = 0);

// Hoisted rotations: the digits of one input (compact, eval domain, ModUp done once)
// read through item z's Galois permutation, times item z's key.
struct KsRotArgs {
    const u64 *c;        // the input polynomial c1 (eval), l limbs: the digits' own limbs
    const u64 *key;      // item's evk base (a then b)
    u64 galois;
    const u64 *compact;  // the converted digit limbs of that input's ModUp
};
void launch_ks_inner_rot(const DevTables &t, const KsRotArgs *args_dev, u64 *acc, u64 acc_zstride, u32 nbatch,
                         u32 l, u32 alpha, u32 nparts, u32 key_limbs, u32 dnum, hipStream_t s);

// Batched ciphertext addition: d = a + b (both polynomials), items of `limbs` limbs.
struct AddArgs {
    const u64 *a0, *a1, *b0, *b1;
    u64 *d0, *d1;
};
void launch_add_batch(const DevTables &t, const AddArgs *args_dev, u32 nbatch, u32 limbs, hipStream_t s);

// Plaintext-matrix x encrypted-vector inner products (diagonal method) for groups
// of queries sharing their plaintext blocks:
//   out_{q,g} = bias_q + ptb_g + sum_{r<R} ct_{q,r} (.) pt_{g,r}
// A group's offs (device, u64 word offsets into P), nq queries and ng blocks:
//   [ct0: nq*R][ct1: nq*R][bias0: nq][bias1: nq][pt: ng*R][ptb: ng][out0: nq*ng][out1: nq*ng]
// (outputs row-major by query); ptb and bias entries may be LIN_NONE.  A unit is
// the queries q0..q0+uq-1 and blocks g0..g0+ug-1 of one group that one grid slice
// evaluates: FP64 (LIN_SHAPE_LDS) uq <= 2*qa, ug <= 8*(8/qa), qa in {2, 4, 8}
// (2 x 8 tiles, 8 per unit, qa of them stacked); FP64 small groups
// (LIN_SHAPE_TILES) uq <= 2, ug <= 8, one wave each; integer (LIN_SHAPE_1x8)
// uq = 1, ug <= 8.
static constexpr u64 LIN_NONE = ~0ull;
static constexpr u32 LIN_NS = 40;  // LDS streams: 2*uq + ug <= 40 for every shape
struct LinUnit {
    const u64 *offs;
    u32 nq, ng, q0, g0, uq, ug, qa, pad;
};
enum : int { LIN_SHAPE_1x8 = 0, LIN_SHAPE_LDS = 1, LIN_SHAPE_TILES = 2 };
// true when the FP64 kernels run (the LDS shape exists only there)
bool linear_sum_f64(const DevTables &t);
void launch_linear_sum(const DevTables &t, u64 *P, const LinUnit *units_dev, u32 n_units, u32 R, u32 limbs,
                       int shape, hipStream_t s);

}  // namespace fhea
