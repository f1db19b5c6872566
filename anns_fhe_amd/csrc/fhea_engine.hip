// fhea_engine.hip -- the C-ABI (include/fhea.h): context, device tables, arenas,
// the op-code dispatcher and the batched composite pipelines.
//
// Replaces fhe_accel_top (hardware_fpga_model/hls/src/fhe_accel_top.cpp:90-328),
// SimMemoryBus (cpp_host_model/src/sim_memory_bus.cpp) and the composite
// sequences of FHEHostDriver (cpp_host_model/src/fhe_host_driver.cpp:248-404).
// Ciphertexts and keys stay resident in HBM; a composite is a fixed chain of
// batched kernels over a device workspace (no burst_load/burst_store copies,
// fhe_accel_top.cpp:36-86).
#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fhea_engine.hpp"

#include <mutex>

using namespace fhea;

// Scratch of fhea_eval_mult_relin_rescale_noalloc: one static array per GPU the
// library runs on (the loader places it when the code object is loaded, and it
// stays until the process ends).  Host code never passes `scratch` itself to a
// kernel: in host code the name is a stand-in, not the device address.  The
// address is looked up once per device with hipGetSymbolAddress (with that
// device current) at context creation and kept in g_scratch_dev.
static __device__ __align__(256) u64 scratch[FHEA_NOALLOC_SCRATCH_WORDS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
u64 *g_scratch_dev[kMaxDevices] = {};
}  // namespace

// Looks up (once per device) the device address of `scratch` on the current
// device `dev`; FHEA_ERR_DEVICE if the runtime cannot resolve it.
static int scratch_init(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return FHEA_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[dev]) return FHEA_OK;
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return FHEA_ERR_DEVICE;
    g_scratch_dev[dev] = (u64 *)ptr;
    return FHEA_OK;
}

static u64 *scratch_ptr(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_dev[dev];
}

namespace fhea {

int check_hip(hipError_t e) { return e == hipSuccess ? FHEA_OK : FHEA_ERR_DEVICE; }

#define HIPCHK(x)                                   \
    do {                                            \
        if ((x) != hipSuccess) return FHEA_ERR_DEVICE; \
    } while (0)

// Drops a device buffer that is being replaced by a larger one.  A recorded graph
// may address it (its kernel nodes hold raw pointers), so while any graph exists the
// buffer is parked in c->retired; otherwise it is freed once the stream is idle.
int retire_buffer(fhea_ctx *c, void *p) {
    if (!p) return FHEA_OK;
    if (!c->graphs.empty()) {
        c->retired.push_back(p);
        return FHEA_OK;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    for (u32 k = 1; k < fhea_ctx::MAX_LANES; ++k)
        if (c->lane_stream[k]) HIPCHK(hipStreamSynchronize(c->lane_stream[k]));
    HIPCHK(hipFree(p));
    return FHEA_OK;
}

int ensure_ws(fhea_ctx *c, u64 words) {
    if (words <= c->ws_words) return FHEA_OK;
    if (c->capturing) return FHEA_ERR_NO_MEM;  // captured kernels may already point at it
    if (c->ws) {
        const int st = retire_buffer(c, c->ws);
        if (st) return st;
        c->ws = nullptr;
        c->ws_words = 0;
    }
    if (hipMalloc((void **)&c->ws, words * 8) != hipSuccess) {
        c->ws = nullptr;
        return FHEA_ERR_NO_MEM;
    }
    c->ws_words = words;
    return FHEA_OK;
}

// Reserves `bytes` of argument staging: the caller writes them at *h (and may
// embed *d-based device addresses), then stage_commit copies them to *d on the
// stream.  While capturing, the graph owns its argument area (pinned host +
// device, allocated before the capture began and never reused) and the copy is
// recorded as a graph node.
static int stage_reserve(fhea_ctx *c, size_t bytes, char **h, char **d) {
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (c->capturing) {
        if (c->cap_used + need > fhea_ctx::CAP_BYTES) return FHEA_ERR_NO_MEM;
        *h = (char *)c->cap_host + c->cap_used;
        *d = (char *)c->cap_dev + c->cap_used;
        c->cap_used += need;
        return FHEA_OK;
    }
    if (c->args_used + need > c->args_bytes) {
        // every earlier copy (on either lane) must have landed before the staging
        // area is reused
        if (hipStreamSynchronize(c->stream) != hipSuccess) return FHEA_ERR_DEVICE;
        for (u32 k = 1; k < fhea_ctx::MAX_LANES; ++k)
            if (c->lane_stream[k] && hipStreamSynchronize(c->lane_stream[k]) != hipSuccess) return FHEA_ERR_DEVICE;
        c->args_used = 0;
        if (need > c->args_bytes) return FHEA_ERR_NO_MEM;
    }
    *h = (char *)c->args_host + c->args_used;
    *d = (char *)c->args + c->args_used;
    c->args_used += need;
    return FHEA_OK;
}

static int stage_commit(fhea_ctx *c, const char *h, char *d, size_t bytes) {
    return hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream) == hipSuccess ? FHEA_OK
                                                                                       : FHEA_ERR_DEVICE;
}

void *stage_args(fhea_ctx *c, const void *host, size_t bytes) {
    char *h = nullptr, *d = nullptr;
    if (stage_reserve(c, bytes, &h, &d)) return nullptr;
    memcpy(h, host, bytes);
    if (stage_commit(c, h, d, bytes)) return nullptr;
    return d;
}

// ---------------------------------------------------------------------------
// Device table blob: all tables in one allocation; ConvDesc pointer fields are
// stored as blob offsets and relocated once the base address is known.
// ---------------------------------------------------------------------------
struct Blob {
    std::vector<uint8_t> d;
    std::vector<size_t> fix;
    size_t add(const void *p, size_t n) {
        size_t off = (d.size() + 255) & ~(size_t)255;
        d.resize(off + std::max<size_t>(n, 1));
        if (n) memcpy(d.data() + off, p, n);
        return off;
    }
    template <class T>
    size_t addv(const std::vector<T> &v) {
        return add(v.data(), v.size() * sizeof(T));
    }
    // ConvDesc whose pointer fields hold blob offsets
    size_t add_descs(const std::vector<ConvDesc> &v) {
        size_t off = addv(v);
        for (size_t i = 0; i < v.size(); ++i) {
            size_t b = off + i * sizeof(ConvDesc);
            fix.push_back(b + offsetof(ConvDesc, in_prime));
            fix.push_back(b + offsetof(ConvDesc, in_cst));
            fix.push_back(b + offsetof(ConvDesc, out_prime));
            fix.push_back(b + offsetof(ConvDesc, out_pos));
            fix.push_back(b + offsetof(ConvDesc, hat));
            if (v[i].in_cstf) fix.push_back(b + offsetof(ConvDesc, in_cstf));
            if (v[i].hatf) fix.push_back(b + offsetof(ConvDesc, hatf));
            if (v[i].in_ncstf) fix.push_back(b + offsetof(ConvDesc, in_ncstf));
        }
        return off;
    }
};

template <class T>
static T *off_ptr(size_t off) {
    return reinterpret_cast<T *>((uintptr_t)off);
}

static u64 mulm(u64 a, u64 b, u64 q) { return (u64)((u128)a * b % q); }

struct LevelOffsets {
    size_t modup_desc, modup_first, modup_ntt_prime, modup_slots, moddown_desc, pinv, rs_qlinv, rs_qlql, rs_sub,
        rs_last, pinvf, rs_qlinvf, rs_qlqlf, rs_pqf, rs_liftf;
};

// (value, value/modulus) pairs for the FP64 kernels, from (value, shoup) pairs
static std::vector<double> f64_pairs(const std::vector<u64> &vs, const std::vector<u64> &mods) {
    std::vector<double> out(2 * mods.size());
    for (size_t i = 0; i < mods.size(); ++i) {
        out[2 * i] = (double)vs[2 * i];
        out[2 * i + 1] = (double)vs[2 * i] / (double)mods[i];
    }
    return out;
}

static int build_device_tables(fhea_ctx *c) {
    const HostTables &h = c->h;
    const u32 N = h.N, L = h.L, K = h.K, T = L + K, A = h.alpha;
    Blob b;
    std::vector<ModParams> mp(T);
    for (u32 i = 0; i < T; ++i) mp[i] = make_mod_params(h.primes[i]);
    size_t o_mp = b.addv(mp);
    std::vector<u64> tw((size_t)T * N * 2), itw((size_t)T * N * 2), ninv(2 * T);
    for (u32 i = 0; i < T; ++i) {
        u64 q = h.primes[i];
        for (u32 j = 0; j < N; ++j) {
            u64 w = h.tw[(size_t)i * N + j], wi = h.itw[(size_t)i * N + j];
            tw[((size_t)i * N + j) * 2] = w;
            tw[((size_t)i * N + j) * 2 + 1] = shoup_companion(w, q);
            itw[((size_t)i * N + j) * 2] = wi;
            itw[((size_t)i * N + j) * 2 + 1] = shoup_companion(wi, q);
        }
        ninv[2 * i] = h.ninv[i];
        ninv[2 * i + 1] = shoup_companion(h.ninv[i], q);
    }
    size_t o_tw = b.addv(tw), o_itw = b.addv(itw), o_ninv = b.addv(ninv);
    tw.clear();
    tw.shrink_to_fit();
    itw.clear();
    itw.shrink_to_fit();
    // FP64 transform tables (w, w/q) when every prime is below 2^50
    // The FP64 bounds need every q < 2^50 and max(q)/min(q) < 4 (a residue
    // reduced mod one prime must stay below 8x any other prime).
    bool use_f64 = true;
    u64 qmin = ~0ull, qmax = 0;
    for (u32 i = 0; i < T; ++i) {
        use_f64 &= h.primes[i] < (1ull << 50);
        qmin = std::min(qmin, h.primes[i]);
        qmax = std::max(qmax, h.primes[i]);
    }
    use_f64 &= qmax / qmin < 4;
    // a chain with primes on both sides of 2^50 runs the integer kernels, except
    // for the transforms of its small-prime limbs (FP64 tables for those)
    std::vector<u64> mask_small(4, 0), mask_big(4, 0);
    for (u32 i = 0; i < T; ++i) (h.primes[i] < (1ull << 50) ? mask_small : mask_big)[i >> 6] |= 1ull << (i & 63);
    const bool mixed = !use_f64 && mask_small != std::vector<u64>(4, 0);
    const size_t o_mask_small = b.addv(mask_small), o_mask_big = b.addv(mask_big);
    size_t o_twf = 0, o_itwf = 0, o_ninvf = 0, o_pmodqf = 0;
    if (use_f64 || mixed) {
        std::vector<double> twf((size_t)T * N * 2), itwf((size_t)T * N * 2), ninvf(2 * T);
        for (u32 i = 0; i < T; ++i) {
            const double qd = (double)h.primes[i];
            for (u32 j = 0; j < N; ++j) {
                const double w = (double)h.tw[(size_t)i * N + j], wi = (double)h.itw[(size_t)i * N + j];
                twf[((size_t)i * N + j) * 2] = w;
                twf[((size_t)i * N + j) * 2 + 1] = w / qd;
                itwf[((size_t)i * N + j) * 2] = wi;
                itwf[((size_t)i * N + j) * 2 + 1] = wi / qd;
            }
            ninvf[2 * i] = (double)h.ninv[i];
            ninvf[2 * i + 1] = (double)h.ninv[i] / qd;
        }
        o_twf = b.addv(twf);
        o_itwf = b.addv(itwf);
        o_ninvf = b.addv(ninvf);
        std::vector<u64> pmq(2 * L), qv(L);
        for (u32 i = 0; i < L; ++i) {
            u64 pr = 1 % h.primes[i];
            for (u32 k = 0; k < K; ++k) pr = mulm(pr, h.primes[L + k] % h.primes[i], h.primes[i]);
            pmq[2 * i] = pr;
            qv[i] = h.primes[i];
        }
        o_pmodqf = b.addv(f64_pairs(pmq, qv));
    }
    std::vector<u32> ident(T);
    for (u32 i = 0; i < T; ++i) ident[i] = i;
    size_t o_ident = b.addv(ident);
    u32 zero = 0;
    size_t o_zero = b.add(&zero, 4);

    std::vector<LevelOffsets> lo(L + 1);
    c->lv.assign(L + 1, LevelTables());
    for (u32 l = 1; l <= L; ++l) {
        LevelTables &lt = c->lv[l];
        lt.l = l;
        lt.nparts = (l + A - 1) / A;
        std::vector<ConvDesc> descs(lt.nparts);
        std::vector<u32> first(lt.nparts), nttp((size_t)lt.nparts * (l + K), 0xFFFFFFFFu);
        for (u32 p = 0; p < lt.nparts; ++p) {
            u32 start = p * A, sz = std::min(A, l - start), n_out = l - sz + K;
            lt.part_start.push_back(start);
            lt.part_size.push_back(sz);
            first[p] = start;
            std::vector<u32> inp(sz), outp;
            std::vector<u64> cst(2 * sz), hat((size_t)sz * n_out);
            for (u32 i = 0; i < l; ++i)
                if (i < start || i >= start + sz) outp.push_back(i);
            for (u32 j = 0; j < K; ++j) outp.push_back(L + j);
            for (u32 i = 0; i < sz; ++i) {
                u32 qi_idx = start + i;
                u64 qi = h.primes[qi_idx];
                inp[i] = qi_idx;
                u64 prod = 1 % qi;
                for (u32 k = 0; k < sz; ++k)
                    if (k != i) prod = mulm(prod, h.primes[start + k] % qi, qi);
                u64 inv = inv_mod(prod, qi);
                cst[2 * i] = inv;
                cst[2 * i + 1] = shoup_companion(inv, qi);
                for (u32 j = 0; j < n_out; ++j) {
                    u64 pj = h.primes[outp[j]], pr = 1 % pj;
                    for (u32 k = 0; k < sz; ++k)
                        if (k != i) pr = mulm(pr, h.primes[start + k] % pj, pj);
                    hat[(size_t)i * n_out + j] = pr;
                }
            }
            for (u32 j = 0; j < n_out; ++j) nttp[(size_t)p * (l + K) + j] = outp[j];
            ConvDesc &D = descs[p];
            D.n_in = sz;
            D.n_out = n_out;
            D.in_prime = off_ptr<const u32>(b.addv(inp));
            D.in_cst = off_ptr<const u64>(b.addv(cst));
            D.out_prime = off_ptr<const u32>(b.addv(outp));
            D.out_pos = off_ptr<const u32>(b.addv(outp));
            D.hat = off_ptr<const u64>(b.addv(hat));
            if (use_f64) {
                std::vector<u64> inm(sz);
                for (u32 i = 0; i < sz; ++i) inm[i] = h.primes[inp[i]];
                std::vector<double> hf(2 * (size_t)sz * n_out);
                for (u32 i = 0; i < sz; ++i)
                    for (u32 j = 0; j < n_out; ++j) {
                        hf[2 * ((size_t)i * n_out + j)] = (double)hat[(size_t)i * n_out + j];
                        hf[2 * ((size_t)i * n_out + j) + 1] =
                            (double)hat[(size_t)i * n_out + j] / (double)h.primes[outp[j]];
                    }
                D.in_cstf = off_ptr<const double>(b.addv(f64_pairs(cst, inm)));
                D.hatf = off_ptr<const double>(b.addv(hf));
                // N^-1 qhat_i^-1 mod q_i: the INTT's last scaling and the conversion's
                // first product in one (fused column passes)
                std::vector<u64> ncst(2 * sz);
                for (u32 i = 0; i < sz; ++i) ncst[2 * i] = mulm(h.ninv[inp[i]], cst[2 * i], inm[i]);
                D.in_ncstf = off_ptr<const double>(b.addv(f64_pairs(ncst, inm)));
            }
        }
        lo[l].modup_desc = b.add_descs(descs);
        lo[l].modup_first = b.addv(first);
        lo[l].modup_ntt_prime = b.addv(nttp);
        {
            std::vector<u32> sl;
            for (u32 k = 0; k < nttp.size(); ++k)
                if (nttp[k] != 0xFFFFFFFFu) sl.push_back(k);
            lt.n_modup_slots = (u32)sl.size();
            lo[l].modup_slots = b.addv(sl);
        }
        // ModDown: P -> Q_l  (PHatInvModp, PHatModq row-major by P index)
        {
            std::vector<u32> inp(K), outp(l);
            std::vector<u64> cst(2 * K), hat((size_t)K * l), pinv(2 * l);
            for (u32 j = 0; j < K; ++j) {
                u64 pj = h.primes[L + j], pr = 1 % pj;
                inp[j] = L + j;
                for (u32 k = 0; k < K; ++k)
                    if (k != j) pr = mulm(pr, h.primes[L + k] % pj, pj);
                u64 inv = inv_mod(pr, pj);
                cst[2 * j] = inv;
                cst[2 * j + 1] = shoup_companion(inv, pj);
                for (u32 i = 0; i < l; ++i) {
                    u64 qi = h.primes[i], hh = 1 % qi;
                    for (u32 k = 0; k < K; ++k)
                        if (k != j) hh = mulm(hh, h.primes[L + k] % qi, qi);
                    hat[(size_t)j * l + i] = hh;
                }
            }
            for (u32 i = 0; i < l; ++i) {
                outp[i] = i;
                u64 qi = h.primes[i], pr = 1 % qi;
                for (u32 k = 0; k < K; ++k) pr = mulm(pr, h.primes[L + k] % qi, qi);
                u64 inv = inv_mod(pr, qi);
                pinv[2 * i] = inv;
                pinv[2 * i + 1] = shoup_companion(inv, qi);
            }
            std::vector<ConvDesc> md(1);
            md[0].n_in = K;
            md[0].n_out = l;
            md[0].in_prime = off_ptr<const u32>(b.addv(inp));
            md[0].in_cst = off_ptr<const u64>(b.addv(cst));
            md[0].out_prime = off_ptr<const u32>(b.addv(outp));
            md[0].out_pos = off_ptr<const u32>(b.addv(outp));
            md[0].hat = off_ptr<const u64>(b.addv(hat));
            std::vector<u64> qm(l), pm(K);
            for (u32 i = 0; i < l; ++i) qm[i] = h.primes[i];
            for (u32 j = 0; j < K; ++j) pm[j] = h.primes[L + j];
            if (use_f64) {
                std::vector<double> hf(2 * (size_t)K * l);
                for (u32 j = 0; j < K; ++j)
                    for (u32 i = 0; i < l; ++i) {
                        hf[2 * ((size_t)j * l + i)] = (double)hat[(size_t)j * l + i];
                        hf[2 * ((size_t)j * l + i) + 1] = (double)hat[(size_t)j * l + i] / (double)qm[i];
                    }
                md[0].in_cstf = off_ptr<const double>(b.addv(f64_pairs(cst, pm)));
                md[0].hatf = off_ptr<const double>(b.addv(hf));
                std::vector<u64> ncst(2 * K);
                for (u32 j = 0; j < K; ++j) ncst[2 * j] = mulm(h.ninv[L + j], cst[2 * j], pm[j]);
                md[0].in_ncstf = off_ptr<const double>(b.addv(f64_pairs(ncst, pm)));
            }
            // (value, value/q) forms for the FP64 fused transform ends (on a mixed
            // chain used for the limbs whose prime is below 2^50 only)
            if (use_f64 || mixed) lo[l].pinvf = b.addv(f64_pairs(pinv, qm));
            lo[l].moddown_desc = b.add_descs(md);
            lo[l].pinv = b.addv(pinv);
        }
        // Rescale l -> l-1 (rescale.cpp:3-72; tables test_vector_generator.cpp:304-325)
        {
            u32 n = l >= 2 ? l - 1 : 0;
            std::vector<u64> qlinv(2 * n + 2), qlql(2 * n + 2), sub(n + 1), pq(2 * n + 2);
            u64 ql = h.primes[l - 1];
            for (u32 i = 0; i < n; ++i) {
                u64 qi = h.primes[i];
                u64 inv = inv_mod(ql % qi, qi);
                u64 t = inv ? qi - inv : 0;
                qlinv[2 * i] = inv;
                qlinv[2 * i + 1] = shoup_companion(inv, qi);
                qlql[2 * i] = t;
                qlql[2 * i + 1] = shoup_companion(t, qi);
                sub[i] = mulm(ql % qi, t, qi);
                // (P q_l)^-1 mod q_i for the ModDown-and-rescale transform
                u64 pr = 1 % qi;
                for (u32 k = 0; k < K; ++k) pr = mulm(pr, h.primes[L + k] % qi, qi);
                pq[2 * i] = mulm(inv_mod(pr, qi), inv, qi);
            }
            std::vector<u32> last = {l - 1, l - 1};
            if (use_f64 || mixed) {
                std::vector<u64> qm(n + 1, 1);
                for (u32 i = 0; i < n; ++i) qm[i] = h.primes[i];
                lo[l].rs_qlinvf = b.addv(f64_pairs(qlinv, qm));
                lo[l].rs_qlqlf = b.addv(f64_pairs(qlql, qm));
                lo[l].rs_pqf = b.addv(f64_pairs(pq, qm));
                // the rescaled limb's ModDown result from coefficient-domain parts
                u64 pl = 1 % ql;
                for (u32 k = 0; k < K; ++k) pl = mulm(pl, h.primes[L + k] % ql, ql);
                const u64 pinv_l = inv_mod(pl, ql), ninv_l = h.ninv[l - 1];
                std::vector<u64> lc = {ninv_l, 0, mulm(pinv_l, ninv_l, ql), 0, pinv_l, 0};
                lo[l].rs_liftf = b.addv(f64_pairs(lc, std::vector<u64>(3, ql)));
            }
            lo[l].rs_qlinv = b.addv(qlinv);
            lo[l].rs_qlql = b.addv(qlql);
            lo[l].rs_sub = b.addv(sub);
            lo[l].rs_last = b.addv(last);
        }
    }

    void *base = nullptr;
    if (hipMalloc(&base, b.d.size()) != hipSuccess) return FHEA_ERR_NO_MEM;
    for (size_t f : b.fix) *(uintptr_t *)(b.d.data() + f) += (uintptr_t)base;
    if (hipMemcpy(base, b.d.data(), b.d.size(), hipMemcpyHostToDevice) != hipSuccess) {
        hipFree(base);
        return FHEA_ERR_DEVICE;
    }
    c->dblob = base;
    auto P = [&](size_t off) { return (const void *)((char *)base + off); };
    c->dt.N = N;
    c->dt.logN = h.logN;
    c->dt.L = L;
    c->dt.K = K;
    c->dt.mp = (const ModParams *)P(o_mp);
    c->dt.tw = (const u64 *)P(o_tw);
    c->dt.itw = (const u64 *)P(o_itw);
    c->dt.ninv = (const u64 *)P(o_ninv);
    c->dt.ident = (const u32 *)P(o_ident);
    c->dt.use_f64 = use_f64 && !c->force_int;
    {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && cus > 0)
            c->dt.max_wg = 4u * (u32)cus;
    }
    c->dt.f64_off = c->f64_off;
    c->dt.cc_var = c->cc_var;
    c->dt.ntt_mixed = mixed && !c->force_int;
    {
        u32 qb = 0;
        while (qb < L && h.primes[qb] >= (1ull << 50)) ++qb;
        bool rest_small = true;
        for (u32 i = qb; i < L; ++i) rest_small &= h.primes[i] < (1ull << 50);
        c->dt.q_big_prefix = qb;
        c->dt.q_mixed_fusable = c->dt.ntt_mixed && rest_small;
    }
    c->dt.pmask_small = (const u64 *)P(o_mask_small);
    c->dt.pmask_big = (const u64 *)P(o_mask_big);
    if (use_f64 || mixed) {
        c->dt.twf = (const double *)P(o_twf);
        c->dt.itwf = (const double *)P(o_itwf);
        c->dt.ninvf = (const double *)P(o_ninvf);
        c->dt.pmodqf = (const double *)P(o_pmodqf);
    }
    c->zero_u32 = (const u32 *)P(o_zero);
    c->modup_desc_dev.assign(L + 1, nullptr);
    c->modup_first_dev.assign(L + 1, nullptr);
    c->moddown_desc_dev.assign(L + 1, nullptr);
    c->rs_last_map.assign(L + 1, nullptr);
    c->rs_sub.assign(L + 1, nullptr);
    for (u32 l = 1; l <= L; ++l) {
        LevelTables &lt = c->lv[l];
        c->modup_desc_dev[l] = (const ConvDesc *)P(lo[l].modup_desc);
        c->modup_first_dev[l] = (const u32 *)P(lo[l].modup_first);
        lt.modup_ntt_prime = (const u32 *)P(lo[l].modup_ntt_prime);
        lt.modup_slots = (const u32 *)P(lo[l].modup_slots);
        c->moddown_desc_dev[l] = (const ConvDesc *)P(lo[l].moddown_desc);
        lt.pinv = (const u64 *)P(lo[l].pinv);
        lt.rs_qlinv = (const u64 *)P(lo[l].rs_qlinv);
        lt.rs_qlql = (const u64 *)P(lo[l].rs_qlql);
        c->rs_sub[l] = (const u64 *)P(lo[l].rs_sub);
        c->rs_last_map[l] = (const u32 *)P(lo[l].rs_last);
        if (use_f64 || mixed) {
            lt.pinvf = (const double *)P(lo[l].pinvf);
            lt.rs_qlinvf = (const double *)P(lo[l].rs_qlinvf);
            lt.rs_qlqlf = (const double *)P(lo[l].rs_qlqlf);
            lt.rs_pqf = (const double *)P(lo[l].rs_pqf);
            lt.rs_liftf = (const double *)P(lo[l].rs_liftf);
        }
    }
    return FHEA_OK;
}

// ---------------------------------------------------------------------------
// Pipelines.  Word counts per batch item are computed first so the workspace
// is sized once per call; regions are then carved from it in order.
// ---------------------------------------------------------------------------
static u64 ks_words(const fhea_ctx *c, u32 l) {
    const u64 N = c->h.N, K = c->h.K, np = c->lv[l].nparts;
    return N * (l + np * (l + K) + 2 * (l + K) + 2 * K + 2 * l);
}
static u64 rescale_words(const fhea_ctx *c, u32 l) { return (u64)c->h.N * (2 + 2 * (l - 1)); }

// Per-item argument arrays of one key-switch chunk already on the device (staged
// by the caller together with its other chunks', one copy for all): the pipeline
// then stages nothing itself.  Null members are staged as usual.
struct PreArgs {
    const TensorArgs *fold = nullptr;
    const GatherArgs *gather = nullptr;
    const KsIpArgs *ip = nullptr;
    const CombineArgs *cb = nullptr;
    const RescaleOut *ro = nullptr;
};

static int ks_moddown(fhea_ctx *c, u32 l, u32 nb, u64 *acc, u64 *pb, u64 *tt,
                      const std::vector<CombineArgs> &comb, bool pb_mid = false,
                      const KsRowsArgs *late = nullptr, const PreArgs *pre = nullptr);

// The rescale that follows a key switch in eval_mult_relin_rescale: the switch's
// add operand d_z = (d_base + z*dz, + ds) is also its output (l limbs, limb l-1
// of both polynomials is overwritten), outs[2z+s] receive the rescaled l-1 limbs.
struct RescaleAfter {
    u64 *d_base;
    u64 dz, ds;
    const std::vector<u64 *> *outs;
};
static int ks_moddown_rescale(fhea_ctx *c, u32 l, u32 nb, u64 *acc, u64 *pb, u64 *tt, u64 *v,
                              const std::vector<CombineArgs> &comb, const RescaleAfter &ra, bool pb_mid,
                              bool folded = false, const KsRowsArgs *late = nullptr,
                              const PreArgs *pre = nullptr);

// Mixed chain whose Q primes past a leading run of >= 2^50 primes are all below
// 2^50: the ModDown / rescale transforms take the FP64 fused ends on those limbs.
static bool mixed_fused(const DevTables &t) { return t.q_mixed_fusable && !(t.f64_off & 17); }

// The ModDown's INTT of the P part of nb accumulator pairs into pb: the whole
// transform from acc, or (pb_mid) only its second pass over the intermediates the
// fused inner product left in pb.
static void moddown_intt(fhea_ctx *c, u32 l, u32 nb, const u64 *acc, u64 *pb, bool pb_mid) {
    const DevTables &t = c->dt;
    const u64 N = t.N, K = t.K;
    if (pb_mid) {
        NttFuse fz;
        fz.only_pass = 2;
        launch_ntt_f64(t, true, pb, pb, K * N, K * N, N, N, 2 * nb, K, t.ident + t.L, fz, c->stream);
        return;
    }
    launch_ntt(t, true, acc + l * N, pb, (l + K) * N, K * N, 2 * nb, K, t.ident + t.L, c->stream);
}

// The fused conversion column kernel applies to the ModUp of level l (digits of
// at most CC_MAX_IN limbs) and to the ModDown (K limbs in); FHEA_F64_OFF bit 512
// keeps the INTT column pass, the conversion and the NTT column pass as three launches.
static bool modup_cc(const fhea_ctx *c, u32 l) { return conv_cols_fusable(c->dt, std::min(l, c->h.alpha)); }
static bool moddown_cc(const fhea_ctx *c) { return conv_cols_fusable(c->dt, c->dt.K); }

// ModUp of npolys inputs through the fused column kernel: coef holds the INTT's
// first-pass intermediates of input z at coef + z*l*N (written by the caller), and
// compact receives every digit's converted limbs after the forward transform's first
// (column) pass, item z at compact + z*cz.
static void modup_cols(fhea_ctx *c, u32 l, u32 npolys, const u64 *coef, u64 *compact) {
    const DevTables &t = c->dt;
    const u64 N = t.N, K = t.K;
    const u32 np = c->lv[l].nparts;
    ConvColsArgs a;
    a.in = (const double *)coef;
    a.in_pstride = l * N;
    a.in_first = c->modup_first_dev[l];
    a.descs = c->modup_desc_dev[l];
    a.out = (double *)compact;
    a.out_pstride = (u64)np * (l + K) * N;
    a.out_dstride = (l + K) * N;
    a.j0 = 0;
    a.nj = l + (u32)K;
    a.npolys = npolys;
    a.ndig = np;
    launch_conv_cols_f64(t, a, c->stream);
}

// The ModDown's conversion P -> Q with its INTT column pass before it and the NTT
// column pass after it: pb holds the first-pass (rows) intermediates of the P part
// of the 2*nb accumulators, tt receives Q slots j0 .. j0+nj-1 (poly at tt + poly*l*N).
// With lift (ModDown + rescale by q_{l-1}): the rows-pass intermediates of acc_{l-1}
// and d_{l-1} of poly p at lift + 2p*N (+ N) (lift_d false: acc_{l-1} alone at
// lift + p*N, d being folded into acc), and the transform input is the ModDown +
// rescale prologue (FUSE_MDRS_IN) with the lift formed in the kernel.
static void moddown_cols(fhea_ctx *c, u32 l, u32 nb, const u64 *pb, u64 *tt, u32 j0, u32 nj,
                         const u64 *lift = nullptr, bool lift_d = true) {
    const DevTables &t = c->dt;
    const u64 N = t.N, K = t.K;
    const LevelTables &lt = c->lv[l];
    ConvColsArgs a;
    a.in = (const double *)pb;
    a.in_pstride = K * N;
    a.in_first = c->zero_u32;
    a.descs = c->moddown_desc_dev[l];
    a.out = (double *)tt;
    a.out_pstride = l * N;
    a.out_dstride = 0;
    a.j0 = j0;
    a.nj = nj;
    a.npolys = 2 * nb;
    a.ndig = 1;
    if (lift) {
        a.lift_in = (const double *)lift;
        a.lift_stride = lift_d ? 2 * N : N;
        a.lift_d = lift_d ? 1 : 0;
        a.lift_prime = l - 1;
        a.lift_j = l - 1;
        a.lift_cst = lt.rs_liftf;
        a.cst_in = lt.rs_qlqlf;
        a.cst_pq = lt.rs_pqf;
    }
    launch_conv_cols_f64(t, a, c->stream);
}

// First (rows) pass of the ModDown INTT of the P part of nb accumulator pairs into pb.
static void moddown_rows(fhea_ctx *c, u32 l, u32 nb, const u64 *acc, u64 *pb) {
    const DevTables &t = c->dt;
    const u64 N = t.N, K = t.K;
    NttFuse fz;
    fz.only_pass = 1;
    launch_ntt_f64(t, true, acc + l * N, pb, (l + K) * N, K * N, N, N, 2 * nb, K, t.ident + t.L, fz, c->stream);
}

// Key switch of nb inputs c_z = c_base + z*c_stride (l limbs, eval domain):
// out0_z = add0_z + ks0, out1_z = add1_z + ks1 (key_switch.cpp:167-439 followed by
// the composite's poly_add).  keys[z] = device pointer of item z's key.
static int ks_core(fhea_ctx *c, u32 l, u32 key_limbs, u32 nb, const u64 *c_base, u64 c_stride,
                   const std::vector<const u64 *> &keys, const std::vector<CombineArgs> &comb,
                   u64 *ws, const RescaleAfter *ra = nullptr, const std::vector<TensorArgs> *fold = nullptr,
                   const std::vector<GatherArgs> *gather = nullptr, bool late_allowed = false,
                   const PreArgs *pre = nullptr) {
    const DevTables &t = c->dt;
    const u64 N = t.N, K = t.K;
    const LevelTables &lt = c->lv[l];
    const u32 np = lt.nparts;
    hipStream_t s = c->stream;
    u64 *coef = ws;
    u64 *compact = coef + nb * l * N;
    u64 *acc = compact + (u64)nb * np * (l + K) * N;
    u64 *pb = acc + nb * 2 * (l + K) * N;
    u64 *tt = pb + nb * 2 * K * N;
    const u64 cz = (u64)np * (l + K) * N;
    // ModUp: INTT of every input limb, convert each digit, NTT the converted limbs
    // (fused: the INTT's rows pass, then one kernel for its column pass, the
    // conversion and the NTT's column pass)
    const bool cc = modup_cc(c, l);
    // the folded HMult tensor (fold): its factors per item, for the fused rows kernel
    // and (with cc) for the ModUp INTT's first pass, which then forms d2 = a1 b1 itself
    const TensorArgs *fold_d = nullptr;
    if (fold) {
        fold_d = (pre && pre->fold) ? pre->fold
                                    : (const TensorArgs *)stage_args(c, fold->data(), nb * sizeof(TensorArgs));
        if (!fold_d) return FHEA_ERR_DEVICE;
    }
    // gather (rotations): input z is sigma_g of (*gather)[z].in, formed while loading
    // (the ModUp INTT's first pass, the rows kernel's own limbs); needs cc and the
    // fused rows kernel (ks_gather_ok)
    const GatherArgs *gather_d = nullptr;
    if (gather) {
        gather_d = (pre && pre->gather) ? pre->gather
                                        : (const GatherArgs *)stage_args(c, gather->data(), nb * sizeof(GatherArgs));
        if (!gather_d) return FHEA_ERR_DEVICE;
    }
    if (cc) {
        NttFuse fz;
        fz.only_pass = 1;
        if (fold) {
            fz.pro = FUSE_PROD_IN;
            fz.ta = fold_d;
        }
        if (gather) {
            fz.pro = FUSE_GATHER_IN;
            fz.ga = gather_d;
        }
        launch_ntt_f64(t, true, c_base, coef, c_stride, l * N, N, N, nb, l, t.ident, fz, s);
        modup_cols(c, l, nb, coef, compact);
    } else {
        launch_ntt(t, true, c_base, coef, c_stride, l * N, nb, l, t.ident, s);
        launch_fbc(t, c->modup_desc_dev[l], c->modup_first_dev[l], np, std::min(l, c->h.alpha), l - 1 + K,
                   coef, l * N, compact, (l + K) * N, cz, nb, s);
    }
    const KsIpArgs *ip_d = pre ? pre->ip : nullptr;
    if (!ip_d) {
        std::vector<KsIpArgs> ip(nb);
        for (u32 z = 0; z < nb; ++z)
            ip[z] = gather ? KsIpArgs{(*gather)[z].in, keys[z], (*gather)[z].galois}
                           : KsIpArgs{c_base + (u64)z * c_stride, keys[z], 0};
        ip_d = (const KsIpArgs *)stage_args(c, ip.data(), nb * sizeof(KsIpArgs));
        if (!ip_d) return FHEA_ERR_DEVICE;
    }
    // FHEA_F64_OFF bit 64 keeps the inner product as its own kernel
    const bool fused = t.use_f64 && !(t.f64_off & (1 | 4 | 64)) && np <= 8 && ks_rows_fusable(t);
    // HMult (the tensor folded in, the rescale after) and rotations (sigma_g read
    // while loading), both conversions fused: the rows kernel runs twice, first for
    // the P slots (and slot l-1 before a rescale), whose sums the ModDown transforms
    // and converts, then for the other Q slots with the last row pass of the ModDown
    // (+ rescale) and its epilogue fused in, so the Q slots' sums never go through
    // HBM (FHEA_F64_OFF bit 2048 keeps one launch).  The second launch reads the
    // items' inputs while it writes their outputs: late_allowed, from the caller,
    // says that no output overlaps an input other than at the same words.
    const bool late_ok = fused && cc && late_allowed && !(t.f64_off & (16 | 128 | 2048)) && moddown_cc(c);
    const bool late = late_ok && ((fold && ra && l >= 2) || (gather && !ra));
    KsRowsArgs ka_late{};
    if (fused) {
        // first forward pass of the converted digit limbs; the second runs inside the
        // inner product, which for the P slots also runs the ModDown INTT's first pass
        if (!cc) {
            NttFuse fz;
            fz.slots = lt.modup_slots;
            fz.nslots = lt.n_modup_slots;
            fz.only_pass = 1;
            launch_ntt_f64(t, false, compact, compact, cz, cz, N, N, nb, np * (l + K), lt.modup_ntt_prime, fz, s);
        }
        KsRowsArgs ka{compact, cz, ip_d, acc, 2 * (l + K) * N, pb, l, (u32)K, t.L, c->h.alpha, np,
                      key_limbs, c->h.dnum, nb};
        if (fold) {
            ka.ta = fold_d;
            ka.pmodq = t.pmodqf;
        }
        ka.gather = gather ? 1 : 0;
        if (late) {
            ka_late = ka;
            if (!(t.f64_off & 8192)) {
                // the P slots (and before the rescale slot l-1) also run the ModDown
                // INTT's row pass on their sums (the lift of limb l-1 into the dead
                // ModUp region)
                ka_late.lift = (double *)coef;
                ka_late.early = 1;
                if (ra) launch_ks_rows_f64_early(t, ka_late, l - 1, (u32)K + 1, s);
                else launch_ks_rows_f64_early(t, ka_late, l, (u32)K, s);
            } else if (ra) {
                launch_ks_rows_f64_q(t, ka, l - 1, (u32)K + 1, s);
            } else {
                launch_ks_rows_f64_q(t, ka, l, (u32)K, s);
            }
        } else {
            launch_ks_rows_f64(t, ka, s);
        }
    } else if (!t.use_f64 && mixed_fused(t) && !(t.f64_off & (4 | 64)) && np <= 8 && ks_rows_fusable(t) &&
               t.q_big_prefix < l) {
        // mixed chain: the converted limbs of primes below 2^50 take the FP64 first
        // pass and the fused second pass + inner product on Q slots qb .. l-1; those of
        // primes >= 2^50 the whole integer transform and the integer inner product on
        // slots 0 .. qb-1 and the P slots
        const u32 qb = t.q_big_prefix;
        NttFuse fz;
        fz.slots = lt.modup_slots;
        fz.nslots = lt.n_modup_slots;
        fz.only_pass = 1;
        fz.pmask = t.pmask_small;
        launch_ntt_f64(t, false, compact, compact, cz, cz, N, N, nb, np * (l + K), lt.modup_ntt_prime, fz, s);
        launch_ntt_int(t, false, compact, compact, cz, cz, nb, np * (l + K), lt.modup_ntt_prime, s, t.pmask_big);
        const KsRowsArgs ka{compact, cz, ip_d, acc, 2 * (l + K) * N, pb, l, (u32)K, t.L, c->h.alpha, np,
                            key_limbs, c->h.dnum, nb};
        launch_ks_rows_f64_q(t, ka, qb, l - qb, s);
        if (qb)
            launch_ks_inner(t, ip_d, compact, cz, acc, 2 * (l + K) * N, nb, l, c->h.alpha, np, key_limbs,
                            c->h.dnum, false, s, 0, qb);
        launch_ks_inner(t, ip_d, compact, cz, acc, 2 * (l + K) * N, nb, l, c->h.alpha, np, key_limbs, c->h.dnum,
                        false, s, l, (u32)K);
    } else {
        if (t.use_f64 && !(t.f64_off & 1)) {
            NttFuse fz;  // only the converted slots of each digit
            fz.slots = lt.modup_slots;
            fz.nslots = lt.n_modup_slots;
            if (cc) fz.only_pass = 2;  // the column pass ran in the fused conversion
            launch_ntt_f64(t, false, compact, compact, cz, cz, N, N, nb, np * (l + K), lt.modup_ntt_prime, fz, s);
        } else {
            launch_ntt(t, false, compact, compact, cz, cz, nb, np * (l + K), lt.modup_ntt_prime, s);
        }
        // inner product with the evaluation key
        bool shared = true;
        for (u32 z = 1; z < nb; ++z) shared &= keys[z] == keys[0];
        launch_ks_inner(t, ip_d, compact, cz, acc, 2 * (l + K) * N, nb, l, c->h.alpha, np, key_limbs,
                        c->h.dnum, shared, s);
    }
    // the ModUp coefficients are dead now: their region holds the rescale's lift
    // the P slots' sums are stored like the Q slots' and the ModDown INTT runs whole
    // in its own launch; FHEA_F64_OFF bit 128 runs its first pass inside the inner
    // product instead (one wave per SIMD there: measured 2.6% slower per HMult)
    const bool pb_mid = fused && (t.f64_off & 128);
    if (ra)
        return ks_moddown_rescale(c, l, nb, acc, pb, tt, coef, comb, *ra, pb_mid, fold != nullptr,
                                  late ? &ka_late : nullptr, pre);
    return ks_moddown(c, l, nb, acc, pb, tt, comb, pb_mid, late ? &ka_late : nullptr, pre);
}

// ModDown of nb accumulator pairs acc_z = [2][l+K][N] (key_switch.cpp:385-427) and the
// composite's add: out_z = add_z + ModDown(acc_z).  pb: 2*nb*K*N words, tt: 2*nb*l*N.
static int ks_moddown(fhea_ctx *c, u32 l, u32 nb, u64 *acc, u64 *pb, u64 *tt,
                      const std::vector<CombineArgs> &comb, bool pb_mid, const KsRowsArgs *late,
                      const PreArgs *pre) {
    const DevTables &t = c->dt;
    const u64 N = t.N, K = t.K;
    const LevelTables &lt = c->lv[l];
    hipStream_t s = c->stream;
    const bool cc = t.use_f64 && !(t.f64_off & 17) && moddown_cc(c);
    if (cc) {
        // INTT rows pass (unless the rows kernel ran it on its sums), then its column
        // pass + conversion + the NTT column pass fused
        if (!pb_mid && !(late && late->early)) moddown_rows(c, l, nb, acc, pb);
        moddown_cols(c, l, nb, pb, tt, 0, l);
    } else {
        moddown_intt(c, l, nb, acc, pb, pb_mid);
        launch_fbc(t, c->moddown_desc_dev[l], c->zero_u32, 1, K, l, pb, K * N, tt, 0, l * N, 2 * nb, s);
    }
    auto *cb_d = (pre && pre->cb) ? pre->cb
                                  : (const CombineArgs *)stage_args(c, comb.data(), nb * sizeof(CombineArgs));
    if (!cb_d) return FHEA_ERR_DEVICE;
    if (late && cc) {
        // the Q slots' rows kernel with the last row pass and its epilogue in place of
        // the stores of their sums
        KsRowsArgs ka = *late;
        ka.tt = (const double *)tt;
        ka.tt_pstride = l * N;
        ka.cb = cb_d;
        ka.cst = lt.pinvf;
        launch_ks_rows_f64_q(t, ka, 0, l, s);
        return FHEA_OK;
    }
    if (t.use_f64 && !(t.f64_off & 17)) {
        // NTT of the converted P part with (acc - t) * P^-1 [+ add] fused into its last pass
        NttFuse fz;
        fz.epi = FUSE_MODDOWN_OUT;
        fz.cst_out = lt.pinvf;
        fz.cb = cb_d;
        fz.acc = acc;
        fz.acc_stride = (l + K) * N;
        if (cc) fz.only_pass = 2;
        launch_ntt_f64(t, false, tt, tt, l * N, l * N, N, N, 2 * nb, l, t.ident, fz, s);
    } else if (mixed_fused(t)) {
        // mixed chain: the fused FP64 transform on the limbs below 2^50, the integer
        // transform and combine on the leading limbs of primes >= 2^50
        const u32 qb = std::min(t.q_big_prefix, l);
        if (qb < l) {
            NttFuse fz;
            fz.epi = FUSE_MODDOWN_OUT;
            fz.cst_out = lt.pinvf;
            fz.cb = cb_d;
            fz.acc = acc;
            fz.acc_stride = (l + K) * N;
            fz.pmask = t.pmask_small;
            launch_ntt_f64(t, false, tt, tt, l * N, l * N, N, N, 2 * nb, l, t.ident, fz, s);
        }
        if (qb) {
            launch_ntt_int(t, false, tt, tt, l * N, l * N, 2 * nb, qb, t.ident, s);
            launch_moddown_combine(t, cb_d, acc, 2 * (l + K) * N, (l + K) * N, tt, 2 * l * N, l * N, lt.pinv,
                                   lt.pinvf, nb, qb, s);
        }
    } else {
        launch_ntt(t, false, tt, tt, l * N, l * N, 2 * nb, l, t.ident, s);
        launch_moddown_combine(t, cb_d, acc, 2 * (l + K) * N, (l + K) * N, tt, 2 * l * N, l * N, lt.pinv,
                               lt.pinvf, nb, l, s);
    }
    return FHEA_OK;
}

// Rescale of nb items of two polynomials each; poly (z, s) at a_base + z*az + s*as,
// outputs outs[2z+s] (l-1 limbs).  rescale.cpp:3-72 applied to c0 and c1 (OP 8).
static int rescale_core(fhea_ctx *c, u32 l, u32 nb, const u64 *a_base, u64 az, u64 as,
                        const std::vector<u64 *> &outs, u64 *ws) {
    const DevTables &t = c->dt;
    const u64 N = t.N;
    const LevelTables &lt = c->lv[l];
    hipStream_t s = c->stream;
    u64 *v = ws;
    u64 *tmp = v + nb * 2 * N;
    launch_ntt(t, true, a_base + (l - 1) * N, v, az, 2 * N, nb, 2, c->rs_last_map[l], s, as, N);
    std::vector<RescaleOut> ro(2 * nb);
    for (u32 z = 0; z < nb; ++z)
        for (u32 k = 0; k < 2; ++k) ro[2 * z + k] = {a_base + z * az + k * as, outs[2 * z + k]};
    auto *ro_d = (const RescaleOut *)stage_args(c, ro.data(), ro.size() * sizeof(RescaleOut));
    if (!ro_d) return FHEA_ERR_DEVICE;
    if (t.use_f64 && !(t.f64_off & 17)) {
        // centred lift * T_i fused into the NTT's first pass, a*q_l^-1 + x into its last
        NttFuse fz;
        fz.pro = FUSE_RESCALE_IN;
        fz.v = v;
        fz.v_stride = N;
        fz.ql = c->h.primes[l - 1];
        fz.cst_in = lt.rs_qlqlf;
        fz.epi = FUSE_RESCALE_OUT;
        fz.cst_out = lt.rs_qlinvf;
        fz.ro = ro_d;
        launch_ntt_f64(t, false, tmp, tmp, (l - 1) * N, (l - 1) * N, N, N, 2 * nb, l - 1, t.ident, fz, s);
        return FHEA_OK;
    }
    if (mixed_fused(t) && t.q_big_prefix < l - 1) {
        // mixed chain, q_{l-1} below 2^50: the fused FP64 transform on the output limbs
        // past the big-prime prefix, the integer route on the prefix
        const u32 qb = t.q_big_prefix;
        NttFuse fz;
        fz.pro = FUSE_RESCALE_IN;
        fz.v = v;
        fz.v_stride = N;
        fz.ql = c->h.primes[l - 1];
        fz.cst_in = lt.rs_qlqlf;
        fz.epi = FUSE_RESCALE_OUT;
        fz.cst_out = lt.rs_qlinvf;
        fz.ro = ro_d;
        fz.pmask = t.pmask_small;
        launch_ntt_f64(t, false, tmp, tmp, (l - 1) * N, (l - 1) * N, N, N, 2 * nb, l - 1, t.ident, fz, s);
        if (qb) {
            launch_rescale_prep(t, v, N, tmp, (l - 1) * N, lt.rs_qlql, c->rs_sub[l], lt.rs_qlqlf, 2 * nb, l, s, qb);
            launch_ntt_int(t, false, tmp, tmp, (l - 1) * N, (l - 1) * N, 2 * nb, qb, t.ident, s);
            launch_rescale_combine(t, ro_d, tmp, (l - 1) * N, lt.rs_qlinv, lt.rs_qlinvf, 2 * nb, l, s, qb);
        }
        return FHEA_OK;
    }
    launch_rescale_prep(t, v, N, tmp, (l - 1) * N, lt.rs_qlql, c->rs_sub[l], lt.rs_qlqlf, 2 * nb, l, s);
    launch_ntt(t, false, tmp, tmp, (l - 1) * N, (l - 1) * N, 2 * nb, l - 1, t.ident, s);
    launch_rescale_combine(t, ro_d, tmp, (l - 1) * N, lt.rs_qlinv, lt.rs_qlinvf, 2 * nb, l, s);
    return FHEA_OK;
}

// ModDown + add (comb[z] = {d0, d1, d0, d1}) + rescale by q_{l-1}
// (fhe_host_driver.cpp:272-306: key_switch.cpp:385-427, poly_add, rescale.cpp:3-72).
// v: 2*nb*N words.  With the FP64 transforms only limb l-1 of the ModDown result
// is formed (the rescale lifts it); the other limbs go straight to the output as
//   out_i = acc_i*(P q_l)^-1 + d_i*q_l^-1 + NTT(c*T_i - conv_i*(P q_l)^-1)  (mod q_i),
// which is ((acc_i - NTT(conv_i))*P^-1 + d_i)*q_l^-1 + NTT(c*T_i) rearranged by the
// linearity of the NTT mod q_i: the same words, one transform of l-1 limbs fewer.
static int ks_moddown_rescale(fhea_ctx *c, u32 l, u32 nb, u64 *acc, u64 *pb, u64 *tt, u64 *v,
                              const std::vector<CombineArgs> &comb, const RescaleAfter &ra, bool pb_mid,
                              bool folded, const KsRowsArgs *late, const PreArgs *pre) {
    const DevTables &t = c->dt;
    const u64 N = t.N, K = t.K;
    const LevelTables &lt = c->lv[l];
    hipStream_t s = c->stream;
    // mixed chain with q_{l-1} below 2^50: as below on the limbs past the
    // big-prime prefix (FP64 fused ends, prime bitmask), ModDown + rescale by the
    // integer kernels on the prefix
    const bool mix = !t.use_f64 && mixed_fused(t) && l >= 2 && t.q_big_prefix < l - 1;
    if ((!t.use_f64 && !mix) || (t.f64_off & 17) || l < 2) {
        int st = ks_moddown(c, l, nb, acc, pb, tt, comb, pb_mid, nullptr, pre);
        if (st) return st;
        return rescale_core(c, l, nb, ra.d_base, ra.dz, ra.ds, *ra.outs, v);
    }
    std::vector<RescaleOut> ro(2 * nb);
    for (u32 z = 0; z < nb; ++z)
        for (u32 k = 0; k < 2; ++k)
            ro[2 * z + k] = {folded ? nullptr : ra.d_base + z * ra.dz + k * ra.ds, (*ra.outs)[2 * z + k]};
    if (!mix && moddown_cc(c)) {
        // With the fused column kernel: the INTT rows passes of the P part and of the
        // last limbs acc_{l-1}, d_{l-1}, then one kernel for the INTT column passes, the
        // conversion, the rescaled limb's ModDown result and lift (in registers, by
        // linearity) and the ModDown + rescale prologue with the forward column pass;
        // then the rows pass with the epilogue.  lift (4*nb*N words) reuses the dead
        // ModUp region at v, which extends into the dead digit region for l < 4.
        NttFuse r1;
        r1.only_pass = 1;
        // (folded: the tensor's d is already in acc, the lift needs acc_{l-1} alone)
        u64 *lift = v;
        if (!(late && late->early)) {
            // (else the rows kernel has run these row passes on its sums)
            if (!pb_mid) moddown_rows(c, l, nb, acc, pb);
            const u64 ls = folded ? N : 2 * N;
            launch_ntt_f64(t, true, acc + (l - 1) * N, lift, (l + K) * N, ls, N, N, 2 * nb, 1, t.ident + (l - 1), r1,
                           s);
            if (!folded)
                launch_ntt_f64(t, true, ra.d_base + (l - 1) * N, lift + N, ra.dz, 4 * N, ra.ds, 2 * N, nb, 2,
                               c->rs_last_map[l], r1, s);
        }
        moddown_cols(c, l, nb, pb, tt, 0, l - 1, lift, !folded);
        auto *ro_d = (pre && pre->ro) ? pre->ro
                                      : (const RescaleOut *)stage_args(c, ro.data(), ro.size() * sizeof(RescaleOut));
        if (!ro_d) return FHEA_ERR_DEVICE;
        if (late) {
            // the Q slots' rows kernel with this row pass and epilogue in place of
            // the stores of their sums
            KsRowsArgs ka = *late;
            ka.tt = (const double *)tt;
            ka.tt_pstride = l * N;
            ka.ro = ro_d;
            ka.cst = lt.rs_pqf;
            launch_ks_rows_f64_q(t, ka, 0, l - 1, s);
            return FHEA_OK;
        }
        NttFuse f2;
        f2.epi = FUSE_MDRS_OUT;
        f2.cst_out = lt.rs_qlinvf;
        f2.cst_pq = lt.rs_pqf;
        f2.ro = ro_d;
        f2.acc = acc;
        f2.acc_stride = (l + K) * N;
        f2.only_pass = 2;
        launch_ntt_f64(t, false, tt, tt, l * N, l * N, N, N, 2 * nb, l - 1, t.ident, f2, s);
        return FHEA_OK;
    }
    moddown_intt(c, l, nb, acc, pb, pb_mid);
    launch_fbc(t, c->moddown_desc_dev[l], c->zero_u32, 1, K, l, pb, K * N, tt, 0, l * N, 2 * nb, s);
    auto *cb_d = (pre && pre->cb) ? pre->cb
                                  : (const CombineArgs *)stage_args(c, comb.data(), nb * sizeof(CombineArgs));
    if (!cb_d) return FHEA_ERR_DEVICE;
    // limb l-1 of the ModDown result, (acc - NTT(conv)) * P^-1 + d, in place in d
    NttFuse f1;
    f1.epi = FUSE_MODDOWN_OUT;
    f1.cst_out = lt.pinvf;
    f1.cb = cb_d;
    f1.acc = acc;
    f1.acc_stride = (l + K) * N;
    f1.slots = t.ident + (l - 1);
    f1.nslots = 1;
    if (mix) f1.pmask = t.pmask_small;
    launch_ntt_f64(t, false, tt, tt, l * N, l * N, N, N, 2 * nb, l, t.ident, f1, s);
    // its coefficients
    launch_ntt(t, true, ra.d_base + (l - 1) * N, v, ra.dz, 2 * N, nb, 2, c->rs_last_map[l], s, ra.ds, N);
    auto *ro_d = (pre && pre->ro) ? pre->ro
                                  : (const RescaleOut *)stage_args(c, ro.data(), ro.size() * sizeof(RescaleOut));
    if (!ro_d) return FHEA_ERR_DEVICE;
    NttFuse f2;
    f2.pro = FUSE_MDRS_IN;
    f2.v = v;
    f2.v_stride = N;
    f2.ql = c->h.primes[l - 1];
    f2.cst_in = lt.rs_qlqlf;
    f2.cst_pq = lt.rs_pqf;
    f2.epi = FUSE_MDRS_OUT;
    f2.cst_out = lt.rs_qlinvf;
    f2.ro = ro_d;
    f2.acc = acc;
    f2.acc_stride = (l + K) * N;
    if (mix) f2.pmask = t.pmask_small;
    launch_ntt_f64(t, false, tt, tt, l * N, l * N, N, N, 2 * nb, l - 1, t.ident, f2, s);
    if (mix && t.q_big_prefix) {
        // the prefix: its ModDown + add in place in d, then its rescale; its converted
        // limbs in tt are consumed by the combine, and tt (dead once the fused
        // transform above has read it) holds the rescale's lifted words
        const u32 qb = t.q_big_prefix;
        launch_ntt_int(t, false, tt, tt, l * N, l * N, 2 * nb, qb, t.ident, s);
        launch_moddown_combine(t, cb_d, acc, 2 * (l + K) * N, (l + K) * N, tt, 2 * l * N, l * N, lt.pinv, lt.pinvf,
                               nb, qb, s);
        launch_rescale_prep(t, v, N, tt, l * N, lt.rs_qlql, c->rs_sub[l], lt.rs_qlqlf, 2 * nb, l, s, qb);
        launch_ntt_int(t, false, tt, tt, l * N, l * N, 2 * nb, qb, t.ident, s);
        launch_rescale_combine(t, ro_d, tt, l * N, lt.rs_qlinv, lt.rs_qlinvf, 2 * nb, l, s, qb);
    }
    return FHEA_OK;
}

struct MultItem {
    const u64 *a0, *a1, *b0, *b1;
    u64 *d0, *d1;
};

// A caller-provided workspace: the pipeline uses it as is and never grows it.
struct FixedWs {
    u64 *p;
    u64 words;
};

// eval_mult_relin_rescale (fhe_host_driver.cpp:272-306) for a batch sharing one
// relinearisation key: tensor -> KS(d2) (+d0, +d1 fused) -> rescale.  With `fixed`
// the chunks run in that workspace (one item per chunk if need be) instead of
// the context's, and nothing is allocated.
static int mult_relin_rescale_lane(fhea_ctx *c, u32 l, const u64 *key, u32 key_limbs,
                                   const std::vector<MultItem> &items, const FixedWs *fixed = nullptr);
// workspace words of one chunk of nb items
static u64 mult_relin_rescale_words(const fhea_ctx *c, u32 l, u32 nb);

// HMult with the tensor folded into the key switch: the tensor kernel writes d2
// only, the fused rows kernel adds P*d0 and P*d1 to the Q slots' accumulators, and
// the ModDown + rescale then yields the composite's sum (FP64 path with the fused
// key-switch rows kernel and the fused ModDown conversion; FHEA_F64_OFF bit 1024
// keeps the whole tensor).
static bool tensor_folded(const fhea_ctx *c, u32 l) {
    const DevTables &t = c->dt;
    return t.use_f64 && !(t.f64_off & (1 | 4 | 8 | 16 | 64 | 1024)) && l >= 2 && c->lv[l].nparts <= 8 &&
           ks_rows_fusable(t) && moddown_cc(c);
}

// Rotations with sigma_g applied while loading (no automorphism launch): the FP64
// path with the fused ModUp, the fused rows kernel and the fused ModDown row pass
// (FHEA_F64_OFF bit 4096 keeps the automorphism kernel).
static bool ks_gather_ok(const fhea_ctx *c, u32 l) {
    const DevTables &t = c->dt;
    return t.use_f64 && !(t.f64_off & (1 | 4 | 16 | 64 | 4096)) && c->lv[l].nparts <= 8 && ks_rows_fusable(t) &&
           modup_cc(c, l) && moddown_cc(c);
}

// Swaps the context's stream and workspace with lane k's.
static void swap_lane(fhea_ctx *c, u32 k) {
    std::swap(c->stream, c->lane_stream[k]);
    std::swap(c->ws, c->lane_ws[k]);
    std::swap(c->ws_words, c->lane_ws_words[k]);
}

static int mult_relin_rescale(fhea_ctx *c, u32 l, const u64 *key, u32 key_limbs,
                              const std::vector<MultItem> &items) {
    const u32 nl = (u32)std::min<size_t>(c->lanes, items.size());
    if (nl < 2 || c->capturing) return mult_relin_rescale_lane(c, l, key, key_limbs, items);
    // parts k = 1 .. nl-1 on their lanes: each starts after everything issued so far
    // on the stream, and the stream continues only after all of them; part 0 on the
    // stream itself, issued last
    if (hipEventRecord(c->ev_fork, c->stream) != hipSuccess) return FHEA_ERR_DEVICE;
    const size_t n = items.size();
    int st = FHEA_OK;
    for (u32 k = nl - 1; k >= 1; --k) {
        if (hipStreamWaitEvent(c->lane_stream[k], c->ev_fork, 0) != hipSuccess) return FHEA_ERR_DEVICE;
        std::vector<MultItem> part(items.begin() + k * n / nl, items.begin() + (k + 1) * n / nl);
        swap_lane(c, k);
        const int sk = mult_relin_rescale_lane(c, l, key, key_limbs, part);
        swap_lane(c, k);
        if (!st) st = sk;
    }
    std::vector<MultItem> part0(items.begin(), items.begin() + n / nl);
    const int s0 = mult_relin_rescale_lane(c, l, key, key_limbs, part0);
    if (!st) st = s0;
    for (u32 k = 1; k < nl; ++k)
        if (hipEventRecord(c->lane_join[k], c->lane_stream[k]) != hipSuccess ||
            hipStreamWaitEvent(c->stream, c->lane_join[k], 0) != hipSuccess)
            return FHEA_ERR_DEVICE;
    return st;
}

static u64 mult_relin_rescale_words(const fhea_ctx *c, u32 l, u32 nb) {
    return (u64)nb * 3 * l * c->h.N + std::max<u64>(nb * ks_words(c, l), nb * rescale_words(c, l));
}

// Host-side argument blocks of several chunks, staged with one copy.  Batches of
// more than ARGPACK_ITEMS items run in segments of that many, one copy each, so a
// block stays well inside the context's staging area.
static constexpr size_t ARGPACK_ITEMS = 4096;
struct ArgPack {
    std::vector<char> h;
    template <class T>
    size_t put(const std::vector<T> &v) {
        const size_t o = (h.size() + 63) & ~(size_t)63;
        h.resize(o + std::max<size_t>(v.size() * sizeof(T), 1));
        if (!v.empty()) memcpy(h.data() + o, v.data(), v.size() * sizeof(T));
        return o;
    }
};

static int mult_relin_rescale_lane(fhea_ctx *c, u32 l, const u64 *key, u32 key_limbs,
                                   const std::vector<MultItem> &items, const FixedWs *fixed) {
    const u64 N = c->h.N;
    u32 chunk = c->max_chunk;
    if (fixed) {
        while (chunk > 1 && mult_relin_rescale_words(c, l, chunk) > fixed->words) --chunk;
        if (mult_relin_rescale_words(c, l, chunk) > fixed->words) return FHEA_ERR_SCRATCH;
    }
    if (items.empty()) return FHEA_OK;
    if (items.size() > ARGPACK_ITEMS) {
        for (size_t off = 0; off < items.size(); off += ARGPACK_ITEMS) {
            const std::vector<MultItem> seg(items.begin() + off,
                                            items.begin() + std::min(items.size(), off + ARGPACK_ITEMS));
            const int st = mult_relin_rescale_lane(c, l, key, key_limbs, seg, fixed);
            if (st) return st;
        }
        return FHEA_OK;
    }
    // one workspace for every chunk (sized for the largest), so that all chunks'
    // argument arrays can be formed, and staged with one copy, before any launch
    u64 *D = nullptr;
    if (fixed) {
        D = fixed->p;
    } else {
        int st = ensure_ws(c, mult_relin_rescale_words(c, l, (u32)std::min<size_t>(chunk, items.size())));
        if (st) return st;
        D = c->ws;
    }
    const bool fold = tensor_folded(c, l);
    // with the fused ModUp as well, the ModUp INTT forms d2 itself: no tensor launch
    const bool prod = fold && modup_cc(c, l);
    struct Chunk {
        u32 off, nb;
        bool late;  // no output overlaps an input other than at the same words
        std::vector<TensorArgs> ta;
        std::vector<CombineArgs> cb;
        std::vector<u64 *> outs;
        size_t o_ta, o_ip, o_cb, o_ro;
    };
    std::vector<Chunk> chunks;
    ArgPack pk;
    for (size_t off = 0; off < items.size(); off += chunk) {
        Chunk ch;
        ch.off = (u32)off;
        ch.nb = (u32)std::min<size_t>(chunk, items.size() - off);
        const u32 nb = ch.nb;
        ch.late = true;
        for (u32 z = 0; z < nb && ch.late; ++z) {
            const MultItem &it = items[off + z];
            const u64 Wi = (u64)l * N, Wo = (u64)(l - 1) * N;
            for (const u64 *o : {(const u64 *)it.d0, (const u64 *)it.d1})
                for (const u64 *in : {it.a0, it.a1, it.b0, it.b1})
                    if (o != in && o < in + Wi && in < o + Wo) ch.late = false;
        }
        ch.ta.resize(nb);
        ch.cb.resize(nb);
        ch.outs.resize(2 * nb);
        std::vector<KsIpArgs> ip(nb);
        std::vector<RescaleOut> ro(2 * nb);
        for (u32 z = 0; z < nb; ++z) {
            const MultItem &it = items[ch.off + z];
            u64 *d0 = D + (u64)z * 3 * l * N, *d1 = d0 + l * N, *d2 = d1 + l * N;
            ch.ta[z] = {it.a0, it.a1, it.b0, it.b1, fold ? nullptr : d0, fold ? nullptr : d1, d2};
            ch.cb[z] = {d0, d1, d0, d1};
            ch.outs[2 * z] = it.d0;
            ch.outs[2 * z + 1] = it.d1;
            // as ks_core and ks_moddown_rescale would form them: the key switch's input
            // d2 (c_base + z*c_stride), the rescale's (d, out) pairs
            ip[z] = KsIpArgs{d2, key, 0};
            for (u32 k = 0; k < 2; ++k) ro[2 * z + k] = {fold ? nullptr : d0 + (u64)k * l * N, ch.outs[2 * z + k]};
        }
        ch.o_ta = pk.put(ch.ta);
        ch.o_ip = pk.put(ip);
        ch.o_cb = pk.put(ch.cb);
        ch.o_ro = pk.put(ro);
        chunks.push_back(std::move(ch));
    }
    char *dev = (char *)stage_args(c, pk.h.data(), pk.h.size());
    if (!dev) return FHEA_ERR_DEVICE;
    for (const Chunk &ch : chunks) {
        const u32 nb = ch.nb;
        u64 *rest = D + (u64)nb * 3 * l * N;
        std::vector<const u64 *> keys(nb, key);
        PreArgs pa;
        pa.fold = (const TensorArgs *)(dev + ch.o_ta);
        pa.ip = (const KsIpArgs *)(dev + ch.o_ip);
        pa.cb = (const CombineArgs *)(dev + ch.o_cb);
        pa.ro = (const RescaleOut *)(dev + ch.o_ro);
        if (!prod) launch_tensor(c->dt, pa.fold, nb, l, c->stream);
        const RescaleAfter ra{D, 3 * l * N, l * N, &ch.outs};
        const int st = ks_core(c, l, key_limbs, nb, D + 2 * l * N, 3 * l * N, keys, ch.cb, rest, &ra,
                               fold ? &ch.ta : nullptr, nullptr, ch.late, &pa);
        if (st) return st;
    }
    return check_hip(hipGetLastError());
}

// mult + relinearise without the rescale (the first half of
// fhe_host_driver.cpp:272-306): tensor -> KS(d2) with d0, d1 added in the ModDown.
static int mult_relin(fhea_ctx *c, u32 l, const u64 *key, u32 key_limbs, const std::vector<MultItem> &items) {
    const u64 N = c->h.N;
    for (size_t off = 0; off < items.size(); off += c->max_chunk) {
        u32 nb = (u32)std::min<size_t>(c->max_chunk, items.size() - off);
        u64 dwords = (u64)nb * 3 * l * N;
        int st = ensure_ws(c, dwords + nb * ks_words(c, l));
        if (st) return st;
        u64 *D = c->ws, *rest = D + dwords;
        std::vector<TensorArgs> ta(nb);
        std::vector<CombineArgs> cb(nb);
        std::vector<const u64 *> keys(nb, key);
        for (u32 z = 0; z < nb; ++z) {
            const MultItem &it = items[off + z];
            u64 *d0 = D + (u64)z * 3 * l * N, *d1 = d0 + l * N, *d2 = d1 + l * N;
            ta[z] = {it.a0, it.a1, it.b0, it.b1, d0, d1, d2};
            cb[z] = {it.d0, it.d1, d0, d1};
        }
        auto *ta_d = (const TensorArgs *)stage_args(c, ta.data(), nb * sizeof(TensorArgs));
        if (!ta_d) return FHEA_ERR_DEVICE;
        launch_tensor(c->dt, ta_d, nb, l, c->stream);
        st = ks_core(c, l, key_limbs, nb, D + 2 * l * N, 3 * l * N, keys, cb, rest);
        if (st) return st;
    }
    return check_hip(hipGetLastError());
}

struct RotItem {
    u64 *d0, *d1;
    u64 galois;
    const u64 *key;
    u32 input;  // index into the call's inputs
};
struct RotInput {
    const u64 *in0, *in1;
};

// Rotations of ciphertexts by several Galois elements each, every input ModUp'd
// once: digits of c1 computed once per input, then per rotation the permuted
// inner product with its key, ModDown, and sigma_g(c0) added (rotation =
// sigma_g(c0) + KS(sigma_g(c1)) with sigma_g applied to the digits instead of c1).
// All ModUps run first (batched); the rotations then run in chunks ordered by key,
// so that a key is read from HBM once for all the inputs that use it.  Outputs
// must not overlap any input (the caller checks).
static int rotations_hoisted(fhea_ctx *c, u32 l, u32 key_limbs, const std::vector<RotInput> &inputs,
                             std::vector<RotItem> items) {
    const DevTables &t = c->dt;
    const u64 N = t.N, K = t.K;
    const LevelTables &lt = c->lv[l];
    const u32 np = lt.nparts;
    hipStream_t s = c->stream;
    const u32 ni = (u32)inputs.size();
    const u64 cz = (u64)np * (l + K) * N;
    const u64 per_item = 2 * l * N + 2 * (l + K) * N + 2 * K * N + 2 * l * N;
    // chunk: up to 256 rotations, within 2 GiB of pipeline workspace
    const u64 cap = std::max<u64>(1, std::min<u64>(256, (2ull << 30) / (8 * per_item)));
    const u32 nmax = (u32)std::min<u64>(cap, items.size());
    int st = ensure_ws(c, ni * (l * N + cz) + nmax * per_item);
    if (st) return st;
    u64 *coef = c->ws, *compact = coef + (u64)ni * l * N, *rest = compact + (u64)ni * cz;
    // ModUp of every input's c1: INTT (one launch when the inputs are evenly spaced)
    bool strided = ni > 1;
    const long long step = ni > 1 ? (long long)(inputs[1].in1 - inputs[0].in1) : 0;
    for (u32 k = 1; k < ni && strided; ++k) strided = (long long)(inputs[k].in1 - inputs[k - 1].in1) == step;
    // (with the fused column kernel: the INTT's rows pass only, then its column pass,
    // the conversion and the NTT's column pass in one launch over all inputs)
    const bool cc = modup_cc(c, l);
    NttFuse rows;
    rows.only_pass = 1;
    if (strided && step > 0) {
        if (cc)
            launch_ntt_f64(t, true, inputs[0].in1, coef, (u64)step, l * N, N, N, ni, l, t.ident, rows, s);
        else
            launch_ntt(t, true, inputs[0].in1, coef, (u64)step, l * N, ni, l, t.ident, s);
    } else {
        for (u32 k = 0; k < ni; ++k) {
            if (cc)
                launch_ntt_f64(t, true, inputs[k].in1, coef + (u64)k * l * N, l * N, l * N, N, N, 1, l, t.ident, rows,
                               s);
            else
                launch_ntt(t, true, inputs[k].in1, coef + (u64)k * l * N, l * N, l * N, 1, l, t.ident, s);
        }
    }
    if (cc)
        modup_cols(c, l, ni, coef, compact);
    else
        launch_fbc(t, c->modup_desc_dev[l], c->modup_first_dev[l], np, std::min(l, c->h.alpha), l - 1 + K, coef,
                   l * N, compact, (l + K) * N, cz, ni, s);
    if (t.use_f64 && !(t.f64_off & 1)) {
        NttFuse fz;
        fz.slots = lt.modup_slots;
        fz.nslots = lt.n_modup_slots;
        if (cc) fz.only_pass = 2;
        launch_ntt_f64(t, false, compact, compact, cz, cz, N, N, ni, np * (l + K), lt.modup_ntt_prime, fz, s);
    } else {
        launch_ntt(t, false, compact, compact, cz, cz, ni, np * (l + K), lt.modup_ntt_prime, s);
    }
    std::stable_sort(items.begin(), items.end(), [](const RotItem &a, const RotItem &b) { return a.key < b.key; });
    // sigma_g(c0) read by the fused ModDown row pass (FP64 path with the fused ModDown
    // conversion) instead of an automorphism launch; outputs never overlap inputs here
    const bool add_gather = t.use_f64 && !(t.f64_off & (16 | 4096)) && moddown_cc(c);
    for (size_t off = 0; off < items.size(); off += nmax) {
        const u32 nb = (u32)std::min<size_t>(nmax, items.size() - off);
        u64 *S = rest, *acc = S + (u64)nb * 2 * l * N, *pb = acc + (u64)nb * 2 * (l + K) * N;
        u64 *tt = pb + (u64)nb * 2 * K * N;
        std::vector<AutoArgs> aa(nb);
        std::vector<KsRotArgs> ra(nb);
        std::vector<CombineArgs> cb(nb);
        for (u32 z = 0; z < nb; ++z) {
            const RotItem &it = items[off + z];
            const RotInput &in = inputs[it.input];
            u64 *s0 = S + (u64)z * 2 * l * N;
            aa[z] = {in.in0, nullptr, s0, nullptr, it.galois};  // sigma_g(c0) only
            ra[z] = {in.in1, it.key, it.galois, compact + (u64)it.input * cz};
            // the fused ModDown row pass reads c0 through sigma_g itself
            cb[z] = add_gather ? CombineArgs{it.d0, it.d1, in.in0, nullptr, it.galois}
                               : CombineArgs{it.d0, it.d1, s0, nullptr, 0};
        }
        auto *ra_d = (const KsRotArgs *)stage_args(c, ra.data(), nb * sizeof(KsRotArgs));
        if (!ra_d) return FHEA_ERR_DEVICE;
        if (!add_gather) {
            auto *aa_d = (const AutoArgs *)stage_args(c, aa.data(), nb * sizeof(AutoArgs));
            if (!aa_d) return FHEA_ERR_DEVICE;
            launch_automorph(t, aa_d, nb, l, s, 1);
        }
        launch_ks_inner_rot(t, ra_d, acc, 2 * (l + K) * N, nb, l, c->h.alpha, np, key_limbs, c->h.dnum, s);
        st = ks_moddown(c, l, nb, acc, pb, tt, cb);
        if (st) return st;
    }
    return check_hip(hipGetLastError());
}

struct AutoItem {
    const u64 *in0, *in1;
    u64 *d0, *d1;
    u64 galois;
    const u64 *key;
};

// eval_rotate_single_step (fhe_host_driver.cpp:331-359) with explicit Galois
// elements: sigma(c0), sigma(c1) -> KS(sigma(c1)) -> (sigma(c0) + ks0, ks1).
static int automorphism_ks(fhea_ctx *c, u32 l, u32 key_limbs, const std::vector<AutoItem> &items) {
    if (items.size() > ARGPACK_ITEMS) {
        for (size_t off = 0; off < items.size(); off += ARGPACK_ITEMS) {
            const std::vector<AutoItem> seg(items.begin() + off,
                                            items.begin() + std::min(items.size(), off + ARGPACK_ITEMS));
            const int st = automorphism_ks(c, l, key_limbs, seg);
            if (st) return st;
        }
        return FHEA_OK;
    }
    const u64 N = c->h.N;
    const bool gok = ks_gather_ok(c, l);
    // the chunks that read sigma_g while loading: their argument arrays first, all
    // staged with one copy
    struct Chunk {
        u32 off, nb;
        bool gather, late;
        std::vector<GatherArgs> ga;
        std::vector<CombineArgs> cb;
        std::vector<const u64 *> keys;
        size_t o_ga, o_ip, o_cb;
    };
    std::vector<Chunk> chunks;
    ArgPack pk;
    u32 gather_nb = 0;
    for (size_t off = 0; off < items.size(); off += c->max_chunk) {
        Chunk ch;
        ch.off = (u32)off;
        ch.nb = (u32)std::min<size_t>(c->max_chunk, items.size() - off);
        const u32 nb = ch.nb;
        // sigma_g while loading, unless an item writes over its own c0 (its last kernel
        // reads c0 through sigma_g while it writes the outputs); late (the rows kernel
        // that writes the outputs also reads c1 through sigma_g) unless an output
        // overlaps c1 as well
        ch.gather = gok;
        ch.late = true;
        for (u32 z = 0; z < nb && ch.gather; ++z) {
            const AutoItem &it = items[off + z];
            const u64 W = (u64)l * N;
            auto ov = [W](const u64 *a, const u64 *b) { return a < b + W && b < a + W; };
            ch.gather = !ov(it.d0, it.in0) && !ov(it.d1, it.in0);
            ch.late = ch.late && !ov(it.d0, it.in1) && !ov(it.d1, it.in1);
        }
        if (ch.gather) {
            ch.ga.resize(nb);
            ch.cb.resize(nb);
            ch.keys.resize(nb);
            std::vector<KsIpArgs> ip(nb);
            for (u32 z = 0; z < nb; ++z) {
                const AutoItem &it = items[off + z];
                ch.ga[z] = {it.in1, it.galois};
                ch.cb[z] = {it.d0, it.d1, it.in0, nullptr, it.galois};
                ch.keys[z] = it.key;
                ip[z] = KsIpArgs{it.in1, it.key, it.galois};  // as ks_core forms them
            }
            ch.o_ga = pk.put(ch.ga);
            ch.o_ip = pk.put(ip);
            ch.o_cb = pk.put(ch.cb);
            gather_nb = std::max(gather_nb, nb);
        }
        chunks.push_back(std::move(ch));
    }
    char *dev = nullptr;
    if (gather_nb) {
        int st = ensure_ws(c, gather_nb * ks_words(c, l));
        if (st) return st;
        dev = (char *)stage_args(c, pk.h.data(), pk.h.size());
        if (!dev) return FHEA_ERR_DEVICE;
    }
    for (const Chunk &ch : chunks) {
        const u32 nb = ch.nb;
        if (ch.gather) {
            int st = ensure_ws(c, nb * ks_words(c, l));
            if (st) return st;
            PreArgs pa;
            pa.gather = (const GatherArgs *)(dev + ch.o_ga);
            pa.ip = (const KsIpArgs *)(dev + ch.o_ip);
            pa.cb = (const CombineArgs *)(dev + ch.o_cb);
            st = ks_core(c, l, key_limbs, nb, nullptr, 0, ch.keys, ch.cb, c->ws, nullptr, nullptr, &ch.ga, ch.late,
                         &pa);
            if (st) return st;
            continue;
        }
        u64 swords = (u64)nb * 2 * l * N;
        int st = ensure_ws(c, swords + nb * ks_words(c, l));
        if (st) return st;
        u64 *S = c->ws, *rest = S + swords;
        std::vector<AutoArgs> aa(nb);
        std::vector<CombineArgs> cb(nb);
        std::vector<const u64 *> keys(nb);
        for (u32 z = 0; z < nb; ++z) {
            const AutoItem &it = items[ch.off + z];
            u64 *s0 = S + (u64)z * 2 * l * N, *s1 = s0 + l * N;
            aa[z] = {it.in0, it.in1, s0, s1, it.galois};
            cb[z] = {it.d0, it.d1, s0, nullptr};
            keys[z] = it.key;
        }
        auto *aa_d = (const AutoArgs *)stage_args(c, aa.data(), nb * sizeof(AutoArgs));
        if (!aa_d) return FHEA_ERR_DEVICE;
        launch_automorph(c->dt, aa_d, nb, l, c->stream);
        st = ks_core(c, l, key_limbs, nb, S + l * N, 2 * l * N, keys, cb, rest);
        if (st) return st;
    }
    return check_hip(hipGetLastError());
}

}  // namespace fhea

// ===========================================================================
// C-ABI
// ===========================================================================
static bool range_ok(u64 off, u64 words, u64 total) { return off <= total && words <= total - off; }

#define CTX_CHECK(c)                               \
    do {                                           \
        if (!(c) || !(c)->poly) return FHEA_ERR_STATE; \
    } while (0)

#define POLY_CHECK(c, off, words)                                   \
    do {                                                            \
        if ((off) & 7) return FHEA_ERR_ALIGN;                       \
        if (!range_ok((off), (words), (c)->poly_words)) return FHEA_ERR_RANGE; \
    } while (0)

static bool overlaps(u64 a, u64 b, u64 W) { return a < b + W && b < a + W; }

// POLY_CHECK as a value: FHEA_OK, FHEA_ERR_ALIGN or FHEA_ERR_RANGE
static int poly_code(const fhea_ctx *c, u64 off, u64 words) {
    if (off & 7) return FHEA_ERR_ALIGN;
    return range_ok(off, words, c->poly_words) ? FHEA_OK : FHEA_ERR_RANGE;
}

static int key_check(const fhea_ctx *c, u64 evk, u32 key_limbs) {
    if (evk & 7) return FHEA_ERR_ALIGN;
    u64 words = 2ull * c->h.dnum * (key_limbs + c->h.K) * c->h.N;
    return range_ok(evk, words, c->key_words) ? FHEA_OK : FHEA_ERR_RANGE;
}

// Items of one fhea_submit group run as one pipeline (several chunks, possibly two
// streams), so an item may read its own inputs in place but no other item's outputs:
// every output range must be disjoint from every other output range and from the
// input ranges of the other items.  ins/outs: (word offset, item index); inputs span
// in_w words, outputs out_w words.
static bool batch_aliasing_ok(std::vector<std::pair<u64, u32>> outs, const std::vector<std::pair<u64, u32>> &ins,
                              u64 in_w, u64 out_w) {
    std::sort(outs.begin(), outs.end());
    for (size_t k = 0; k + 1 < outs.size(); ++k)
        if (outs[k + 1].first - outs[k].first < out_w) return false;
    for (const auto &x : ins) {
        // first output whose range ends after the input starts
        auto it = std::lower_bound(outs.begin(), outs.end(), x.first >= out_w - 1 ? x.first - (out_w - 1) : 0,
                                   [](const std::pair<u64, u32> &a, u64 v) { return a.first < v; });
        for (; it != outs.end() && it->first < x.first + in_w; ++it)
            if (it->second != x.second) return false;
    }
    return true;
}

extern "C" {

const char *fhea_status_string(int st) {
    switch (st) {
    case FHEA_OK: return "ok";
    case FHEA_ERR_BAD_OPCODE: return "unknown op code";
    case FHEA_ERR_RANGE: return "offset or size outside the arena";
    case FHEA_ERR_ALIGN: return "offset not a multiple of 8 words";
    case FHEA_ERR_ARG: return "invalid argument";
    case FHEA_ERR_NO_KEY: return "missing evaluation key";
    case FHEA_ERR_NO_MEM: return "device allocation failed";
    case FHEA_ERR_DEVICE: return "HIP runtime error";
    case FHEA_ERR_STATE: return "invalid context";
    case FHEA_ERR_SCRATCH: return "work exceeds the static scratch area of the no-allocation entry point";
    }
    return "unknown status";
}

int fhea_ctx_create(const fhea_params *p, fhea_ctx **out) {
    if (!p || !out) return FHEA_ERR_ARG;
    *out = nullptr;
    fhea_ctx *c = new fhea_ctx();
    {
        const int hs = host_tables_from_params(p, 1024, c->h);
        if (hs != FHEA_OK) {
            delete c;
            return hs;
        }
    }
    {
        const char *fi = getenv("FHEA_FORCE_INT");
        c->force_int = fi && atoi(fi) != 0;
        const char *fo = getenv("FHEA_F64_OFF");
        c->f64_off = fo ? (u32)strtoul(fo, nullptr, 0) : 0;
        const char *cv = getenv("FHEA_CC_VARIANT");
        c->cc_var = cv ? (u32)strtoul(cv, nullptr, 0) : 0;
    }
    c->device = (int)p->device;
    int st = FHEA_OK;
    if (hipSetDevice(c->device) != hipSuccess) st = FHEA_ERR_DEVICE;
    if (!st && hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) st = FHEA_ERR_DEVICE;
    c->own_stream = true;
    if (!st && (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess))
        st = FHEA_ERR_DEVICE;
    c->poly_words = p->poly_words ? (p->poly_words + 7) & ~7ull : 8ull * 1024 * 1024;
    c->key_words = p->key_words ? (p->key_words + 7) & ~7ull : 8ull * 1024 * 1024;
    if (!st && hipMalloc((void **)&c->poly, c->poly_words * 8) != hipSuccess) st = FHEA_ERR_NO_MEM;
    if (!st && hipMalloc((void **)&c->key, c->key_words * 8) != hipSuccess) st = FHEA_ERR_NO_MEM;
    if (!st && (hipMemset(c->poly, 0, c->poly_words * 8) != hipSuccess ||
                hipMemset(c->key, 0, c->key_words * 8) != hipSuccess))
        st = FHEA_ERR_DEVICE;
    c->args_bytes = 4u << 20;
    if (!st && hipMalloc(&c->args, c->args_bytes) != hipSuccess) st = FHEA_ERR_NO_MEM;
    if (!st && hipHostMalloc(&c->args_host, c->args_bytes, hipHostMallocDefault) != hipSuccess)
        st = FHEA_ERR_NO_MEM;
    if (!st) st = build_device_tables(c);
    if (!st) st = scratch_init(c->device);  // device c->device is current (hipSetDevice above)
    if (st) {
        fhea_ctx_destroy(c);
        return st;
    }
    *out = c;
    return FHEA_OK;
}

int fhea_ctx_destroy(fhea_ctx *c) {
    if (!c) return FHEA_ERR_STATE;
    hipSetDevice(c->device);
    if (c->capturing) {
        hipGraph_t g = nullptr;
        hipStreamEndCapture(c->stream, &g);
        if (g) hipGraphDestroy(g);
        c->capturing = false;
    }
    if (c->cap_dev) hipFree(c->cap_dev);
    if (c->cap_host) hipHostFree(c->cap_host);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->copy_stream) {
        hipStreamSynchronize(c->copy_stream);
        hipStreamDestroy(c->copy_stream);
        for (hipEvent_t &e : c->read_done)
            if (e) hipEventDestroy(e);
        if (c->read_dep) hipEventDestroy(c->read_dep);
    }
    for (auto &kv : c->graphs) {
        if (kv.second.exec) hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) hipGraphDestroy(kv.second.graph);
        if (kv.second.dev) hipFree(kv.second.dev);
        if (kv.second.host) hipHostFree(kv.second.host);
    }
    c->graphs.clear();
    for (void *p : c->retired) hipFree(p);
    c->retired.clear();
    if (c->poly) hipFree(c->poly);
    if (c->key) hipFree(c->key);
    if (c->ws) hipFree(c->ws);
    if (c->pp) hipFree(c->pp);
    if (c->dblob) hipFree(c->dblob);
    if (c->args) hipFree(c->args);
    if (c->args_host) hipHostFree(c->args_host);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    for (u32 k = 1; k < fhea_ctx::MAX_LANES; ++k) {
        if (c->lane_stream[k]) {
            hipStreamSynchronize(c->lane_stream[k]);
            hipStreamDestroy(c->lane_stream[k]);
        }
        if (c->lane_ws[k]) hipFree(c->lane_ws[k]);
        if (c->lane_join[k]) hipEventDestroy(c->lane_join[k]);
    }
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->stream && c->own_stream) hipStreamDestroy(c->stream);
    delete c;
    return FHEA_OK;
}

int fhea_ctx_info(const fhea_ctx *c, uint32_t *info) {
    if (!c || !info) return FHEA_ERR_STATE;
    info[0] = c->h.N;
    info[1] = c->h.L;
    info[2] = c->h.K;
    info[3] = c->h.dnum;
    info[4] = c->h.alpha;
    return FHEA_OK;
}

int fhea_ctx_primes(const fhea_ctx *c, uint64_t *out) {
    if (!c || !out) return FHEA_ERR_STATE;
    std::copy(c->h.primes.begin(), c->h.primes.end(), out);
    return FHEA_OK;
}

int fhea_ctx_table(const fhea_ctx *c, int which, uint64_t *out, uint64_t count) {
    if (!c || !out) return FHEA_ERR_STATE;
    const std::vector<u64> *v = which == 0 ? &c->h.psi : which == 1 ? &c->h.ninv
                              : which == 2 ? &c->h.tw : which == 3 ? &c->h.itw : nullptr;
    if (!v) return FHEA_ERR_ARG;
    if (count != v->size()) return FHEA_ERR_RANGE;
    std::copy(v->begin(), v->end(), out);
    return FHEA_OK;
}

int fhea_ctx_set_stream(fhea_ctx *c, void *stream) {
    CTX_CHECK(c);
    if (c->capturing) return FHEA_ERR_STATE;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return FHEA_ERR_DEVICE;
    if (stream) {
        if (c->own_stream) hipStreamDestroy(c->stream);
        c->stream = (hipStream_t)stream;
        c->own_stream = false;
    } else if (!c->own_stream) {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return FHEA_ERR_DEVICE;
        c->own_stream = true;
    }
    return FHEA_OK;
}

void *fhea_ctx_stream(const fhea_ctx *c) { return c ? (void *)c->stream : nullptr; }

uint64_t *fhea_poly_arena(fhea_ctx *c) { return c ? c->poly : nullptr; }
uint64_t *fhea_key_arena(fhea_ctx *c) { return c ? c->key : nullptr; }
uint64_t fhea_poly_words(const fhea_ctx *c) { return c ? c->poly_words : 0; }
uint64_t fhea_key_words(const fhea_ctx *c) { return c ? c->key_words : 0; }

static int arena_copy(fhea_ctx *c, u64 *base, u64 total, u64 off, void *host, u64 count, bool write) {
    CTX_CHECK(c);
    if (c->capturing) return FHEA_ERR_STATE;  // host copies synchronise the stream
    if (!host && count) return FHEA_ERR_ARG;
    if (!range_ok(off, count, total)) return FHEA_ERR_RANGE;
    if (!count) return FHEA_OK;
    hipError_t e;
    if (write)
        e = hipMemcpyAsync(base + off, host, count * 8, hipMemcpyHostToDevice, c->stream);
    else
        e = hipMemcpyAsync(host, base + off, count * 8, hipMemcpyDeviceToHost, c->stream);
    if (e != hipSuccess) return FHEA_ERR_DEVICE;
    return check_hip(hipStreamSynchronize(c->stream));
}

int fhea_write_poly(fhea_ctx *c, uint64_t off, const uint64_t *h, uint64_t n) {
    return c ? arena_copy(c, c->poly, c->poly_words, off, (void *)h, n, true) : FHEA_ERR_STATE;
}
int fhea_read_poly(fhea_ctx *c, uint64_t off, uint64_t *h, uint64_t n) {
    return c ? arena_copy(c, c->poly, c->poly_words, off, h, n, false) : FHEA_ERR_STATE;
}
int fhea_read_polys(fhea_ctx *c, uint32_t n, const uint64_t *off, const uint64_t *count, uint64_t *h) {
    CTX_CHECK(c);
    if (c->capturing) return FHEA_ERR_STATE;
    if (!off || !count || (!h && n)) return FHEA_ERR_ARG;
    for (u32 k = 0; k < n; ++k)
        if (!range_ok(off[k], count[k], c->poly_words)) return FHEA_ERR_RANGE;
    u64 at = 0;
    for (u32 k = 0; k < n; ++k) {
        if (count[k] &&
            hipMemcpyAsync(h + at, c->poly + off[k], count[k] * 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            return FHEA_ERR_DEVICE;
        at += count[k];
    }
    return check_hip(hipStreamSynchronize(c->stream));
}
int fhea_write_polys(fhea_ctx *c, uint32_t n, const uint64_t *off, const uint64_t *count, const uint64_t *h) {
    CTX_CHECK(c);
    if (c->capturing) return FHEA_ERR_STATE;
    if (!off || !count || (!h && n)) return FHEA_ERR_ARG;
    for (u32 k = 0; k < n; ++k)
        if (!range_ok(off[k], count[k], c->poly_words)) return FHEA_ERR_RANGE;
    u64 at = 0;
    for (u32 k = 0; k < n; ++k) {
        if (count[k] &&
            hipMemcpyAsync(c->poly + off[k], h + at, count[k] * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            return FHEA_ERR_DEVICE;
        at += count[k];
    }
    return check_hip(hipStreamSynchronize(c->stream));
}

int fhea_read_polys_async(fhea_ctx *c, uint32_t n, const uint64_t *off, const uint64_t *count, uint64_t *h,
                          uint32_t *ticket) {
    CTX_CHECK(c);
    if (c->capturing) return FHEA_ERR_STATE;
    if (!off || !count || (!h && n) || !ticket) return FHEA_ERR_ARG;
    for (u32 k = 0; k < n; ++k)
        if (!range_ok(off[k], count[k], c->poly_words)) return FHEA_ERR_RANGE;
    if (!c->copy_stream) {
        if (hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->read_dep, hipEventDisableTiming) != hipSuccess)
            return FHEA_ERR_DEVICE;
        for (hipEvent_t &e : c->read_done)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return FHEA_ERR_DEVICE;
    }
    const u32 seq = c->read_seq;
    hipEvent_t done = c->read_done[seq % fhea_ctx::READ_RING];
    // a ticket's event is reused READ_RING reads later: that read must be complete
    if (seq >= fhea_ctx::READ_RING && check_hip(hipEventSynchronize(done)) != FHEA_OK) return FHEA_ERR_DEVICE;
    if (hipEventRecord(c->read_dep, c->stream) != hipSuccess ||
        hipStreamWaitEvent(c->copy_stream, c->read_dep, 0) != hipSuccess)
        return FHEA_ERR_DEVICE;
    u64 at = 0;
    for (u32 k = 0; k < n; ++k) {
        if (count[k] && hipMemcpyAsync(h + at, c->poly + off[k], count[k] * 8, hipMemcpyDeviceToHost,
                                       c->copy_stream) != hipSuccess)
            return FHEA_ERR_DEVICE;
        at += count[k];
    }
    if (hipEventRecord(done, c->copy_stream) != hipSuccess) return FHEA_ERR_DEVICE;
    *ticket = seq;
    c->read_seq = seq + 1;
    return FHEA_OK;
}

int fhea_read_wait(fhea_ctx *c, uint32_t ticket) {
    CTX_CHECK(c);
    if (ticket >= c->read_seq) return FHEA_ERR_ARG;
    // a ticket more than READ_RING reads old has completed: fhea_read_polys_async
    // waited for its event before handing the event to a later read
    if (c->read_seq - ticket > fhea_ctx::READ_RING) return FHEA_OK;
    return check_hip(hipEventSynchronize(c->read_done[ticket % fhea_ctx::READ_RING]));
}

int fhea_write_key(fhea_ctx *c, uint64_t off, const uint64_t *h, uint64_t n) {
    return c ? arena_copy(c, c->key, c->key_words, off, (void *)h, n, true) : FHEA_ERR_STATE;
}
int fhea_read_key(fhea_ctx *c, uint64_t off, uint64_t *h, uint64_t n) {
    return c ? arena_copy(c, c->key, c->key_words, off, h, n, false) : FHEA_ERR_STATE;
}

int fhea_ctx_set_tuning(fhea_ctx *c, int knob, uint32_t v) {
    CTX_CHECK(c);
    if (c->capturing) return FHEA_ERR_STATE;
    const u32 lg = c->dt.logN;
    switch (knob) {
    case FHEA_TUNE_NTT_SPLIT_FWD:
    case FHEA_TUNE_NTT_SPLIT_INV:
        // v = lg: one pass over the whole limb (N = 2^13, 2^14)
        if (v != 0 && !(v == lg && (lg == 13 || lg == 14)) &&
            (lg <= 12 || v < 6 || v > 12 || lg - v < 6 || lg - v > 12))
            return FHEA_ERR_ARG;
        (knob == FHEA_TUNE_NTT_SPLIT_FWD ? c->dt.k1_fwd : c->dt.k1_inv) = v;
        return FHEA_OK;
    case FHEA_TUNE_NTT_WG_PER_CU:
        if (v > 4) return FHEA_ERR_ARG;
        c->dt.wg_per_cu = v;
        return FHEA_OK;
    case FHEA_TUNE_CHUNK:
        if (v < 1 || v > 64) return FHEA_ERR_ARG;
        c->max_chunk = v;
        return FHEA_OK;
    case FHEA_TUNE_STREAMS:
        if (v < 1 || v > fhea_ctx::MAX_LANES) return FHEA_ERR_ARG;
        if (v > 1 && !c->ev_fork && hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess)
            return FHEA_ERR_DEVICE;
        for (u32 k = 1; k < v; ++k) {
            if (!c->lane_stream[k] && hipStreamCreateWithFlags(&c->lane_stream[k], hipStreamNonBlocking) != hipSuccess)
                return FHEA_ERR_DEVICE;
            if (!c->lane_join[k] && hipEventCreateWithFlags(&c->lane_join[k], hipEventDisableTiming) != hipSuccess)
                return FHEA_ERR_DEVICE;
        }
        c->lanes = v;
        return FHEA_OK;
    default:
        return FHEA_ERR_ARG;
    }
}

int fhea_ctx_get_tuning(const fhea_ctx *c, int knob, uint32_t *v) {
    if (!c || !c->poly) return FHEA_ERR_STATE;
    if (!v) return FHEA_ERR_ARG;
    switch (knob) {
    case FHEA_TUNE_NTT_SPLIT_FWD: *v = c->dt.k1_fwd; return FHEA_OK;
    case FHEA_TUNE_NTT_SPLIT_INV: *v = c->dt.k1_inv; return FHEA_OK;
    case FHEA_TUNE_NTT_WG_PER_CU: *v = c->dt.wg_per_cu; return FHEA_OK;
    case FHEA_TUNE_CHUNK: *v = c->max_chunk; return FHEA_OK;
    case FHEA_TUNE_STREAMS: *v = c->lanes; return FHEA_OK;
    default: return FHEA_ERR_ARG;
    }
}

int fhea_mem_info(uint32_t device, uint64_t *free_bytes, uint64_t *total_bytes) {
    if (!free_bytes || !total_bytes) return FHEA_ERR_ARG;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return FHEA_ERR_DEVICE;
    if ((int)device != cur && hipSetDevice((int)device) != hipSuccess) return FHEA_ERR_DEVICE;
    size_t f = 0, t = 0;
    const hipError_t e = hipMemGetInfo(&f, &t);
    if ((int)device != cur) hipSetDevice(cur);
    if (e != hipSuccess) return FHEA_ERR_DEVICE;
    *free_bytes = f;
    *total_bytes = t;
    return FHEA_OK;
}

void *fhea_host_alloc(uint64_t bytes) {
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 8, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

int fhea_host_free(void *p) {
    if (!p) return FHEA_OK;
    return hipHostFree(p) == hipSuccess ? FHEA_OK : FHEA_ERR_DEVICE;
}

int fhea_copy_poly(fhea_ctx *c, uint64_t dst, uint64_t src, uint64_t n) {
    CTX_CHECK(c);
    if (!range_ok(dst, n, c->poly_words) || !range_ok(src, n, c->poly_words)) return FHEA_ERR_RANGE;
    if (!n || dst == src) return FHEA_OK;
    return check_hip(hipMemcpyAsync(c->poly + dst, c->poly + src, n * 8, hipMemcpyDeviceToDevice, c->stream));
}

int fhea_op(fhea_ctx *c, uint32_t op, uint64_t src_a, uint64_t src_b, uint64_t dst, uint64_t evk,
            uint32_t limbs, uint32_t key_limbs) {
    CTX_CHECK(c);
    if (op == FHEA_OP_LOAD_CONTEXT) return FHEA_OK;
    if (op < FHEA_OP_NTT || op > FHEA_OP_RESCALE) return FHEA_ERR_BAD_OPCODE;
    if ((src_a | src_b | dst | evk) & 7) return FHEA_ERR_ALIGN;
    if (limbs == 0 || limbs > c->h.L) return FHEA_ERR_ARG;
    const u64 N = c->h.N, W = (u64)limbs * N;
    const DevTables &t = c->dt;
    POLY_CHECK(c, src_a, W);
    u64 *P = c->poly;
    int st = FHEA_OK;
    switch (op) {
    case FHEA_OP_NTT:
    case FHEA_OP_INTT:
        POLY_CHECK(c, dst, W);
        launch_ntt(t, op == FHEA_OP_INTT, P + src_a, P + dst, W, W, 1, limbs, t.ident, c->stream);
        break;
    case FHEA_OP_POLY_ADD:
    case FHEA_OP_POLY_SUB:
    case FHEA_OP_POLY_MUL_NTT:
        POLY_CHECK(c, src_b, W);
        POLY_CHECK(c, dst, W);
        launch_binop(t, (int)op - FHEA_OP_POLY_ADD, P + src_a, P + src_b, P + dst, limbs, c->stream);
        break;
    case FHEA_OP_AUTOMORPHISM: {
        POLY_CHECK(c, src_b, N / 2);
        POLY_CHECK(c, dst, W);
        // the gather needs distinct input and output: stage through the workspace
        if ((st = ensure_ws(c, W))) return st;
        launch_automorph_map(t, P + src_a, c->ws, P + src_b, limbs, c->stream);
        if (hipMemcpyAsync(P + dst, c->ws, W * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
            return FHEA_ERR_DEVICE;
        break;
    }
    case FHEA_OP_KEY_SWITCH: {
        u32 kl = key_limbs ? key_limbs : limbs;
        if (kl < limbs || kl > c->h.L) return FHEA_ERR_ARG;
        POLY_CHECK(c, dst, 2 * W);
        if ((st = key_check(c, evk, kl))) return st;
        if ((st = ensure_ws(c, ks_words(c, limbs)))) return st;
        std::vector<const u64 *> keys = {c->key + evk};
        std::vector<CombineArgs> cb = {{P + dst, P + dst + W, nullptr, nullptr}};
        st = ks_core(c, limbs, kl, 1, P + src_a, 0, keys, cb, c->ws);
        break;
    }
    case FHEA_OP_RESCALE: {
        if (limbs < 2) return FHEA_ERR_ARG;
        const u64 W1 = (u64)(limbs - 1) * N;
        POLY_CHECK(c, src_b, W);
        POLY_CHECK(c, dst, 2 * W1);
        // inputs are staged so outputs may overlap them (fhe_accel_top.cpp:282-296
        // loads both operands into on-chip banks first)
        if ((st = ensure_ws(c, 2 * W + rescale_words(c, limbs)))) return st;
        u64 *A = c->ws;
        if (hipMemcpyAsync(A, P + src_a, W * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(A + W, P + src_b, W * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
            return FHEA_ERR_DEVICE;
        std::vector<u64 *> outs = {P + dst, P + dst + W1};
        st = rescale_core(c, limbs, 1, A, 2 * W, W, outs, A + 2 * W);
        break;
    }
    }
    if (st) return st;
    return check_hip(hipGetLastError());
}

int fhea_eval_add(fhea_ctx *c, uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1, uint64_t d0,
                  uint64_t d1, uint32_t limbs) {
    int st = fhea_op(c, FHEA_OP_POLY_ADD, a0, b0, d0, 0, limbs, 0);
    if (st) return st;
    return fhea_op(c, FHEA_OP_POLY_ADD, a1, b1, d1, 0, limbs, 0);
}

int fhea_eval_sub_plain(fhea_ctx *c, uint64_t a0, uint64_t a1, uint64_t pt, uint64_t d0, uint64_t d1,
                        uint32_t limbs) {
    CTX_CHECK(c);
    if (limbs == 0 || limbs > c->h.L) return FHEA_ERR_ARG;
    const u64 W = (u64)limbs * c->h.N;
    POLY_CHECK(c, a1, W);
    POLY_CHECK(c, d1, W);
    int st = fhea_op(c, FHEA_OP_POLY_SUB, a0, pt, d0, 0, limbs, 0);
    if (st) return st;
    return fhea_copy_poly(c, d1, a1, W);  // fhe_host_driver.cpp:264-269
}

int fhea_set_relin_key(fhea_ctx *c, uint64_t evk, uint32_t key_limbs) {
    CTX_CHECK(c);
    u32 kl = key_limbs ? key_limbs : c->h.L;
    if (kl > c->h.L) return FHEA_ERR_ARG;
    int st = key_check(c, evk, kl);
    if (st) return st;
    c->has_relin = true;
    c->relin_off = evk;
    c->relin_limbs = kl;
    return FHEA_OK;
}

static int check_mult_args(fhea_ctx *c, u64 a0, u64 a1, u64 b0, u64 b1, u64 d0, u64 d1, u32 l) {
    if (l < 2 || l > c->h.L) return FHEA_ERR_ARG;
    const u64 W = (u64)l * c->h.N, W1 = (u64)(l - 1) * c->h.N;
    POLY_CHECK(c, a0, W);
    POLY_CHECK(c, a1, W);
    POLY_CHECK(c, b0, W);
    POLY_CHECK(c, b1, W);
    POLY_CHECK(c, d0, W1);
    POLY_CHECK(c, d1, W1);
    return FHEA_OK;
}

int fhea_eval_mult_relin_rescale(fhea_ctx *c, uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1,
                                 uint64_t d0, uint64_t d1, uint32_t l) {
    CTX_CHECK(c);
    int st = check_mult_args(c, a0, a1, b0, b1, d0, d1, l);
    if (st) return st;
    if (!c->has_relin) return FHEA_ERR_NO_KEY;
    u32 kl = c->relin_limbs ? c->relin_limbs : l;
    if (kl < l) return FHEA_ERR_ARG;
    u64 *P = c->poly;
    std::vector<MultItem> it = {{P + a0, P + a1, P + b0, P + b1, P + d0, P + d1}};
    return mult_relin_rescale(c, l, c->key + c->relin_off, kl, it);
}

int fhea_eval_mult_relin_rescale_noalloc(fhea_ctx *c, uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1,
                                         uint64_t d0, uint64_t d1, uint32_t l) {
    CTX_CHECK(c);
    if (c->capturing) return FHEA_ERR_STATE;
    int st = check_mult_args(c, a0, a1, b0, b1, d0, d1, l);
    if (st) return st;
    if (!c->has_relin) return FHEA_ERR_NO_KEY;
    u32 kl = c->relin_limbs ? c->relin_limbs : l;
    if (kl < l) return FHEA_ERR_ARG;
    if (mult_relin_rescale_words(c, l, 1) > FHEA_NOALLOC_SCRATCH_WORDS) return FHEA_ERR_SCRATCH;
    u64 *ptr = scratch_ptr(c->device);
    if (!ptr) return FHEA_ERR_STATE;
    const FixedWs ws{ptr, FHEA_NOALLOC_SCRATCH_WORDS};
    u64 *P = c->poly;
    std::vector<MultItem> it = {{P + a0, P + a1, P + b0, P + b1, P + d0, P + d1}};
    return mult_relin_rescale_lane(c, l, c->key + c->relin_off, kl, it, &ws);
}

int fhea_eval_mult_plain_rescale(fhea_ctx *c, uint64_t a0, uint64_t a1, uint64_t pt, uint64_t d0,
                                 uint64_t d1, uint32_t l) {
    CTX_CHECK(c);
    int st = check_mult_args(c, a0, a1, pt, pt, d0, d1, l);
    if (st) return st;
    const u64 W = (u64)l * c->h.N;
    if ((st = ensure_ws(c, 2 * W + rescale_words(c, l)))) return st;
    u64 *A = c->ws, *P = c->poly;
    launch_binop(c->dt, 2, P + a0, P + pt, A, l, c->stream);
    launch_binop(c->dt, 2, P + a1, P + pt, A + W, l, c->stream);
    std::vector<u64 *> outs = {P + d0, P + d1};
    st = rescale_core(c, l, 1, A, 2 * W, W, outs, A + 2 * W);
    if (st) return st;
    return check_hip(hipGetLastError());
}

int fhea_add_rotation_key(fhea_ctx *c, int32_t step, uint64_t galois, uint64_t evk, uint32_t key_limbs) {
    CTX_CHECK(c);
    u32 kl = key_limbs ? key_limbs : c->h.L;
    if (kl > c->h.L || (galois & 1) == 0 || galois >= 2ull * c->h.N) return FHEA_ERR_ARG;
    int st = key_check(c, evk, kl);
    if (st) return st;
    c->rot[step] = RotKey{step, galois, evk, kl};
    return FHEA_OK;
}

int fhea_eval_automorphism_ks(fhea_ctx *c, uint64_t galois, uint64_t evk, uint32_t key_limbs,
                              uint64_t in0, uint64_t in1, uint64_t d0, uint64_t d1, uint32_t l) {
    CTX_CHECK(c);
    if (l == 0 || l > c->h.L) return FHEA_ERR_ARG;
    if ((galois & 1) == 0 || galois >= 2ull * c->h.N) return FHEA_ERR_ARG;
    u32 kl = key_limbs ? key_limbs : l;
    if (kl < l || kl > c->h.L) return FHEA_ERR_ARG;
    const u64 W = (u64)l * c->h.N;
    POLY_CHECK(c, in0, W);
    POLY_CHECK(c, in1, W);
    POLY_CHECK(c, d0, W);
    POLY_CHECK(c, d1, W);
    int st = key_check(c, evk, kl);
    if (st) return st;
    u64 *P = c->poly;
    std::vector<AutoItem> it = {{P + in0, P + in1, P + d0, P + d1, galois, c->key + evk}};
    return automorphism_ks(c, l, kl, it);
}

int fhea_eval_rotate(fhea_ctx *c, uint64_t in0, uint64_t in1, uint64_t d0, uint64_t d1,
                     int32_t total_rot, uint32_t l) {
    CTX_CHECK(c);
    if (l == 0 || l > c->h.L) return FHEA_ERR_ARG;
    const u64 W = (u64)l * c->h.N;
    POLY_CHECK(c, in0, W);
    POLY_CHECK(c, in1, W);
    POLY_CHECK(c, d0, W);
    POLY_CHECK(c, d1, W);
    // decomposition and key lookup first, so a missing key leaves dst untouched
    std::vector<fhea::RotKey> steps;
    int64_t abs_rot = total_rot < 0 ? -(int64_t)total_rot : total_rot;
    int sign = total_rot > 0 ? 1 : -1;
    while (abs_rot > 0) {
        int64_t step_size = 1;
        while ((step_size << 1) <= abs_rot) step_size <<= 1;
        auto it = c->rot.find((int32_t)(sign * step_size));
        if (it == c->rot.end()) return FHEA_ERR_NO_KEY;
        if (it->second.key_limbs < l) return FHEA_ERR_ARG;
        steps.push_back(it->second);
        abs_rot -= step_size;
    }
    u64 *P = c->poly;
    if (steps.empty()) {
        int st = fhea_copy_poly(c, d0, in0, W);
        return st ? st : fhea_copy_poly(c, d1, in1, W);
    }
    // ping-pong ciphertext buffers live in a separate allocation from the
    // pipeline workspace so that a workspace resize cannot invalidate them
    u64 *pp = nullptr;
    if (steps.size() > 1) {
        if (c->pp_words < 4 * W) {
            if (c->capturing) return FHEA_ERR_NO_MEM;
            if (c->pp) {
                if (retire_buffer(c, c->pp)) return FHEA_ERR_DEVICE;
                c->pp = nullptr;
                c->pp_words = 0;
            }
            if (hipMalloc((void **)&c->pp, 4 * W * 8) != hipSuccess) return FHEA_ERR_NO_MEM;
            c->pp_words = 4 * W;
        }
        pp = c->pp;
    }
    const u64 *cur0 = P + in0, *cur1 = P + in1;
    int st = FHEA_OK;
    for (size_t k = 0; k < steps.size() && !st; ++k) {
        bool last = k + 1 == steps.size();
        u64 *n0 = last ? P + d0 : pp + (k & 1) * 2 * W;
        u64 *n1 = last ? P + d1 : n0 + W;
        std::vector<AutoItem> it = {{cur0, cur1, n0, n1, steps[k].galois, c->key + steps[k].evk_off}};
        st = automorphism_ks(c, l, steps[k].key_limbs, it);
        cur0 = n0;
        cur1 = n1;
    }
    return st;
}

int fhea_submit(fhea_ctx *c, const fhea_work *items, uint32_t n) {
    CTX_CHECK(c);
    if (n && !items) return FHEA_ERR_ARG;
    u64 *P = c->poly;
    // group consecutive items of the same kind / level / key level
    uint32_t i = 0;
    while (i < n) {
        const fhea_work &w0 = items[i];
        uint32_t j = i;
        while (j < n && items[j].kind == w0.kind && items[j].limbs == w0.limbs &&
               items[j].key_limbs == w0.key_limbs)
            ++j;
        const u32 l = w0.limbs;
        int st = FHEA_OK;
        if (w0.kind == FHEA_WORK_MULT_RELIN_RESCALE) {
            if (!c->has_relin) return FHEA_ERR_NO_KEY;
            u32 kl = c->relin_limbs ? c->relin_limbs : l;
            if (kl < l) return FHEA_ERR_ARG;
            std::vector<MultItem> v;
            std::vector<std::pair<u64, u32>> ins, outs;
            for (uint32_t k = i; k < j; ++k) {
                const fhea_work &w = items[k];
                if ((st = check_mult_args(c, w.a0, w.a1, w.b0, w.b1, w.d0, w.d1, l))) return st;
                v.push_back({P + w.a0, P + w.a1, P + w.b0, P + w.b1, P + w.d0, P + w.d1});
                for (u64 o : {w.a0, w.a1, w.b0, w.b1}) ins.push_back({o, k});
                outs.push_back({w.d0, k});
                outs.push_back({w.d1, k});
            }
            if (!batch_aliasing_ok(std::move(outs), ins, (u64)l * c->h.N, (u64)(l - 1) * c->h.N))
                return FHEA_ERR_ARG;
            st = mult_relin_rescale(c, l, c->key + c->relin_off, kl, v);
        } else if (w0.kind == FHEA_WORK_AUTOMORPHISM_KS) {
            if (l == 0 || l > c->h.L) return FHEA_ERR_ARG;
            u32 kl = w0.key_limbs ? w0.key_limbs : l;
            if (kl < l || kl > c->h.L) return FHEA_ERR_ARG;
            const u64 W = (u64)l * c->h.N;
            std::vector<AutoItem> v;
            std::vector<std::pair<u64, u32>> ins, outs;
            for (uint32_t k = i; k < j; ++k) {
                const fhea_work &w = items[k];
                POLY_CHECK(c, w.a0, W);
                POLY_CHECK(c, w.a1, W);
                POLY_CHECK(c, w.d0, W);
                POLY_CHECK(c, w.d1, W);
                if ((w.galois & 1) == 0 || w.galois >= 2ull * c->h.N) return FHEA_ERR_ARG;
                if ((st = key_check(c, w.evk, kl))) return st;
                v.push_back({P + w.a0, P + w.a1, P + w.d0, P + w.d1, w.galois, c->key + w.evk});
                ins.push_back({w.a0, k});
                ins.push_back({w.a1, k});
                outs.push_back({w.d0, k});
                outs.push_back({w.d1, k});
            }
            if (!batch_aliasing_ok(std::move(outs), ins, W, W)) return FHEA_ERR_ARG;
            st = automorphism_ks(c, l, kl, v);
        } else if (w0.kind == FHEA_WORK_MULT_RELIN) {
            if (!c->has_relin) return FHEA_ERR_NO_KEY;
            u32 kl = c->relin_limbs ? c->relin_limbs : l;
            if (kl < l) return FHEA_ERR_ARG;
            const u64 W = (u64)l * c->h.N;
            std::vector<MultItem> v;
            std::vector<std::pair<u64, u32>> ins, outs;
            for (uint32_t k = i; k < j; ++k) {
                const fhea_work &w = items[k];
                if (l == 0 || l > c->h.L) return FHEA_ERR_ARG;
                POLY_CHECK(c, w.a0, W);
                POLY_CHECK(c, w.a1, W);
                POLY_CHECK(c, w.b0, W);
                POLY_CHECK(c, w.b1, W);
                POLY_CHECK(c, w.d0, W);
                POLY_CHECK(c, w.d1, W);
                v.push_back({P + w.a0, P + w.a1, P + w.b0, P + w.b1, P + w.d0, P + w.d1});
                for (u64 o : {w.a0, w.a1, w.b0, w.b1}) ins.push_back({o, k});
                outs.push_back({w.d0, k});
                outs.push_back({w.d1, k});
            }
            if (!batch_aliasing_ok(std::move(outs), ins, W, W)) return FHEA_ERR_ARG;
            st = mult_relin(c, l, c->key + c->relin_off, kl, v);
        } else if (w0.kind == FHEA_WORK_ROTATE_HOISTED) {
            // runs of items with the same input share one ModUp; when no output of
            // the whole run overlaps an input of the run, all its inputs go through
            // one pass (ModUps first, rotations key by key)
            if (l == 0 || l > c->h.L) return FHEA_ERR_ARG;
            u32 kl = w0.key_limbs ? w0.key_limbs : l;
            if (kl < l || kl > c->h.L) return FHEA_ERR_ARG;
            const u64 W = (u64)l * c->h.N;
            std::vector<std::vector<RotItem>> groups;
            std::vector<RotInput> ins;
            std::vector<u64> in_offs, out_offs;
            uint32_t a = i;
            while (a < j) {
                uint32_t b = a;
                while (b < j && items[b].a0 == items[a].a0 && items[b].a1 == items[a].a1) ++b;
                POLY_CHECK(c, items[a].a0, W);
                POLY_CHECK(c, items[a].a1, W);
                std::vector<RotItem> v;
                for (uint32_t k = a; k < b; ++k) {
                    const fhea_work &w = items[k];
                    POLY_CHECK(c, w.d0, W);
                    POLY_CHECK(c, w.d1, W);
                    if ((w.galois & 1) == 0 || w.galois >= 2ull * c->h.N) return FHEA_ERR_ARG;
                    if ((st = key_check(c, w.evk, kl))) return st;
                    // outputs are written while later rotations still read the input
                    if (overlaps(w.d0, items[a].a0, W) || overlaps(w.d0, items[a].a1, W) ||
                        overlaps(w.d1, items[a].a0, W) || overlaps(w.d1, items[a].a1, W))
                        return FHEA_ERR_ARG;
                    v.push_back({P + w.d0, P + w.d1, w.galois, c->key + w.evk, (u32)groups.size()});
                    out_offs.push_back(w.d0);
                    out_offs.push_back(w.d1);
                }
                ins.push_back({P + items[a].a0, P + items[a].a1});
                in_offs.push_back(items[a].a0);
                in_offs.push_back(items[a].a1);
                groups.push_back(std::move(v));
                a = b;
            }
            std::sort(out_offs.begin(), out_offs.end());
            bool merge = groups.size() > 1;
            for (u64 x : in_offs) {
                auto it = std::lower_bound(out_offs.begin(), out_offs.end(), x >= W - 1 ? x - (W - 1) : 0);
                if (it != out_offs.end() && overlaps(x, *it, W)) merge = false;
            }
            if (merge) {
                std::vector<RotItem> all;
                for (auto &g : groups) all.insert(all.end(), g.begin(), g.end());
                if ((st = rotations_hoisted(c, l, kl, ins, std::move(all)))) return st;
            } else {
                for (size_t g = 0; g < groups.size(); ++g) {
                    for (RotItem &it : groups[g]) it.input = 0;
                    if ((st = rotations_hoisted(c, l, kl, {ins[g]}, std::move(groups[g])))) return st;
                }
            }
        } else if (w0.kind == FHEA_WORK_ADD) {
            if (l == 0 || l > c->h.L) return FHEA_ERR_ARG;
            const u64 W = (u64)l * c->h.N;
            std::vector<AddArgs> v;
            std::vector<std::pair<u64, u32>> ins, outs;
            for (uint32_t k = i; k < j; ++k) {
                const fhea_work &w = items[k];
                POLY_CHECK(c, w.a0, W);
                POLY_CHECK(c, w.a1, W);
                POLY_CHECK(c, w.b0, W);
                POLY_CHECK(c, w.b1, W);
                POLY_CHECK(c, w.d0, W);
                POLY_CHECK(c, w.d1, W);
                v.push_back({P + w.a0, P + w.a1, P + w.b0, P + w.b1, P + w.d0, P + w.d1});
                for (u64 o : {w.a0, w.a1, w.b0, w.b1}) ins.push_back({o, k});
                outs.push_back({w.d0, k});
                outs.push_back({w.d1, k});
            }
            if (!batch_aliasing_ok(std::move(outs), ins, W, W)) return FHEA_ERR_ARG;
            for (size_t o = 0; o < v.size(); o += 65535) {
                const u32 nb = (u32)std::min<size_t>(65535, v.size() - o);
                auto *ad = (const AddArgs *)stage_args(c, v.data() + o, nb * sizeof(AddArgs));
                if (!ad) return FHEA_ERR_DEVICE;
                launch_add_batch(c->dt, ad, nb, l, c->stream);
            }
            st = check_hip(hipGetLastError());
        } else if (w0.kind == FHEA_WORK_NTT || w0.kind == FHEA_WORK_INTT) {
            for (uint32_t k = i; k < j && !st; ++k)
                st = fhea_op(c, w0.kind == FHEA_WORK_NTT ? FHEA_OP_NTT : FHEA_OP_INTT, items[k].a0, 0,
                             items[k].d0, 0, l, 0);
        } else {
            return FHEA_ERR_ARG;
        }
        if (st) return st;
        i = j;
    }
    return FHEA_OK;
}

// Diagonal-method inner products over groups of queries sharing plaintext blocks
// (see fhea.h).  All units of all groups run in one launch per unit shape, so
// every output is checked against every input and every other output (intervals
// of W words: outputs sorted, every input looked up among them).
// Unit shape for a group of nq queries x ng blocks (FP64 kernel): qa tiles of 2
// queries stacked, 8/qa tiles of 8 blocks side by side (16 x 8, 8 x 16 or 4 x 32),
// the one that reads the fewest words: every unit reads its queries' two
// ciphertext streams and its blocks' plaintext stream once (ties: the taller).
static u32 lin_qa(u32 nq, u32 ng) {
    u32 best = 8;
    u64 best_streams = ~0ull;
    for (u32 qa : {8u, 4u, 2u}) {
        const u64 uq = 2 * qa, ug = 64 / qa;
        const u64 streams = 2ull * nq * ((ng + ug - 1) / ug) + (u64)ng * ((nq + uq - 1) / uq);
        if (streams < best_streams) {
            best_streams = streams;
            best = qa;
        }
    }
    return best;
}

// Groups of at most LIN_TILED_NQ queries go to the tile kernel (one wave per 2 x 8
// tile, eight tiles per workgroup): as LDS units most of their waves would idle.
static constexpr u32 LIN_TILED_NQ = 4;
static bool lin_tiled(u32 nq) { return nq <= LIN_TILED_NQ; }

struct LinGroupHost {
    u32 nq, ng;
    const u64 *ct0, *ct1, *bias0, *bias1;  // nq*R, nq*R, nq (or null), nq (or null)
    const u64 *pt, *ptb;                   // ng*R, ng (or null)
    const u64 *out0, *out1;                // nq*ng each, row-major by query
};

static int linear_sum_groups(fhea_ctx *c, u32 R, u32 limbs, const std::vector<LinGroupHost> &groups) {
    if (R == 0 || limbs == 0 || limbs > c->h.L || groups.empty()) return FHEA_ERR_ARG;
    const u64 W = (u64)limbs * c->h.N;
    const bool f64 = linear_sum_f64(c->dt);
    std::vector<u64> ins, outs;
    size_t words = 0, n4 = 0, nt = 0, n8 = 0;
    int st = FHEA_OK;
    auto in = [&](u64 o) {
        if (!st) st = poly_code(c, o, W);
        ins.push_back(o);
    };
    for (const LinGroupHost &g : groups) {
        if (g.nq == 0 || g.ng == 0) return FHEA_ERR_ARG;
        if ((g.bias0 == nullptr) != (g.bias1 == nullptr)) return FHEA_ERR_ARG;
        for (size_t k = 0; k < (size_t)g.nq * R; ++k) in(g.ct0[k]);
        for (size_t k = 0; k < (size_t)g.nq * R; ++k) in(g.ct1[k]);
        for (u32 k = 0; g.bias0 && k < g.nq; ++k) {
            if ((g.bias0[k] == FHEA_NONE) != (g.bias1[k] == FHEA_NONE)) return FHEA_ERR_ARG;
            if (g.bias0[k] != FHEA_NONE) {
                in(g.bias0[k]);
                in(g.bias1[k]);
            }
        }
        for (size_t k = 0; k < (size_t)g.ng * R; ++k) in(g.pt[k]);
        for (u32 k = 0; g.ptb && k < g.ng; ++k)
            if (g.ptb[k] != FHEA_NONE) in(g.ptb[k]);
        for (size_t k = 0; k < (size_t)g.nq * g.ng; ++k) {
            if (!st) st = poly_code(c, g.out0[k], W);
            if (!st) st = poly_code(c, g.out1[k], W);
            outs.push_back(g.out0[k]);
            outs.push_back(g.out1[k]);
        }
        words += 2 * (size_t)g.nq * (R + 1) + (size_t)g.ng * (R + 1) + 2 * (size_t)g.nq * g.ng;
        if (f64 && lin_tiled(g.nq)) {
            nt += (size_t)((g.nq + 1) / 2) * ((g.ng + 7) / 8);
        } else if (f64) {
            const u32 qa = lin_qa(g.nq, g.ng), uq = 2 * qa, ug = 64 / qa;
            n4 += (size_t)((g.nq + uq - 1) / uq) * ((g.ng + ug - 1) / ug);
        } else {
            n8 += (size_t)g.nq * ((g.ng + 7) / 8);
        }
    }
    if (st) return st;
    if (n4 > 65535 || n8 > 65535 || nt > 65535ull * 8) return FHEA_ERR_ARG;
    std::sort(outs.begin(), outs.end());
    for (size_t k = 0; k + 1 < outs.size(); ++k)
        if (outs[k + 1] - outs[k] < W) return FHEA_ERR_ARG;
    for (u64 x : ins) {
        auto it = std::lower_bound(outs.begin(), outs.end(), x >= W - 1 ? x - (W - 1) : 0);
        if (it != outs.end() && overlaps(x, *it, W)) return FHEA_ERR_ARG;
    }
    // staged: every group's offsets, then the LDS-shape units, the tiles, the 1x8 units
    const size_t bytes = words * 8 + (n4 + nt + n8) * sizeof(LinUnit);
    char *h = nullptr, *d = nullptr;
    if ((st = stage_reserve(c, bytes, &h, &d))) return st;
    u64 *hw = (u64 *)h;
    LinUnit *u4 = (LinUnit *)(h + words * 8), *ut = u4 + n4, *u8 = ut + nt;
    size_t w = 0;
    for (const LinGroupHost &g : groups) {
        const u64 *dev = (const u64 *)(d + w * 8);
        const size_t nqR = (size_t)g.nq * R, ngR = (size_t)g.ng * R, nqg = (size_t)g.nq * g.ng;
        u64 *x = hw + w;
        memcpy(x, g.ct0, nqR * 8);
        memcpy(x + nqR, g.ct1, nqR * 8);
        x += 2 * nqR;
        for (u32 k = 0; k < g.nq; ++k) {
            x[k] = g.bias0 ? g.bias0[k] : FHEA_NONE;
            x[g.nq + k] = g.bias1 ? g.bias1[k] : FHEA_NONE;
        }
        x += 2 * (size_t)g.nq;
        memcpy(x, g.pt, ngR * 8);
        x += ngR;
        for (u32 k = 0; k < g.ng; ++k) x[k] = g.ptb ? g.ptb[k] : FHEA_NONE;
        x += g.ng;
        memcpy(x, g.out0, nqg * 8);
        memcpy(x + nqg, g.out1, nqg * 8);
        w += 2 * nqR + 2 * (size_t)g.nq + ngR + g.ng + 2 * nqg;
        if (f64 && lin_tiled(g.nq)) {
            for (u32 q0 = 0; q0 < g.nq; q0 += 2)
                for (u32 g0 = 0; g0 < g.ng; g0 += 8)
                    *ut++ = LinUnit{dev, g.nq, g.ng, q0, g0, std::min(2u, g.nq - q0), std::min(8u, g.ng - g0), 0, 0};
        } else if (f64) {
            const u32 qa = lin_qa(g.nq, g.ng), uq = 2 * qa, ug = 64 / qa;
            for (u32 q0 = 0; q0 < g.nq; q0 += uq)
                for (u32 g0 = 0; g0 < g.ng; g0 += ug)
                    *u4++ = LinUnit{dev, g.nq, g.ng, q0, g0, std::min(uq, g.nq - q0), std::min(ug, g.ng - g0), qa,
                                    0};
        } else {
            for (u32 q0 = 0; q0 < g.nq; ++q0)
                for (u32 g0 = 0; g0 < g.ng; g0 += 8)
                    *u8++ = LinUnit{dev, g.nq, g.ng, q0, g0, 1, std::min(8u, g.ng - g0), 1, 0};
        }
    }
    if ((st = stage_commit(c, h, d, bytes))) return st;
    const LinUnit *d4 = (const LinUnit *)(d + words * 8);
    launch_linear_sum(c->dt, c->poly, d4, (u32)n4, R, limbs, LIN_SHAPE_LDS, c->stream);
    launch_linear_sum(c->dt, c->poly, d4 + n4, (u32)nt, R, limbs, LIN_SHAPE_TILES, c->stream);
    launch_linear_sum(c->dt, c->poly, d4 + n4 + nt, (u32)n8, R, limbs, LIN_SHAPE_1x8, c->stream);
    return check_hip(hipGetLastError());
}

int fhea_eval_linear_sum_groups(fhea_ctx *c, uint32_t n_groups, uint32_t R, const uint32_t *nq,
                                const uint32_t *ng, const uint64_t *ct0, const uint64_t *ct1,
                                const uint64_t *bias0, const uint64_t *bias1, const uint64_t *pt,
                                const uint64_t *ptb, const uint64_t *out0, const uint64_t *out1, uint32_t limbs) {
    CTX_CHECK(c);
    if (!nq || !ng || !ct0 || !ct1 || !pt || !out0 || !out1 || n_groups == 0) return FHEA_ERR_ARG;
    if ((bias0 == nullptr) != (bias1 == nullptr)) return FHEA_ERR_ARG;
    std::vector<LinGroupHost> gs(n_groups);
    size_t aq = 0, ag = 0, ao = 0;
    for (u32 k = 0; k < n_groups; ++k) {
        gs[k] = {nq[k], ng[k], ct0 + aq * R, ct1 + aq * R, bias0 ? bias0 + aq : nullptr,
                 bias1 ? bias1 + aq : nullptr, pt + ag * R, ptb ? ptb + ag : nullptr, out0 + ao, out1 + ao};
        aq += nq[k];
        ag += ng[k];
        ao += (size_t)nq[k] * ng[k];
    }
    return linear_sum_groups(c, R, limbs, gs);
}

int fhea_eval_linear_sum_batch(fhea_ctx *c, uint32_t n_jobs, uint32_t R, const uint64_t *ct0,
                               const uint64_t *ct1, const uint32_t *G, const uint64_t *pt,
                               const uint64_t *ptb, const uint64_t *bias0, const uint64_t *bias1,
                               const uint64_t *out0, const uint64_t *out1, uint32_t limbs) {
    CTX_CHECK(c);
    if (!ct0 || !ct1 || !G || !pt || !out0 || !out1 || n_jobs == 0) return FHEA_ERR_ARG;
    if ((bias0 == nullptr) != (bias1 == nullptr)) return FHEA_ERR_ARG;
    std::vector<LinGroupHost> gs(n_jobs);
    size_t ag = 0;
    for (u32 j = 0; j < n_jobs; ++j) {
        gs[j] = {1, G[j], ct0 + (size_t)j * R, ct1 + (size_t)j * R, bias0 ? bias0 + j : nullptr,
                 bias1 ? bias1 + j : nullptr, pt + ag * R, ptb ? ptb + ag : nullptr, out0 + ag, out1 + ag};
        ag += G[j];
    }
    return linear_sum_groups(c, R, limbs, gs);
}

int fhea_eval_linear_sum(fhea_ctx *c, const uint64_t *ct0, const uint64_t *ct1, uint32_t R,
                         const uint64_t *pt, const uint64_t *ptb, uint32_t G, uint64_t bias0,
                         uint64_t bias1, const uint64_t *out0, const uint64_t *out1, uint32_t limbs) {
    CTX_CHECK(c);
    if ((bias0 == FHEA_NONE) != (bias1 == FHEA_NONE)) return FHEA_ERR_ARG;
    return fhea_eval_linear_sum_batch(c, 1, R, ct0, ct1, &G, pt, ptb, &bias0, &bias1, out0, out1, limbs);
}

int fhea_rescale_scatter(fhea_ctx *c, uint64_t src, const uint64_t *dst, uint32_t n, uint32_t limbs) {
    CTX_CHECK(c);
    if (!dst || limbs < 2 || limbs > c->h.L || n == 0) return FHEA_ERR_ARG;
    const u64 N = c->h.N, W = (u64)limbs * N, W1 = (u64)(limbs - 1) * N;
    POLY_CHECK(c, src, 2 * W * n);
    std::vector<u64> d(dst, dst + n);
    for (u64 o : d) POLY_CHECK(c, o, 2 * W1);
    std::sort(d.begin(), d.end());
    for (u32 z = 0; z < n; ++z) {
        if (z + 1 < n && d[z + 1] - d[z] < 2 * W1) return FHEA_ERR_ARG;
        if (src < d[z] + 2 * W1 && d[z] < src + 2 * W * n) return FHEA_ERR_ARG;
    }
    // as many items per launch as 1 GiB of workspace holds (at most 1024): at small N
    // and few limbs, batches of 64 left most of the device idle in each transform
    const u32 cap = (u32)std::max<u64>(1, std::min<u64>(1024, (1ull << 27) / rescale_words(c, limbs)));
    for (u32 off = 0; off < n; off += cap) {
        const u32 nb = std::min<u32>(cap, n - off);
        int st = ensure_ws(c, nb * rescale_words(c, limbs));
        if (st) return st;
        std::vector<u64 *> outs(2 * nb);
        for (u32 z = 0; z < nb; ++z) {
            outs[2 * z] = c->poly + dst[off + z];
            outs[2 * z + 1] = outs[2 * z] + W1;
        }
        st = rescale_core(c, limbs, nb, c->poly + src + (u64)off * 2 * W, 2 * W, W, outs, c->ws);
        if (st) return st;
    }
    return check_hip(hipGetLastError());
}

int fhea_rescale_batch(fhea_ctx *c, uint64_t src, uint64_t dst, uint32_t n, uint32_t limbs) {
    CTX_CHECK(c);
    if (limbs < 2 || limbs > c->h.L || n == 0) return FHEA_ERR_ARG;
    const u64 W1 = (u64)(limbs - 1) * c->h.N;
    POLY_CHECK(c, dst, 2 * W1 * n);
    std::vector<u64> d(n);
    for (u32 z = 0; z < n; ++z) d[z] = dst + (u64)z * 2 * W1;
    return fhea_rescale_scatter(c, src, d.data(), n, limbs);
}

int fhea_ntt_batch(fhea_ctx *c, uint64_t src, uint64_t dst, uint32_t npolys, uint32_t limbs,
                   int inverse) {
    CTX_CHECK(c);
    if (limbs == 0 || limbs > c->h.L || npolys == 0) return FHEA_ERR_ARG;
    const u64 W = (u64)limbs * c->h.N;
    POLY_CHECK(c, src, W * npolys);
    POLY_CHECK(c, dst, W * npolys);
    launch_ntt(c->dt, inverse != 0, c->poly + src, c->poly + dst, W, W, npolys, limbs, c->dt.ident,
               c->stream);
    return check_hip(hipGetLastError());
}

int fhea_sync(fhea_ctx *c) {
    CTX_CHECK(c);
    if (c->capturing) return FHEA_ERR_STATE;
    int st = check_hip(hipStreamSynchronize(c->stream));
    if (!st) c->args_used = 0;
    return st;
}

int fhea_timer_start(fhea_ctx *c) {
    CTX_CHECK(c);
    if (c->capturing) return FHEA_ERR_STATE;
    return check_hip(hipEventRecord(c->ev0, c->stream));
}

int fhea_timer_stop(fhea_ctx *c, float *ms) {
    CTX_CHECK(c);
    if (c->capturing) return FHEA_ERR_STATE;
    if (hipEventRecord(c->ev1, c->stream) != hipSuccess) return FHEA_ERR_DEVICE;
    if (hipEventSynchronize(c->ev1) != hipSuccess) return FHEA_ERR_DEVICE;
    float v = 0;
    if (hipEventElapsedTime(&v, c->ev0, c->ev1) != hipSuccess) return FHEA_ERR_DEVICE;
    if (ms) *ms = v;
    return FHEA_OK;
}

// ---- HIP graph capture of engine work
int fhea_graph_begin(fhea_ctx *c) {
    CTX_CHECK(c);
    if (c->capturing) return FHEA_ERR_STATE;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return FHEA_ERR_DEVICE;
    c->args_used = 0;
    if (!c->cap_dev && hipMalloc(&c->cap_dev, fhea_ctx::CAP_BYTES) != hipSuccess) return FHEA_ERR_NO_MEM;
    if (!c->cap_host && hipHostMalloc(&c->cap_host, fhea_ctx::CAP_BYTES, hipHostMallocDefault) != hipSuccess)
        return FHEA_ERR_NO_MEM;
    c->cap_used = 0;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return FHEA_ERR_DEVICE;
    c->capturing = true;
    return FHEA_OK;
}

int fhea_graph_end(fhea_ctx *c, uint32_t *graph_id) {
    CTX_CHECK(c);
    if (!c->capturing) return FHEA_ERR_STATE;
    c->capturing = false;
    fhea_ctx::Graph g;
    int st = FHEA_OK;
    if (hipStreamEndCapture(c->stream, &g.graph) != hipSuccess || !g.graph) st = FHEA_ERR_DEVICE;
    if (!st && hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0) != hipSuccess) st = FHEA_ERR_DEVICE;
    if (st) {
        if (g.graph) hipGraphDestroy(g.graph);
        return st;  // the argument area stays with the context for the next capture
    }
    // the argument area now belongs to the graph
    g.dev = c->cap_dev;
    g.host = c->cap_host;
    c->cap_dev = c->cap_host = nullptr;
    const uint32_t id = c->next_graph++;
    c->graphs[id] = std::move(g);
    if (graph_id) *graph_id = id;
    return FHEA_OK;
}

int fhea_graph_launch(fhea_ctx *c, uint32_t graph_id) {
    CTX_CHECK(c);
    if (c->capturing) return FHEA_ERR_STATE;
    auto it = c->graphs.find(graph_id);
    if (it == c->graphs.end()) return FHEA_ERR_ARG;
    return check_hip(hipGraphLaunch(it->second.exec, c->stream));
}

int fhea_graph_destroy(fhea_ctx *c, uint32_t graph_id) {
    CTX_CHECK(c);
    if (c->capturing) return FHEA_ERR_STATE;
    auto it = c->graphs.find(graph_id);
    if (it == c->graphs.end()) return FHEA_ERR_ARG;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return FHEA_ERR_DEVICE;
    hipGraphExecDestroy(it->second.exec);
    hipGraphDestroy(it->second.graph);
    hipFree(it->second.dev);
    hipHostFree(it->second.host);
    c->graphs.erase(it);
    if (c->graphs.empty()) {
        // no recorded kernel addresses the replaced workspaces any more
        for (void *p : c->retired) hipFree(p);
        c->retired.clear();
    }
    return FHEA_OK;
}

}  // extern "C"
