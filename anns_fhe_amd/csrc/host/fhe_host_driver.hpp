// fhe_host_driver.hpp -- C++ FHEHostDriver over the libfhea C-ABI (header only).
//
// Keeps the reference class's names, argument order and meaning
// (cpp_host_model/src/fhe_host_driver.h:55-123) so that callers such as
// InteractiveQueryManager (cpp_host_model/src/interactive_query_manager.cpp:85-274)
// and main_sim (cpp_host_model/src/main_sim.cpp:33-45) compile against it
// unchanged: FheDevice takes the place of SimMemoryBus (two word-addressed HBM
// arenas, poly_gmem and key_gmem, of the same default sizes, with the same
// read/write/load/save calls), and FHEHostDriver::initialize_context_from_directory
// reads the same test-vector directory (fhe_host_driver.cpp:18-175).  Errors are
// thrown as std::runtime_error / std::out_of_range like the reference
// (fhe_host_driver.cpp:336-339, sim_memory_bus.cpp:13-15).
#pragma once
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/fhea.h"

namespace fhea_host {

inline void throw_on(int st, const char *what) {
    if (st == FHEA_OK) return;
    std::string msg = std::string(what) + ": " + fhea_status_string(st);
    if (st == FHEA_ERR_RANGE) throw std::out_of_range(msg);
    throw std::runtime_error(msg);
}

// Whole little-endian file of T words; empty if the file is absent.
template <class T>
inline std::vector<T> read_words(const std::string &path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return {};
    const std::streamsize bytes = f.tellg();
    if (bytes < 0 || bytes % (std::streamsize)sizeof(T)) throw std::runtime_error("bad file size: " + path);
    std::vector<T> v((size_t)bytes / sizeof(T));
    f.seekg(0);
    f.read(reinterpret_cast<char *>(v.data()), bytes);
    if (!f) throw std::runtime_error("cannot read " + path);
    return v;
}

inline bool file_exists(const std::string &path) { return (bool)std::ifstream(path); }

// Owns one context and its two HBM arenas; plays the role of SimMemoryBus
// (sim_memory_bus.h:10-34).  Constructed from a parameter block, the context
// exists at once; constructed from arena sizes only (the SimMemoryBus defaults:
// 8M poly words, 64M key words), the context is created by
// FHEHostDriver::initialize_context_from_directory from the directory's primes.
class FheDevice {
public:
    explicit FheDevice(size_t poly_words = 8388608, size_t key_words = 67108864)
        : poly_words_(poly_words), key_words_(key_words) {}
    explicit FheDevice(const fhea_params &p) { create(p); }
    ~FheDevice() {
        if (ctx_) fhea_ctx_destroy(ctx_);
    }
    FheDevice(const FheDevice &) = delete;
    FheDevice &operator=(const FheDevice &) = delete;

    // Creates the context (once); arena sizes of p default to this device's.
    void create(fhea_params p) {
        if (ctx_) throw std::runtime_error("FheDevice: context already created");
        if (!p.poly_words) p.poly_words = poly_words_;
        if (!p.key_words) p.key_words = key_words_;
        throw_on(fhea_ctx_create(&p, &ctx_), "fhea_ctx_create");
        poly_words_ = fhea_poly_words(ctx_);
        key_words_ = fhea_key_words(ctx_);
    }
    bool has_context() const { return ctx_ != nullptr; }
    fhea_ctx *ctx() const {
        if (!ctx_) throw std::runtime_error("FheDevice: no context (initialize_context_from_directory first)");
        return ctx_;
    }
    uint64_t *poly_gmem_ptr() { return fhea_poly_arena(ctx()); }
    uint64_t *key_gmem_ptr() { return fhea_key_arena(ctx()); }
    void write_poly(uint64_t off, const uint64_t *d, size_t n) { throw_on(fhea_write_poly(ctx(), off, d, n), "write_poly"); }
    void read_poly(uint64_t off, uint64_t *d, size_t n) const { throw_on(fhea_read_poly(ctx(), off, d, n), "read_poly"); }
    void write_key(uint64_t off, const uint64_t *d, size_t n) { throw_on(fhea_write_key(ctx(), off, d, n), "write_key"); }
    void read_key(uint64_t off, uint64_t *d, size_t n) const { throw_on(fhea_read_key(ctx(), off, d, n), "read_key"); }
    // SimMemoryBus::load_poly_file / load_key_file / save_poly_file (sim_memory_bus.cpp:36-64):
    // the whole file (or exactly expected_words of it when given) at the word offset.
    void load_poly_file(uint64_t off, const std::string &path, size_t expected_words = 0) {
        const std::vector<uint64_t> v = load_file(path, expected_words);
        write_poly(off, v.data(), v.size());
    }
    void load_key_file(uint64_t off, const std::string &path, size_t expected_words = 0) {
        const std::vector<uint64_t> v = load_file(path, expected_words);
        write_key(off, v.data(), v.size());
    }
    void save_poly_file(uint64_t off, const std::string &path, size_t count) const {
        std::vector<uint64_t> v(count);
        read_poly(off, v.data(), count);
        std::ofstream f(path, std::ios::binary);
        if (!f) throw std::runtime_error("cannot write " + path);
        f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(count * 8));
    }
    size_t poly_size_words() const { return poly_words_; }
    size_t key_size_words() const { return key_words_; }

private:
    static std::vector<uint64_t> load_file(const std::string &path, size_t expected_words) {
        std::vector<uint64_t> v = read_words<uint64_t>(path);
        if (v.empty() && !file_exists(path)) throw std::runtime_error("cannot open " + path);
        if (expected_words && v.size() < expected_words) throw std::runtime_error("short file: " + path);
        if (expected_words) v.resize(expected_words);
        return v;
    }
    fhea_ctx *ctx_ = nullptr;
    size_t poly_words_ = 8388608, key_words_ = 67108864;
};

// RotationKeyMeta (fhe_host_driver.h:10-15): step, the Galois element as the key
// directory states it (auto_idx), where the packed map sits in poly_gmem, where the
// key sits in key_gmem.  `galois` is the element whose closed-form permutation the
// engine applies (auto_idx itself, or its inverse when the directory's map file is
// written in the opposite convention), key_limbs the stored Q limbs of the key.
struct RotationKeyMeta {
    int32_t step;
    uint32_t auto_idx;
    uint32_t auto_map_offset;  // ~0u: no map stored (the engine needs none)
    uint64_t evk_offset;
    uint64_t galois;
    uint32_t key_limbs;
};

// FHEContextConfig (fhe_host_driver.h:17-53): the parameters a loaded directory
// fixes.  The reference's table offsets into key_gmem have no counterpart: the
// engine keeps its tables for every level beside the arenas.
struct FHEContextConfig {
    uint32_t N = 0, sizeQl = 0, sizeP = 0, numPartQ = 0, alpha = 0;
    std::vector<uint64_t> rns_primes, p_primes, n_inv_q, n_inv_p;
    uint64_t relin_key_offset = 0;
    uint32_t relin_key_limbs = 0;
};

class FHEHostDriver {
public:
    // Scratch slots of poly_gmem (fhe_host_driver.h:63-69); maps of the loaded
    // rotation keys go at SCRATCH_AUTO_MAP + k*N as in the reference.
    static constexpr uint32_t SCRATCH_D0 = 0x00400000;
    static constexpr uint32_t SCRATCH_D1 = 0x00440000;
    static constexpr uint32_t SCRATCH_D2 = 0x00480000;
    static constexpr uint32_t SCRATCH_KS_C0 = 0x004C0000;
    static constexpr uint32_t SCRATCH_KS_C1 = 0x00500000;
    static constexpr uint32_t SCRATCH_TMP = 0x00540000;
    static constexpr uint32_t SCRATCH_AUTO_MAP = 0x00580000;

    explicit FHEHostDriver(FheDevice *dev) : m_dev(dev) {
        if (dev->has_context()) bind();
    }

    uint32_t N = 0;

    // ---- context & keys (fhe_host_driver.cpp:18-175, 206-208)
    //
    // Reads the directory the reference's generator writes
    // (integration_tools/generate_interactive_test_vectors.cpp:246-326):
    //   rns_primes_Q.bin, rns_primes_P.bin  -> L, K and the primes;
    //   twiddles.bin / inv_twiddles.bin     -> N ((L+K)*N words) and each prime's root
    //                                          psi = twiddles[i*N + N/2];
    //   barrett_constants.bin, n_inv.bin    -> checked against their definitions;
    //   evalkey_mult_{a,b}_{d}.bin          -> dnum (files present) and the relin key
    //                                          ([a digits | b digits], each key_limbs+K limbs);
    //   rotation_keys_meta.txt, rotkey_step{s}_{a,b}_{d}.bin, auto_map_step{s}.bin
    //                                       -> the rotation-key directory.
    // The context is created from those primes and roots (or, if the device already
    // has one, must have the same primes), so the engine's twiddles are the files'
    // (checked word for word).  Keys go into key_gmem at 512-word-aligned offsets,
    // relin key first, as the reference lays them out.  A map file must be the
    // closed-form sigma_g map of auto_idx or of its inverse; the engine then applies
    // exactly that permutation (out[i] = in[map[i]], automorphism.cpp:3-12).
    // dnum = 0 takes it from the number of evalkey_mult_a_{d}.bin files.
    void initialize_context_from_directory(const std::string &dir, uint32_t dnum = 0) {
        const std::vector<uint64_t> q = read_words<uint64_t>(dir + "/rns_primes_Q.bin");
        const std::vector<uint64_t> p = read_words<uint64_t>(dir + "/rns_primes_P.bin");
        if (q.empty() || p.empty()) throw std::runtime_error("no rns_primes_Q.bin / rns_primes_P.bin in " + dir);
        const size_t L = q.size(), K = p.size(), T = L + K;
        std::vector<uint64_t> primes(q);
        primes.insert(primes.end(), p.begin(), p.end());
        const std::vector<uint64_t> tw = read_words<uint64_t>(dir + "/twiddles.bin");
        const std::vector<uint64_t> itw = read_words<uint64_t>(dir + "/inv_twiddles.bin");
        if (tw.empty() || tw.size() % T || itw.size() != tw.size())
            throw std::runtime_error("twiddles.bin / inv_twiddles.bin do not hold (L+K)*N words");
        const uint32_t n = (uint32_t)(tw.size() / T);
        std::vector<uint64_t> psi(T);
        for (size_t i = 0; i < T; ++i) psi[i] = tw[i * n + n / 2];  // tw[br(1)] = psi
        check_barrett(dir, primes);
        const std::vector<uint64_t> ninv = read_words<uint64_t>(dir + "/n_inv.bin");
        if (!ninv.empty()) {
            if (ninv.size() < T) throw std::runtime_error("n_inv.bin is short");
            for (size_t i = 0; i < T; ++i)
                if (mulmod(ninv[i], n, primes[i]) != 1) throw std::runtime_error("n_inv.bin: wrong N^-1");
        }
        if (dnum == 0)
            while (file_exists(dir + "/evalkey_mult_a_" + std::to_string(dnum) + ".bin")) ++dnum;
        if (dnum == 0) dnum = 1;
        if (!m_dev->has_context()) {
            fhea_params prm{};
            prm.N = n;
            prm.L = (uint32_t)L;
            prm.K = (uint32_t)K;
            prm.dnum = dnum;
            prm.q_primes = q.data();
            prm.p_primes = p.data();
            prm.psi = psi.data();
            m_dev->create(prm);
        }
        bind();
        if (N != n || m_cfg.sizeQl != L || m_cfg.sizeP != K || m_primes != primes)
            throw std::runtime_error("the device's context has other parameters than " + dir);
        {
            std::vector<uint64_t> mine(T * (size_t)n);
            throw_on(fhea_ctx_table(m_dev->ctx(), 2, mine.data(), mine.size()), "fhea_ctx_table");
            if (mine != tw) throw std::runtime_error("engine twiddles differ from twiddles.bin");
            throw_on(fhea_ctx_table(m_dev->ctx(), 3, mine.data(), mine.size()), "fhea_ctx_table");
            if (mine != itw) throw std::runtime_error("engine inverse twiddles differ from inv_twiddles.bin");
        }
        // relinearisation key (fhe_host_driver.cpp:105-117)
        uint64_t cur = 0;
        const uint32_t relin_limbs = load_switch_key(dir, "evalkey_mult", cur);
        if (relin_limbs) {
            set_relin_key(cur, relin_limbs);
            m_cfg.relin_key_offset = cur;
            m_cfg.relin_key_limbs = relin_limbs;
            cur += fhea_key_size_words(m_dev->ctx(), relin_limbs);
        }
        // rotation keys (fhe_host_driver.cpp:120-172)
        std::ifstream meta(dir + "/rotation_keys_meta.txt");
        std::string line;
        uint64_t map_off = SCRATCH_AUTO_MAP;
        while (meta && std::getline(meta, line)) {
            int step = 0;
            unsigned long long idx = 0;
            if (std::sscanf(line.c_str(), " step=%d auto_idx=%llu", &step, &idx) != 2) continue;
            const std::string stem = "rotkey_step" + std::to_string(step);
            cur = (cur + 511) & ~511ull;
            const uint32_t kl = load_switch_key(dir, stem, cur);
            if (!kl) throw std::runtime_error("rotation key files missing for step " + std::to_string(step));
            uint64_t galois = idx;
            uint32_t moff = ~0u;
            const std::vector<uint32_t> amap = read_words<uint32_t>(dir + "/auto_map_step" + std::to_string(step) + ".bin");
            if (!amap.empty()) {
                if (amap.size() != n) throw std::runtime_error("auto_map_step file is not N words");
                galois = map_galois(amap, idx);
                if (map_off + n <= m_dev->poly_size_words()) {
                    std::vector<uint64_t> packed(n / 2);
                    for (uint32_t i = 0; i < n / 2; ++i)
                        packed[i] = (uint64_t)amap[2 * i] | ((uint64_t)amap[2 * i + 1] << 32);
                    m_dev->write_poly(map_off, packed.data(), packed.size());
                    moff = (uint32_t)map_off;
                    map_off += n;
                }
            }
            add_rotation_key(step, galois, cur, kl);
            m_rot[step].auto_idx = (uint32_t)idx;
            m_rot[step].auto_map_offset = moff;
            cur += fhea_key_size_words(m_dev->ctx(), kl);
        }
    }
    void load_context_to_hardware() { dispatch(FHEA_OP_LOAD_CONTEXT, 0, 0, 0, 0, m_cfg.sizeQl); }
    void set_relin_key(uint64_t evk_off, uint32_t key_limbs) {
        throw_on(fhea_set_relin_key(m_dev->ctx(), evk_off, key_limbs), "set_relin_key");
    }
    void add_rotation_key(int32_t step, uint64_t galois, uint64_t evk_off, uint32_t key_limbs) {
        throw_on(fhea_add_rotation_key(m_dev->ctx(), step, galois, evk_off, key_limbs), "add_rotation_key");
        m_rot[step] = RotationKeyMeta{step, (uint32_t)galois, ~0u, evk_off, galois, key_limbs};
    }
    const std::map<int32_t, RotationKeyMeta> &rotation_keys() const { return m_rot; }
    const FHEContextConfig &get_config() const { return m_cfg; }

    // ---- primitive wrappers (fhe_host_driver.cpp:210-242)
    void ntt(uint64_t src, uint64_t dst, uint32_t limbs) { dispatch(FHEA_OP_NTT, src, 0, dst, 0, limbs); }
    void intt(uint64_t src, uint64_t dst, uint32_t limbs) { dispatch(FHEA_OP_INTT, src, 0, dst, 0, limbs); }
    void poly_add(uint64_t a, uint64_t b, uint64_t dst, uint32_t limbs) { dispatch(FHEA_OP_POLY_ADD, a, b, dst, 0, limbs); }
    void poly_sub(uint64_t a, uint64_t b, uint64_t dst, uint32_t limbs) { dispatch(FHEA_OP_POLY_SUB, a, b, dst, 0, limbs); }
    void poly_mul_ntt(uint64_t a, uint64_t b, uint64_t dst, uint32_t limbs) {
        dispatch(FHEA_OP_POLY_MUL_NTT, a, b, dst, 0, limbs);
    }
    void automorphism(uint64_t src, uint64_t map, uint64_t dst, uint32_t limbs, uint32_t /*galois_elt*/) {
        dispatch(FHEA_OP_AUTOMORPHISM, src, map, dst, 0, limbs);
    }
    void rescale(uint64_t c0, uint64_t c1, uint64_t out_c0, uint64_t /*out_c1*/, uint32_t limbs) {
        dispatch(FHEA_OP_RESCALE, c0, c1, out_c0, 0, limbs);
    }
    void key_switch(uint64_t c1, uint64_t evk, uint64_t dst_c0, uint64_t /*dst_c1*/, uint32_t sizeQl) {
        dispatch(FHEA_OP_KEY_SWITCH, c1, 0, dst_c0, evk, sizeQl);
    }

    // ---- composites (fhe_host_driver.cpp:248-404)
    void eval_add(uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1, uint64_t d0, uint64_t d1, uint32_t limbs) {
        throw_on(fhea_eval_add(m_dev->ctx(), a0, a1, b0, b1, d0, d1, limbs), "eval_add");
    }
    void eval_sub_plain(uint64_t a0, uint64_t a1, uint64_t pt, uint64_t d0, uint64_t d1, uint32_t limbs) {
        throw_on(fhea_eval_sub_plain(m_dev->ctx(), a0, a1, pt, d0, d1, limbs), "eval_sub_plain");
    }
    void eval_mult_relin_rescale(uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1, uint64_t d0, uint64_t d1,
                                 uint32_t in_limbs) {
        throw_on(fhea_eval_mult_relin_rescale(m_dev->ctx(), a0, a1, b0, b1, d0, d1, in_limbs),
                 "eval_mult_relin_rescale");
    }
    void eval_mult_plain_rescale(uint64_t a0, uint64_t a1, uint64_t pt, uint64_t d0, uint64_t d1,
                                 uint32_t in_limbs) {
        throw_on(fhea_eval_mult_plain_rescale(m_dev->ctx(), a0, a1, pt, d0, d1, in_limbs),
                 "eval_mult_plain_rescale");
    }
    void eval_rotate_single_step(uint64_t in0, uint64_t in1, uint64_t d0, uint64_t d1, int32_t step,
                                 uint32_t limbs) {
        auto it = m_rot.find(step);
        if (it == m_rot.end()) throw std::runtime_error("Rotation key not found for step: " + std::to_string(step));
        throw_on(fhea_eval_automorphism_ks(m_dev->ctx(), it->second.galois, it->second.evk_offset,
                                           it->second.key_limbs, in0, in1, d0, d1, limbs),
                 "eval_rotate_single_step");
    }
    void eval_rotate(uint64_t in0, uint64_t in1, uint64_t d0, uint64_t d1, int32_t total_rot, uint32_t limbs) {
        throw_on(fhea_eval_rotate(m_dev->ctx(), in0, in1, d0, d1, total_rot, limbs), "eval_rotate");
    }

    void sync() { throw_on(fhea_sync(m_dev->ctx()), "fhea_sync"); }
    FheDevice *get_bus() { return m_dev; }

private:
    static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) {
        return (uint64_t)((unsigned __int128)a * b % m);
    }
    static uint64_t powmod(uint64_t a, uint64_t e, uint64_t m) {
        uint64_t r = 1 % m;
        for (a %= m; e; e >>= 1, a = mulmod(a, a, m))
            if (e & 1) r = mulmod(r, a, m);
        return r;
    }
    static uint32_t bitrev(uint32_t x, uint32_t bits) {
        uint32_t r = 0;
        for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
        return r;
    }
    // The element g' in {g, g^-1 mod 2N} whose closed-form evaluation-domain map
    // br(((g' (2 br(i) + 1)) mod 2N - 1) / 2) equals the file's map.
    uint64_t map_galois(const std::vector<uint32_t> &amap, uint64_t g) const {
        const uint64_t two_n = 2ull * N;
        uint32_t bits = 0;
        while ((1u << bits) < N) ++bits;
        uint64_t ginv = powmod(g, N - 1, two_n);  // the odd residues mod 2N form a group of order N
        for (uint64_t cand : {g, ginv}) {
            bool same = true;
            for (uint32_t i = 0; i < N && same; ++i) {
                const uint64_t e = (cand * (2ull * bitrev(i, bits) + 1)) % two_n;
                same = amap[i] == bitrev((uint32_t)((e - 1) / 2), bits);
            }
            if (same) return cand;
        }
        throw std::runtime_error("auto_map file is not the sigma_g map of its auto_idx or of its inverse");
    }
    // barrett_constants.bin: (L+K) unpadded 20-byte records {m_lo, m_hi, k} with
    // k = 2*bitlen(q) + 3, m = floor(2^k / q) (generate_interactive_test_vectors.cpp:248-254)
    static void check_barrett(const std::string &dir, const std::vector<uint64_t> &primes) {
        std::ifstream f(dir + "/barrett_constants.bin", std::ios::binary);
        if (!f) return;
        for (uint64_t qv : primes) {
            uint64_t lo = 0, hi = 0;
            uint32_t k = 0;
            f.read(reinterpret_cast<char *>(&lo), 8);
            f.read(reinterpret_cast<char *>(&hi), 8);
            f.read(reinterpret_cast<char *>(&k), 4);
            if (!f) throw std::runtime_error("barrett_constants.bin is short");
            const uint32_t bits = 64 - (uint32_t)__builtin_clzll(qv);
            if (k != 2 * bits + 3) throw std::runtime_error("barrett_constants.bin: wrong k");
            // m = floor(2^k / q), k <= 125: long division of 2^k by q over 128 bits
            const unsigned __int128 m = (((unsigned __int128)1) << k) / qv;
            if ((uint64_t)m != lo || (uint64_t)(m >> 64) != hi) throw std::runtime_error("barrett_constants.bin: wrong m");
        }
    }
    // {stem}_a_{d}.bin then {stem}_b_{d}.bin into key_gmem at off as [a digits | b digits]
    // (each key_limbs+K limbs); returns key_limbs, 0 if the files are absent.
    uint32_t load_switch_key(const std::string &dir, const std::string &stem, uint64_t off) {
        const uint32_t dn = m_info[3];
        const std::string a0 = dir + "/" + stem + "_a_0.bin";
        if (!file_exists(a0)) return 0;
        const std::vector<uint64_t> first = read_words<uint64_t>(a0);
        if (first.empty() || first.size() % N) throw std::runtime_error("bad key file " + a0);
        const uint64_t R = first.size();
        const uint32_t limbs = (uint32_t)(R / N);
        if (limbs <= m_cfg.sizeP || limbs - m_cfg.sizeP > m_cfg.sizeQl)
            throw std::runtime_error("key file " + a0 + " does not hold key_limbs + K limbs");
        for (uint32_t part = 0; part < 2; ++part)
            for (uint32_t d = 0; d < dn; ++d) {
                const std::string path = dir + "/" + stem + (part ? "_b_" : "_a_") + std::to_string(d) + ".bin";
                m_dev->load_key_file(off + (uint64_t)(part * dn + d) * R, path, R);
            }
        return limbs - m_cfg.sizeP;
    }
    void bind() {
        throw_on(fhea_ctx_info(m_dev->ctx(), m_info), "fhea_ctx_info");
        N = m_info[0];
        m_cfg.N = N;
        m_cfg.sizeQl = m_info[1];
        m_cfg.sizeP = m_info[2];
        m_cfg.alpha = m_info[4];
        m_cfg.numPartQ = (m_info[1] + m_info[4] - 1) / m_info[4];
        m_primes.assign(m_info[1] + m_info[2], 0);
        throw_on(fhea_ctx_primes(m_dev->ctx(), m_primes.data()), "fhea_ctx_primes");
        m_cfg.rns_primes.assign(m_primes.begin(), m_primes.begin() + m_info[1]);
        m_cfg.p_primes.assign(m_primes.begin() + m_info[1], m_primes.end());
        std::vector<uint64_t> ninv(m_primes.size());
        throw_on(fhea_ctx_table(m_dev->ctx(), 1, ninv.data(), ninv.size()), "fhea_ctx_table");
        m_cfg.n_inv_q.assign(ninv.begin(), ninv.begin() + m_info[1]);
        m_cfg.n_inv_p.assign(ninv.begin() + m_info[1], ninv.end());
    }
    void dispatch(uint32_t op, uint64_t a, uint64_t b, uint64_t dst, uint64_t evk, uint32_t limbs) {
        throw_on(fhea_op(m_dev->ctx(), op, a, b, dst, evk, limbs, 0), "fhe_accel_top");
    }
    FheDevice *m_dev;
    uint32_t m_info[5] = {};
    std::vector<uint64_t> m_primes;
    FHEContextConfig m_cfg;
    std::map<int32_t, RotationKeyMeta> m_rot;
};

}  // namespace fhea_host
