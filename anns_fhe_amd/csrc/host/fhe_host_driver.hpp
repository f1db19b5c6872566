// fhe_host_driver.hpp -- C++ FHEHostDriver over the libfhea C-ABI (header only).
//
// Keeps the reference class's names, argument order and meaning
// (cpp_host_model/src/fhe_host_driver.h:55-123) so that callers such as
// InteractiveQueryManager (cpp_host_model/src/interactive_query_manager.cpp:85-274)
// compile against it unchanged apart from the constructor: the SimMemoryBus
// pointer becomes an FheDevice that owns a libfhea context with its two HBM
// arenas (poly_gmem / key_gmem, word-addressed as before).  Errors are thrown as
// std::runtime_error / std::out_of_range like the reference
// (fhe_host_driver.cpp:336-339, sim_memory_bus.cpp:13-15).
#pragma once
#include <cstdint>
#include <map>
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

// Owns one context; plays the role of SimMemoryBus (sim_memory_bus.h:10-34).
class FheDevice {
public:
    explicit FheDevice(const fhea_params &p) { throw_on(fhea_ctx_create(&p, &ctx_), "fhea_ctx_create"); }
    ~FheDevice() { fhea_ctx_destroy(ctx_); }
    FheDevice(const FheDevice &) = delete;
    FheDevice &operator=(const FheDevice &) = delete;

    fhea_ctx *ctx() const { return ctx_; }
    uint64_t *poly_gmem_ptr() { return fhea_poly_arena(ctx_); }
    uint64_t *key_gmem_ptr() { return fhea_key_arena(ctx_); }
    void write_poly(uint64_t off, const uint64_t *d, size_t n) { throw_on(fhea_write_poly(ctx_, off, d, n), "write_poly"); }
    void read_poly(uint64_t off, uint64_t *d, size_t n) const { throw_on(fhea_read_poly(ctx_, off, d, n), "read_poly"); }
    void write_key(uint64_t off, const uint64_t *d, size_t n) { throw_on(fhea_write_key(ctx_, off, d, n), "write_key"); }
    void read_key(uint64_t off, uint64_t *d, size_t n) const { throw_on(fhea_read_key(ctx_, off, d, n), "read_key"); }
    size_t poly_size_words() const { return fhea_poly_words(ctx_); }
    size_t key_size_words() const { return fhea_key_words(ctx_); }

private:
    fhea_ctx *ctx_ = nullptr;
};

struct RotationKeyMeta {  // fhe_host_driver.h:10-15 (auto_map_offset: closed-form map on device)
    int32_t step;
    uint64_t galois;
    uint64_t evk_offset;
    uint32_t key_limbs;
};

class FHEHostDriver {
public:
    explicit FHEHostDriver(FheDevice *dev) : m_dev(dev), m_ctx(dev->ctx()) {
        uint32_t info[5];
        throw_on(fhea_ctx_info(m_ctx, info), "fhea_ctx_info");
        N = info[0];
        m_sizeQl = info[1];
    }

    uint32_t N;

    // ---- context & keys (initialize_context_from_directory / load_context_to_hardware)
    void load_context_to_hardware() { dispatch(FHEA_OP_LOAD_CONTEXT, 0, 0, 0, 0, m_sizeQl); }
    void set_relin_key(uint64_t evk_off, uint32_t key_limbs) {
        throw_on(fhea_set_relin_key(m_ctx, evk_off, key_limbs), "set_relin_key");
    }
    void add_rotation_key(int32_t step, uint64_t galois, uint64_t evk_off, uint32_t key_limbs) {
        throw_on(fhea_add_rotation_key(m_ctx, step, galois, evk_off, key_limbs), "add_rotation_key");
        m_rot[step] = RotationKeyMeta{step, galois, evk_off, key_limbs};
    }

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
        throw_on(fhea_eval_add(m_ctx, a0, a1, b0, b1, d0, d1, limbs), "eval_add");
    }
    void eval_sub_plain(uint64_t a0, uint64_t a1, uint64_t pt, uint64_t d0, uint64_t d1, uint32_t limbs) {
        throw_on(fhea_eval_sub_plain(m_ctx, a0, a1, pt, d0, d1, limbs), "eval_sub_plain");
    }
    void eval_mult_relin_rescale(uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1, uint64_t d0, uint64_t d1,
                                 uint32_t in_limbs) {
        throw_on(fhea_eval_mult_relin_rescale(m_ctx, a0, a1, b0, b1, d0, d1, in_limbs), "eval_mult_relin_rescale");
    }
    void eval_mult_plain_rescale(uint64_t a0, uint64_t a1, uint64_t pt, uint64_t d0, uint64_t d1,
                                 uint32_t in_limbs) {
        throw_on(fhea_eval_mult_plain_rescale(m_ctx, a0, a1, pt, d0, d1, in_limbs), "eval_mult_plain_rescale");
    }
    void eval_rotate_single_step(uint64_t in0, uint64_t in1, uint64_t d0, uint64_t d1, int32_t step,
                                 uint32_t limbs) {
        auto it = m_rot.find(step);
        if (it == m_rot.end()) throw std::runtime_error("Rotation key not found for step: " + std::to_string(step));
        throw_on(fhea_eval_automorphism_ks(m_ctx, it->second.galois, it->second.evk_offset, it->second.key_limbs,
                                           in0, in1, d0, d1, limbs),
                 "eval_rotate_single_step");
    }
    void eval_rotate(uint64_t in0, uint64_t in1, uint64_t d0, uint64_t d1, int32_t total_rot, uint32_t limbs) {
        throw_on(fhea_eval_rotate(m_ctx, in0, in1, d0, d1, total_rot, limbs), "eval_rotate");
    }

    void sync() { throw_on(fhea_sync(m_ctx), "fhea_sync"); }
    FheDevice *get_bus() { return m_dev; }

private:
    void dispatch(uint32_t op, uint64_t a, uint64_t b, uint64_t dst, uint64_t evk, uint32_t limbs) {
        throw_on(fhea_op(m_ctx, op, a, b, dst, evk, limbs, 0), "fhe_accel_top");
    }
    FheDevice *m_dev;
    fhea_ctx *m_ctx;
    uint32_t m_sizeQl;
    std::map<int32_t, RotationKeyMeta> m_rot;
};

}  // namespace fhea_host
