// fhea_engine.hpp -- the context object behind the C-ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "../../include/fhea.h"
#include "fhea_internal.hpp"
#include "fhea_kernels.hpp"

struct fhea_ctx {
    fhea::HostTables h;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    u64 *poly = nullptr, *key = nullptr;
    u64 poly_words = 0, key_words = 0;

    void *dblob = nullptr;  // device tables
    fhea::DevTables dt;
    std::vector<fhea::LevelTables> lv;       // index l = 1..L
    std::vector<const fhea::ConvDesc *> modup_desc_dev;  // per level
    std::vector<const u32 *> modup_first_dev;            // per level
    std::vector<const fhea::ConvDesc *> moddown_desc_dev;
    std::vector<const u32 *> rs_last_map;                // per level: {l-1, l-1}
    std::vector<const u64 *> rs_sub;                     // per level: q_l * T_i mod q_i
    const u32 *zero_u32 = nullptr;                       // {0}

    // workspace
    u64 *ws = nullptr;
    u64 ws_words = 0;
    // per-item pointer lists: bump-allocated in a pinned host buffer and copied
    // to the same offset of a device buffer; reset only after a stream sync
    void *args = nullptr;
    void *args_host = nullptr;
    size_t args_bytes = 0;
    size_t args_used = 0;

    bool has_relin = false;
    u64 relin_off = 0;
    u32 relin_limbs = 0;
    std::map<int32_t, fhea::RotKey> rot;

    // second workspace: ping-pong ciphertexts of multi-step rotations
    u64 *pp = nullptr;
    u64 pp_words = 0;

    // HIP graph capture (fhea_graph_*): while capturing, staged arguments go to
    // device buffers owned by the graph and workspaces cannot grow
    bool capturing = false;
    void *cap_dev = nullptr, *cap_host = nullptr;  // argument area of the graph being captured
    size_t cap_used = 0;
    static constexpr size_t CAP_BYTES = 4u << 20;
    struct Graph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        void *dev = nullptr, *host = nullptr;
    };
    std::map<uint32_t, Graph> graphs;
    uint32_t next_graph = 1;
    // device buffers (workspace, rotation ping-pong) replaced while graphs existed:
    // recorded kernel nodes may still address them, so they are freed only when the
    // last graph is destroyed (or with the context)
    std::vector<void *> retired;

    // lanes (FHEA_TUNE_STREAMS = n <= MAX_LANES): a batched HMult runs its items in n
    // parts, part k > 0 on lane_stream[k] with its own workspace lane_ws[k], forked
    // from and joined back into `stream` with events, so that the parts' kernels
    // overlap on the device (part 0 runs on `stream` in `ws`)
    static constexpr u32 MAX_LANES = 4;
    u32 lanes = 1;
    hipStream_t lane_stream[MAX_LANES] = {};
    u64 *lane_ws[MAX_LANES] = {};
    u64 lane_ws_words[MAX_LANES] = {};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t lane_join[MAX_LANES] = {};

    // reads in flight (fhea_read_polys_async): copies on their own stream, after an
    // event of `stream`, each ticket's completion event in a ring
    static constexpr u32 READ_RING = 16;
    hipStream_t copy_stream = nullptr;
    hipEvent_t read_dep = nullptr;
    hipEvent_t read_done[READ_RING] = {};
    u32 read_seq = 0;

    u32 max_chunk = 16;
    bool force_int = false;  // FHEA_FORCE_INT=1: integer kernels even for < 2^50 primes
    u32 f64_off = 0;         // FHEA_F64_OFF: per-component switch (DevTables::f64_off)
    u32 cc_var = 0;          // FHEA_CC_VARIANT: fused conversion column kernel variant
};

namespace fhea {
int check_hip(hipError_t e);
int ensure_ws(fhea_ctx *c, u64 words);
// stage a host array into the device args buffer, returns device pointer
void *stage_args(fhea_ctx *c, const void *host, size_t bytes);
}  // namespace fhea
