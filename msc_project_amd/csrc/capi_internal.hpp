// Internal object layouts behind the opaque handles of include/lgf_flow.h.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "kernels.hpp"
#include "lattice.hpp"

struct lgf_batch;

struct lgf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int stage_variant[2] = {0, 0};
    int tile_block = 4;       // traversal order of new batches: see batch_tile_order
    void* staging = nullptr;  // device staging chunk for uploads/downloads
    void* pinned = nullptr;   // pinned host staging chunk
    void* lines = nullptr;    // line products of the clover measurement (grown on demand)
    size_t lines_bytes = 0;
    size_t lines_cap = size_t(1) << 30;  // fields are measured in chunks of about this many line bytes
    size_t group_bytes = 0;   // cache groups of new batches (lgf_ctx_set_batch_group_bytes); 0 = one group
    int live_fields = 0;      // fields and user batches not yet destroyed
    int live_batches = 0;
};

struct lgf_field {
    lgf_ctx* ctx = nullptr;
    int shape[4] = {0, 0, 0, 0};
    int prec = 0;
    int vol = 0;
    int64_t stride = 0;
    int ntiles = 0;
    void* links[2] = {nullptr, nullptr};
    void* acc = nullptr;
    double* qx = nullptr;  // lazily allocated per-site charge density
    int cur = 0;           // links[cur] holds the current configuration
    lgf_batch* solo = nullptr;
    int in_batches = 0;    // user batches that reference this field
    double flow_t = 0.0;   // flow time, t += eps per RK3 step (GradientFlow's m_flowTime)
};

struct lgf_batch {
    lgf_ctx* ctx = nullptr;
    int prec = 0;
    std::vector<lgf_field*> fields;
    std::vector<lgf::FieldDesc> h_desc;
    lgf::FieldDesc* d_desc = nullptr;
    int2* d_tiles = nullptr;
    int ntiles = 0;
    double* d_partials = nullptr;
    double* d_out = nullptr;
    double* h_out = nullptr;
    int ping = 0;
    // cache groups: consecutive fields whose footprint fits ctx->group_bytes,
    // as (first tile, tiles) of the tile list; the flow runs each group's
    // steps back to back
    std::vector<int2> groups;
    // per group, two RK3 steps captured as a graph, by starting ping
    std::vector<hipGraphExec_t> graph;  // [2 * group + ping]
    std::vector<int> graph_variant;
    bool is_solo = false;  // a field's private one-field batch
    // deferred measurement (lgf_batch_measure_start / _finish): own partials
    // and result buffers, completion event
    hipEvent_t ev_done = nullptr;
    double* d_partials_m = nullptr;
    double* d_out_m = nullptr;
    double* h_out_m = nullptr;
    bool pending = false;
    // on-device flow-scale tracking (lgf_batch_track_t0): per field
    // {t, t^2 E} of the previous measurement and t0 (NaN until reached)
    double* d_t0 = nullptr;
    double t0_target = 0.0;
    bool track_t0 = false;
};

int lgf_fail(lgf_ctx* ctx, int code, const std::string& msg);
int lgf_hip_fail(lgf_ctx* ctx, hipError_t e, const char* what);
int batch_sync_desc(lgf_batch* b, const double* eps = nullptr);
int ctx_staging(lgf_ctx* ctx, void** dev, void** host);
// one field's clover sums (k_reduce record) -> obs[LGF_NUM_OBS]; LGF_E_NAN if
// any is not finite
int observables_of_sums(lgf_ctx* ctx, const double* h, int vol, double* obs);
