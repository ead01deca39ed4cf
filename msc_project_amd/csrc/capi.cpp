// C-ABI of the gfx950 Wilson-flow path (declarations and the reference
// interface each call replaces: include/lgf_flow.h).
//
// Ownership: a context owns a HIP stream and a device staging buffer; a field
// owns its two ping-pong link buffers and the RK3 accumulator; a batch owns a
// device table of field descriptors, the tile map and the reduction buffers
// but not its fields.  Nothing here falls back to the CPU: every compute
// call launches the kernels of kernels.hip or returns an error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/lgf_flow.h"
#include "capi_internal.hpp"
#include "kernels.hpp"

using namespace lgf;

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr size_t kStageBytes = size_t(64) << 20;  // host<->device staging chunk

int64_t round_stride(int64_t vol) { return (vol + 63) / 64 * 64; }

int check_shape(const int* shape, int64_t* vol) {
    if (!shape) return LGF_E_ARG;
    int64_t v = 1;
    for (int i = 0; i < 4; ++i) {
        if (shape[i] <= 0 || (shape[i] & 1)) return LGF_E_SHAPE;  // even extents (LGF/Lattice/Lattice.cpp:42-49)
        v *= shape[i];
    }
    // kernels address a vector plane with 32-bit byte offsets of 16 B per site
    if (v > (int64_t(1) << 28)) return LGF_E_SHAPE;
    *vol = v;
    return LGF_OK;
}

size_t elem_size(int prec) { return prec == LGF_FP64 ? sizeof(double) : sizeof(float); }

}  // namespace

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
int lgf_fail(lgf_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int lgf_hip_fail(lgf_ctx* ctx, hipError_t e, const char* what) {
    return lgf_fail(ctx, LGF_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPTRY(ctx, expr)                                    \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return lgf_hip_fail(ctx, _e, #expr); \
    } while (0)

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int lgf_ctx_create(int device, lgf_ctx** out) {
    if (!out) return LGF_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return LGF_E_HIP;
    if (device < 0 || device >= n) return LGF_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return LGF_E_HIP;
    lgf_ctx* c = new (std::nothrow) lgf_ctx();
    if (!c) return LGF_E_NOMEM;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return LGF_E_HIP;
    }
    *out = c;
    return LGF_OK;
}

extern "C" int lgf_ctx_destroy(lgf_ctx* ctx) {
    if (!ctx) return LGF_E_ARG;
    if (ctx->live_fields || ctx->live_batches)
        return lgf_fail(ctx, LGF_E_STATE, "destroy the context's fields and batches first");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->staging) (void)hipFree(ctx->staging);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->lines) (void)hipFree(ctx->lines);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return LGF_OK;
}

extern "C" const char* lgf_ctx_last_error(const lgf_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int lgf_ctx_stream(lgf_ctx* ctx, void** s) {
    if (!ctx || !s) return LGF_E_ARG;
    *s = (void*)ctx->stream;
    return LGF_OK;
}

extern "C" int lgf_ctx_set_measure_scratch(lgf_ctx* ctx, size_t bytes) {
    if (!ctx) return LGF_E_ARG;
    ctx->lines_cap = bytes;
    return LGF_OK;
}

extern "C" int lgf_ctx_synchronize(lgf_ctx* ctx) {
    if (!ctx) return LGF_E_ARG;
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    return LGF_OK;
}

extern "C" int lgf_num_stage_variants(int precision, int* n) {
    if (!n || (precision != LGF_FP64 && precision != LGF_FP32)) return LGF_E_ARG;
    *n = num_stage_variants(precision);
    return LGF_OK;
}

extern "C" int lgf_ctx_set_stage_variant(lgf_ctx* ctx, int precision, int variant) {
    if (!ctx || (precision != LGF_FP64 && precision != LGF_FP32)) return LGF_E_ARG;
    if (variant < 0 || variant >= num_stage_variants(precision))
        return lgf_fail(ctx, LGF_E_ARG, "stage variant out of range");
    ctx->stage_variant[precision] = variant;
    return LGF_OK;
}

extern "C" int lgf_ctx_get_stage_variant(lgf_ctx* ctx, int precision, int* variant) {
    if (!ctx || !variant || (precision != LGF_FP64 && precision != LGF_FP32)) return LGF_E_ARG;
    *variant = ctx->stage_variant[precision];
    return LGF_OK;
}

int ctx_staging(lgf_ctx* ctx, void** dev, void** host) {
    if (!ctx->staging) HIPTRY(ctx, hipMalloc(&ctx->staging, kStageBytes));
    if (!ctx->pinned) HIPTRY(ctx, hipHostMalloc(&ctx->pinned, kStageBytes, hipHostMallocDefault));
    *dev = ctx->staging;
    *host = ctx->pinned;
    return LGF_OK;
}

// ---------------------------------------------------------------------------
// fields
// ---------------------------------------------------------------------------
extern "C" int lgf_field_footprint(const int shape[4], int precision, size_t* bytes) {
    int64_t vol;
    int rc = check_shape(shape, &vol);
    if (rc) return rc;
    if (!bytes || (precision != LGF_FP64 && precision != LGF_FP32)) return LGF_E_ARG;
    *bytes = size_t(2 * 72 + 36) * size_t(round_stride(vol)) * elem_size(precision);
    return LGF_OK;
}

extern "C" int lgf_field_create(lgf_ctx* ctx, const int shape[4], int precision, lgf_field** out) {
    if (!ctx || !out) return LGF_E_ARG;
    *out = nullptr;
    int64_t vol;
    int rc = check_shape(shape, &vol);
    if (rc) return lgf_fail(ctx, rc, "lattice extents must be positive and even");
    if (precision != LGF_FP64 && precision != LGF_FP32) return lgf_fail(ctx, LGF_E_ARG, "unknown precision");
    lgf_field* f = new (std::nothrow) lgf_field();
    if (!f) return LGF_E_NOMEM;
    f->ctx = ctx;
    for (int i = 0; i < 4; ++i) f->shape[i] = shape[i];
    f->prec = precision;
    f->vol = (int)vol;
    f->stride = round_stride(vol);
    f->ntiles = (int)((vol + kTileSites - 1) / kTileSites);
    // two link buffers of 72 reals per site, and the accumulator of 36
    // (9 per link: traceless anti-Hermitian, see lattice.hpp)
    const size_t bytes = size_t(72) * size_t(f->stride) * elem_size(precision);
    HIPTRY(ctx, hipSetDevice(ctx->device));
    void* p[3] = {nullptr, nullptr, nullptr};
    for (int i = 0; i < 3; ++i) {
        if (hipMalloc(&p[i], i < 2 ? bytes : bytes / 2) != hipSuccess) {
            for (int j = 0; j < i; ++j) (void)hipFree(p[j]);
            delete f;
            (void)hipGetLastError();
            return lgf_fail(ctx, LGF_E_NOMEM, "device allocation of a field failed");
        }
    }
    f->links[0] = p[0];
    f->links[1] = p[1];
    f->acc = p[2];
    f->cur = 0;
    ctx->live_fields++;
    rc = lgf_field_set_cold(f);
    if (rc) {
        lgf_field_destroy(f);
        return rc;
    }
    *out = f;
    return LGF_OK;
}

extern "C" int lgf_field_destroy(lgf_field* f) {
    if (!f) return LGF_E_ARG;
    lgf_ctx* ctx = f->ctx;
    if (f->in_batches) return lgf_fail(ctx, LGF_E_STATE, "destroy the batches holding this field first");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (f->solo) lgf_batch_destroy(f->solo);
    ctx->live_fields--;
    (void)hipFree(f->links[0]);
    (void)hipFree(f->links[1]);
    (void)hipFree(f->acc);
    if (f->qx) (void)hipFree(f->qx);
    delete f;
    return LGF_OK;
}

extern "C" int lgf_field_shape(const lgf_field* f, int shape[4], int* precision) {
    if (!f) return LGF_E_ARG;
    if (shape)
        for (int i = 0; i < 4; ++i) shape[i] = f->shape[i];
    if (precision) *precision = f->prec;
    return LGF_OK;
}

extern "C" int lgf_field_flow_time(const lgf_field* f, double* t) {
    if (!f || !t) return LGF_E_ARG;
    *t = f->flow_t;
    return LGF_OK;
}

extern "C" int lgf_field_set_flow_time(lgf_field* f, double t) {
    if (!f || !(t == t)) return LGF_E_ARG;
    f->flow_t = t;
    return LGF_OK;
}

extern "C" int lgf_field_set_cold(lgf_field* f) {
    if (!f) return LGF_E_ARG;
    lgf_ctx* ctx = f->ctx;
    HIPTRY(ctx, launch_cold(f->prec, f->links[f->cur], f->stride, f->vol, ctx->stream));
    HIPTRY(ctx, hipMemsetAsync(f->acc, 0, size_t(36) * size_t(f->stride) * elem_size(f->prec), ctx->stream));
    return LGF_OK;
}

extern "C" int lgf_field_set_random(lgf_field* f, uint64_t seed) {
    if (!f) return LGF_E_ARG;
    lgf_ctx* ctx = f->ctx;
    HIPTRY(ctx, launch_hot(f->prec, f->links[f->cur], f->stride, f->vol, seed, ctx->stream));
    return LGF_OK;
}

// Stream `vol` sites of file-order links through the staging buffers.
// `fill(host_chunk, s0, n)` writes n sites of file-order data starting at s0.
template <typename Fill>
static int upload_stream(lgf_field* f, Fill fill) {
    lgf_ctx* ctx = f->ctx;
    void *dev, *host;
    int rc = ctx_staging(ctx, &dev, &host);
    if (rc) return rc;
    const int per = (int)(kStageBytes / (72 * sizeof(double)));
    for (int s0 = 0; s0 < f->vol; s0 += per) {
        const int n = std::min(per, f->vol - s0);
        HIPTRY(ctx, hipStreamSynchronize(ctx->stream));  // host staging buffer free again
        rc = fill((double*)host, s0, n);
        if (rc) return rc;
        HIPTRY(ctx, hipMemcpyAsync(dev, host, size_t(n) * 72 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HIPTRY(ctx, launch_to_soa(f->prec, (const double*)dev, f->links[f->cur], f->stride, s0, n, ctx->stream));
    }
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    return LGF_OK;
}

template <typename Drain>
static int download_stream(lgf_field* f, Drain drain) {
    lgf_ctx* ctx = f->ctx;
    void *dev, *host;
    int rc = ctx_staging(ctx, &dev, &host);
    if (rc) return rc;
    const int per = (int)(kStageBytes / (72 * sizeof(double)));
    for (int s0 = 0; s0 < f->vol; s0 += per) {
        const int n = std::min(per, f->vol - s0);
        HIPTRY(ctx, launch_from_soa(f->prec, f->links[f->cur], (double*)dev, f->stride, s0, n, ctx->stream));
        HIPTRY(ctx, hipMemcpyAsync(host, dev, size_t(n) * 72 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
        rc = drain((const double*)host, s0, n);
        if (rc) return rc;
    }
    return LGF_OK;
}

extern "C" int lgf_field_upload(lgf_field* f, const double* links) {
    if (!f || !links) return LGF_E_ARG;
    return upload_stream(f, [&](double* h, int s0, int n) {
        std::memcpy(h, links + size_t(s0) * 72, size_t(n) * 72 * sizeof(double));
        return LGF_OK;
    });
}

extern "C" int lgf_field_download(lgf_field* f, double* links) {
    if (!f || !links) return LGF_E_ARG;
    return download_stream(f, [&](const double* h, int s0, int n) {
        std::memcpy(links + size_t(s0) * 72, h, size_t(n) * 72 * sizeof(double));
        return LGF_OK;
    });
}

// The reference's format: record (g*4+mu) of 144 bytes at byte offset
// (g*4+mu)*144, i.e. the whole file is the file-order array.  Offsets are
// 64-bit here (the reference computes them in int, LGF/Field/SU3_field.cpp:331).
extern "C" int lgf_field_load_file(lgf_field* f, const char* path) {
    if (!f || !path) return LGF_E_ARG;
    lgf_ctx* ctx = f->ctx;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return lgf_fail(ctx, LGF_E_IO, std::string("cannot open configuration file ") + path);
    std::fseek(fp, 0, SEEK_END);
    const long long size = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    const long long want = (long long)f->vol * 576;
    if (size != want) {
        std::fclose(fp);
        return lgf_fail(ctx, LGF_E_IO, std::string("configuration file ") + path + " has " + std::to_string(size) +
                                           " bytes, expected " + std::to_string(want));
    }
    int rc = upload_stream(f, [&](double* h, int, int n) {
        size_t got = std::fread(h, sizeof(double), size_t(n) * 72, fp);
        return got == size_t(n) * 72 ? LGF_OK : lgf_fail(ctx, LGF_E_IO, std::string("short read from ") + path);
    });
    std::fclose(fp);
    return rc;
}

extern "C" int lgf_field_save_file(lgf_field* f, const char* path) {
    if (!f || !path) return LGF_E_ARG;
    lgf_ctx* ctx = f->ctx;
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return lgf_fail(ctx, LGF_E_IO, std::string("cannot write configuration file ") + path);
    int rc = download_stream(f, [&](const double* h, int, int n) {
        size_t put = std::fwrite(h, sizeof(double), size_t(n) * 72, fp);
        return put == size_t(n) * 72 ? LGF_OK : lgf_fail(ctx, LGF_E_IO, std::string("short write to ") + path);
    });
    if (std::fclose(fp) != 0 && rc == LGF_OK) rc = lgf_fail(ctx, LGF_E_IO, std::string("close failed: ") + path);
    return rc;
}

extern "C" int lgf_field_copy(lgf_field* dst, const lgf_field* src) {
    if (!dst || !src || dst->ctx != src->ctx) return LGF_E_ARG;
    lgf_ctx* ctx = dst->ctx;
    for (int i = 0; i < 4; ++i)
        if (dst->shape[i] != src->shape[i]) return lgf_fail(ctx, LGF_E_SHAPE, "copy between different shapes");
    HIPTRY(ctx, launch_convert(src->prec, src->links[src->cur], dst->prec, dst->links[dst->cur], dst->stride,
                               dst->vol, ctx->stream));
    return LGF_OK;
}

static int solo(lgf_field* f, lgf_batch** b) {
    if (!f->solo) {
        int rc = lgf_batch_create(f->ctx, &f, 1, &f->solo);
        if (rc) return rc;
        f->solo->is_solo = true;
        f->in_batches--;
        f->ctx->live_batches--;
    }
    *b = f->solo;
    return LGF_OK;
}

extern "C" int lgf_flow_steps(lgf_field* f, double eps, int nsteps) {
    if (!f || nsteps < 0) return LGF_E_ARG;
    lgf_batch* b;
    int rc = solo(f, &b);
    if (rc) return rc;
    return lgf_batch_flow_steps(b, &eps, nsteps, 0);
}

extern "C" int lgf_measure_observables(lgf_field* f, int flags, double* obs, double* qraw) {
    if (!f || !obs) return LGF_E_ARG;
    lgf_ctx* ctx = f->ctx;
    lgf_batch* b;
    int rc = solo(f, &b);
    if (rc) return rc;
    if (qraw && !f->qx) HIPTRY(ctx, hipMalloc(&f->qx, size_t(f->vol) * sizeof(double)));
    rc = lgf_batch_measure_observables(b, flags, obs);
    if (rc && rc != LGF_E_NAN) return rc;
    if (qraw) {
        HIPTRY(ctx, hipMemcpyAsync(qraw, f->qx, size_t(f->vol) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    return rc;
}

// Seeded configuration generator (generator.hip): nsweeps iterations of one
// heatbath sweep plus or_per_hb overrelaxation sweeps, iteration indices
// first_sweep.. (so a chain can be continued), links reunitarised after each
// iteration.  Start from lgf_field_set_cold / lgf_field_set_random.
extern "C" int lgf_field_heatbath(lgf_field* f, double beta, uint64_t seed, int first_sweep, int nsweeps,
                                  int or_per_hb) {
    if (!f || nsweeps < 0 || or_per_hb < 0 || first_sweep < 0 || !(beta > 0)) return LGF_E_ARG;
    lgf_ctx* ctx = f->ctx;
    lgf_batch* b;
    int rc = solo(f, &b);
    if (rc) return rc;
    rc = batch_sync_desc(b);
    if (rc) return rc;
    for (int it = 0; it < nsweeps; ++it)
        HIPTRY(ctx, launch_heatbath_iteration(f->prec, b->d_desc, f->vol, b->ping, beta, seed,
                                              (uint32_t)(first_sweep + it), or_per_hb, ctx->stream));
    return LGF_OK;
}

extern "C" int lgf_flow_scale_t0(const double* t, const double* E, int n, double target, double* t0) {
    if (!t || !E || !t0 || n < 2) return LGF_E_ARG;
    for (int i = 0; i + 1 < n; ++i) {
        if (!(t[i + 1] > t[i])) return LGF_E_ARG;
        const double a = t[i] * t[i] * E[i] - target, b = t[i + 1] * t[i + 1] * E[i + 1] - target;
        if (a < 0.0 && b >= 0.0) {
            *t0 = t[i] + (t[i + 1] - t[i]) * (-a) / (b - a);
            return LGF_OK;
        }
    }
    return LGF_E_STATE;
}

extern "C" int lgf_diagnostics(lgf_field* f, double* plaq, double* umax, double* umean) {
    if (!f) return LGF_E_ARG;
    lgf_ctx* ctx = f->ctx;
    lgf_batch* b;
    int rc = solo(f, &b);
    if (rc) return rc;
    rc = batch_sync_desc(b);
    if (rc) return rc;
    HIPTRY(ctx, launch_diag(f->prec, b->d_desc, b->d_tiles, b->ntiles, b->ping, b->d_partials, 1, b->d_out,
                            ctx->stream));
    HIPTRY(ctx, hipMemcpyAsync(b->h_out, b->d_out, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    if (plaq) *plaq = b->h_out[0] / (3.0 * 6.0 * (double)f->vol);
    if (umean) *umean = b->h_out[1] / (4.0 * (double)f->vol);
    if (umax) *umax = b->h_out[2];
    return LGF_OK;
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
// Processing order of a field's 64-site tiles.  Memory stays lexicographic
// (x3 fastest); the order in which workgroups visit tiles decides how far
// apart in time a link and its neighbours are touched.  In lexicographic
// order the x0-neighbour of a site is a whole 3-d slab (N1*N2*N3 sites)
// away, beyond an XCD's L2.  With block size B > 0 tiles are visited by the
// key (x0 / B, x1, x0 % B, x2, x3): the x0-neighbour is then N2*N3 sites
// away for all but one x0 in B, and the x1-neighbour B*N2*N3 sites away.
// B = 0 keeps lexicographic order.  Reductions are indexed by the tile,
// not by its place in this order, so results do not depend on B.
// Tile order code: 0 = lexicographic; B in 2..64 = x0 blocked by B; 100 B2 + B0
// (B0 in 1..64, B2 in 2..64) = x0 blocked by B0 and x2 by B2.  A block of
// x0 (and x2) values is swept along x1, so the neighbour rows a tile reads
// along x0, x1 (and x2) were read a few tiles earlier: with both blocked the
// sweep's window stays within an XCD's L2 also when one x2-x3 plane does not.
static bool tile_order_valid(int code) {
    if (code >= 0 && code <= 64) return true;
    const int b0 = code % 100, b2 = code / 100;
    return code > 100 && b0 >= 1 && b0 <= 64 && b2 >= 2 && b2 <= 64;
}

static std::vector<int> batch_tile_order(const lgf_field* f, int code) {
    std::vector<int> order(f->ntiles);
    for (int t = 0; t < f->ntiles; ++t) order[t] = t;
    if (code <= 1) return order;
    const int64_t B0 = code > 100 ? code % 100 : code, B2 = code > 100 ? code / 100 : 1;
    const int64_t n1 = f->shape[1], n2 = f->shape[2], n3 = f->shape[3];
    const int64_t nb2 = (n2 + B2 - 1) / B2;
    auto key = [&](int t) {
        int64_t s = (int64_t)t * kTileSites;
        const int64_t x3 = s % n3;
        s /= n3;
        const int64_t x2 = s % n2;
        s /= n2;
        const int64_t x1 = s % n1;
        const int64_t x0 = s / n1;
        return (((((x0 / B0) * nb2 + x2 / B2) * n1 + x1) * B0 + x0 % B0) * B2 + x2 % B2) * n3 + x3;
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key(a) < key(b); });
    return order;
}

extern "C" int lgf_ctx_set_tile_order(lgf_ctx* ctx, int block) {
    if (!ctx || !tile_order_valid(block)) return LGF_E_ARG;
    ctx->tile_block = block;
    return LGF_OK;
}

extern "C" int lgf_ctx_get_tile_order(lgf_ctx* ctx, int* block) {
    if (!ctx || !block) return LGF_E_ARG;
    *block = ctx->tile_block;
    return LGF_OK;
}
extern "C" int lgf_batch_create(lgf_ctx* ctx, lgf_field* const* fields, int nfields, lgf_batch** out) {
    if (!ctx || !fields || nfields <= 0 || !out) return LGF_E_ARG;
    *out = nullptr;
    const int prec = fields[0]->prec;
    for (int i = 0; i < nfields; ++i) {
        if (!fields[i] || fields[i]->ctx != ctx) return lgf_fail(ctx, LGF_E_ARG, "field of another context");
        if (fields[i]->prec != prec) return lgf_fail(ctx, LGF_E_ARG, "a batch needs one precision");
    }
    lgf_batch* b = new (std::nothrow) lgf_batch();
    if (!b) return LGF_E_NOMEM;
    b->ctx = ctx;
    b->prec = prec;
    b->fields.assign(fields, fields + nfields);
    std::vector<int2> tiles;
    b->h_desc.resize(nfields);
    for (int i = 0; i < nfields; ++i) {
        lgf_field* f = fields[i];
        FieldDesc& d = b->h_desc[i];
        std::memset(&d, 0, sizeof d);
        d.links[0] = f->links[0];
        d.links[1] = f->links[1];
        d.acc = f->acc;
        d.qx = f->qx;
        for (int k = 0; k < 4; ++k) {
            d.n[k] = f->shape[k];
            d.rcp[k] = (1.0 + 1.0 / 1099511627776.0) / (double)f->shape[k];
        }
        d.vol = f->vol;
        d.ntiles = f->ntiles;
        d.stride = f->stride;
        d.eps = 0.0;
        d.tile0 = (int)tiles.size();
        d.flip = f->cur;
        for (int t : batch_tile_order(f, ctx->tile_block)) tiles.push_back(make_int2(i, t));
    }
    b->ntiles = (int)tiles.size();
    // cache groups: a field that alone exceeds the cap forms a group of its own
    {
        int first = 0, nt = 0;
        size_t bytes = 0;
        for (int i = 0; i < nfields; ++i) {
            size_t fb = 0;
            (void)lgf_field_footprint(fields[i]->shape, prec, &fb);
            if (ctx->group_bytes && nt > 0 && bytes + fb > ctx->group_bytes) {
                b->groups.push_back(make_int2(first, nt));
                first += nt;
                nt = 0;
                bytes = 0;
            }
            nt += fields[i]->ntiles;
            bytes += fb;
        }
        b->groups.push_back(make_int2(first, nt));
    }
    b->graph.assign(2 * b->groups.size(), nullptr);
    b->graph_variant.assign(2 * b->groups.size(), -1);
    b->ping = 0;
    hipError_t e = hipMalloc(&b->d_desc, sizeof(FieldDesc) * nfields);
    if (e == hipSuccess) e = hipMalloc(&b->d_tiles, sizeof(int2) * tiles.size());
    if (e == hipSuccess) e = hipMalloc(&b->d_partials, sizeof(double) * kCloverParts * tiles.size());
    if (e == hipSuccess) e = hipMalloc(&b->d_out, sizeof(double) * kOutStride * nfields);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&b->h_out, sizeof(double) * kOutStride * nfields, hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_desc, b->h_desc.data(), sizeof(FieldDesc) * nfields, hipMemcpyHostToDevice);
    for (lgf_field* f : b->fields) f->in_batches++;
    ctx->live_batches++;
    if (e != hipSuccess) {
        lgf_batch_destroy(b);
        return lgf_hip_fail(ctx, e, "batch allocation");
    }
    *out = b;
    return LGF_OK;
}

extern "C" int lgf_batch_destroy(lgf_batch* b) {
    if (!b) return LGF_E_ARG;
    if (!b->is_solo) {
        for (lgf_field* f : b->fields) f->in_batches--;
        b->ctx->live_batches--;
    }
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    for (hipGraphExec_t g : b->graph)
        if (g) (void)hipGraphExecDestroy(g);
    if (b->d_desc) (void)hipFree(b->d_desc);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->d_partials) (void)hipFree(b->d_partials);
    if (b->d_out) (void)hipFree(b->d_out);
    if (b->h_out) (void)hipHostFree(b->h_out);
    if (b->d_partials_m) (void)hipFree(b->d_partials_m);
    if (b->d_out_m) (void)hipFree(b->d_out_m);
    if (b->h_out_m) (void)hipHostFree(b->h_out_m);
    if (b->d_t0) (void)hipFree(b->d_t0);
    if (b->ev_done) (void)hipEventDestroy(b->ev_done);
    delete b;
    return LGF_OK;
}

// Bring the device descriptor table in line with the fields (current
// buffer, qx output) and the requested step sizes.
int batch_sync_desc(lgf_batch* b, const double* eps) {
    bool dirty = false;
    for (size_t i = 0; i < b->fields.size(); ++i) {
        lgf_field* f = b->fields[i];
        FieldDesc& d = b->h_desc[i];
        const int flip = f->cur ^ b->ping;
        if (d.flip != flip || d.qx != f->qx || (eps && d.eps != eps[i])) {
            d.flip = flip;
            d.qx = f->qx;
            if (eps) d.eps = eps[i];
            dirty = true;
        }
    }
    if (dirty) {
        lgf_ctx* ctx = b->ctx;
        HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
        HIPTRY(ctx, hipMemcpy(b->d_desc, b->h_desc.data(), sizeof(FieldDesc) * b->h_desc.size(),
                              hipMemcpyHostToDevice));
    }
    return LGF_OK;
}

// one RK3 step of the fields of tiles [t0, t0 + nt) of the tile list
static int enqueue_step(lgf_batch* b, int ping, int t0, int nt) {
    lgf_ctx* ctx = b->ctx;
    const int var = ctx->stage_variant[b->prec];
    for (int stage = 0; stage < 3; ++stage) {
        hipError_t e = launch_rk3_stage_variant(b->prec, var, stage, b->d_desc, b->d_tiles + t0, nt, ping ^ (stage & 1),
                                                ctx->stream);
        if (e != hipSuccess) return lgf_hip_fail(ctx, e, "rk3 stage launch");
    }
    return LGF_OK;
}

// t += eps once per step, as the reference accumulates its flow time
// (LGF/Gradient_Flow/GradientFlow.cpp:52-58)
static void add_flow_time(lgf_batch* b, const double* eps, int steps) {
    for (size_t i = 0; i < b->fields.size(); ++i)
        for (int k = 0; k < steps; ++k) b->fields[i]->flow_t += eps[i];
}

static void advance(lgf_batch* b, int stages) {
    if (stages & 1) {
        b->ping ^= 1;
        for (lgf_field* f : b->fields) f->cur ^= 1;
    }
}

// Each cache group runs all nsteps before the next group starts: fields are
// independent, so the order of groups changes no result, and a group whose
// links, next links and accumulator fit the last-level cache keeps them
// there across its stages instead of streaming the whole batch from HBM
// every stage.
extern "C" int lgf_batch_flow_steps(lgf_batch* b, const double* eps, int nsteps, int use_graph) {
    if (!b || !eps || nsteps < 0) return LGF_E_ARG;
    lgf_ctx* ctx = b->ctx;
    int rc = batch_sync_desc(b, eps);
    if (rc) return rc;
    const int var = ctx->stage_variant[b->prec];
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        const int t0 = b->groups[gi].x, nt = b->groups[gi].y;
        int p = b->ping, n = nsteps;
        while (n > 0) {
            if (use_graph && n >= 2) {
                hipGraphExec_t& gx = b->graph[2 * gi + p];
                int& gv = b->graph_variant[2 * gi + p];
                if (gx && gv != var) {
                    (void)hipGraphExecDestroy(gx);
                    gx = nullptr;
                }
                if (!gx) {
                    hipGraph_t graph;
                    HIPTRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                    int rc1 = enqueue_step(b, p, t0, nt);
                    int rc2 = enqueue_step(b, p ^ 1, t0, nt);
                    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
                    if (rc1 || rc2) return rc1 ? rc1 : rc2;
                    if (e != hipSuccess) return lgf_hip_fail(ctx, e, "graph capture");
                    e = hipGraphInstantiate(&gx, graph, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(graph);
                    if (e != hipSuccess) return lgf_hip_fail(ctx, e, "graph instantiate");
                    gv = var;
                }
                HIPTRY(ctx, hipGraphLaunch(gx, ctx->stream));
                n -= 2;  // six stages: the group is back at ping p
            } else {
                rc = enqueue_step(b, p, t0, nt);
                if (rc) return rc;
                p ^= 1;
                n -= 1;
            }
        }
    }
    advance(b, 3 * nsteps);
    add_flow_time(b, eps, nsteps);
    return LGF_OK;
}

extern "C" int lgf_measure(lgf_field* f, int flags, double* E, double* Q, double* Qabs, double* qraw) {
    double obs[LGF_NUM_OBS];
    int rc = lgf_measure_observables(f, flags, obs, qraw);
    if (rc && rc != LGF_E_NAN) return rc;
    if (E) *E = obs[LGF_OBS_E];
    if (Q) *Q = obs[LGF_OBS_Q];
    if (Qabs) *Qabs = obs[LGF_OBS_QABS];
    return rc;
}

extern "C" int lgf_batch_measure(lgf_batch* b, int flags, double* E, double* Q, double* Qabs) {
    if (!b) return LGF_E_ARG;
    std::vector<double> obs(size_t(LGF_NUM_OBS) * b->fields.size());
    int rc = lgf_batch_measure_observables(b, flags, obs.data());
    if (rc && rc != LGF_E_NAN) return rc;
    for (size_t i = 0; i < b->fields.size(); ++i) {
        if (E) E[i] = obs[LGF_NUM_OBS * i + LGF_OBS_E];
        if (Q) Q[i] = obs[LGF_NUM_OBS * i + LGF_OBS_Q];
        if (Qabs) Qabs[i] = obs[LGF_NUM_OBS * i + LGF_OBS_QABS];
    }
    return rc;
}

// Groups of consecutive fields whose line products fit ctx->lines_cap (a
// single larger field forms a group of its own): (first tile, tiles), and
// the line-product bytes the largest group needs.
static std::vector<std::pair<int, int>> measure_chunks(const lgf_batch* b, int compute, size_t* need) {
    const size_t per_tile = line_bytes_per_site(compute) * kTileSites;
    const int n = (int)b->fields.size();
    std::vector<std::pair<int, int>> chunks;
    *need = 0;
    for (int i = 0; i < n;) {
        int t0 = b->h_desc[i].tile0, nt = 0;
        while (i < n && (nt == 0 || size_t(nt + b->h_desc[i].ntiles) * per_tile <= b->ctx->lines_cap))
            nt += b->h_desc[i++].ntiles;
        chunks.emplace_back(t0, nt);
        *need = std::max(*need, size_t(nt) * per_tile);
    }
    return chunks;
}

// device buffer of at least `need` bytes, grown (after a wait on `stream`)
static int grow_buffer(lgf_ctx* ctx, void** buf, size_t* have, size_t need, hipStream_t stream, const char* what) {
    if (need <= *have) return LGF_OK;
    HIPTRY(ctx, hipStreamSynchronize(stream));
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *have = 0;
    if (hipMalloc(buf, need) != hipSuccess) {
        (void)hipGetLastError();
        *buf = nullptr;
        return lgf_fail(ctx, LGF_E_NOMEM, std::string("no device memory for ") + what);
    }
    *have = need;
    return LGF_OK;
}

// On-device t0 tracking after a measurement whose per-field sums are in
// `out`: the fields' flow times travel as kernel arguments (captured at
// enqueue), kT0Chunk fields per launch.
static int enqueue_t0_track(lgf_batch* b, const double* out) {
    if (!b->track_t0) return LGF_OK;
    lgf_ctx* ctx = b->ctx;
    const int n = (int)b->fields.size();
    for (int first = 0; first < n; first += kT0Chunk) {
        T0Args a;
        a.first = first;
        a.n = std::min(kT0Chunk, n - first);
        a.target = b->t0_target;
        for (int k = 0; k < a.n; ++k) a.t[k] = b->fields[first + k]->flow_t;
        HIPTRY(ctx, launch_t0_track(b->d_desc, out, a, b->d_t0, ctx->stream));
    }
    return LGF_OK;
}

extern "C" int lgf_batch_track_t0(lgf_batch* b, double target) {
    if (!b || !(target > 0.0)) return LGF_E_ARG;
    lgf_ctx* ctx = b->ctx;
    if (b->pending) return lgf_fail(ctx, LGF_E_STATE, "a measurement of this batch is still outstanding");
    if (!b->d_t0) {
        hipError_t e = hipMalloc(&b->d_t0, sizeof(double) * 3 * b->fields.size());
        if (e != hipSuccess) return lgf_hip_fail(ctx, e, "t0 records");
    }
    HIPTRY(ctx, launch_t0_reset(b->d_t0, (int)b->fields.size(), ctx->stream));
    b->t0_target = target;
    b->track_t0 = true;
    return LGF_OK;
}

extern "C" int lgf_batch_t0(lgf_batch* b, double* t0) {
    if (!b || !t0) return LGF_E_ARG;
    lgf_ctx* ctx = b->ctx;
    if (!b->track_t0) return lgf_fail(ctx, LGF_E_STATE, "t0 is not tracked on this batch (lgf_batch_track_t0)");
    std::vector<double> rec(3 * b->fields.size());
    HIPTRY(ctx, hipMemcpyAsync(rec.data(), b->d_t0, sizeof(double) * rec.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < b->fields.size(); ++i) t0[i] = rec[3 * i + 2];
    return LGF_OK;
}

// one field's reduction record (the sums of k_reduce) -> lgf_observable values
static bool sums_to_observables(const double* h, int vol, double* o) {
    const double rq = 1.0 / (4.0 * kPi * kPi);  // LGF/Observables/TopologicalCharge.cpp:21
    const double rv = 1.0 / (double)vol;
    o[LGF_OBS_E] = h[0] * rv;
    o[LGF_OBS_Q] = h[1] * rq;
    o[LGF_OBS_QABS] = h[2] * rq;
    o[LGF_OBS_E_IMPROVED] = h[3] * rv;
    o[LGF_OBS_Q_UNIMPROVED] = h[4] * rq;
    bool finite = true;
    for (int k = 0; k < LGF_NUM_OBS; ++k) finite = finite && std::isfinite(o[k]);
    return finite;
}

int observables_of_sums(lgf_ctx* ctx, const double* h, int vol, double* obs) {
    if (!sums_to_observables(h, vol, obs)) return lgf_fail(ctx, LGF_E_NAN, "non-finite energy density or charge");
    return LGF_OK;
}

// per-field reduction records -> lgf_observable values
static int observables_from(lgf_batch* b, const double* h_out, double* obs) {
    bool finite = true;
    for (size_t i = 0; i < b->fields.size(); ++i)
        finite = sums_to_observables(h_out + kOutStride * i, b->fields[i]->vol, obs + LGF_NUM_OBS * i) && finite;
    if (!finite) return lgf_fail(b->ctx, LGF_E_NAN, "non-finite energy density or charge");
    return LGF_OK;
}

extern "C" int lgf_batch_measure_observables(lgf_batch* b, int flags, double* obs) {
    if (!b || !obs) return LGF_E_ARG;
    lgf_ctx* ctx = b->ctx;
    int rc = batch_sync_desc(b);
    if (rc) return rc;
    const int compute = (flags & LGF_MEASURE_FP64) ? LGF_FP64 : b->prec;
    const int n = (int)b->fields.size();
    size_t need;
    const auto chunks = measure_chunks(b, compute, &need);
    rc = grow_buffer(ctx, &ctx->lines, &ctx->lines_bytes, need, ctx->stream, "the clover line products");
    if (rc) return rc;
    for (const auto& c : chunks)
        HIPTRY(ctx, launch_clover(b->prec, compute, b->d_desc, b->d_tiles, c.first, c.second, b->ping, ctx->lines,
                                  b->d_partials, ctx->stream));
    HIPTRY(ctx, launch_reduce(b->d_desc, n, b->d_partials, kCloverParts, kCloverParts, b->d_out, ctx->stream));
    rc = enqueue_t0_track(b, b->d_out);
    if (rc) return rc;
    HIPTRY(ctx, hipMemcpyAsync(b->h_out, b->d_out, sizeof(double) * kOutStride * n, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPTRY(ctx, hipStreamSynchronize(ctx->stream));
    return observables_from(b, b->h_out, obs);
}

// Deferred measurement.  The sweep is enqueued on the context's stream after
// all work enqueued so far, so it sees the links of the current flow time;
// flow steps enqueued afterwards follow it in stream order, and the caller
// collects the result later, so the stream never waits for the host between
// a measurement and the next flow steps.  One measurement may be outstanding
// per batch; it has its own partials and result buffers.
extern "C" int lgf_batch_measure_start(lgf_batch* b, int flags) {
    if (!b) return LGF_E_ARG;
    lgf_ctx* ctx = b->ctx;
    if (b->pending) return lgf_fail(ctx, LGF_E_STATE, "a measurement of this batch is still outstanding");
    int rc = batch_sync_desc(b);
    if (rc) return rc;
    const int n = (int)b->fields.size();
    if (!b->ev_done) {
        HIPTRY(ctx, hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming));
        hipError_t e = hipMalloc(&b->d_partials_m, sizeof(double) * kCloverParts * b->ntiles);
        if (e == hipSuccess) e = hipMalloc(&b->d_out_m, sizeof(double) * kOutStride * n);
        if (e == hipSuccess) e = hipHostMalloc((void**)&b->h_out_m, sizeof(double) * kOutStride * n, hipHostMallocDefault);
        if (e != hipSuccess) return lgf_hip_fail(ctx, e, "measurement buffers");
    }
    const int compute = (flags & LGF_MEASURE_FP64) ? LGF_FP64 : b->prec;
    size_t need;
    const auto chunks = measure_chunks(b, compute, &need);
    rc = grow_buffer(ctx, &ctx->lines, &ctx->lines_bytes, need, ctx->stream, "the clover line products");
    if (rc) return rc;
    for (const auto& c : chunks)
        HIPTRY(ctx, launch_clover(b->prec, compute, b->d_desc, b->d_tiles, c.first, c.second, b->ping, ctx->lines,
                                  b->d_partials_m, ctx->stream));
    HIPTRY(ctx, launch_reduce(b->d_desc, n, b->d_partials_m, kCloverParts, kCloverParts, b->d_out_m, ctx->stream));
    rc = enqueue_t0_track(b, b->d_out_m);
    if (rc) return rc;
    HIPTRY(ctx, hipMemcpyAsync(b->h_out_m, b->d_out_m, sizeof(double) * kOutStride * n, hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPTRY(ctx, hipEventRecord(b->ev_done, ctx->stream));
    b->pending = true;
    return LGF_OK;
}

extern "C" int lgf_batch_measure_finish(lgf_batch* b, double* obs) {
    if (!b || !obs) return LGF_E_ARG;
    lgf_ctx* ctx = b->ctx;
    if (!b->pending) return lgf_fail(ctx, LGF_E_STATE, "no measurement of this batch is outstanding");
    b->pending = false;
    HIPTRY(ctx, hipEventSynchronize(b->ev_done));
    return observables_from(b, b->h_out_m, obs);
}

extern "C" int lgf_measure_start(lgf_field* f, int flags) {
    if (!f) return LGF_E_ARG;
    lgf_batch* b;
    int rc = solo(f, &b);
    if (rc) return rc;
    return lgf_batch_measure_start(b, flags);
}

extern "C" int lgf_measure_finish(lgf_field* f, double* obs) {
    if (!f || !f->solo) return LGF_E_ARG;
    return lgf_batch_measure_finish(f->solo, obs);
}

// ---------------------------------------------------------------------------
// autotuner
// ---------------------------------------------------------------------------
// Times every (tile order, stage-kernel variant) pair on seeded random
// lattices of the given shapes, flowed together as one batch (nsteps RK3
// steps per timing, after one warm-up step, HIP events on the context's
// stream), and makes the fastest pair (after the re-timing below) the
// context's tile order and variant for that precision.  The pairs are searched jointly: the fastest order
// depends on the variant (non-temporal stores change which neighbours stay
// in cache).  ms_per_step (optional) receives one time per variant, at the
// chosen order.
// each candidate is timed kAutotuneReps times and its fastest run counts
constexpr int kAutotuneReps = 3;
// The search's trials are short (a pair's timings differ by a few per cent
// from run to run, as much as the variants differ), so its kRefinePairs
// fastest pairs are timed again with kRefineSteps times the steps, and then
// every variant at the (cache cap, order) of the fastest of those; that
// row's fastest variant is chosen.
constexpr int kRefinePairs = 4;
constexpr int kRefineSteps = 3;

static int time_steps(lgf_ctx* ctx, lgf_batch* b, const double* eps, int nsteps, hipEvent_t e0, hipEvent_t e1,
                      float* ms) {
    if (hipEventRecord(e0, ctx->stream) != hipSuccess) return lgf_fail(ctx, LGF_E_HIP, "event record");
    int rc = lgf_batch_flow_steps(b, eps, nsteps, 1);
    if (rc) return rc;
    if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(ms, e0, e1) != hipSuccess)
        return lgf_fail(ctx, LGF_E_HIP, "event timing");
    return LGF_OK;
}

// fastest of kAutotuneReps timings of nsteps graph-replayed steps, as the
// flow runs them (after two untimed steps, which capture the graphs), of a
// batch of `fields` created with the context's current settings
static int time_batch(lgf_ctx* ctx, const std::vector<lgf_field*>& fields, const std::vector<double>& eps,
                      int nsteps, hipEvent_t e0, hipEvent_t e1, double* ms_step) {
    lgf_batch* b = nullptr;
    int rc = lgf_batch_create(ctx, fields.data(), (int)fields.size(), &b);
    if (!rc) rc = lgf_batch_flow_steps(b, eps.data(), 2, 1);  // captures the batch's step graphs
    *ms_step = 1e300;
    for (int r = 0; r < kAutotuneReps && !rc; ++r) {
        float m = 0.f;
        rc = time_steps(ctx, b, eps.data(), nsteps, e0, e1, &m);
        if (!rc) *ms_step = std::min(*ms_step, (double)m / nsteps);
    }
    if (b) lgf_batch_destroy(b);
    return rc;
}

// Cache-group caps for batches of several lattices.  The (order, variant)
// search runs with the batch as one group and with groups of kGroupSearch
// bytes (stage kernels with non-temporal stores gain nothing from groups,
// the others a quarter on batches that overflow the 256 MB last-level
// cache, so the best pair differs); the cap is then refined at the best pair.
static const size_t kGroupSearch = size_t(224) << 20;
static const size_t kGroupCaps[] = {0, size_t(128) << 20, size_t(192) << 20, size_t(224) << 20,
                                    size_t(256) << 20};

// FP64 variants of another group than the context's current one sum the
// staples in another order: the autotuners leave them out, so that tuning
// never changes an FP64 result
static bool tunable(int precision, int current, int v) {
    return precision != LGF_FP64 || stage_variant_group(precision, v) == stage_variant_group(precision, current);
}

static int autotune_shapes(lgf_ctx* ctx, const int* shapes, int nshapes, int precision, int nsteps, int* best,
                           double* ms_per_step) {
    static const int orders[] = LGF_AUTOTUNE_TILE_ORDERS;
    constexpr int norders = sizeof(orders) / sizeof(orders[0]);
    const int nv = num_stage_variants(precision);
    const int keep_order = ctx->tile_block, keep_variant = ctx->stage_variant[precision];
    const size_t keep_groups = ctx->group_bytes;
    std::vector<lgf_field*> fields;
    int rc = LGF_OK;
    for (int i = 0; i < nshapes && !rc; ++i) {
        lgf_field* f = nullptr;
        rc = lgf_field_create(ctx, shapes + 4 * i, precision, &f);
        if (!rc) {
            fields.push_back(f);
            rc = lgf_field_set_random(f, 0x5eedULL + (uint64_t)i);
        }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = lgf_fail(ctx, LGF_E_HIP, "event creation");
    const std::vector<double> eps(fields.size(), 0.01);
    // tile order and stage variant, jointly, as one group and (several
    // lattices) in cache groups
    const int ncaps = fields.size() > 1 ? 2 : 1;
    const size_t search_caps[2] = {0, kGroupSearch};
    std::vector<double> ms(size_t(ncaps) * norders * nv, HUGE_VAL);
    for (int ci = 0; ci < ncaps && !rc; ++ci) {
        ctx->group_bytes = search_caps[ci];
        for (int oi = 0; oi < norders && !rc; ++oi) {
            ctx->tile_block = orders[oi];
            for (int v = 0; v < nv && !rc; ++v) {
                if (!tunable(precision, keep_variant, v)) continue;
                ctx->stage_variant[precision] = v;
                rc = time_batch(ctx, fields, eps, nsteps, e0, e1, &ms[(size_t(ci) * norders + oi) * nv + v]);
            }
        }
    }
    int bo = 0, bv = 0, bc = 0;
    size_t bg = keep_groups;
    std::vector<double> row(nv, HUGE_VAL);  // every variant at the chosen cap and order, re-timed
    const int long_steps = nsteps <= INT_MAX / kRefineSteps ? nsteps * kRefineSteps : nsteps;
    if (!rc) {
        std::vector<size_t> idx(ms.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
        const size_t k = std::min<size_t>(kRefinePairs, idx.size());
        std::partial_sort(idx.begin(), idx.begin() + k, idx.end(), [&](size_t a, size_t b) { return ms[a] < ms[b]; });
        size_t arg = idx[0];
        double best_ms = HUGE_VAL;
        for (size_t j = 0; j < k && !rc && ms[idx[j]] < HUGE_VAL; ++j) {
            const size_t a = idx[j];
            ctx->group_bytes = search_caps[a / (size_t(norders) * nv)];
            ctx->tile_block = orders[a / nv % norders];
            ctx->stage_variant[precision] = int(a % nv);
            double m = 0;
            rc = time_batch(ctx, fields, eps, long_steps, e0, e1, &m);
            if (!rc && m < best_ms) {
                best_ms = m;
                arg = a;
            }
        }
        bc = int(arg / (size_t(norders) * nv));
        bo = int(arg / nv % norders);
        ctx->group_bytes = search_caps[bc];
        ctx->tile_block = orders[bo];
        for (int v = 0; v < nv && !rc; ++v) {
            if (!tunable(precision, keep_variant, v)) continue;
            ctx->stage_variant[precision] = v;
            rc = time_batch(ctx, fields, eps, long_steps, e0, e1, &row[v]);
        }
        bv = int(std::min_element(row.begin(), row.end()) - row.begin());
        ctx->stage_variant[precision] = bv;
        if (!rc && fields.size() > 1) {
            // refine the cache-group cap at the chosen pair
            best_ms = row[bv];
            bg = search_caps[bc];
            for (size_t cap : kGroupCaps) {
                if (cap == search_caps[0] || cap == search_caps[1]) continue;
                ctx->group_bytes = cap;
                double m = 0;
                rc = time_batch(ctx, fields, eps, long_steps, e0, e1, &m);
                if (rc) break;
                if (m < best_ms) {
                    best_ms = m;
                    bg = cap;
                }
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (lgf_field* f : fields) lgf_field_destroy(f);
    if (rc) {
        ctx->tile_block = keep_order;
        ctx->stage_variant[precision] = keep_variant;
        ctx->group_bytes = keep_groups;
        return rc;
    }
    ctx->tile_block = orders[bo];
    ctx->stage_variant[precision] = bv;
    ctx->group_bytes = bg;
    if (best) *best = bv;
    if (ms_per_step)
        for (int v = 0; v < nv; ++v) ms_per_step[v] = row[v];
    return LGF_OK;
}

extern "C" int lgf_ctx_set_batch_group_bytes(lgf_ctx* ctx, size_t bytes) {
    if (!ctx) return LGF_E_ARG;
    ctx->group_bytes = bytes;
    return LGF_OK;
}

extern "C" int lgf_ctx_get_batch_group_bytes(lgf_ctx* ctx, size_t* bytes) {
    if (!ctx || !bytes) return LGF_E_ARG;
    *bytes = ctx->group_bytes;
    return LGF_OK;
}

extern "C" int lgf_autotune(lgf_ctx* ctx, const int shape[4], int precision, int nsteps, int* best,
                            double* ms_per_step) {
    if (!ctx || !shape || nsteps <= 0 || (precision != LGF_FP64 && precision != LGF_FP32)) return LGF_E_ARG;
    return autotune_shapes(ctx, shape, 1, precision, nsteps, best, ms_per_step);
}

extern "C" int lgf_autotune_batch(lgf_ctx* ctx, const int* shapes, int nshapes, int precision, int nsteps, int* best,
                                  double* ms_per_step) {
    if (!ctx || !shapes || nshapes <= 0 || nsteps <= 0 || (precision != LGF_FP64 && precision != LGF_FP32))
        return LGF_E_ARG;
    return autotune_shapes(ctx, shapes, nshapes, precision, nsteps, best, ms_per_step);
}
