// Allocation-free variant of lgf_measure_observables (include/lgf_flow.h,
// lgf_measure_observables_noalloc), for callers that cannot allocate device
// memory: code under graph capture, tight serving loops, harnesses that
// forbid hipMalloc.
//
// The measurement's scratch -- the field descriptor and tile list its kernels
// read, the clover line products (k_lines), the per-tile partial sums
// (k_clover) and the per-field sums (k_reduce) -- lives in one file-scope
// device array of kScratchBytes instead of buffers the library allocates.
// Each GPU holds its own copy of the array for the life of the process.  Its
// device address is looked up once per device with hipGetSymbolAddress, with
// that device current, and kept in g_scratch[device]; host code passes only
// that address to kernels and copies, never the array's name (which in host
// code is a host-side stand-in).
//
// Nothing in the array has to start at zero: every call writes the
// descriptor and the whole tile list (k_noalloc_setup), k_lines writes every
// line product and k_clover every tile's partial record before k_reduce reads
// them.  Calls share the array, so calls on one GPU are serialised by a
// per-device mutex held until the call's results are back on the host;
// another process has its own copy of the array.
//
// The tiles are taken in lexicographic order.  Partial sums are stored and
// reduced by tile index, not by traversal order, and line products by site,
// so the results are bitwise those of lgf_measure_observables on the same
// field, whatever tile order its context uses.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/lgf_flow.h"
#include "capi_internal.hpp"
#include "kernels.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(256) << 20;
constexpr int kMaxDevices = 64;

// offsets within the scratch array, each 256-byte aligned
constexpr size_t kDescOff = 0;
constexpr size_t kOutOff = 256;
constexpr size_t kTilesOff = 512;

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

struct Layout {
    size_t partials, lines, end;
};

Layout layout(int ntiles, int compute) {
    Layout l;
    l.partials = align256(kTilesOff + sizeof(int2) * size_t(ntiles));
    l.lines = align256(l.partials + sizeof(double) * lgf::kCloverParts * size_t(ntiles));
    l.end = l.lines + lgf::line_bytes_per_site(compute) * lgf::kTileSites * size_t(ntiles);
    return l;
}

std::mutex g_mu[kMaxDevices];
void* g_scratch[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) char scratch[kScratchBytes];

// the descriptor of the one field measured and its tile list, tile i = (0, i)
__global__ __launch_bounds__(256) void k_noalloc_setup(lgf::FieldDesc d, lgf::FieldDesc* __restrict__ dst,
                                                       int2* __restrict__ tiles, int ntiles) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *dst = d;
    if (i < ntiles) tiles[i] = make_int2(0, i);
}

// the largest field (sites) the scratch array can measure with clover
// arithmetic in `compute`
static int64_t max_sites(int compute) {
    int64_t lo = 0, hi = int64_t(1) << 28;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) / 2;
        const int nt = (int)((mid + lgf::kTileSites - 1) / lgf::kTileSites);
        if (layout(nt, compute).end <= kScratchBytes) lo = mid;
        else hi = mid - 1;
    }
    return lo / lgf::kTileSites * lgf::kTileSites;
}

extern "C" int lgf_measure_observables_noalloc(lgf_field* f, int flags, double* obs) {
    if (!f || !obs) return LGF_E_ARG;
    lgf_ctx* ctx = f->ctx;
    const int dev = ctx->device;
    if (dev < 0 || dev >= kMaxDevices) return lgf_fail(ctx, LGF_E_UNSUPPORTED, "device index beyond the scratch table");
    const int compute = (flags & LGF_MEASURE_FP64) ? LGF_FP64 : f->prec;
    const Layout lay = layout(f->ntiles, compute);
    if (lay.end > kScratchBytes)
        return lgf_fail(ctx, LGF_E_ARG,
                        "field of " + std::to_string(f->vol) + " sites needs " + std::to_string(lay.end) +
                            " bytes of measurement scratch; the allocation-free measurement has " +
                            std::to_string(kScratchBytes) + " (at most " + std::to_string(max_sites(compute)) +
                            " sites with " + (compute == LGF_FP64 ? "FP64" : "FP32") + " clover arithmetic)");
    std::lock_guard<std::mutex> lk(g_mu[dev]);
    hipError_t e = hipSetDevice(dev);
    if (e != hipSuccess) return lgf_hip_fail(ctx, e, "hipSetDevice");
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return lgf_hip_fail(ctx, e, "hipGetSymbolAddress(scratch)");
        g_scratch[dev] = p;
    }
    char* base = (char*)g_scratch[dev];
    lgf::FieldDesc d{};
    d.links[0] = f->links[0];
    d.links[1] = f->links[1];
    d.acc = f->acc;
    d.qx = nullptr;
    for (int k = 0; k < 4; ++k) {
        d.n[k] = f->shape[k];
        d.rcp[k] = (1.0 + 1.0 / 1099511627776.0) / (double)f->shape[k];
    }
    d.vol = f->vol;
    d.ntiles = f->ntiles;
    d.stride = f->stride;
    d.eps = 0.0;
    d.tile0 = 0;
    d.flip = f->cur;  // launched with ping 0: the current links are links[cur]
    lgf::FieldDesc* ddesc = (lgf::FieldDesc*)(base + kDescOff);
    double* dout = (double*)(base + kOutOff);
    int2* dtiles = (int2*)(base + kTilesOff);
    double* dparts = (double*)(base + lay.partials);
    void* dlines = base + lay.lines;
    hipLaunchKernelGGL(k_noalloc_setup, dim3((f->ntiles + 255) / 256), dim3(256), 0, ctx->stream, d, ddesc, dtiles,
                       f->ntiles);
    e = hipGetLastError();
    if (e != hipSuccess) return lgf_hip_fail(ctx, e, "measurement set-up launch");
    e = lgf::launch_clover(f->prec, compute, ddesc, dtiles, 0, f->ntiles, 0, dlines, dparts, ctx->stream);
    if (e != hipSuccess) return lgf_hip_fail(ctx, e, "clover launch");
    e = lgf::launch_reduce(ddesc, 1, dparts, lgf::kCloverParts, lgf::kCloverParts, dout, ctx->stream);
    if (e != hipSuccess) return lgf_hip_fail(ctx, e, "reduction launch");
    double h[lgf::kOutStride];
    e = hipMemcpyAsync(h, dout, sizeof h, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return lgf_hip_fail(ctx, e, "measurement result copy");
    return observables_of_sums(ctx, h, f->vol, obs);
}
