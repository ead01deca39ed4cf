// Host-callable launchers of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lgf_flow.h"
#include "lattice.hpp"

namespace lgf {

constexpr int kPrecFP64 = 0;  // FP64 links, accumulator and arithmetic
constexpr int kPrecFP32 = 1;  // FP32 links/arithmetic, FP64 reductions and reunitarisation
constexpr int kCloverParts = 5;  // clover partials per tile: e, q, |q|, improved e, 1x1 q
constexpr int kOutStride = 8;    // doubles per field in reduction outputs

int num_stage_variants(int prec);
// stage variants of one group compute bitwise identical links; groups agree
// to rounding in the path's precision
int stage_variant_group(int prec, int variant);
hipError_t launch_rk3_stage_variant(int prec, int variant, int stage, const FieldDesc* fds, const int2* tiles,
                                    int ntiles, int ping, hipStream_t stream);
hipError_t launch_rk3_stage(int prec, int stage, const FieldDesc* fds, const int2* tiles, int ntiles, int ping,
                            hipStream_t stream);
// bytes of line products (k_lines) per site for clover arithmetic in `compute`
size_t line_bytes_per_site(int compute);
// clover observables of the tiles [t0, t0 + ntiles) of a batch's tile list;
// those tiles must cover whole fields, and `lines` must hold
// ntiles * 64 * line_bytes_per_site(compute) bytes
hipError_t launch_clover(int prec, int compute, const FieldDesc* fds, const int2* tiles, int t0, int ntiles, int ping,
                         void* lines, double* partials, hipStream_t stream);
// per-field sums of tile partials (records of pstride doubles, components
// 0..ncomp-1) into out[kOutStride * field + component]
hipError_t launch_reduce(const FieldDesc* fds, int nfields, const double* partials, int pstride, int ncomp, double* out,
                         hipStream_t stream);
// On-device t0 tracking: for each field of a batch, after a measurement
// whose per-field sums are in `out` (kOutStride doubles per field, E sum
// first), with flow times t[i]: t0 = the linear interpolation of t^2 E = target
// between the previous and this measurement, at the first crossing.  rec holds
// 3 doubles per field: previous t, previous t^2 E, t0 (NaN = none yet).
constexpr int kT0Chunk = 32;  // fields per launch (their flow times travel as kernel arguments)
struct T0Args {
    double t[kT0Chunk];
    double target;
    int first, n;
};
hipError_t launch_t0_track(const FieldDesc* fds, const double* out, const T0Args& a, double* rec, hipStream_t stream);
hipError_t launch_t0_reset(double* rec, int nfields, hipStream_t stream);
hipError_t launch_diag(int prec, const FieldDesc* fds, const int2* tiles, int ntiles, int ping, double* partials,
                       int nfields, double* out, hipStream_t stream);
hipError_t launch_to_soa(int prec, const double* src, void* dst, int64_t stride, int s0, int n, hipStream_t stream);
hipError_t launch_from_soa(int prec, const void* src, double* dst, int64_t stride, int s0, int n,
                           hipStream_t stream);
hipError_t launch_convert(int prec_in, const void* src, int prec_out, void* dst, int64_t stride, int vol,
                          hipStream_t stream);
hipError_t launch_heatbath_iteration(int prec, const FieldDesc* fds, int vol, int ping, double beta, uint64_t seed,
                                     uint32_t sweep, int or_per_hb, hipStream_t stream);
hipError_t launch_hot(int prec, void* dst, int64_t stride, int vol, uint64_t seed, hipStream_t stream);
hipError_t launch_cold(int prec, void* dst, int64_t stride, int vol, hipStream_t stream);

}  // namespace lgf
