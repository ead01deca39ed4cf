// Device-side lattice geometry and the per-field descriptor shared by all
// kernels.
//
// Site index (identical to the reference's global index, LGF/Lattice/
// Lattice.cpp:423-436): s = ((x0*N1 + x1)*N2 + x2)*N3 + x3, last coordinate
// fastest; periodic in every direction (the reference's "torus").
//
// Link storage is structure-of-arrays, one plane per (mu, real component):
//   links[(mu*18 + c)*stride + s],  c = 2*(3*row+col) + {0: Re, 1: Im}
// so that the 64 lanes of a wavefront, which hold 64 consecutive sites, read
// each component with one coalesced 256/512-byte access.  `stride` is the
// volume rounded up to a multiple of 64.  The accumulator of the low-storage
// RK3 scheme uses the same layout.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "su3.hpp"

namespace lgf {

constexpr int kTileSites = 64;  // sites per tile = one wavefront

struct FieldDesc {
    void* links[2];  // ping-pong link buffers
    void* acc;       // RK3 accumulator X
    double* qx;      // optional per-site charge density output (may be null)
    int n[4];
    int vol;
    int ntiles;
    int64_t stride;
    double eps;      // flow step of this field
    int tile0;       // first global tile (block) index of this field
    int flip;        // the current link buffer is links[ping ^ flip]
};

// Coordinates, extents and index strides of one site, as scalars (the
// direction arguments below are wavefront-uniform, so the selects compile to
// scalar code and nothing is indexed dynamically).
struct Geo {
    int x0, x1, x2, x3;
    int n0, n1, n2, n3;
    int s;
};

LGF_DEV Geo make_geo(const FieldDesc& d, int s) {
    Geo g;
    g.s = s;
    g.n0 = d.n[0];
    g.n1 = d.n[1];
    g.n2 = d.n[2];
    g.n3 = d.n[3];
    int t = s;
    g.x3 = t % g.n3;
    t /= g.n3;
    g.x2 = t % g.n2;
    t /= g.n2;
    g.x1 = t % g.n1;
    g.x0 = t / g.n1;
    return g;
}

LGF_DEV int geo_x(const Geo& g, int d) { return d == 0 ? g.x0 : d == 1 ? g.x1 : d == 2 ? g.x2 : g.x3; }
LGF_DEV int geo_n(const Geo& g, int d) { return d == 0 ? g.n0 : d == 1 ? g.n1 : d == 2 ? g.n2 : g.n3; }
LGF_DEV int geo_str(const Geo& g, int d) {
    return d == 3 ? 1 : d == 2 ? g.n3 : d == 1 ? g.n3 * g.n2 : g.n3 * g.n2 * g.n1;
}

// offset of the neighbour one step forward (+1) or back (-1) along d
LGF_DEV int nb_off(const Geo& g, int d, int sgn) {
    const int xd = geo_x(g, d), nd = geo_n(g, d), sd = geo_str(g, d);
    if (sgn > 0) return (xd == nd - 1) ? -(nd - 1) * sd : sd;
    return (xd == 0) ? (nd - 1) * sd : -sd;
}

// move g one step along +-d (periodic)
LGF_DEV void geo_move(Geo& g, int d, int sgn) {
    g.s += nb_off(g, d, sgn);
    const int nd = geo_n(g, d);
    int v = geo_x(g, d) + sgn;
    v = v < 0 ? nd - 1 : (v == nd ? 0 : v);
    if (d == 0) g.x0 = v;
    else if (d == 1) g.x1 = v;
    else if (d == 2) g.x2 = v;
    else g.x3 = v;
}

template <typename S, typename C>
LGF_DEV M3<C> load_link(const S* __restrict__ base, int64_t stride, int mu, int s) {
    M3<C> m;
    const S* p = base + (int64_t)mu * 18 * stride + s;
#pragma unroll
    for (int c = 0; c < 18; ++c) m.v[c] = (C)p[c * stride];
    return m;
}

template <typename S, typename C>
LGF_DEV void store_link(S* __restrict__ base, int64_t stride, int mu, int s, const M3<C>& m) {
    S* p = base + (int64_t)mu * 18 * stride + s;
#pragma unroll
    for (int c = 0; c < 18; ++c) p[c * stride] = (S)m.v[c];
}

// Block -> tile mapping that keeps each XCD on a contiguous run of tiles.
// Workgroups are dispatched to the 8 XCDs round-robin by blockIdx.x, and each
// XCD has its own L2, so neighbouring tiles (which share halo links) should
// land on the same XCD.
LGF_DEV int xcd_remap(int i, int n) {
    const int nx = 8;
    int q = n / nx, r = n % nx;
    int x = i % nx, k = i / nx;
    return x * q + (x < r ? x : r) + k;
}

}  // namespace lgf
