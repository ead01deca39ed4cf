// Device-side lattice geometry and the per-field descriptor shared by all
// kernels.
//
// Site index (identical to the reference's global index, LGF/Lattice/
// Lattice.cpp:423-436): s = ((x0*N1 + x1)*N2 + x2)*N3 + x3, last coordinate
// fastest; periodic in every direction (the reference's "torus").
//
// Link storage is structure-of-arrays of short vectors.  For each direction
// mu the 18 reals c = 2*(3*row+col) + {0: Re, 1: Im} of all sites form a
// block of 18*stride elements, split into planes of vectors so that each
// lane of a wavefront (64 consecutive sites) fetches a link with few wide
// loads and the wavefront's access to a plane is one contiguous run:
//   FP32: planes 0..3 hold float4 (components 4k..4k+3), plane 4 float2
//         (components 16, 17): 5 loads per link;
//   FP64: planes 0..8 hold double2 (one complex element each): 9 loads.
// `stride` is the volume rounded up to a multiple of 64.  The accumulator of
// the low-storage RK3 scheme uses the same layout.
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
    double rcp[4];   // (1 + 2^-40) / n[d], for division-free coordinate decomposition
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

// t / n for 0 <= t < 2^31 and 2 <= n < 2^16 with one double multiply: rcp
// is (1 + 2^-40)/n rounded; the bias keeps exact quotients from rounding
// down and is far too small (t/n * 2^-40 < 1/n) to lift t/n across the next
// integer.
LGF_DEV int fdiv(int t, double rcp) { return (int)((double)t * rcp); }

LGF_DEV Geo make_geo(const FieldDesc& d, int s) {
    Geo g;
    g.s = s;
    g.n0 = d.n[0];
    g.n1 = d.n[1];
    g.n2 = d.n[2];
    g.n3 = d.n[3];
    int t = s, q = fdiv(t, d.rcp[3]);
    g.x3 = t - q * g.n3;
    t = q;
    q = fdiv(t, d.rcp[2]);
    g.x2 = t - q * g.n2;
    t = q;
    q = fdiv(t, d.rcp[1]);
    g.x1 = t - q * g.n1;
    g.x0 = q;
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

// Link arrays live in global memory; loads and stores go through
// address-space-1 pointers built from a wavefront-uniform plane base (SGPRs)
// plus the lane's 32-bit byte offset, which the compiler maps onto global
// loads with a scalar base.
template <typename S>
using gptr = __attribute__((address_space(1))) S*;

// native clang vector types (the HIP vector wrappers do not convert across
// address spaces)
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef float vf2 __attribute__((ext_vector_type(2)));
typedef double vd2 __attribute__((ext_vector_type(2)));

template <typename V, typename S>
LGF_DEV V ld(const S* plane, uint32_t byte_off) {
    return *(const gptr<const V>)((const __attribute__((address_space(1))) char*)plane + byte_off);
}
template <typename V, typename S>
LGF_DEV void st_(S* plane, uint32_t byte_off, const V& v) {
    *(gptr<V>)((__attribute__((address_space(1))) char*)plane + byte_off) = v;
}
// the same store marked non-temporal: data written once per sweep that no
// other wave of the sweep reads (streaming, so it does not displace reused
// neighbour links from L2)
template <typename V, typename S>
LGF_DEV void st_nt(S* plane, uint32_t byte_off, const V& v) {
    __builtin_nontemporal_store(v, (gptr<V>)((__attribute__((address_space(1))) char*)plane + byte_off));
}

// element offset of component c of link (mu, s): the layout described above
__host__ __device__ inline int64_t link_elem_offset(bool fp32, int64_t stride, int mu, int64_t s, int c) {
    const int64_t b = (int64_t)mu * 18 * stride;
    if (fp32) return c < 16 ? b + (int64_t)(c >> 2) * 4 * stride + s * 4 + (c & 3) : b + 16 * stride + s * 2 + (c - 16);
    return b + (int64_t)(c >> 1) * 2 * stride + s * 2 + (c & 1);
}

template <typename S, typename C>
LGF_DEV M3<C> load_link(const S* __restrict__ base, int64_t stride, int mu, int s) {
    M3<C> m;
    const S* p = base + (int64_t)mu * 18 * stride;
    if constexpr (sizeof(S) == 4) {
        const uint32_t o4 = (uint32_t)s * 16u, o2 = (uint32_t)s * 8u;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const vf4 v = ld<vf4>(p + (int64_t)k * 4 * stride, o4);
            m.v[4 * k] = (C)v.x;
            m.v[4 * k + 1] = (C)v.y;
            m.v[4 * k + 2] = (C)v.z;
            m.v[4 * k + 3] = (C)v.w;
        }
        const vf2 w = ld<vf2>(p + 16 * stride, o2);
        m.v[16] = (C)w.x;
        m.v[17] = (C)w.y;
    } else {
        const uint32_t o = (uint32_t)s * 16u;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const vd2 v = ld<vd2>(p + (int64_t)k * 2 * stride, o);
            m.v[2 * k] = (C)v.x;
            m.v[2 * k + 1] = (C)v.y;
        }
    }
    return m;
}

// Rows 0 and 1 of a link (the first 12 reals, 2/3 of the bytes); the
// caller rebuilds row 2 where it needs it (su3_third_row, exact for SU(3)).
template <typename S, typename C>
LGF_DEV M3<C> load_link_rows01(const S* __restrict__ base, int64_t stride, int mu, int s) {
    M3<C> m;
    const S* p = base + (int64_t)mu * 18 * stride;
    if constexpr (sizeof(S) == 4) {
        const uint32_t o4 = (uint32_t)s * 16u;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const vf4 v = ld<vf4>(p + (int64_t)k * 4 * stride, o4);
            m.v[4 * k] = (C)v.x;
            m.v[4 * k + 1] = (C)v.y;
            m.v[4 * k + 2] = (C)v.z;
            m.v[4 * k + 3] = (C)v.w;
        }
    } else {
        const uint32_t o = (uint32_t)s * 16u;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const vd2 v = ld<vd2>(p + (int64_t)k * 2 * stride, o);
            m.v[2 * k] = (C)v.x;
            m.v[2 * k + 1] = (C)v.y;
        }
    }
    return m;
}

// Columns 0 and 1 of an FP32 link: components 0..15, the four float4 planes
// (the float2 plane holds only column 2 of row 2).
template <typename C>
LGF_DEV M3<C> load_link_cols01(const float* __restrict__ base, int64_t stride, int mu, int s) {
    M3<C> m;
    const float* p = base + (int64_t)mu * 18 * stride;
    const uint32_t o4 = (uint32_t)s * 16u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const vf4 v = ld<vf4>(p + (int64_t)k * 4 * stride, o4);
        m.v[4 * k] = (C)v.x;
        m.v[4 * k + 1] = (C)v.y;
        m.v[4 * k + 2] = (C)v.z;
        m.v[4 * k + 3] = (C)v.w;
    }
    return m;
}

template <typename S, typename C>
LGF_DEV void store_link_rows01(S* __restrict__ base, int64_t stride, int mu, int s, const M3<C>& m) {
    S* p = base + (int64_t)mu * 18 * stride;
    if constexpr (sizeof(S) == 4) {
        const uint32_t o4 = (uint32_t)s * 16u;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            st_(p + (int64_t)k * 4 * stride, o4,
                vf4{(float)m.v[4 * k], (float)m.v[4 * k + 1], (float)m.v[4 * k + 2], (float)m.v[4 * k + 3]});
    } else {
        const uint32_t o = (uint32_t)s * 16u;
#pragma unroll
        for (int k = 0; k < 6; ++k)
            st_(p + (int64_t)k * 2 * stride, o, vd2{(double)m.v[2 * k], (double)m.v[2 * k + 1]});
    }
}

template <typename S, typename C, bool NT = false>
LGF_DEV void store_link(S* __restrict__ base, int64_t stride, int mu, int s, const M3<C>& m) {
    S* p = base + (int64_t)mu * 18 * stride;
    if constexpr (sizeof(S) == 4 && NT) {
        const uint32_t o4 = (uint32_t)s * 16u, o2 = (uint32_t)s * 8u;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            st_nt(p + (int64_t)k * 4 * stride, o4,
                  vf4{(float)m.v[4 * k], (float)m.v[4 * k + 1], (float)m.v[4 * k + 2], (float)m.v[4 * k + 3]});
        st_nt(p + 16 * stride, o2, vf2{(float)m.v[16], (float)m.v[17]});
    } else if constexpr (sizeof(S) == 4) {
        const uint32_t o4 = (uint32_t)s * 16u, o2 = (uint32_t)s * 8u;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            st_(p + (int64_t)k * 4 * stride, o4,
                vf4{(float)m.v[4 * k], (float)m.v[4 * k + 1], (float)m.v[4 * k + 2], (float)m.v[4 * k + 3]});
        st_(p + 16 * stride, o2, vf2{(float)m.v[16], (float)m.v[17]});
    } else if constexpr (NT) {
        const uint32_t o = (uint32_t)s * 16u;
#pragma unroll
        for (int k = 0; k < 9; ++k)
            st_nt(p + (int64_t)k * 2 * stride, o, vd2{(double)m.v[2 * k], (double)m.v[2 * k + 1]});
    } else {
        const uint32_t o = (uint32_t)s * 16u;
#pragma unroll
        for (int k = 0; k < 9; ++k)
            st_(p + (int64_t)k * 2 * stride, o, vd2{(double)m.v[2 * k], (double)m.v[2 * k + 1]});
    }
}

// The RK3 accumulator holds traceless anti-Hermitian matrices whose lower
// triangle is computed as -conj of the upper one and whose diagonal real
// parts are exactly zero (su3.hpp, ta_scaled, and linear combinations of
// such matrices), so 9 reals per link hold it exactly:
//   a = (Im X00, Im X11, Im X22, Re X01, Im X01, Re X02, Im X02, Re X12, Im X12)
// stored per direction mu in a block of 9*stride elements:
//   FP32: two float4 planes (a0..a3, a4..a7) and one float plane (a8);
//   FP64: four double2 planes and one double plane.
template <typename S, typename C>
LGF_DEV void load_ta_raw(const S* __restrict__ base, int64_t stride, int mu, int s, C* a) {
    const S* p = base + (int64_t)mu * 9 * stride;
    if constexpr (sizeof(S) == 4) {
        const vf4 v0 = ld<vf4>(p, (uint32_t)s * 16u), v1 = ld<vf4>(p + 4 * stride, (uint32_t)s * 16u);
        a[0] = v0.x, a[1] = v0.y, a[2] = v0.z, a[3] = v0.w;
        a[4] = v1.x, a[5] = v1.y, a[6] = v1.z, a[7] = v1.w;
        a[8] = ld<float>(p + 8 * stride, (uint32_t)s * 4u);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const vd2 v = ld<vd2>(p + (int64_t)k * 2 * stride, (uint32_t)s * 16u);
            a[2 * k] = (C)v.x;
            a[2 * k + 1] = (C)v.y;
        }
        a[8] = (C)ld<double>(p + 8 * stride, (uint32_t)s * 8u);
    }
}

template <typename C>
LGF_DEV M3<C> expand_ta(const C* a) {
    M3<C> m;
    m.v[0] = C(0), m.v[1] = a[0], m.v[8] = C(0), m.v[9] = a[1], m.v[16] = C(0), m.v[17] = a[2];
    m.v[2] = a[3], m.v[3] = a[4], m.v[6] = -a[3], m.v[7] = a[4];      // (0,1), (1,0)
    m.v[4] = a[5], m.v[5] = a[6], m.v[12] = -a[5], m.v[13] = a[6];    // (0,2), (2,0)
    m.v[10] = a[7], m.v[11] = a[8], m.v[14] = -a[7], m.v[15] = a[8];  // (1,2), (2,1)
    return m;
}

template <typename S, typename C>
LGF_DEV M3<C> load_ta(const S* __restrict__ base, int64_t stride, int mu, int s) {
    C a[9];
    load_ta_raw<S, C>(base, stride, mu, s, a);
    return expand_ta(a);
}

template <typename S, typename C, bool NT = false>
LGF_DEV void store_ta(S* __restrict__ base, int64_t stride, int mu, int s, const M3<C>& m) {
    S* p = base + (int64_t)mu * 9 * stride;
    if constexpr (sizeof(S) == 4 && NT) {
        st_nt(p, (uint32_t)s * 16u, vf4{(float)m.v[1], (float)m.v[9], (float)m.v[17], (float)m.v[2]});
        st_nt(p + 4 * stride, (uint32_t)s * 16u, vf4{(float)m.v[3], (float)m.v[4], (float)m.v[5], (float)m.v[10]});
        st_nt(p + 8 * stride, (uint32_t)s * 4u, (float)m.v[11]);
    } else if constexpr (sizeof(S) == 4) {
        st_(p, (uint32_t)s * 16u, vf4{(float)m.v[1], (float)m.v[9], (float)m.v[17], (float)m.v[2]});
        st_(p + 4 * stride, (uint32_t)s * 16u, vf4{(float)m.v[3], (float)m.v[4], (float)m.v[5], (float)m.v[10]});
        st_(p + 8 * stride, (uint32_t)s * 4u, (float)m.v[11]);
    } else if constexpr (NT) {
        st_nt(p, (uint32_t)s * 16u, vd2{(double)m.v[1], (double)m.v[9]});
        st_nt(p + 2 * stride, (uint32_t)s * 16u, vd2{(double)m.v[17], (double)m.v[2]});
        st_nt(p + 4 * stride, (uint32_t)s * 16u, vd2{(double)m.v[3], (double)m.v[4]});
        st_nt(p + 6 * stride, (uint32_t)s * 16u, vd2{(double)m.v[5], (double)m.v[10]});
        st_nt(p + 8 * stride, (uint32_t)s * 8u, (double)m.v[11]);
    } else {
        st_(p, (uint32_t)s * 16u, vd2{(double)m.v[1], (double)m.v[9]});
        st_(p + 2 * stride, (uint32_t)s * 16u, vd2{(double)m.v[17], (double)m.v[2]});
        st_(p + 4 * stride, (uint32_t)s * 16u, vd2{(double)m.v[3], (double)m.v[4]});
        st_(p + 6 * stride, (uint32_t)s * 16u, vd2{(double)m.v[5], (double)m.v[10]});
        st_(p + 8 * stride, (uint32_t)s * 8u, (double)m.v[11]);
    }
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
