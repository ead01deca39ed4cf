// gfx950 kernels of the Wilson-flow / clover-observable path.
//
// k_rk3_stage  one fused Luescher RK3 stage for every link of every field in
//              a batch: staple sum, traceless anti-Hermitian projection
//              (flow gradient Z), low-storage accumulator update, SU(3)
//              exponential and left-multiplication into the other ping-pong
//              buffer.  Follows LGF/Gradient_Flow/GradientFlow.cpp:22-60,
//              LGF/Action/Gauge/Wilson.cpp:55-71, LGF/Field/SU3_field.cpp:519-530.
//              The reference keeps Z0, Z1, Z2; with X0 = Z0,
//              X1 = 8/9 Z1 - 17/36 X0 the three exponents are 1/4 eps X0,
//              eps X1 and eps (3/4 Z2 - X1), which is the same integrator
//              with one accumulator (also the form of the thesis algorithm,
//              doc/thesis.tex:2569-2600).
// k_rk3_stage_pl  the same stage with each site's six plaquettes computed
//              once for its four links (the four direction waves of a
//              workgroup share them in LDS), in both precisions;
//              k_rk3_stage_pb the mixed-precision form that also shares the
//              backward links.
// k_clover     1x1 clover F (energy density) and the improved F (19/9 1x1 +
//              1/36 2x2 - 32/45 1x2 + 1/15 1x3, LGF/Field/SU3_field.cpp:533-797)
//              for the plane pairs of the topological charge; per-site q(x),
//              per-tile FP64 partial sums.
// k_reduce     deterministic per-field FP64 sums of the tile partials.
// plus layout conversion (file order <-> SoA), cold start, plaquette and
// unitarity diagnostics.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "expm.hpp"
#include "kernels.hpp"
#include "lattice.hpp"
#include "philox.hpp"
#include "staple.hpp"
#include "su3.hpp"

namespace lgf {

// ---------------------------------------------------------------------------
// RK3 stage
// ---------------------------------------------------------------------------
// One link update of one RK3 stage, the direction MU a compile-time
// constant (the kernel switches on the wave-uniform direction once).
template <typename S, typename C, int STAGE, int EXPMODE, int MU>
LGF_DEV void stage_link(const S* __restrict__ U, S* __restrict__ Un, S* __restrict__ Xa, int64_t st, const Geo& g,
                        C eps) {
    const M3<C> A = staple_ct<S, C, MU>(U, st, g);
    // the link and the accumulator are loaded after the staple sum (their
    // addresses depend on it), so they do not hold registers through it
    const int s = after_value(g.s, A.v[17]);
    const M3<C> Umu = load_link<S, C>(U, st, MU, s);
    const M3<C> Z = ta_of_product(Umu, A, C(-0.5));

    M3<C> E;
    if (STAGE == 0) {
        store_ta<S, C>(Xa, st, MU, s, Z);
#pragma unroll
        for (int i = 0; i < 18; ++i) E.v[i] = (C(0.25) * eps) * Z.v[i];
    } else if (STAGE == 1) {
        const M3<C> X0 = load_ta<S, C>(Xa, st, MU, after_value(s, Z.v[17]));
        M3<C> X1;
#pragma unroll
        for (int i = 0; i < 18; ++i) X1.v[i] = (C)(8.0 / 9.0) * Z.v[i] - (C)(17.0 / 36.0) * X0.v[i];
        store_ta<S, C>(Xa, st, MU, s, X1);
#pragma unroll
        for (int i = 0; i < 18; ++i) E.v[i] = eps * X1.v[i];
    } else {
        const M3<C> X1 = load_ta<S, C>(Xa, st, MU, after_value(s, Z.v[17]));
#pragma unroll
        for (int i = 0; i < 18; ++i) E.v[i] = eps * ((C)(3.0 / 4.0) * Z.v[i] - X1.v[i]);
    }

    // the exponential by its power series: 13 terms in the FP64 path
    // (truncation below 1e-20), 7 in the mixed-precision one (below 1e-9)
    const M3<C> ex = exp_ah_series<C, EXPMODE == 0 ? 13 : 7>(E);
    M3<C> Unew = su3_nn(ex, Umu);
    if constexpr (EXPMODE == 1 && STAGE == 2) {
        // once per RK3 step: project back onto SU(3) in double precision
        M3<double> Ud;
#pragma unroll
        for (int i = 0; i < 18; ++i) Ud.v[i] = (double)Unew.v[i];
        reunitarize<true>(Ud);
        store_link<S, double>(Un, st, MU, s, Ud);
    } else {
        store_link<S, C>(Un, st, MU, s, Unew);
    }
}

// Launch shapes.  BY = waves per workgroup along y, MPT = directions each
// thread updates, one after the other:
//   BY 4, MPT 1: block 64x4, mu = threadIdx.y (the four links of a site share
//                their neighbours in L1/L2);
//   BY 1, MPT 1: block 64, mu = blockIdx.y;
//   BY 2, MPT 2 and BY 1, MPT 4: a thread updates 2 (4) directions of its
//                site in turn, so an XCD keeps 2 (4) times fewer sites in
//                flight.
// WPE = minimum waves per SIMD the register budget must allow.
template <typename S, typename C, int STAGE, int EXPMODE, int WPE, int BY, int MPT>
__global__ __launch_bounds__(64 * BY) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_rk3_stage(
    const FieldDesc* __restrict__ fds, const int2* __restrict__ tiles, int ping) {
    static_assert(MPT == 1 || BY * MPT == 4, "a workgroup covers the four directions of its sites");
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int2 tm = tiles[t];
    const FieldDesc& d = fds[tm.x];
    const int s = tm.y * kTileSites + threadIdx.x;
    const int mu0 = MPT > 1 ? __builtin_amdgcn_readfirstlane(threadIdx.y) * MPT
                            : (BY == 4 ? __builtin_amdgcn_readfirstlane(threadIdx.y) : (int)blockIdx.y);
    if (s >= d.vol) return;
    const int cur = ping ^ d.flip;
    const S* __restrict__ U = (const S*)d.links[cur];
    S* __restrict__ Un = (S*)d.links[cur ^ 1];
    S* __restrict__ Xa = (S*)d.acc;
    const int64_t st = d.stride;
    const Geo g = make_geo(d, s);
    const C eps = (C)d.eps;
#pragma unroll 1
    for (int m = 0; m < MPT; ++m) {
        switch (mu0 + m) {
            case 0: stage_link<S, C, STAGE, EXPMODE, 0>(U, Un, Xa, st, g, eps); break;
            case 1: stage_link<S, C, STAGE, EXPMODE, 1>(U, Un, Xa, st, g, eps); break;
            case 2: stage_link<S, C, STAGE, EXPMODE, 2>(U, Un, Xa, st, g, eps); break;
            default: stage_link<S, C, STAGE, EXPMODE, 3>(U, Un, Xa, st, g, eps); break;
        }
    }
}

// ---------------------------------------------------------------------------
// RK3 stage with shared plaquettes
// ---------------------------------------------------------------------------
// The staple sum of link (x,mu) has an upper term in each plane (mu,nu),
// U_mu(x) * upper = P_{mu nu}(x) = U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag,
// and P_{nu mu}(x) = P_{mu nu}(x)^dag, whose traceless anti-Hermitian part is
// the negative of P_{mu nu}(x)'s.  The projection is linear, so
//   Z_mu(x) = -1/2 TA(U_mu(x) sum_nu lower_nu) - 1/2 sum_nu TA(P_{mu nu}(x)),
// and the six plaquettes of a site serve the upper terms of its four links.
// A workgroup holds the four direction waves of 64 sites: each wave computes
// its share of the six plaquettes (two, two, one, one), leaves their
// projections in LDS, forms its link's lower terms, and after one barrier
// adds the three projections of its planes.  Per link this replaces three
// half-staple products and three accumulated products (612 FMAs) by 1.5
// plaquettes of 282 FMAs: 1125 instead of 1314 FMAs for Z.
//
// Plane index of (a, b), a < b: (0,1) 0, (0,2) 1, (0,3) 2, (1,2) 3, (1,3) 4, (2,3) 5.
constexpr int plane_index(int a, int b) { return a == 0 ? b - 1 : a == 1 ? b + 1 : 5; }

template <typename C>
struct PlaqLds {
    C v[6][9][kTileSites];  // [plane][folded component][lane]: lane-contiguous rows, no bank conflicts
};

// The four links U_d(x) of a tile's sites in LDS, in the plane layout of
// global memory: FP32 float4 planes of components 0..15 and a float2 plane
// of 16, 17; FP64 nine double2 planes.  With LOC each direction wave loads
// its own link once and leaves it here; the plaquettes read the rows they
// need (0 and 1 of the left factor of each half-plaquette in FP32, all of it
// in FP64) and the wave rereads its whole link for the update.  In FP64
// (LGF_FP64_BACK3) a fifth slot holds U_3(x - e3), which the lower terms of
// the links of directions 0, 1 and 2 in their (mu,3) planes all read: the
// direction-3 wave loads it once instead of three waves each loading it
// (the workgroup's LDS grows from 64 to 74 KB, two workgroups per CU still).
#ifndef LGF_FP64_BACK3
#define LGF_FP64_BACK3 1
#endif
template <typename C>
struct LocLds;
template <>
struct LocLds<float> {
    vf4 q[4][4][kTileSites];
    vf2 r[4][kTileSites];
};
template <>
struct LocLds<double> {
    vd2 q[4 + LGF_FP64_BACK3][9][kTileSites];
};

// Everything of link (x,mu) after its flow gradient is known (nine-real
// form z): accumulator update, exponential, left multiplication, store.
template <typename S, typename C, int STAGE, bool NTS>
LGF_DEV void stage_finish(const C* z, const M3<C>& Umu, S* __restrict__ Un, S* __restrict__ Xa, int64_t st, int mu,
                          int s, bool live, C eps) {
    const M3<C> Z = expand_ta(z);

    M3<C> E;
    if (STAGE == 0) {
        if (live) store_ta<S, C, NTS>(Xa, st, mu, s, Z);
#pragma unroll
        for (int i = 0; i < 18; ++i) E.v[i] = (C(0.25) * eps) * Z.v[i];
    } else if (STAGE == 1) {
        const M3<C> X0 = load_ta<S, C>(Xa, st, mu, s);
        M3<C> X1;
#pragma unroll
        for (int i = 0; i < 18; ++i) X1.v[i] = (C)(8.0 / 9.0) * Z.v[i] - (C)(17.0 / 36.0) * X0.v[i];
        if (live) store_ta<S, C, NTS>(Xa, st, mu, s, X1);
#pragma unroll
        for (int i = 0; i < 18; ++i) E.v[i] = eps * X1.v[i];
    } else {
        const M3<C> X1 = load_ta<S, C>(Xa, st, mu, s);
#pragma unroll
        for (int i = 0; i < 18; ++i) E.v[i] = eps * ((C)(3.0 / 4.0) * Z.v[i] - X1.v[i]);
    }
    // the exponential by its power series: 13 terms with FP64 links (the
    // parity path), 7 in the mixed-precision one
    const M3<C> ex = exp_ah_series<C, sizeof(S) == 8 ? 13 : 7>(E);
    const M3<C> Unew = su3_nn(ex, Umu);
    if constexpr (STAGE == 2 && sizeof(S) == 4) {
        // once per RK3 step: project back onto SU(3) in double precision
        M3<double> Ud;
#pragma unroll
        for (int i = 0; i < 18; ++i) Ud.v[i] = (double)Unew.v[i];
        reunitarize<true>(Ud);
        if (live) store_link<S, double, NTS>(Un, st, mu, s, Ud);
    } else {
        if (live) store_link<S, C, NTS>(Un, st, mu, s, Unew);
    }
}

template <bool ON, typename C>
struct LocSlot {
    LocLds<C> b;
    LGF_DEV LocLds<C>* ptr() { return &b; }
};
template <typename C>
struct LocSlot<false, C> {
    LGF_DEV LocLds<C>* ptr() { return nullptr; }
};

LGF_DEV void loc_store(LocLds<float>& l, int d, int lane, const M3<float>& m) {
#pragma unroll
    for (int k = 0; k < 4; ++k) l.q[d][k][lane] = vf4{m.v[4 * k], m.v[4 * k + 1], m.v[4 * k + 2], m.v[4 * k + 3]};
    l.r[d][lane] = vf2{m.v[16], m.v[17]};
}

LGF_DEV void loc_store(LocLds<double>& l, int d, int lane, const M3<double>& m) {
#pragma unroll
    for (int k = 0; k < 9; ++k) l.q[d][k][lane] = vd2{m.v[2 * k], m.v[2 * k + 1]};
}

LGF_DEV M3<double> loc_load(const LocLds<double>& l, int d, int lane) {
    M3<double> m;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const vd2 v = l.q[d][k][lane];
        m.v[2 * k] = v.x, m.v[2 * k + 1] = v.y;
    }
    return m;
}

LGF_DEV M3<float> loc_load(const LocLds<float>& l, int d, int lane) {
    M3<float> m;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const vf4 v = l.q[d][k][lane];
        m.v[4 * k] = v.x, m.v[4 * k + 1] = v.y, m.v[4 * k + 2] = v.z, m.v[4 * k + 3] = v.w;
    }
    const vf2 w = l.r[d][lane];
    m.v[16] = w.x, m.v[17] = w.y;
    return m;
}

// U_d(x) from global memory or, with LOC, from LDS
template <bool LOC, typename S, typename C>
LGF_DEV M3<C> local_link(const S* __restrict__ U, int64_t st, int d, int s, const LocLds<C>* loc, int lane) {
    if constexpr (LOC) return loc_load(*loc, d, lane);
    else return load_link<S, C>(U, st, d, s);
}

// The projections are summed in the folded form of su3.hpp (Ta9: the
// imaginary diagonal and the upper triangle of W - W^dag, each product term
// going into one running sum), and the trace removal and the -1/2 scale are
// applied once to the link's sum (minus_half_ta).

// the nine reals of -1/2 TA(W) from the traceless folded form of W
template <typename C>
LGF_DEV void minus_half_ta(C* z) {
#pragma unroll
    for (int d = 0; d < 3; ++d) z[d] = -z[d];
#pragma unroll
    for (int k = 3; k < 9; ++k) z[k] *= C(-0.5);
}

// the folded P_ab(x) of one plane into LDS; returns the last value computed
template <typename S, typename C, int A, int B, bool LOC>
LGF_DEV C plaquette_to_lds(PlaqLds<C>& sp, const S* __restrict__ U, int64_t st, const Geo& g, int lane,
                           const LocLds<C>* loc) {
    const int s = g.s;
    const int fa = nb_off(g, A, +1), fb = nb_off(g, B, +1);
    const M3<C> h1 = su3_nn(local_link<LOC, S, C>(U, st, A, s, loc, lane), load_link<S, C>(U, st, B, s + fa));
    const M3<C> h2 = su3_nn(local_link<LOC, S, C>(U, st, B, s, loc, lane), load_link<S, C>(U, st, A, s + fb));
    Ta9<C> w;
    set_zero(w);
    mul_nd_acc_ta(w, h1, h2);
    constexpr int p = plane_index(A, B);
#pragma unroll
    for (int k = 0; k < 9; ++k) sp.v[p][k][lane] = w.v[k];
    return w.v[8];
}

// Sequencing of the sections of a link update.  Loads of read-only link
// arrays may be scheduled across an empty asm with a memory clobber, so with
// ORDERED the site index is passed through an asm that also takes a result
// of the previous section: the next section's addresses, and so its loads,
// come after that result.  This bounds the registers in flight to one
// section's links.
template <bool ORDERED, typename C>
LGF_DEV void after(Geo& g, C dep) {
    if constexpr (ORDERED) asm volatile("" : "+v"(g.s) : "v"(dep));
    else asm volatile("" ::: "memory");
}

// this wave's share of the site's six plaquettes; returns the last value
// computed
template <typename S, typename C, int MU, bool ORDERED, bool LOC>
LGF_DEV C plaquettes_of(PlaqLds<C>& sp, const S* __restrict__ U, int64_t st, Geo g, int lane, const LocLds<C>* loc) {
    C dep;
    if constexpr (MU == 0) {
        dep = plaquette_to_lds<S, C, 0, 1, LOC>(sp, U, st, g, lane, loc);
        after<ORDERED>(g, dep);
        dep = plaquette_to_lds<S, C, 0, 2, LOC>(sp, U, st, g, lane, loc);
    } else if constexpr (MU == 1) {
        dep = plaquette_to_lds<S, C, 1, 2, LOC>(sp, U, st, g, lane, loc);
        after<ORDERED>(g, dep);
        dep = plaquette_to_lds<S, C, 1, 3, LOC>(sp, U, st, g, lane, loc);
    } else if constexpr (MU == 2) {
        dep = plaquette_to_lds<S, C, 2, 3, LOC>(sp, U, st, g, lane, loc);
    } else {
        dep = plaquette_to_lds<S, C, 0, 3, LOC>(sp, U, st, g, lane, loc);
    }
    return dep;
}

// Everything of link (x,MU) before the barrier: this wave's plaquettes and
// -1/2 TA(U_mu sum_nu lower_nu) in the nine-real form.
template <typename S, typename C, int MU, bool ORDERED, bool LOC>
LGF_DEV void stage_pre_pl(PlaqLds<C>& sp, const S* __restrict__ U, int64_t st, Geo g, int lane, C* z, M3<C>& Umu,
                          const LocLds<C>* loc) {
    const C dep = plaquettes_of<S, C, MU, ORDERED, LOC>(sp, U, st, g, lane, loc);
    after<ORDERED>(g, dep);
    constexpr int N0 = MU == 0 ? 1 : 0, N1 = MU <= 1 ? 2 : 1, N2 = MU <= 2 ? 3 : 2;
    const int fmu = nb_off(g, MU, +1);
    M3<C> A;
    set_zero(A);
    lower_term<S, C, MU, N0>(A, U, st, g, fmu);
    after<ORDERED>(g, A.v[17]);
    lower_term<S, C, MU, N1>(A, U, st, g, fmu);
    after<ORDERED>(g, A.v[17]);
    if constexpr (N2 == 3 && LOC && sizeof(S) == 8 && LGF_FP64_BACK3) {
        // lower_term<S, C, MU, 3> with U_3(x - e3) from the LDS slot
        const int bnu = nb_off(g, 3, -1);
        const M3<C> tt =
            mul_halfstaple(load_link<S, C>(U, st, MU, g.s + bnu), load_link<S, C>(U, st, 3, g.s + bnu + fmu));
        mul_dn_acc(A, tt, loc_load(*loc, 4, lane));
    } else {
        lower_term<S, C, MU, N2>(A, U, st, g, fmu);
    }
    Umu = local_link<LOC, S, C>(U, st, MU, g.s, loc, lane);
    Ta9<C> w;
    set_zero(w);
    mul_nn_acc_ta(w, Umu, A);
#pragma unroll
    for (int k = 0; k < 9; ++k) z[k] = w.v[k];
}

// z += the projections of link mu's three planes (sign - where mu is the
// plane's second direction)
template <typename C>
LGF_DEV void add_planes(const PlaqLds<C>& sp, int mu, int lane, C* z) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int nu = k + (k >= mu ? 1 : 0);
        const int p = mu < nu ? plane_index(mu, nu) : plane_index(nu, mu);
        const C sg = mu < nu ? C(1) : C(-1);
#pragma unroll
        for (int c = 0; c < 9; ++c) z[c] += sg * sp.v[p][c][lane];
    }
}

template <typename S, typename C, int STAGE, int WPE, bool ORDERED, bool LOC, bool NTS = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_rk3_stage_pl(
    const FieldDesc* __restrict__ fds, const int2* __restrict__ tiles, int ping) {
    static_assert(std::is_same<S, C>::value, "arithmetic in the precision of the links");
    __shared__ PlaqLds<C> sp;
    __shared__ LocSlot<LOC, C> lslot;
    LocLds<C>* loc = lslot.ptr();
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int2 tm = tiles[t];
    const FieldDesc& d = fds[tm.x];
    const int lane = threadIdx.x;
    const int s0 = tm.y * kTileSites + lane;
    const bool live = s0 < d.vol;
    const int s = live ? s0 : d.vol - 1;  // lanes past the end of a ragged tile compute a copy, store nothing
    const int mu = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int cur = ping ^ d.flip;
    const S* __restrict__ U = (const S*)d.links[cur];
    S* __restrict__ Un = (S*)d.links[cur ^ 1];
    S* __restrict__ Xa = (S*)d.acc;
    const int64_t st = d.stride;
    const Geo g = make_geo(d, s);
    const C eps = (C)d.eps;

    if constexpr (LOC) {
        loc_store(*loc, mu, lane, load_link<S, C>(U, st, mu, s));
        if constexpr (sizeof(S) == 8 && LGF_FP64_BACK3) {
            if (mu == 3) loc_store(*loc, 4, lane, load_link<S, C>(U, st, 3, s + nb_off(g, 3, -1)));
        }
        __syncthreads();
    }
    C z[9];
    M3<C> Umu;
    switch (mu) {
        case 0: stage_pre_pl<S, C, 0, ORDERED, LOC>(sp, U, st, g, lane, z, Umu, loc); break;
        case 1: stage_pre_pl<S, C, 1, ORDERED, LOC>(sp, U, st, g, lane, z, Umu, loc); break;
        case 2: stage_pre_pl<S, C, 2, ORDERED, LOC>(sp, U, st, g, lane, z, Umu, loc); break;
        default: stage_pre_pl<S, C, 3, ORDERED, LOC>(sp, U, st, g, lane, z, Umu, loc); break;
    }
    __syncthreads();
    add_planes(sp, mu, lane, z);
    Ta9<C> zt;
#pragma unroll
    for (int k = 0; k < 9; ++k) zt.v[k] = z[k];
    make_traceless(zt);
#pragma unroll
    for (int k = 0; k < 9; ++k) z[k] = zt.v[k];
    minus_half_ta(z);
    stage_finish<S, C, STAGE, NTS>(z, Umu, Un, Xa, st, mu, s, live, eps);
}

// ---------------------------------------------------------------------------
// RK3 stage with shared plaquettes and shared backward links (mixed precision)
// ---------------------------------------------------------------------------
// For lattices that stream from HBM the shared-plaquette stage waits on L2
// misses, so this form also shares the backward links U_nu(x - nu), which the
// lower terms of three links of a site read, and keeps the workgroup's LDS
// within a quarter of a CU's (four workgroups resident):
//   - the plaquettes folded and made traceless, as eight reals: the third
//     imaginary diagonal entry is minus the sum of the other two;
//   - rows 0 and 1 of the site's own links and of its backward links; row 2
//     is rebuilt as the conjugated cross product where a whole matrix is
//     needed (exact for SU(3), a rounding-level difference for the FP32
//     links between projections).
// 236 -> 180 link loads per site, about 110 more VALU instructions per wave.
struct CompactLds {
    float p[6][8][kTileSites];     // P_ab(x) folded and traceless: components 0, 1, 3..8
    vf4 loc[4][3][kTileSites];     // rows 0, 1 of U_d(x)
    vf4 back[4][3][kTileSites];    // rows 0, 1 of U_d(x - d)
};

LGF_DEV void rows01_store(vf4 (*q)[kTileSites], int lane, const M3<float>& m) {
#pragma unroll
    for (int k = 0; k < 3; ++k) q[k][lane] = vf4{m.v[4 * k], m.v[4 * k + 1], m.v[4 * k + 2], m.v[4 * k + 3]};
}

// rows 0, 1 from LDS; row 2 left zero (a left factor of su3_nn / su3_nd)
// or rebuilt (WHOLE)
template <bool WHOLE>
LGF_DEV M3<float> rows01_load(const vf4 (*q)[kTileSites], int lane) {
    M3<float> m;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const vf4 v = q[k][lane];
        m.v[4 * k] = v.x, m.v[4 * k + 1] = v.y, m.v[4 * k + 2] = v.z, m.v[4 * k + 3] = v.w;
    }
    if constexpr (WHOLE) {
        su3_third_row(m);
    } else {
#pragma unroll
        for (int i = 12; i < 18; ++i) m.v[i] = 0.f;
    }
    return m;
}

template <typename S, int A, int B>
LGF_DEV float plaquette_compact(CompactLds& L, const S* __restrict__ U, int64_t st, const Geo& g, int lane) {
    const int s = g.s;
    const int fa = nb_off(g, A, +1), fb = nb_off(g, B, +1);
    const M3<float> h1 = su3_nn(rows01_load<false>(L.loc[A], lane), load_link<S, float>(U, st, B, s + fa));
    const M3<float> h2 = su3_nn(rows01_load<false>(L.loc[B], lane), load_link<S, float>(U, st, A, s + fb));
    Ta9<float> w;
    set_zero(w);
    mul_nd_acc_ta(w, h1, h2);
    make_traceless(w);
    constexpr int p = plane_index(A, B);
    L.p[p][0][lane] = w.v[0];
    L.p[p][1][lane] = w.v[1];
#pragma unroll
    for (int k = 3; k < 9; ++k) L.p[p][k - 1][lane] = w.v[k];
    return w.v[8];
}

// lower term of the (MU,NU) plane with the backward link from LDS
template <typename S, int MU, int NU>
LGF_DEV void lower_term_compact(M3<float>& A, const CompactLds& L, const S* __restrict__ U, int64_t st, const Geo& g,
                                int fmu, int lane) {
    const int s = g.s;
    const int bnu = nb_off(g, NU, -1);
    const M3<float> tt = su3_nn(load_link<S, float>(U, st, MU, s + bnu), load_link<S, float>(U, st, NU, s + bnu + fmu));
    mul_dn_acc(A, tt, rows01_load<true>(L.back[NU], lane));
}

template <typename S, int MU>
LGF_DEV void stage_pre_pb(CompactLds& L, const S* __restrict__ U, int64_t st, Geo g, int lane, float* z, M3<float>& Umu) {
    float dep;
    if constexpr (MU == 0) {
        dep = plaquette_compact<S, 0, 1>(L, U, st, g, lane);
        after<true>(g, dep);
        dep = plaquette_compact<S, 0, 2>(L, U, st, g, lane);
    } else if constexpr (MU == 1) {
        dep = plaquette_compact<S, 1, 2>(L, U, st, g, lane);
        after<true>(g, dep);
        dep = plaquette_compact<S, 1, 3>(L, U, st, g, lane);
    } else if constexpr (MU == 2) {
        dep = plaquette_compact<S, 2, 3>(L, U, st, g, lane);
    } else {
        dep = plaquette_compact<S, 0, 3>(L, U, st, g, lane);
    }
    after<true>(g, dep);
    constexpr int N0 = MU == 0 ? 1 : 0, N1 = MU <= 1 ? 2 : 1, N2 = MU <= 2 ? 3 : 2;
    const int fmu = nb_off(g, MU, +1);
    M3<float> A;
    set_zero(A);
    lower_term_compact<S, MU, N0>(A, L, U, st, g, fmu, lane);
    after<true>(g, A.v[17]);
    lower_term_compact<S, MU, N1>(A, L, U, st, g, fmu, lane);
    after<true>(g, A.v[17]);
    lower_term_compact<S, MU, N2>(A, L, U, st, g, fmu, lane);
    Umu = rows01_load<true>(L.loc[MU], lane);
    Ta9<float> w;
    set_zero(w);
    mul_nn_acc_ta(w, Umu, A);
    make_traceless(w);
#pragma unroll
    for (int k = 0; k < 9; ++k) z[k] = w.v[k];
}

template <typename S, typename C, int STAGE, int WPE, bool NTS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_rk3_stage_pb(
    const FieldDesc* __restrict__ fds, const int2* __restrict__ tiles, int ping) {
    static_assert(std::is_same<S, float>::value && std::is_same<C, float>::value, "mixed-precision path only");
    __shared__ CompactLds L;
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int2 tm = tiles[t];
    const FieldDesc& d = fds[tm.x];
    const int lane = threadIdx.x;
    const int s0 = tm.y * kTileSites + lane;
    const bool live = s0 < d.vol;
    const int s = live ? s0 : d.vol - 1;  // lanes past the end of a ragged tile compute a copy, store nothing
    const int mu = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int cur = ping ^ d.flip;
    const S* __restrict__ U = (const S*)d.links[cur];
    S* __restrict__ Un = (S*)d.links[cur ^ 1];
    S* __restrict__ Xa = (S*)d.acc;
    const int64_t st = d.stride;
    const Geo g = make_geo(d, s);
    const C eps = (C)d.eps;

    // each direction wave stages rows 0, 1 of its link and of its backward link
    rows01_store(L.loc[mu], lane, load_link<S, C>(U, st, mu, s));
    rows01_store(L.back[mu], lane, load_link<S, C>(U, st, mu, s + nb_off(g, mu, -1)));
    __syncthreads();
    C z[9];
    M3<C> Umu;
    switch (mu) {
        case 0: stage_pre_pb<S, 0>(L, U, st, g, lane, z, Umu); break;
        case 1: stage_pre_pb<S, 1>(L, U, st, g, lane, z, Umu); break;
        case 2: stage_pre_pb<S, 2>(L, U, st, g, lane, z, Umu); break;
        default: stage_pre_pb<S, 3>(L, U, st, g, lane, z, Umu); break;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int nu = k + (k >= mu ? 1 : 0);
        const int p = mu < nu ? plane_index(mu, nu) : plane_index(nu, mu);
        const C sg = mu < nu ? C(1) : C(-1);
        z[0] += sg * L.p[p][0][lane];
        z[1] += sg * L.p[p][1][lane];
#pragma unroll
        for (int c = 3; c < 9; ++c) z[c] += sg * L.p[p][c - 1][lane];
    }
    z[2] = -(z[0] + z[1]);  // the sum is traceless
    minus_half_ta(z);
    stage_finish<S, C, STAGE, NTS>(z, Umu, Un, Xa, st, mu, s, live, eps);
}

// ---------------------------------------------------------------------------
// clovers, E, Q, q(x)
// ---------------------------------------------------------------------------
// Compile-time knobs of the clover kernel:
//   LGF_CLOVER_WPE      minimum waves per SIMD of the register budget (FP32)
//   LGF_CLOVER_WPE64    the same for FP64 arithmetic on FP32 links (the
//                       mixed-precision path's final sweep)
//   LGF_CLOVER_WPE64_PARITY  the same for FP64 links
//   LGF_CLOVER_LDS_REAL 1: per-plane field strengths meet in LDS in the
//                       arithmetic precision (float for FP32 clovers), which
//                       halves LDS use and lets more workgroups share a CU
//   LGF_CLOVER_ORDERED  1: the addresses (and so the loads) of a pair of
//                       leaves depend on the previous pair's result, so few
//                       lines are in flight at a time (fewer registers, more
//                       waves)
//   LGF_CLOVER_ORDERED_INNER 1: also between the two leaves of a pair
// Measured on 24^4 (three sweep precisions): ORDERED 0 spills and doubles
// the sweep times; ORDERED_INNER 1 costs the FP64-arithmetic sweep 4 %;
// three waves per SIMD for that sweep 30 %.
#ifndef LGF_CLOVER_WPE
#define LGF_CLOVER_WPE 4
#endif
#ifndef LGF_CLOVER_WPE64
#define LGF_CLOVER_WPE64 2
#endif
#ifndef LGF_CLOVER_WPE64_PARITY
#define LGF_CLOVER_WPE64_PARITY 2
#endif
#ifndef LGF_CLOVER_LDS_REAL
#define LGF_CLOVER_LDS_REAL 1
#endif
#ifndef LGF_CLOVER_ORDERED
#define LGF_CLOVER_ORDERED 1
#endif
#ifndef LGF_CLOVER_ORDERED_INNER
#define LGF_CLOVER_ORDERED_INNER 0
#endif
// LGF_CLOVER_PAIRS 1: the FP32 sweep runs k_clover_pairs (two planes per
// workgroup) and k_clover_combine instead of k_clover;
// LGF_CLOVER_WPE_PAIRS is the pair kernel's minimum waves per SIMD
#ifndef LGF_CLOVER_PAIRS
#define LGF_CLOVER_PAIRS 1
#endif
#ifndef LGF_CLOVER_WPE_PAIRS
#define LGF_CLOVER_WPE_PAIRS 4
#endif

// Line products.  L^d_l(y) = U_d(y) U_d(y+d) ... U_d(y+(l-1)d), the product
// of l links along a straight line.  k_lines stores L_2 and L_3 of every
// site and direction of the fields being measured (precision C, the vector
// plane layout of the links, line index lbase + site); L_1 is the link.
//
// With them every clover leaf of extent m along mu and n along nu is a
// product of four lines.  The reference's four quadrant leaves of the (mu,nu)
// plane (LGF/Field/SU3_field.cpp:533-584 for 1x1, 598-792 for 2x2, 1x2 and
// 1x3), each a closed loop from x, are:
//   quadrant (+nu,+mu): W = P(x) R(x)^dag
//   quadrant (-nu,-mu): W = R(c)^dag P(c),            c = x - n nu - m mu
//     P(c) = L^nu_n(c) L^mu_m(c + n nu),  R(c) = L^mu_m(c) L^nu_n(c + m mu)
//   quadrant (+mu,-nu): W = L^mu_m(x) H^dag L^nu_n(z), z = x - n nu,
//     H = L^mu_m(z) L^nu_n(z + m mu)
//   quadrant (-mu,+nu): W^dag = L^nu_n(x) H'^dag L^mu_m(z'), z' = x - m mu,
//     H' = L^nu_n(z') L^mu_m(z' + n nu)
// (e.g. the first 1x1 leaf is U_nu(x) U_mu(x+nu) U_nu(x+mu)^dag U_mu(x)^dag,
// LGF/Field/SU3_field.cpp:550-560).  The field strength depends on a clover
// sum only through C - C^dag, to which W and -W^dag contribute alike.  Leaf
// shapes come in both orientations (m x n and n x m) with equal weights.
// XC: the lines that start at the workgroup's own sites are read from its
// LDS cache (xl) rather than from the line buffer
template <typename S, typename C, bool XC>
struct Lines {
    const S* U;   // current links: lines of length 1
    int64_t st;   // link stride
    const C* L2;  // lines of length 2
    const C* L3;  // lines of length 3
    int64_t G;    // line stride
    int lbase;    // line index of site 0 of this field
    const C* xl;  // LDS copy of the 12 lines that start at the workgroup's sites, or null
    int lane;
};

// With FP64 links (the parity path) the twelve lines that start at a site
// (three lengths, four directions) are loaded once per workgroup into LDS,
// where the six plane waves of the site read them: a fifth of the sweep's
// line loads (a_n, R1 and the first factor of T3 in plane_F_factored).  That
// kernel is bound by its line loads (VALU about a third busy) and holds one
// workgroup per CU at its register budget, so the LDS (110 KB, 157 KB with
// the backward lines below) costs no occupancy; the cache shares its LDS
// with the per-plane field strengths, which are written after every wave is
// done with it.  Without it the sweep is 9 % slower (24^4).
//   LGF_CLOVER_XLINES_ROWS: the same cache for the six-plane FP32 kernel,
//   holding rows 0 and 1 of each line (37 KB, within the LDS that kernel
//   already uses for its field strengths, so it costs no residency): that
//   kernel was bound by the bytes its loads return through the vector
//   memory path, and the cache made it 2-3 % faster (24^4); the FP32 sweep
//   now runs k_clover_pairs, which has no cache.  The FP64-arithmetic sweep over FP32
//   links does not use it: with it that kernel needs fewer registers and two
//   workgroups share a CU, which thrashes its caches (33 % slower), and at
//   its old residency (LDS padded) it is no faster.
#ifndef LGF_CLOVER_XLINES
#define LGF_CLOVER_XLINES 1
#endif
#ifndef LGF_CLOVER_XLINES_ROWS
#define LGF_CLOVER_XLINES_ROWS 1
#endif
template <typename S, typename C>
constexpr bool kCloverXLines = sizeof(S) == 8 ? LGF_CLOVER_XLINES : sizeof(C) == 4 && LGF_CLOVER_XLINES_ROWS;

// Products of lines.  With FP32 links (the mixed-precision path, whose links
// are projected onto SU(3) every flow step) they are SU(3) products: rows 0
// and 1 and their conjugated cross product.  Every line is then read as rows
// 0 and 1 (kCloverRows, and k_lines stores only those), and where a product
// needs more of it (a right factor, or columns 0, 1 of a left factor) its
// third row is rebuilt: a rebuilt row costs 24 FMAs against 24 of 72 bytes
// (FP32) or 48 of 144 (FP64) through the vector memory path.  With whole
// lines the leaf-by-leaf FP32 sweep was bound by those bytes (24^4: texture
// data unit 88-91 % busy, VALU 54 %); the factored sweep is near both limits
// (VALU 96 %, texture data unit 89 % busy).  With FP64 links
// (the parity path) the products are full, as the reference's: its links are
// not projected, and rebuilding rows would move the result by the links'
// drift from SU(3).
#ifndef LGF_CLOVER_ROWS2
#define LGF_CLOVER_ROWS2 1
#endif
template <typename S>
constexpr bool kCloverSu3 = sizeof(S) == 4;
template <typename S, typename C>
constexpr bool kCloverRows = LGF_CLOVER_ROWS2 && kCloverSu3<S>;
// reals per line in the LDS line cache
template <typename S, typename C>
constexpr int kXLen = kCloverRows<S, C> ? 12 : 18;
// With FP64 links the cache also holds five backward lines L^d_k(x - k d),
// each staged by the lane of site x: (d, 1) for every d at slots 12 + d and
// (3, 2) at slot 16 (157 KB in all).  A site's six planes read them 60 times
// per site (R4 and the left factor of K2 along mu, g_n along nu): the most
// used of the backward lines, direction 3 being the first direction of
// three planes.
#ifndef LGF_CLOVER_XBACK
#define LGF_CLOVER_XBACK 1
#endif
template <typename S, typename C>
constexpr bool kXBack = sizeof(S) == 8 && LGF_CLOVER_XLINES && LGF_CLOVER_XBACK;
template <typename S, typename C>
constexpr int kXSlots = kXBack<S, C> ? 17 : 12;

// a line that starts at the workgroup's own site x, from the LDS cache:
// whole, or (ROWS = 2) rows 0 and 1 where the cache holds only those
template <int LEN, int ROWS, typename S, typename C, bool XC>
LGF_DEV M3<C> xline(const Lines<S, C, XC>& L, int dir) {
    constexpr int X = kXLen<S, C>;
    M3<C> m;
    const C* q = L.xl + (size_t)(4 * (LEN - 1) + dir) * X * kTileSites + L.lane;
#pragma unroll
    for (int c = 0; c < X; ++c) m.v[c] = q[c * kTileSites];
    if constexpr (X == 12 && ROWS != 2) su3_third_row(m);
    return m;
}

// ROWS = 2: rows 0, 1; 3: whole; 4: columns 0, 1 (a left factor of
// mul_dd_r01 or a right factor of mul_nn_c01), which for whole FP32 lines
// are the four float4 planes.
template <int LEN, int ROWS, typename S, typename C, bool XC>
LGF_DEV M3<C> line(const Lines<S, C, XC>& L, int dir, int site) {
    if constexpr (kCloverSu3<S> && (ROWS == 2 || kCloverRows<S, C>)) {
        M3<C> m = LEN == 1 ? load_link_rows01<S, C>(L.U, L.st, dir, site)
                           : load_link_rows01<C, C>(LEN == 2 ? L.L2 : L.L3, L.G, dir, L.lbase + site);
        if constexpr (ROWS != 2) su3_third_row(m);
        return m;
    } else if constexpr (kCloverSu3<S> && ROWS == 4) {
        return LEN == 1 ? load_link_cols01<C>(L.U, L.st, dir, site)
                        : load_link_cols01<C>(LEN == 2 ? L.L2 : L.L3, L.G, dir, L.lbase + site);
    } else if constexpr (LEN == 1) {
        return load_link<S, C>(L.U, L.st, dir, site);
    } else {
        return load_link<C, C>(LEN == 2 ? L.L2 : L.L3, L.G, dir, L.lbase + site);
    }
}

// site offset of x + steps*dir (|steps| <= 3, periodic); rcp = the field's
// (1 + 2^-40)/n reciprocals, so the wrap is one double multiply (fdiv)
// instead of an integer remainder
LGF_DEV int wrap_offset(const Geo& g, int dir, int steps, const double* rcp) {
    const int xd = geo_x(g, dir), nd = geo_n(g, dir);
    const int t = xd + steps + 4 * nd;  // >= 0: nd >= 2
    const int v = t - fdiv(t, rcp[dir]) * nd;
    return (v - xd) * geo_str(g, dir);
}

// site offsets of x + k mu (m[3 + k]) and x + k nu (n[3 + k]), |k| <= 3;
// offsets along different axes add, offsets along one axis do not (wrap)
struct PlaneOff {
    int m[7], n[7];
};

// Factored clover.  Every leaf of the plane is an outer line, a middle
// product and another outer line, and the four quadrants pair up on a
// shared outer factor, so the leaves of one n (extent along nu) are summed
// as 3x3 matrices before one projected product with that factor:
//   (+nu,+mu):  W = a_n Mid1 R1^dag,   a_n = L^nu_n(x), R1 = L^mu_m(x),
//               Mid1 = L^mu_m(x+n nu) L^nu_n(x+m mu)^dag
//   (-mu,+nu):  W^dag = a_n Mid4 R4,   R4 = L^mu_m(x-m mu),
//               Mid4 = L^mu_m(x-m mu+n nu)^dag L^nu_n(x-m mu)^dag
//   (+mu,-nu):  W = T3 g_n,            g_n = L^nu_n(x-n nu), T3 = L^mu_m(x) H3^dag,
//               H3 = L^mu_m(x-n nu) L^nu_n(x-n nu+m mu)
//   (-nu,-mu):  W^dag = T2 g_n,        T2 = K2^dag L^mu_m(c), c = x-n nu-m mu,
//               K2 = L^nu_n(c) L^mu_m(x-m mu)
// (the quadrant leaves above regrouped; TA(W^dag) = -TA(W)).
// So TA(sum of leaves) = TA(a_n Y_n) + TA(Z_n g_n) summed over n, with
//   Y_n = sum_m w_mn (Mid1 R1^dag - Mid4 R4),  Z_n = sum_m w_mn (T3 - T2).
// Per leaf that is two products instead of three, the third product shared
// by the leaves of one n.  With FP32 links (SU(3) products) the middle
// product is needed only in the rows or columns that the second product
// reads, 72 FMAs.

// a whole line that starts at the workgroup's own site x
template <int LEN, typename S, typename C, bool XC>
LGF_DEV M3<C> own_line(const Lines<S, C, XC>& L, int dir, int s) {
    if constexpr (XC) return xline<LEN, 3>(L, dir);
    else return line<LEN, 3>(L, dir, s);
}

// the whole line L^dir_LEN(x - LEN dir), site = x - LEN dir: from the cache
// where it holds that line (dir is wave-uniform; used for g_n along nu)
template <int LEN, typename S, typename C, bool XC>
LGF_DEV M3<C> back_line(const Lines<S, C, XC>& L, int dir, int site) {
    if constexpr (XC && kXBack<S, C> && LEN <= 2) {
        if (LEN == 1 || dir == 3) {
            M3<C> m;
            const C* q = L.xl + (size_t)(LEN == 1 ? 12 + dir : 16) * 18 * kTileSites + L.lane;
#pragma unroll
            for (int c = 0; c < 18; ++c) m.v[c] = q[c * kTileSites];
            return m;
        }
    }
    return line<LEN, 3>(L, dir, site);
}

// the same along mu in the planes whose mu is 3 (MU3), where the cache holds
// L^3_1(x - e3) and L^3_2(x - 2 e3): a compile-time choice, since a branch
// around the leaves' loads breaks their ordering and spills
template <int LEN, bool MU3, typename S, typename C, bool XC>
LGF_DEV M3<C> mu_back_line(const Lines<S, C, XC>& L, int mu, int site) {
    if constexpr (MU3 && LEN <= 2) {
        M3<C> m;
        const C* q = L.xl + (size_t)(LEN == 1 ? 15 : 16) * 18 * kTileSites + L.lane;
#pragma unroll
        for (int c = 0; c < 18; ++c) m.v[c] = q[c * kTileSites];
        return m;
    } else {
        return line<LEN, 3>(L, mu, site);
    }
}

template <int M, int N, bool MU3, typename S, typename C, bool XC>
LGF_DEV void leaves_a(M3<C>& Y, C w, bool first, C dep, const Lines<S, C, XC>& L, int s, const PlaneOff& o, int mu,
                      int nu) {
#if LGF_CLOVER_ORDERED
    asm volatile("" : "+v"(s) : "v"(dep));
#endif
    const int sm = s + o.m[3 - M];  // x - m mu
    M3<C> p1, p4;
    if constexpr (kCloverSu3<S>) {
        const M3<C> mid1 = mul_nd_r01(line<M, 2>(L, mu, s + o.n[3 + N]), line<N, 3>(L, nu, s + o.m[3 + M]));
        p1 = mul_nd_su3(mid1, own_line<M>(L, mu, s));
    } else {
        const M3<C> mid1 = mul_nd(line<M, 3>(L, mu, s + o.n[3 + N]), line<N, 3>(L, nu, s + o.m[3 + M]));
        p1 = mul_nd(mid1, own_line<M>(L, mu, s));
    }
    if (first) set_scaled(Y, w, p1);
    else axpy(Y, w, p1);
#if LGF_CLOVER_ORDERED && LGF_CLOVER_ORDERED_INNER
    int s4 = sm;
    asm volatile("" : "+v"(s4) : "v"(Y.v[17]));
#else
    const int s4 = sm;
#endif
    if constexpr (kCloverSu3<S>) {
        const M3<C> mid4 = mul_dd_r01(line<M, 4>(L, mu, s4 + o.n[3 + N]), line<N, 3>(L, nu, s4));
        p4 = mul_nn_su3(mid4, line<M, 3>(L, mu, s4));
    } else {
        const M3<C> h4 = mul_nn(line<N, 3>(L, nu, s4), line<M, 3>(L, mu, s4 + o.n[3 + N]));
        p4 = mul_dn(h4, mu_back_line<M, MU3>(L, mu, s4));
    }
    axpy(Y, -w, p4);
}

template <int M, int N, bool MU3, typename S, typename C, bool XC>
LGF_DEV void leaves_b(M3<C>& Z, C w, bool first, C dep, const Lines<S, C, XC>& L, int s, const PlaneOff& o, int mu,
                      int nu) {
#if LGF_CLOVER_ORDERED
    asm volatile("" : "+v"(s) : "v"(dep));
#endif
    const int sn = s + o.n[3 - N];  // x - n nu
    if constexpr (kCloverSu3<S>) {
        const M3<C> h3 = mul_nn_r01(line<M, 2>(L, mu, sn), line<N, 3>(L, nu, sn + o.m[3 + M]));
        const M3<C> t3d = mul_nd_su3(h3, own_line<M>(L, mu, s));  // T3^dag
        if (first) set_scaled_dag(Z, w, t3d);
        else axpy_dag(Z, w, t3d);
    } else {
        const M3<C> h3 = mul_nn(line<M, 3>(L, mu, sn), line<N, 3>(L, nu, sn + o.m[3 + M]));
        const M3<C> t3 = mul_nd(own_line<M>(L, mu, s), h3);
        if (first) set_scaled(Z, w, t3);
        else axpy(Z, w, t3);
    }
#if LGF_CLOVER_ORDERED && LGF_CLOVER_ORDERED_INNER
    int sc = sn + o.m[3 - M];  // c
    asm volatile("" : "+v"(sc) : "v"(Z.v[17]));
#else
    const int sc = sn + o.m[3 - M];
#endif
    const int smm = s + o.m[3 - M];  // x - m mu
    M3<C> t2;
    if constexpr (kCloverSu3<S>) {
        const M3<C> k2 = mul_nn_c01(line<N, 3>(L, nu, sc), line<M, 4>(L, mu, smm));
        t2 = mul_dn_su3(k2, line<M, 3>(L, mu, sc));
    } else {
        const M3<C> k2 = mul_nn(line<N, 3>(L, nu, sc), mu_back_line<M, MU3>(L, mu, smm));
        t2 = mul_dn(k2, line<M, 3>(L, mu, sc));
    }
    axpy(Z, -w, t2);
}

// F = 1/2 (-i) [ (C - C^dag) - 1/3 Tr(C - C^dag) ] (LGF/Field/SU3_field.cpp:575-579)
// from the folded sum: the diagonal of C - C^dag is 2i t[d], so
// F_dd = t[d] - (t[0] + t[1] + t[2]) / 3, and above it F_rc = -i/2 (C - C^dag)_rc.
template <typename C>
LGF_DEV M3<C> project_F(const Ta9<C>& t) {
    M3<C> F;
    const C m = (t.v[0] + t.v[1] + t.v[2]) * (C(1) / C(3));
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        F.v[8 * d] = t.v[d] - m;
        F.v[8 * d + 1] = C(0);
    }
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const int r = p == 2 ? 1 : 0, c = p == 0 ? 1 : 2;
        const C re = C(0.5) * t.v[4 + 2 * p], im = C(-0.5) * t.v[3 + 2 * p];
        F.v[2 * (3 * r + c)] = re;
        F.v[2 * (3 * r + c) + 1] = im;
        F.v[2 * (3 * c + r)] = re;
        F.v[2 * (3 * c + r) + 1] = -im;
    }
    return F;
}

// Plane (mu,nu): the plain 1x1 F (written packed to f11[0..8], one entry
// every kTileSites doubles: LDS) and the improved
// F = 19/9 F11 + 1/36 F22 - 32/45 F12 + 1/15 F13 (LGF/Field/SU3_field.cpp:
// 588-597), where each F is the projection of its clover average: 1/4 of
// four square leaves, 1/8 of eight rectangle leaves.  The projection is
// linear, so it is applied once to the weighted sum.  Returns Re Tr(F11 F11).
template <typename C, typename T>
LGF_DEV void herm_pack(const M3<C>& F, T* h, int step);

template <bool MU3, typename S, typename C, typename T, bool XC>
LGF_DEV C plane_F_factored(const Lines<S, C, XC>& L, const PlaneOff& o, int s, int mu, int nu, M3<C>& Fimp, T* f11,
                           int f11_step) {
    const C k1 = C(19.0 / 9.0), k2 = C(1.0 / 36.0), k3 = C(-32.0 / 45.0), k4 = C(1.0 / 15.0);
    // per-leaf weights of the improved sum: k1/4 of the four square 1x1
    // leaves, k2/4 of the 2x2, k3/8 and k4/8 of the eight 1x2 and 1x3
    const C w22 = C(0.25) * k2, w12 = C(0.125) * k3, w13 = C(0.125) * k4;
    Ta9<C> a11, acc;  // the 1x1 sum (plain F) and the weighted sum of the larger leaves
    set_zero(a11);
    set_zero(acc);
    M3<C> Y;
    // quadrants (+nu,+mu), (-mu,+nu): TA(a_n Y_n)
    leaves_a<1, 1, MU3>(Y, C(1), true, acc.v[8], L, s, o, mu, nu);
    mul_nn_acc_ta(a11, own_line<1>(L, nu, s), Y);
    leaves_a<2, 1, MU3>(Y, w12, true, a11.v[8], L, s, o, mu, nu);
    leaves_a<3, 1, MU3>(Y, w13, false, Y.v[17], L, s, o, mu, nu);
    mul_nn_acc_ta(acc, own_line<1>(L, nu, s), Y);
    leaves_a<1, 2, MU3>(Y, w12, true, acc.v[8], L, s, o, mu, nu);
    leaves_a<2, 2, MU3>(Y, w22, false, Y.v[17], L, s, o, mu, nu);
    mul_nn_acc_ta(acc, own_line<2>(L, nu, s), Y);
    leaves_a<1, 3, MU3>(Y, w13, true, acc.v[8], L, s, o, mu, nu);
    mul_nn_acc_ta(acc, own_line<3>(L, nu, s), Y);
    // quadrants (+mu,-nu), (-nu,-mu): TA(Z_n g_n)
    leaves_b<1, 1, MU3>(Y, C(1), true, acc.v[8], L, s, o, mu, nu);
    mul_nn_acc_ta(a11, Y, back_line<1>(L, nu, s + o.n[2]));
    leaves_b<2, 1, MU3>(Y, w12, true, a11.v[8], L, s, o, mu, nu);
    leaves_b<3, 1, MU3>(Y, w13, false, Y.v[17], L, s, o, mu, nu);
    mul_nn_acc_ta(acc, Y, back_line<1>(L, nu, s + o.n[2]));
    leaves_b<1, 2, MU3>(Y, w12, true, acc.v[8], L, s, o, mu, nu);
    leaves_b<2, 2, MU3>(Y, w22, false, Y.v[17], L, s, o, mu, nu);
    mul_nn_acc_ta(acc, Y, back_line<2>(L, nu, s + o.n[1]));
    leaves_b<1, 3, MU3>(Y, w13, true, acc.v[8], L, s, o, mu, nu);
    mul_nn_acc_ta(acc, Y, line<3, 3>(L, nu, s + o.n[0]));
    C e;
    {
        Ta9<C> avg = a11;
        scale(avg, C(0.25));
        const M3<C> F11 = project_F(avg);
        e = retr_mul(F11, F11);
        herm_pack(F11, f11, f11_step);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) acc.v[i] += (C(0.25) * k1) * a11.v[i];
    Fimp = project_F(acc);
    return e;
}

template <typename S, typename C, typename T, bool XC>
LGF_DEV C plane_F(const Lines<S, C, XC>& L, const Geo& g, const double* rcp, int mu, int nu, M3<C>& Fimp, T* f11,
                  int f11_step) {
    PlaneOff o;
#pragma unroll
    for (int k = -3; k <= 3; ++k) {
        o.m[3 + k] = k == 0 ? 0 : wrap_offset(g, mu, k, rcp);
        o.n[3 + k] = k == 0 ? 0 : wrap_offset(g, nu, k, rcp);
    }
    if constexpr (kXBack<S, C> && XC) {
        if (mu == 3) return plane_F_factored<true, S, C, T>(L, o, g.s, mu, nu, Fimp, f11, f11_step);
    }
    return plane_F_factored<false, S, C, T>(L, o, g.s, mu, nu, Fimp, f11, f11_step);
}

// Re Tr(A B) for Hermitian A, B given as 9 reals each:
// (A00, A11, A22, Re A01, Im A01, Re A02, Im A02, Re A12, Im A12)
template <typename T>
LGF_DEV double herm_dot(const T* a, const T* b) {
    double x[9], y[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) x[k] = (double)a[k], y[k] = (double)b[k];
    return x[0] * y[0] + x[1] * y[1] + x[2] * y[2] +
           2.0 * (x[3] * y[3] + x[4] * y[4] + x[5] * y[5] + x[6] * y[6] + x[7] * y[7] + x[8] * y[8]);
}

template <typename C, typename T>
LGF_DEV void herm_pack(const M3<C>& F, T* h, int step) {
    h[0 * step] = (T)F.v[0];
    h[1 * step] = (T)F.v[8];
    h[2 * step] = (T)F.v[16];
    h[3 * step] = (T)F.v[2];
    h[4 * step] = (T)F.v[3];
    h[5 * step] = (T)F.v[4];
    h[6 * step] = (T)F.v[5];
    h[7 * step] = (T)F.v[10];
    h[8 * step] = (T)F.v[11];
}

// L_2 and L_3 of one (site, direction) of a chunk of the batch's tiles
// (tiles = the chunk's part of the tile list, t0 = its first tile).
template <typename S, typename C>
__global__ __launch_bounds__(256) void k_lines(const FieldDesc* __restrict__ fds, const int2* __restrict__ tiles,
                                               int t0, int ping, C* __restrict__ lines, int64_t G) {
    const int tix = xcd_remap(blockIdx.x, gridDim.x);
    const int2 tm = tiles[tix];
    const FieldDesc& d = fds[tm.x];
    const int s = tm.y * kTileSites + threadIdx.x;
    const int mu = __builtin_amdgcn_readfirstlane(threadIdx.y);
    if (s >= d.vol) return;
    const S* __restrict__ U = (const S*)d.links[ping ^ d.flip];
    const Geo g = make_geo(d, s);
    const int li = (d.tile0 - t0) * kTileSites + s;
    if constexpr (kCloverRows<S, C>) {
        // the clover reads these lines as rows 0 and 1 only
        M3<C> U1 = load_link_rows01<S, C>(U, d.stride, mu, s + nb_off(g, mu, +1));
        su3_third_row(U1);
        const M3<C> L2 = mul_nn_su3(load_link_rows01<S, C>(U, d.stride, mu, s), U1);
        store_link_rows01<C, C>(lines, G, mu, li, L2);
        M3<C> U2 = load_link_rows01<S, C>(U, d.stride, mu, s + wrap_offset(g, mu, 2, d.rcp));
        su3_third_row(U2);
        store_link_rows01<C, C>(lines + 4 * 18 * G, G, mu, li, mul_nn_su3(L2, U2));
    } else {
        const M3<C> L2 =
            su3_nn(load_link<S, C>(U, d.stride, mu, s), load_link<S, C>(U, d.stride, mu, s + nb_off(g, mu, +1)));
        store_link<C, C>(lines, G, mu, li, L2);
        store_link<C, C>(lines + 4 * 18 * G, G, mu, li,
                         su3_nn(L2, load_link<S, C>(U, d.stride, mu, s + wrap_offset(g, mu, 2, d.rcp))));
    }
}

// One thread per (site, plane).  The six planes are taken in the order of
// the reference's charge, (3,2),(0,1) | (3,1),(2,0) | (3,0),(1,2)
// (LGF/Observables/TopologicalCharge.cpp:12-16); they visit each plane
// once, which the energy densities sum over (EnergyDensity.cpp:13-20,
// ImprovedEnergyDensity.cpp:13-20).  Per site:
//   e(x)   = sum_planes Re Tr(F11 F11)            (EnergyDensity)
//   q(x)   = sum_pairs  Re Tr(Fa Fb), improved F  (TopologicalCharge)
//   ei(x)  = sum_planes Re Tr(Fimp Fimp)          (ImprovedEnergyDensity)
//   q11(x) = sum_pairs  Re Tr(Fa Fb), 1x1 F       (UnImprovedTopologicalCharge,
//                                                  UnImprovedTopologicalCharge.cpp:13-15)
// Tiles write the kCloverParts partials (e, q, |q|, ei, q11) and, when
// requested, q(x).  Reads the line products of the same chunk from k_lines.

template <typename S, typename C, int WPE>
__global__ __launch_bounds__(384) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_clover(
    const FieldDesc* __restrict__ fds, const int2* __restrict__ tiles, int t0, int ping, const C* __restrict__ lines,
    int64_t G, double* __restrict__ partials) {
    // F values are computed in C; keeping them as C in LDS loses nothing
    using T = typename std::conditional<LGF_CLOVER_LDS_REAL != 0, C, double>::type;
    constexpr bool XL = kCloverXLines<S, C>;
    struct Out {
        T F[6][9][kTileSites];
        T F11[6][9][kTileSites];
        double E[6][kTileSites];
        double Ei[6][kTileSites];
    };
    union Shared {
        Out o;
        C xl[XL ? kXSlots<S, C> : 1][kXLen<S, C>][kTileSites];
    };
    __shared__ Shared sh;
    T (*sF)[9][kTileSites] = sh.o.F;
    T (*sF11)[9][kTileSites] = sh.o.F11;
    double (*sE)[kTileSites] = sh.o.E;
    double (*sEi)[kTileSites] = sh.o.Ei;
    const int tix = xcd_remap(blockIdx.x, gridDim.x);
    const int2 tm = tiles[tix];
    const FieldDesc& d = fds[tm.x];
    const int s = tm.y * kTileSites + threadIdx.x;
    const int p = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const bool live = s < d.vol;
    Lines<S, C, XL> L;
    L.U = (const S*)d.links[ping ^ d.flip];
    L.st = d.stride;
    L.L2 = lines;
    L.L3 = lines + 4 * 18 * G;
    L.G = G;
    L.lbase = (d.tile0 - t0) * kTileSites;
    L.xl = XL ? &sh.xl[0][0][0] : nullptr;
    L.lane = threadIdx.x;
    if constexpr (XL) {
        // wave p stages lines 2p and 2p+1 (line k: length k/4 + 1, direction k%4)
        if (live) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k = 2 * p + h, dir = k & 3;
                M3<C> m;
                if constexpr (kCloverRows<S, C>)
                    m = k < 4 ? load_link_rows01<S, C>(L.U, L.st, dir, s)
                              : load_link_rows01<C, C>(k < 8 ? L.L2 : L.L3, G, dir, L.lbase + s);
                else
                    m = k < 4 ? load_link<S, C>(L.U, L.st, dir, s)
                              : load_link<C, C>(k < 8 ? L.L2 : L.L3, G, dir, L.lbase + s);
#pragma unroll
                for (int c = 0; c < kXLen<S, C>; ++c) sh.xl[k][c][threadIdx.x] = m.v[c];
            }
            if constexpr (kXBack<S, C>) {
                // wave p < 5 stages backward line 12 + p
                if (p < 5) {
                    const Geo g = make_geo(d, s);
                    const int dir = p < 4 ? p : 3, len = p < 4 ? 1 : 2;
                    const int sb = s + wrap_offset(g, dir, -len, d.rcp);
                    const M3<C> m = len == 1 ? load_link<S, C>(L.U, L.st, dir, sb)
                                             : load_link<C, C>(L.L2, G, dir, L.lbase + sb);
#pragma unroll
                    for (int c = 0; c < 18; ++c) sh.xl[12 + p][c][threadIdx.x] = m.v[c];
                }
            }
        }
        __syncthreads();
    }
    double e = 0.0, ei = 0.0;
    T f11r[9], fir[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) f11r[i] = fir[i] = T(0);
    if (live) {
        const Geo g = make_geo(d, s);
        // plane p: (3,2),(0,1),(3,1),(2,0),(3,0),(1,2)
        const int mu = (p & 1) ? (p == 1 ? 0 : (p == 3 ? 2 : 1)) : 3;
        const int nu = (p & 1) ? (p == 1 ? 1 : (p == 3 ? 0 : 2)) : (p == 0 ? 2 : (p == 2 ? 1 : 0));
        M3<C> Fi;
        if constexpr (XL) {
            e = (double)plane_F<S, C, T>(L, g, d.rcp, mu, nu, Fi, f11r, 1);
            herm_pack(Fi, fir, 1);
        } else {
            e = (double)plane_F<S, C, T>(L, g, d.rcp, mu, nu, Fi, &sF11[p][0][threadIdx.x], kTileSites);
            herm_pack(Fi, &sF[p][0][threadIdx.x], kTileSites);
        }
        ei = (double)retr_mul(Fi, Fi);
    } else if constexpr (!XL) {
#pragma unroll
        for (int i = 0; i < 9; ++i) sF[p][i][threadIdx.x] = sF11[p][i][threadIdx.x] = T(0);
    }
    if constexpr (XL) {
        __syncthreads();  // every wave is done with the line cache
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            sF[p][i][threadIdx.x] = fir[i];
            sF11[p][i][threadIdx.x] = f11r[i];
        }
    }
    sE[p][threadIdx.x] = e;
    sEi[p][threadIdx.x] = ei;
    __syncthreads();
    if (p == 0) {
        T a[9], b[9];
        double qs = 0.0, q11 = 0.0, es = 0.0, eis = 0.0;
#pragma unroll 1
        for (int k = 0; k < 3; ++k) {
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                a[i] = sF[2 * k][i][threadIdx.x];
                b[i] = sF[2 * k + 1][i][threadIdx.x];
            }
            qs += herm_dot(a, b);
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                a[i] = sF11[2 * k][i][threadIdx.x];
                b[i] = sF11[2 * k + 1][i][threadIdx.x];
            }
            q11 += herm_dot(a, b);
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            es += sE[k][threadIdx.x];
            eis += sEi[k][threadIdx.x];
        }
        if (live && d.qx) d.qx[s] = qs;
        double qa = fabs(qs);
        // wavefront reduction (64 lanes), fixed order
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            es += __shfl_down(es, off, 64);
            qs += __shfl_down(qs, off, 64);
            qa += __shfl_down(qa, off, 64);
            eis += __shfl_down(eis, off, 64);
            q11 += __shfl_down(q11, off, 64);
        }
        if (threadIdx.x == 0) {
            double* o = partials + kCloverParts * (d.tile0 + tm.y);
            o[0] = es;
            o[1] = qs;
            o[2] = qa;
            o[3] = eis;
            o[4] = q11;
        }
    }
}

// The FP32 sweep takes the planes in pairs: a workgroup of two waves
// computes planes 2k and 2k+1 of one tile, the pairs whose products make
// q(x), so that the field strengths of a pair still meet in LDS.  Its two
// waves need no line cache and hold 11 KB of LDS, so a CU runs twenty of
// these waves (90 VGPRs) instead of the eighteen of three six-wave
// workgroups, and a barrier joins two waves instead of six: the 24^4 sweep
// takes 0.453-0.464 ms instead of 0.479-0.485.  (The FP64-arithmetic sweep
// over FP32 links gains nothing from the pairs, 0.92 ms either way, and
// keeps k_clover.)  Each workgroup leaves its pair's q(x)
// per site and its tile sums of e, ei and q11; k_clover_combine adds the
// three pairs of every site in the order k_clover does and writes the same
// tile partials (e, ei and q11 are summed per pair first, so their tile
// sums can differ from k_clover's in the last bits).
//
// Both go in the part of the line buffer that lines stored as rows 0 and 1
// leave free: elements [12 G, 18 G) of each direction block of 18 G (the
// third-row planes), 24 bytes per site in FP32; the pair charges (3 G
// doubles) take that of L_2 direction 0, the tile sums (9 doubles per tile)
// that of L_2 direction 1.
// the free planes of a direction block hold 6 G reals of C: 3 G doubles of
// pair charges, and 9 doubles per tile (9 G / 64) of tile sums
static_assert(3 * sizeof(double) <= 6 * sizeof(float) && 9 * sizeof(double) <= 6 * sizeof(float) * kTileSites,
              "pair charges and tile sums fit the free third-row planes");
template <typename C>
LGF_DEV double* pair_charges(C* lines, int64_t G) {
    return (double*)(lines + 12 * G);
}
template <typename C>
LGF_DEV const double* pair_charges(const C* lines, int64_t G) {
    return (const double*)(lines + 12 * G);
}
template <typename C>
LGF_DEV double* pair_sums(C* lines, int64_t G) {
    return (double*)(lines + 30 * G);
}
template <typename C>
LGF_DEV const double* pair_sums(const C* lines, int64_t G) {
    return (const double*)(lines + 30 * G);
}

template <typename S, typename C, int WPE>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_clover_pairs(
    const FieldDesc* __restrict__ fds, const int2* __restrict__ tiles, int t0, int ping, C* __restrict__ lines,
    int64_t G) {
    static_assert(kCloverRows<S, C>, "the pair charges use the planes that rows-0-and-1 lines leave free");
    using T = typename std::conditional<LGF_CLOVER_LDS_REAL != 0, C, double>::type;
    __shared__ T sF[2][9][kTileSites];
    __shared__ T sF11[2][9][kTileSites];
    __shared__ double sE[2][kTileSites];
    __shared__ double sEi[2][kTileSites];
    const int item = xcd_remap(blockIdx.x, gridDim.x);
    const int tix = item / 3, k = item - 3 * tix;
    const int2 tm = tiles[tix];
    const FieldDesc& d = fds[tm.x];
    const int s = tm.y * kTileSites + threadIdx.x;
    const int h = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int p = 2 * k + h;
    const bool live = s < d.vol;
    Lines<S, C, false> L;
    L.U = (const S*)d.links[ping ^ d.flip];
    L.st = d.stride;
    L.L2 = lines;
    L.L3 = lines + 4 * 18 * G;
    L.G = G;
    L.lbase = (d.tile0 - t0) * kTileSites;
    L.xl = nullptr;
    L.lane = threadIdx.x;
    double e = 0.0, ei = 0.0;
    if (live) {
        const Geo g = make_geo(d, s);
        // plane p: (3,2),(0,1),(3,1),(2,0),(3,0),(1,2)
        const int mu = (p & 1) ? (p == 1 ? 0 : (p == 3 ? 2 : 1)) : 3;
        const int nu = (p & 1) ? (p == 1 ? 1 : (p == 3 ? 0 : 2)) : (p == 0 ? 2 : (p == 2 ? 1 : 0));
        M3<C> Fi;
        e = (double)plane_F<S, C, T>(L, g, d.rcp, mu, nu, Fi, &sF11[h][0][threadIdx.x], kTileSites);
        herm_pack(Fi, &sF[h][0][threadIdx.x], kTileSites);
        ei = (double)retr_mul(Fi, Fi);
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) sF[h][i][threadIdx.x] = sF11[h][i][threadIdx.x] = T(0);
    }
    sE[h][threadIdx.x] = e;
    sEi[h][threadIdx.x] = ei;
    __syncthreads();
    if (h == 0) {
        T a[9], b[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) a[i] = sF[0][i][threadIdx.x], b[i] = sF[1][i][threadIdx.x];
        const double q = herm_dot(a, b);
#pragma unroll
        for (int i = 0; i < 9; ++i) a[i] = sF11[0][i][threadIdx.x], b[i] = sF11[1][i][threadIdx.x];
        double q11 = herm_dot(a, b);
        double es = sE[0][threadIdx.x] + sE[1][threadIdx.x];
        double eis = sEi[0][threadIdx.x] + sEi[1][threadIdx.x];
        pair_charges(lines, G)[k * G + L.lbase + s] = live ? q : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            es += __shfl_down(es, off, 64);
            eis += __shfl_down(eis, off, 64);
            q11 += __shfl_down(q11, off, 64);
        }
        if (threadIdx.x == 0) {
            double* o = pair_sums(lines, G) + 9 * (d.tile0 - t0 + tm.y) + 3 * k;
            o[0] = es;
            o[1] = eis;
            o[2] = q11;
        }
    }
}

// q(x) = the three pairs' charges of each site, |q|, and the tile partials
// (e, q, |q|, ei, q11) from the pair sums
template <typename C>
__global__ __launch_bounds__(64) void k_clover_combine(const FieldDesc* __restrict__ fds, const int2* __restrict__ tiles,
                                                       int t0, const C* __restrict__ lines, int64_t G,
                                                       double* __restrict__ partials) {
    const int tix = xcd_remap(blockIdx.x, gridDim.x);
    const int2 tm = tiles[tix];
    const FieldDesc& d = fds[tm.x];
    const int s = tm.y * kTileSites + threadIdx.x;
    const bool live = s < d.vol;
    const double* qp = pair_charges(lines, G) + (d.tile0 - t0) * kTileSites + s;
    double qs = 0.0;
    if (live) {
#pragma unroll
        for (int k = 0; k < 3; ++k) qs += qp[k * G];
    }
    if (live && d.qx) d.qx[s] = qs;
    double qa = fabs(qs);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        qs += __shfl_down(qs, off, 64);
        qa += __shfl_down(qa, off, 64);
    }
    if (threadIdx.x == 0) {
        const double* ps = pair_sums(lines, G) + 9 * (d.tile0 - t0 + tm.y);
        double* o = partials + kCloverParts * (d.tile0 + tm.y);
        o[0] = ps[0] + ps[3] + ps[6];
        o[1] = qs;
        o[2] = qa;
        o[3] = ps[1] + ps[4] + ps[7];
        o[4] = ps[2] + ps[5] + ps[8];
    }
}

// One workgroup per field: sum that field's tile partials (records of
// `pstride` doubles, components 0..ncomp-1) in a fixed order, so that a
// field's result does not depend on the other fields of its batch.
// out[kOutStride*f + c]; for the clover partials c = {sum e, sum q, sum |q|,
// sum ei, sum q11}.
constexpr int kReduceThreads = 1024;  // 16 waves: the tile loop is latency bound

__global__ __launch_bounds__(kReduceThreads) void k_reduce(const FieldDesc* __restrict__ fds,
                                                           const double* __restrict__ partials, int pstride,
                                                           int ncomp, double* __restrict__ out) {
    constexpr int kWaves = kReduceThreads / 64;
    __shared__ double red[8][kWaves];
    const FieldDesc& d = fds[blockIdx.x];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double acc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = 0.0;
    for (int t = threadIdx.x; t < d.ntiles; t += kReduceThreads) {
        const double* rec = partials + (int64_t)pstride * (d.tile0 + t);
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (c < ncomp) acc[c] += rec[c];
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_down(acc[c], off, 64);
        if (lane == 0) red[c][w] = acc[c];
    }
    __syncthreads();
    if ((int)threadIdx.x < ncomp) {
        const int c = threadIdx.x;
        double v[kWaves];
#pragma unroll
        for (int k = 0; k < kWaves; ++k) v[k] = red[c][k];
        // pairwise, in a fixed order
#pragma unroll
        for (int h = kWaves / 2; h > 0; h >>= 1)
#pragma unroll
            for (int k = 0; k < h; ++k) v[k] += v[k + h];
        out[kOutStride * blockIdx.x + c] = v[0];
    }
}

// ---------------------------------------------------------------------------
// flow-scale t0 on the device
// ---------------------------------------------------------------------------
// The same interpolation as lgf_flow_scale_t0 (the reference leaves t0 to its
// analysis scripts, DataAnalysis_and_Plotting/t0_extrapolation.py; t0 is
// defined by t^2 E(t0) = 0.3, doc/thesis.tex:1214-1216), applied one
// measurement at a time: E = sum e / V from the reduction that just ran.
__global__ __launch_bounds__(64) void k_t0_track(const FieldDesc* __restrict__ fds, const double* __restrict__ out,
                                                 T0Args a, double* __restrict__ rec) {
    const int k = threadIdx.x;
    if (k >= a.n) return;
    const int i = a.first + k;
    double* r = rec + 3 * i;
    const double t = a.t[k];
    const double y = t * t * (out[kOutStride * i] / (double)fds[i].vol);
    const double tp = r[0], yp = r[1];
    // a < 0 <= b with a, b the distances of the two t^2 E from the target
    if (isnan(r[2]) && !isnan(yp) && t > tp && yp - a.target < 0.0 && y - a.target >= 0.0)
        r[2] = tp + (t - tp) * (a.target - yp) / (y - yp);
    r[0] = t;
    r[1] = y;
}

__global__ __launch_bounds__(256) void k_t0_reset(double* __restrict__ rec, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3 * n) rec[i] = __builtin_nan("");
}

// ---------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------
// average plaquette (LGF/Field/SU3_field.cpp:485-499) partials: sum over the
// six planes of Re Tr U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag; and
// unitarity deviation max / sum of ||U^dag U - 1||_F over the 4 links.
template <typename S>
__global__ __launch_bounds__(64) void k_diag(const FieldDesc* __restrict__ fds, const int2* __restrict__ tiles,
                                             int ping, double* __restrict__ partials) {
    const int tix = xcd_remap(blockIdx.x, gridDim.x);
    const int2 tm = tiles[tix];
    const FieldDesc& d = fds[tm.x];
    const int s = tm.y * kTileSites + threadIdx.x;
    double pl = 0.0, umax = 0.0, usum = 0.0;
    if (s < d.vol) {
        const S* __restrict__ U = (const S*)d.links[ping ^ d.flip];
        const Geo g = make_geo(d, s);
        for (int mu = 0; mu < 4; ++mu) {
            M3<double> Um = load_link<S, double>(U, d.stride, mu, s);
            M3<double> P = mul_dn(Um, Um);
            P.v[0] -= 1.0;
            P.v[8] -= 1.0;
            P.v[16] -= 1.0;
            double n = 0.0;
#pragma unroll
            for (int i = 0; i < 18; ++i) n += P.v[i] * P.v[i];
            n = sqrt(n);
            umax = fmax(umax, n);
            usum += n;
            for (int nu = mu + 1; nu < 4; ++nu) {
                M3<double> a = mul_nn(Um, load_link<S, double>(U, d.stride, nu, s + nb_off(g, mu, +1)));
                M3<double> b = mul_nn(load_link<S, double>(U, d.stride, nu, s),
                                      load_link<S, double>(U, d.stride, mu, s + nb_off(g, nu, +1)));
                pl += retr(mul_nd(a, b));
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        pl += __shfl_down(pl, off, 64);
        usum += __shfl_down(usum, off, 64);
        umax = fmax(umax, __shfl_down(umax, off, 64));
    }
    if (threadIdx.x == 0) {
        const int gt = d.tile0 + tm.y;
        partials[3 * gt + 0] = pl;
        partials[3 * gt + 1] = usum;
        partials[3 * gt + 2] = umax;
    }
}

__global__ __launch_bounds__(256) void k_reduce_max(const FieldDesc* __restrict__ fds,
                                                    const double* __restrict__ partials, int ncomp, int comp,
                                                    double* __restrict__ out) {
    __shared__ double red[256];
    const FieldDesc& d = fds[blockIdx.x];
    double acc = 0.0;
    for (int t = threadIdx.x; t < d.ntiles; t += 256) acc = fmax(acc, partials[ncomp * (d.tile0 + t) + comp]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[kOutStride * blockIdx.x + comp] = red[0];
}

// ---------------------------------------------------------------------------
// layout conversion and initialisation
// ---------------------------------------------------------------------------
// file order (site-major, 4 links x 18 doubles) for sites [s0, s0+n) -> device layout
template <typename S>
__global__ __launch_bounds__(256) void k_fileorder_to_soa(const double* __restrict__ src, S* __restrict__ dst,
                                                          int64_t stride, int s0, int n) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // link within chunk
    if (l >= (int64_t)n * 4) return;
    M3<double> m;
#pragma unroll
    for (int c = 0; c < 18; ++c) m.v[c] = src[l * 18 + c];
    store_link<S, double>(dst, stride, (int)(l & 3), s0 + (int)(l >> 2), m);
}

template <typename S>
__global__ __launch_bounds__(256) void k_soa_to_fileorder(const S* __restrict__ src, double* __restrict__ dst,
                                                          int64_t stride, int s0, int n) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (int64_t)n * 4) return;
    const M3<double> m = load_link<S, double>(src, stride, (int)(l & 3), s0 + (int)(l >> 2));
#pragma unroll
    for (int c = 0; c < 18; ++c) dst[l * 18 + c] = m.v[c];
}

template <typename SI, typename SO>
__global__ __launch_bounds__(256) void k_convert(const SI* __restrict__ src, SO* __restrict__ dst, int64_t stride,
                                                 int vol) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (int64_t)vol * 4) return;
    const int s = (int)(l >> 2), mu = (int)(l & 3);
    store_link<SO, double>(dst, stride, mu, s, load_link<SI, double>(src, stride, mu, s));
}

// Seeded Haar-random SU(3) links: rows 0 and 1 from complex Gaussians
// (Box-Muller on Philox uniforms, counter (link, i, ~0, 3) as in the CPU
// oracle's hot start), orthonormalised, row 2 = conj(row0 x row1).
template <typename S>
__global__ __launch_bounds__(256) void k_hot(S* __restrict__ dst, int64_t stride, int vol, uint64_t seed) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // link = 4*site + mu
    if (l >= (int64_t)vol * 4) return;
    const int s = (int)(l >> 2), mu = (int)(l & 3);
    M3<double> m;
    set_zero(m);
    for (int i = 0; i < 6; ++i) {
        double u[2];
        uniform2(seed, (uint32_t)l, (uint32_t)i, 0xFFFFFFFFu, 3u, u);
        const double r = sqrt(-2.0 * log(u[0]));
        m.v[2 * i] = r * cos(2.0 * M_PI * u[1]);
        m.v[2 * i + 1] = r * sin(2.0 * M_PI * u[1]);
    }
    reunitarize(m);
    store_link<S, double>(dst, stride, mu, s, m);
}

template <typename S>
__global__ __launch_bounds__(256) void k_cold(S* __restrict__ dst, int64_t stride, int vol) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (int64_t)stride * 4) return;  // padding sites too: every slot holds a valid number
    M3<double> m;
    set_unit(m);
    store_link<S, double>(dst, stride, (int)(l & 3), (int)(l >> 2), m);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename S, typename C, int EXPMODE, int WPE, int BY, int MPT = 1>
static hipError_t stage_launch(int stage, const FieldDesc* fds, const int2* tiles, int ntiles, int ping,
                               hipStream_t stream) {
    dim3 grid(ntiles, MPT > 1 ? 1 : 4 / BY), block(kTileSites, BY);
    if (stage == 0)
        hipLaunchKernelGGL((k_rk3_stage<S, C, 0, EXPMODE, WPE, BY, MPT>), grid, block, 0, stream, fds, tiles, ping);
    else if (stage == 1)
        hipLaunchKernelGGL((k_rk3_stage<S, C, 1, EXPMODE, WPE, BY, MPT>), grid, block, 0, stream, fds, tiles, ping);
    else
        hipLaunchKernelGGL((k_rk3_stage<S, C, 2, EXPMODE, WPE, BY, MPT>), grid, block, 0, stream, fds, tiles, ping);
    return hipGetLastError();
}

template <bool NTS>
static hipError_t stage_launch_pb(int stage, const FieldDesc* fds, const int2* tiles, int ntiles, int ping,
                                  hipStream_t stream) {
    dim3 grid(ntiles), block(kTileSites, 4);
    if (stage == 0)
        hipLaunchKernelGGL((k_rk3_stage_pb<float, float, 0, 4, NTS>), grid, block, 0, stream, fds, tiles, ping);
    else if (stage == 1)
        hipLaunchKernelGGL((k_rk3_stage_pb<float, float, 1, 4, NTS>), grid, block, 0, stream, fds, tiles, ping);
    else
        hipLaunchKernelGGL((k_rk3_stage_pb<float, float, 2, 4, NTS>), grid, block, 0, stream, fds, tiles, ping);
    return hipGetLastError();
}

template <int WPE, bool ORDERED, bool LOC = false, bool NTS = false, typename S = float>
static hipError_t stage_launch_pl(int stage, const FieldDesc* fds, const int2* tiles, int ntiles, int ping,
                                  hipStream_t stream) {
    dim3 grid(ntiles), block(kTileSites, 4);
    if (stage == 0)
        hipLaunchKernelGGL((k_rk3_stage_pl<S, S, 0, WPE, ORDERED, LOC, NTS>), grid, block, 0, stream, fds, tiles, ping);
    else if (stage == 1)
        hipLaunchKernelGGL((k_rk3_stage_pl<S, S, 1, WPE, ORDERED, LOC, NTS>), grid, block, 0, stream, fds, tiles, ping);
    else
        hipLaunchKernelGGL((k_rk3_stage_pl<S, S, 2, WPE, ORDERED, LOC, NTS>), grid, block, 0, stream, fds, tiles, ping);
    return hipGetLastError();
}

// Stage-kernel variants: (register budget as minimum waves per SIMD, launch
// shape).  Mixed-precision variants 0..5 are the shared-plaquette kernel
// (0, 1 and 5 with the site's links in LDS, 5 with non-temporal stores),
// 6 and 7 the shared-plaquette kernel that also shares the backward links
// (7 with non-temporal stores), 8..13 the per-link kernel.  Variant 0 is the default (a budget the kernel
// fits without spilling); lgf_autotune picks the fastest.
// FP64 variants 0..2 are the shared-plaquette kernel (0 and 2 with the
// site's links in LDS, 2 with non-temporal stores), 3..7 the per-link
// kernel, which sums each link's staples in the reference's order.
int num_stage_variants(int prec) { return prec == kPrecFP64 ? 8 : 14; }

int stage_variant_group(int prec, int variant) {
    if (prec == kPrecFP64) return variant < LGF_FP64_SHARED_PLAQUETTE_VARIANTS ? 0 : 1;
    return variant < LGF_FP32_COMPACT_VARIANTS ? 0 : variant < LGF_FP32_SHARED_PLAQUETTE_VARIANTS ? 1 : 2;
}

hipError_t launch_rk3_stage_variant(int prec, int variant, int stage, const FieldDesc* fds, const int2* tiles,
                                    int ntiles, int ping, hipStream_t stream) {
    if (prec == kPrecFP64) {
        switch (variant) {
            case 1: return stage_launch_pl<2, true, false, false, double>(stage, fds, tiles, ntiles, ping, stream);
            case 2: return stage_launch_pl<2, true, true, true, double>(stage, fds, tiles, ntiles, ping, stream);
            case 3: return stage_launch<double, double, 0, 2, 4>(stage, fds, tiles, ntiles, ping, stream);
            case 4: return stage_launch<double, double, 0, 3, 4>(stage, fds, tiles, ntiles, ping, stream);
            case 5: return stage_launch<double, double, 0, 4, 4>(stage, fds, tiles, ntiles, ping, stream);
            case 6: return stage_launch<double, double, 0, 3, 1>(stage, fds, tiles, ntiles, ping, stream);
            case 7: return stage_launch<double, double, 0, 2, 1>(stage, fds, tiles, ntiles, ping, stream);
            default: return stage_launch_pl<2, true, true, false, double>(stage, fds, tiles, ntiles, ping, stream);
        }
    }
    switch (variant) {
        case 1: return stage_launch_pl<4, false, true>(stage, fds, tiles, ntiles, ping, stream);
        case 2: return stage_launch_pl<4, true>(stage, fds, tiles, ntiles, ping, stream);
        case 3: return stage_launch_pl<4, false>(stage, fds, tiles, ntiles, ping, stream);
        case 4: return stage_launch_pl<5, true>(stage, fds, tiles, ntiles, ping, stream);
        case 5: return stage_launch_pl<4, true, true, true>(stage, fds, tiles, ntiles, ping, stream);
        case 6: return stage_launch_pb<false>(stage, fds, tiles, ntiles, ping, stream);
        case 7: return stage_launch_pb<true>(stage, fds, tiles, ntiles, ping, stream);
        case 8: return stage_launch<float, float, 1, 4, 4>(stage, fds, tiles, ntiles, ping, stream);
        case 9: return stage_launch<float, float, 1, 3, 4>(stage, fds, tiles, ntiles, ping, stream);
        case 10: return stage_launch<float, float, 1, 5, 4>(stage, fds, tiles, ntiles, ping, stream);
        case 11: return stage_launch<float, float, 1, 6, 4>(stage, fds, tiles, ntiles, ping, stream);
        case 12: return stage_launch<float, float, 1, 5, 1>(stage, fds, tiles, ntiles, ping, stream);
        case 13: return stage_launch<float, float, 1, 6, 1>(stage, fds, tiles, ntiles, ping, stream);
        default: return stage_launch_pl<4, true, true>(stage, fds, tiles, ntiles, ping, stream);
    }
}

hipError_t launch_rk3_stage(int prec, int stage, const FieldDesc* fds, const int2* tiles, int ntiles, int ping,
                            hipStream_t stream) {
    return launch_rk3_stage_variant(prec, 0, stage, fds, tiles, ntiles, ping, stream);
}

size_t line_bytes_per_site(int compute) { return size_t(2 * 4 * 18) * (compute == kPrecFP64 ? 8 : 4); }

hipError_t launch_clover(int prec, int compute, const FieldDesc* fds, const int2* tiles, int t0, int ntiles, int ping,
                         void* lines, double* partials, hipStream_t stream) {
    const int2* ct = tiles + t0;
    const int64_t G = (int64_t)ntiles * kTileSites;
    dim3 grid(ntiles), lblock(kTileSites, 4), block(kTileSites, 6), pgrid(3 * ntiles), pblock(kTileSites, 2);
    if (prec == kPrecFP64) {
        hipLaunchKernelGGL((k_lines<double, double>), grid, lblock, 0, stream, fds, ct, t0, ping, (double*)lines, G);
        hipLaunchKernelGGL((k_clover<double, double, LGF_CLOVER_WPE64_PARITY>), grid, block, 0, stream, fds, ct, t0, ping,
                           (const double*)lines, G, partials);
    } else if (compute == kPrecFP64) {
        hipLaunchKernelGGL((k_lines<float, double>), grid, lblock, 0, stream, fds, ct, t0, ping, (double*)lines, G);
        hipLaunchKernelGGL((k_clover<float, double, LGF_CLOVER_WPE64>), grid, block, 0, stream, fds, ct, t0, ping,
                           (const double*)lines, G, partials);
    } else {
        hipLaunchKernelGGL((k_lines<float, float>), grid, lblock, 0, stream, fds, ct, t0, ping, (float*)lines, G);
        if (LGF_CLOVER_PAIRS) {
            hipLaunchKernelGGL((k_clover_pairs<float, float, LGF_CLOVER_WPE_PAIRS>), pgrid, pblock, 0, stream, fds, ct,
                               t0, ping, (float*)lines, G);
            hipLaunchKernelGGL((k_clover_combine<float>), grid, dim3(kTileSites), 0, stream, fds, ct, t0,
                               (const float*)lines, G, partials);
        } else {
            hipLaunchKernelGGL((k_clover<float, float, LGF_CLOVER_WPE>), grid, block, 0, stream, fds, ct, t0, ping,
                               (const float*)lines, G, partials);
        }
    }
    return hipGetLastError();
}

hipError_t launch_reduce(const FieldDesc* fds, int nfields, const double* partials, int pstride, int ncomp, double* out,
                         hipStream_t stream) {
    hipLaunchKernelGGL(k_reduce, dim3(nfields), dim3(kReduceThreads), 0, stream, fds, partials, pstride, ncomp, out);
    return hipGetLastError();
}

hipError_t launch_diag(int prec, const FieldDesc* fds, const int2* tiles, int ntiles, int ping, double* partials,
                       int nfields, double* out, hipStream_t stream) {
    if (prec == kPrecFP64)
        hipLaunchKernelGGL((k_diag<double>), dim3(ntiles), dim3(64), 0, stream, fds, tiles, ping, partials);
    else
        hipLaunchKernelGGL((k_diag<float>), dim3(ntiles), dim3(64), 0, stream, fds, tiles, ping, partials);
    hipLaunchKernelGGL(k_reduce, dim3(nfields), dim3(kReduceThreads), 0, stream, fds, partials, 3, 2, out);
    hipLaunchKernelGGL(k_reduce_max, dim3(nfields), dim3(256), 0, stream, fds, partials, 3, 2, out);
    return hipGetLastError();
}

hipError_t launch_t0_track(const FieldDesc* fds, const double* out, const T0Args& a, double* rec, hipStream_t stream) {
    hipLaunchKernelGGL(k_t0_track, dim3(1), dim3(64), 0, stream, fds, out, a, rec);
    return hipGetLastError();
}

hipError_t launch_t0_reset(double* rec, int nfields, hipStream_t stream) {
    hipLaunchKernelGGL(k_t0_reset, dim3((3 * nfields + 255) / 256), dim3(256), 0, stream, rec, nfields);
    return hipGetLastError();
}

hipError_t launch_to_soa(int prec, const double* src, void* dst, int64_t stride, int s0, int n, hipStream_t stream) {
    int64_t total = (int64_t)n * 4;
    dim3 grid((unsigned)((total + 255) / 256)), block(256);
    if (prec == kPrecFP64)
        hipLaunchKernelGGL(k_fileorder_to_soa<double>, grid, block, 0, stream, src, (double*)dst, stride, s0, n);
    else
        hipLaunchKernelGGL(k_fileorder_to_soa<float>, grid, block, 0, stream, src, (float*)dst, stride, s0, n);
    return hipGetLastError();
}

hipError_t launch_from_soa(int prec, const void* src, double* dst, int64_t stride, int s0, int n,
                           hipStream_t stream) {
    int64_t total = (int64_t)n * 4;
    dim3 grid((unsigned)((total + 255) / 256)), block(256);
    if (prec == kPrecFP64)
        hipLaunchKernelGGL(k_soa_to_fileorder<double>, grid, block, 0, stream, (const double*)src, dst, stride, s0, n);
    else
        hipLaunchKernelGGL(k_soa_to_fileorder<float>, grid, block, 0, stream, (const float*)src, dst, stride, s0, n);
    return hipGetLastError();
}

hipError_t launch_convert(int prec_in, const void* src, int prec_out, void* dst, int64_t stride, int vol,
                          hipStream_t stream) {
    dim3 grid((unsigned)(((int64_t)vol * 4 + 255) / 256)), block(256);
    if (prec_in == kPrecFP64 && prec_out == kPrecFP64)
        hipLaunchKernelGGL((k_convert<double, double>), grid, block, 0, stream, (const double*)src, (double*)dst,
                           stride, vol);
    else if (prec_in == kPrecFP64)
        hipLaunchKernelGGL((k_convert<double, float>), grid, block, 0, stream, (const double*)src, (float*)dst,
                           stride, vol);
    else if (prec_out == kPrecFP64)
        hipLaunchKernelGGL((k_convert<float, double>), grid, block, 0, stream, (const float*)src, (double*)dst,
                           stride, vol);
    else
        hipLaunchKernelGGL((k_convert<float, float>), grid, block, 0, stream, (const float*)src, (float*)dst,
                           stride, vol);
    return hipGetLastError();
}

hipError_t launch_hot(int prec, void* dst, int64_t stride, int vol, uint64_t seed, hipStream_t stream) {
    dim3 grid((unsigned)(((int64_t)vol * 4 + 255) / 256)), block(256);
    if (prec == kPrecFP64)
        hipLaunchKernelGGL(k_hot<double>, grid, block, 0, stream, (double*)dst, stride, vol, seed);
    else
        hipLaunchKernelGGL(k_hot<float>, grid, block, 0, stream, (float*)dst, stride, vol, seed);
    return hipGetLastError();
}

hipError_t launch_cold(int prec, void* dst, int64_t stride, int vol, hipStream_t stream) {
    int64_t total = (int64_t)4 * stride;
    dim3 grid((unsigned)((total + 255) / 256)), block(256);
    if (prec == kPrecFP64)
        hipLaunchKernelGGL(k_cold<double>, grid, block, 0, stream, (double*)dst, stride, vol);
    else
        hipLaunchKernelGGL(k_cold<float>, grid, block, 0, stream, (float*)dst, stride, vol);
    return hipGetLastError();
}

}  // namespace lgf
