// Staple sum of the Wilson gauge action (LGF/Field/SU3_field.cpp:519-530):
//   A_mu(x) = sum_{nu != mu} U_nu(x+mu) (U_nu(x) U_mu(x+nu))^dag
//                          + (U_mu(x-nu) U_nu(x-nu+mu))^dag U_nu(x-nu)
// so that U_mu(x) A_mu(x) is the sum of the six plaquettes through U_mu(x).
// Shared by the flow-stage and the heatbath kernels.
#pragma once
#include "lattice.hpp"
#include "su3.hpp"

namespace lgf {

template <typename S, typename C>
LGF_DEV M3<C> staple(const S* __restrict__ U, int64_t st, const Geo& g, int mu) {
    const int s = g.s;
    const int fmu = nb_off(g, mu, +1);
    M3<C> A;
    set_zero(A);
#pragma unroll 1
    for (int k = 0; k < 3; ++k) {
        const int nu = k + (k >= mu ? 1 : 0);
        const int fnu = nb_off(g, nu, +1), bnu = nb_off(g, nu, -1);
        {
            M3<C> a = load_link<S, C>(U, st, nu, s);
            M3<C> b = load_link<S, C>(U, st, mu, s + fnu);
            M3<C> tt = mul_nn(a, b);
            M3<C> c = load_link<S, C>(U, st, nu, s + fmu);
            mul_nd_acc(A, c, tt);
        }
        {
            M3<C> a = load_link<S, C>(U, st, mu, s + bnu);
            M3<C> b = load_link<S, C>(U, st, nu, s + bnu + fmu);
            M3<C> tt = mul_nn(a, b);
            M3<C> c = load_link<S, C>(U, st, nu, s + bnu);
            mul_dn_acc(A, tt, c);
        }
    }
    return A;
}

// Half-staple product U V of two links.  The mixed-precision path (FP32
// arithmetic, links projected onto SU(3) every step) forms it from two rows
// and their cross product, 96 instead of 108 FMAs; the FP64 path keeps the
// full product, as the reference, whose links are never projected and
// drift from SU(3) by up to ~1e-14 over a flow.
template <typename R>
LGF_DEV M3<R> mul_halfstaple(const M3<R>& a, const M3<R>& b) {
    return su3_nn(a, b);
}

// lower term of the (MU,NU) plane: A += (U_mu(x-nu) U_nu(x-nu+mu))^dag U_nu(x-nu)
template <typename S, typename C, int MU, int NU>
LGF_DEV void lower_term(M3<C>& A, const S* __restrict__ U, int64_t st, const Geo& g, int fmu) {
    const int s = g.s;
    const int bnu = nb_off(g, NU, -1);
    M3<C> a = load_link<S, C>(U, st, MU, s + bnu);
    M3<C> b = load_link<S, C>(U, st, NU, s + bnu + fmu);
    M3<C> tt = mul_halfstaple(a, b);
    M3<C> c = load_link<S, C>(U, st, NU, s + bnu);
    mul_dn_acc(A, tt, c);
}

// The same sum with the direction a compile-time constant: every neighbour
// offset and plane base is then a constant expression of the site's
// coordinates, with no selects on a run-time direction.  The terms are added
// in the order of staple() (nu ascending, upper before lower).
template <typename S, typename C, int MU, int NU>
LGF_DEV void staple_term(M3<C>& A, const S* __restrict__ U, int64_t st, const Geo& g, int fmu) {
    const int s = g.s;
    const int fnu = nb_off(g, NU, +1);
    {
        M3<C> a = load_link<S, C>(U, st, NU, s);
        M3<C> b = load_link<S, C>(U, st, MU, s + fnu);
        M3<C> tt = mul_halfstaple(a, b);
        M3<C> c = load_link<S, C>(U, st, NU, s + fmu);
        mul_nd_acc(A, c, tt);
    }
    lower_term<S, C, MU, NU>(A, U, st, g, fmu);
}

// Site index s passed through an asm that takes `dep`: loads addressed from
// the result cannot be issued before dep is computed, so the links of a
// chain of products are in flight one step at a time.
template <typename C>
LGF_DEV int after_value(int s, C dep) {
    asm volatile("" : "+v"(s) : "v"(dep));
    return s;
}

// The staple term of the (MU,NU) plane as chains of products with the third
// link of each half loaded after the first product:
//   upper  U_nu(x+mu) (U_nu(x) U_mu(x+nu))^dag = (U_nu(x+mu) U_mu(x+nu)^dag) U_nu(x)^dag
//   lower  (U_mu(x-nu) U_nu(x-nu+mu))^dag U_nu(x-nu) = U_nu(x-nu+mu)^dag (U_mu(x-nu)^dag U_nu(x-nu))
// The same products as staple_term regrouped, so at most two links and one
// partial product are live besides the sum: an FP64 matrix less at the
// register peak.  Terms follow each other the same way.
template <typename S, typename C, int MU, int NU>
LGF_DEV void staple_term_chain(M3<C>& A, const S* __restrict__ U, int64_t st, const Geo& g, int fmu) {
    const int s = after_value(g.s, A.v[17]);
    const int fnu = nb_off(g, NU, +1), bnu = nb_off(g, NU, -1);
    {
        const M3<C> M = su3_nd(load_link<S, C>(U, st, NU, s + fmu), load_link<S, C>(U, st, MU, s + fnu));
        mul_nd_acc(A, M, load_link<S, C>(U, st, NU, after_value(s, M.v[17])));
    }
    {
        const int s2 = after_value(s, A.v[17]);
        const M3<C> M = su3_dn(load_link<S, C>(U, st, MU, s2 + bnu), load_link<S, C>(U, st, NU, s2 + bnu));
        mul_dn_acc(A, load_link<S, C>(U, st, NU, after_value(s2, M.v[17]) + bnu + fmu), M);
    }
}

template <typename S, typename C, int MU>
LGF_DEV M3<C> staple_ct(const S* __restrict__ U, int64_t st, const Geo& g) {
    constexpr int N0 = MU == 0 ? 1 : 0, N1 = MU <= 1 ? 2 : 1, N2 = MU <= 2 ? 3 : 2;
    const int fmu = nb_off(g, MU, +1);
    M3<C> A;
    set_zero(A);
    staple_term_chain<S, C, MU, N0>(A, U, st, g, fmu);
    staple_term_chain<S, C, MU, N1>(A, U, st, g, fmu);
    staple_term_chain<S, C, MU, N2>(A, U, st, g, fmu);
    return A;
}

}  // namespace lgf
