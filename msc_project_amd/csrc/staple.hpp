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

}  // namespace lgf
