// Seeded SU(3) configuration generator: Cabibbo-Marinari heatbath over the
// three SU(2) subgroups with the Kennedy-Pendleton SU(2) sampler, followed by
// microcanonical overrelaxation, on a checkerboard (all links of one
// direction and one site parity at once).  The reference's generator
// (LGF/SU3_Heatbath/SU3_Heatbath.cpp:16-224) is the same algorithm but draws
// from an unseeded std::random_device stream; here every random number comes
// from Philox4x32-10 at a counter fixed by (site, direction, subgroup, draw,
// sweep), so a configuration is a pure function of (seed, sweeps), whatever
// the launch shape.  The CPU oracle (oracle/lgf_oracle.c, orc_generate) uses
// the same counters; the two agree to rounding.
//
// Weight of a link U with staple sum A: exp(beta/3 Re Tr(U A)).  For the
// SU(2) block (p,q) of W = U A with quaternion projection k*v (v in SU(2)),
// the update R = y v^dag is drawn with y0 ~ sqrt(1-y0^2) exp(2 beta k y0 / 3).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "expm.hpp"
#include "kernels.hpp"
#include "lattice.hpp"
#include "philox.hpp"
#include "staple.hpp"
#include "su3.hpp"

namespace lgf {

namespace {

struct Quat {
    double x0, x1, x2, x3;
};

LGF_DEV Quat qmul(const Quat& x, const Quat& y) {
    return Quat{x.x0 * y.x0 - x.x1 * y.x1 - x.x2 * y.x2 - x.x3 * y.x3,
                x.x0 * y.x1 + x.x1 * y.x0 - (x.x2 * y.x3 - x.x3 * y.x2),
                x.x0 * y.x2 + x.x2 * y.x0 - (x.x3 * y.x1 - x.x1 * y.x3),
                x.x0 * y.x3 + x.x3 * y.x0 - (x.x1 * y.x2 - x.x2 * y.x1)};
}

// rows (p,q) of u <- r * rows (p,q), r = [[a, b], [-conj(b), conj(a)]],
// a = r0 + i r3, b = r2 + i r1
LGF_DEV void su2_left(M3<double>& u, int p, int q, const Quat& r) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double pr = u.v[2 * (3 * p + c)], pi = u.v[2 * (3 * p + c) + 1];
        const double qr = u.v[2 * (3 * q + c)], qi = u.v[2 * (3 * q + c) + 1];
        u.v[2 * (3 * p + c)] = r.x0 * pr - r.x3 * pi + r.x2 * qr - r.x1 * qi;
        u.v[2 * (3 * p + c) + 1] = r.x0 * pi + r.x3 * pr + r.x2 * qi + r.x1 * qr;
        u.v[2 * (3 * q + c)] = -r.x2 * pr - r.x1 * pi + r.x0 * qr + r.x3 * qi;
        u.v[2 * (3 * q + c) + 1] = -r.x2 * pi + r.x1 * pr + r.x0 * qi - r.x3 * qr;
    }
}

}  // namespace

// One checkerboard pass: every link U_mu(x) with parity(x) == par gets the
// three subgroup updates (heatbath when hb, else overrelaxation).
template <typename S>
__global__ __launch_bounds__(256) void k_heatbath(const FieldDesc* __restrict__ fds, int ping, int mu, int par, int hb,
                                                  double beta, uint64_t seed, uint32_t sweep) {
    const FieldDesc& d = fds[0];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;  // t-th site of this parity
    if (2 * t >= d.vol) return;
    // extents are even, so sites 2t and 2t+1 share their x3-row and have
    // opposite parities
    Geo g = make_geo(d, 2 * t);
    if (((g.x0 + g.x1 + g.x2 + g.x3) & 1) != par) {
        g.s += 1;
        g.x3 += 1;
    }
    const int s = g.s;
    S* __restrict__ U = (S*)d.links[ping ^ d.flip];
    const int64_t st = d.stride;
    const M3<double> A = staple<S, double>(U, st, g, mu);
    M3<double> u = load_link<S, double>(U, st, mu, s);
    for (int j = 0; j < 3; ++j) {
        const int p = j == 2 ? 1 : 0, q = j == 0 ? 1 : 2;
        const M3<double> W = mul_nn(u, A);
        const double a0 = 0.5 * (W.v[2 * (3 * p + p)] + W.v[2 * (3 * q + q)]);
        const double a3 = 0.5 * (W.v[2 * (3 * p + p) + 1] - W.v[2 * (3 * q + q) + 1]);
        const double a2 = 0.5 * (W.v[2 * (3 * p + q)] - W.v[2 * (3 * q + p)]);
        const double a1 = 0.5 * (W.v[2 * (3 * p + q) + 1] + W.v[2 * (3 * q + p) + 1]);
        const double k = sqrt(a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3);
        if (k < 1e-12) continue;
        const Quat v{a0 / k, -a1 / k, -a2 / k, -a3 / k};
        Quat r;
        if (!hb) {
            r = qmul(v, v);
        } else {
            const double alpha = 2.0 * beta * k / 3.0;
            double uu[2], u2[2], delta = 0.0;
            uint32_t call = 0;
            for (;;) {
                uniform2(seed, (uint32_t)s, (uint32_t)mu | (uint32_t)j << 2 | call << 4, sweep, 1u, uu);
                call++;
                uniform2(seed, (uint32_t)s, (uint32_t)mu | (uint32_t)j << 2 | call << 4, sweep, 1u, u2);
                call++;
                const double cs = cos(2.0 * M_PI * u2[0]);
                delta = (-log(uu[0]) - cs * cs * log(uu[1])) / alpha;
                if (u2[1] * u2[1] <= 1.0 - 0.5 * delta) break;
                if (call > (1u << 26)) {  // bounded: acceptance is > 0.75 for any alpha
                    delta = 0.0;
                    break;
                }
            }
            uniform2(seed, (uint32_t)s, (uint32_t)mu | (uint32_t)j << 2 | call << 4, sweep, 2u, uu);
            const double ct = 2.0 * uu[0] - 1.0, sn = sqrt(fmax(0.0, 1.0 - ct * ct)), phi = 2.0 * M_PI * uu[1];
            const double y0 = 1.0 - delta, ay = sqrt(fmax(0.0, 1.0 - y0 * y0));
            r = qmul(Quat{y0, ay * sn * cos(phi), ay * sn * sin(phi), ay * ct}, v);
        }
        su2_left(u, p, q, r);
    }
    store_link<S, double>(U, st, mu, s, u);
}

template <typename S>
__global__ __launch_bounds__(256) void k_reunitarize(const FieldDesc* __restrict__ fds, int ping) {
    const FieldDesc& d = fds[0];
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (int64_t)d.vol * 4) return;
    const int s = (int)(l >> 2), mu = (int)(l & 3);
    S* __restrict__ U = (S*)d.links[ping ^ d.flip];
    M3<double> u = load_link<S, double>(U, d.stride, mu, s);
    reunitarize(u);
    store_link<S, double>(U, d.stride, mu, s, u);
}

hipError_t launch_heatbath_iteration(int prec, const FieldDesc* fds, int vol, int ping, double beta, uint64_t seed,
                                     uint32_t sweep, int or_per_hb, hipStream_t stream) {
    const int half = vol / 2;
    dim3 grid((unsigned)((half + 255) / 256)), block(256);
    for (int pass = 0; pass <= or_per_hb; ++pass) {
        const int hb = pass == 0;
        for (int par = 0; par < 2; ++par)
            for (int mu = 0; mu < 4; ++mu) {
                if (prec == kPrecFP64)
                    hipLaunchKernelGGL(k_heatbath<double>, grid, block, 0, stream, fds, ping, mu, par, hb, beta, seed,
                                       sweep);
                else
                    hipLaunchKernelGGL(k_heatbath<float>, grid, block, 0, stream, fds, ping, mu, par, hb, beta, seed,
                                       sweep);
            }
    }
    dim3 g2((unsigned)(((int64_t)vol * 4 + 255) / 256));
    if (prec == kPrecFP64)
        hipLaunchKernelGGL(k_reunitarize<double>, g2, block, 0, stream, fds, ping);
    else
        hipLaunchKernelGGL(k_reunitarize<float>, g2, block, 0, stream, fds, ping);
    return hipGetLastError();
}

}  // namespace lgf
