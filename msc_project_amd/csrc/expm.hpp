// exp(X) for X anti-Hermitian and traceless (an su(3) element), as
// f0 + f1 Q + f2 Q^2 with Q = -iX (Cayley-Hamilton: every power of Q is a
// combination of 1, Q, Q^2), the f_j summed from the exponential series with
// the recurrence Q^{n+1} = c0 d_n + (a_n + c1 d_n) Q + b_n Q^2 for
// Q^n = a_n + b_n Q + d_n Q^2 (c0 = det Q, c1 = Tr Q^2 / 2).  No
// transcendental functions and no data-dependent branches but the rare
// scaling of a large exponent.  The reference computes the same exponential
// in closed form (HermTrLessExp, LGF/Math/SU3_mat.cpp:216-289); both are
// accurate to rounding, and the FP64 flow with this series agrees with the
// reference's to ~1e-14 in E over its golden runs.
#pragma once
#include "su3.hpp"

namespace lgf {

// Q^2 and Re Tr(Q^3) for Hermitian Q (the exponent's -iX, X anti-Hermitian
// with exactly mirrored lower triangle): Q^2 is Hermitian with a real
// diagonal, so only its diagonal and upper triangle are computed (54 FMAs
// instead of 108) and the lower triangle is their conjugate.
// Q's imaginary diagonal is zero and never read: the exponents here are
// anti-Hermitian with an exactly zero real diagonal, and x + a * 0 is not
// folded away without fast-math.  The terms that remain are summed in the
// order of the full product, so the result is the same to the last bit.
template <typename R>
LGF_DEV void herm_square(const M3<R>& Q, M3<R>& Q2, R& tr3) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        R d = R(0);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            d += Q.v[2 * (3 * r + k)] * Q.v[2 * (3 * r + k)];
            if (k != r) d += Q.v[2 * (3 * r + k) + 1] * Q.v[2 * (3 * r + k) + 1];
        }
        Q2.v[8 * r] = d;
        Q2.v[8 * r + 1] = R(0);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = r + 1; c < 3; ++c) {
            R x = R(0), y = R(0);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const R ar = Q.v[2 * (3 * r + k)], ai = Q.v[2 * (3 * r + k) + 1];
                const R br = Q.v[2 * (3 * k + c)], bi = Q.v[2 * (3 * k + c) + 1];
                if (k == r) {  // a real
                    x += ar * br;
                    y += ar * bi;
                } else if (k == c) {  // b real
                    x += ar * br;
                    y += ai * br;
                } else {
                    LGF_CMAC(x, y, ar, ai, br, bi)
                }
            }
            Q2.v[2 * (3 * r + c)] = x;
            Q2.v[2 * (3 * r + c) + 1] = y;
            Q2.v[2 * (3 * c + r)] = x;
            Q2.v[2 * (3 * c + r) + 1] = -y;
        }
    // Tr(Q Q^2) = sum_r Q_rr Q2_rr + 2 sum_{r<c} Re(Q_rc conj(Q2_rc))
    R t = Q.v[0] * Q2.v[0];
    t += Q.v[8] * Q2.v[8];
    t += Q.v[16] * Q2.v[16];
    R u = R(0);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = r + 1; c < 3; ++c) {
            u += Q.v[2 * (3 * r + c)] * Q2.v[2 * (3 * r + c)];
            u += Q.v[2 * (3 * r + c) + 1] * Q2.v[2 * (3 * r + c) + 1];
        }
    tr3 = t + R(2) * u;
}

// Power series through the Cayley-Hamilton recurrence, real type R.
// Truncated after Q^(NTERMS-1): with c1 <= C1MAX every eigenvalue satisfies
// |lambda| <= sqrt(4 c1 / 3) and the first omitted term is below
// |lambda|^NTERMS / NTERMS!.  With c1 <= 0.02 (|lambda| <= 0.164) that is
// 7e-10 for the default 7 terms, far below FP32 rounding, and 1e-20 for the
// 13 terms of the FP64 path; a flow step's exponent eps*Z has c1 ~
// 1e-4..1e-3 at eps = 0.01.  Larger exponents are scaled by 2^-k into that
// range and the result squared k times, all in precision R.
template <typename R, int NTERMS = 7>
LGF_DEV M3<R> exp_ah_series(const M3<R>& X) {
    constexpr float C1MAX = 0.02f;
    // Q = -i X is Hermitian; X's real diagonal is zero, so Q's imaginary
    // diagonal is left zero and never read
    M3<R> Q;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        Q.v[2 * i] = X.v[2 * i + 1];
        Q.v[2 * i + 1] = i % 4 == 0 ? R(0) : -X.v[2 * i];
    }
    M3<R> Q2;
    R tr3;
    herm_square(Q, Q2, tr3);
    R c1 = retr(Q2) * R(0.5);
    int k = 0;
    if (c1 > R(C1MAX)) {  // scaling: c1 ~ |Q|^2 shrinks by 4 per halving
        R scale = R(1);
        while (c1 > R(C1MAX) && k < 30) {
            c1 *= R(0.25);
            scale *= R(0.5);
            ++k;
        }
#pragma unroll
        for (int i = 0; i < 18; ++i) Q.v[i] *= scale;
        herm_square(Q, Q2, tr3);
        c1 = retr(Q2) * R(0.5);
    }
    R c0 = tr3 * (R(1) / R(3));
    // f_j = sum_n i^n/n! * coef_j(Q^n); real part from n = 0,2 (mod 4) with
    // signs +,-, imaginary part from n = 1,3 (mod 4) with signs +,-.
    R a = R(1), b = R(0), d = R(0);
    R f0r = R(1), f0i = R(0), f1r = R(0), f1i = R(0), f2r = R(0), f2i = R(0);
    R fact = R(1);
#pragma unroll
    for (int n = 1; n < NTERMS; ++n) {
        R an = c0 * d, bn = a + c1 * d, dn = b;
        a = an;
        b = bn;
        d = dn;
        fact = fact / R(n);  // folded at compile time
        switch (n & 3) {
            case 0: f0r += fact * a; f1r += fact * b; f2r += fact * d; break;
            case 1: f0i += fact * a; f1i += fact * b; f2i += fact * d; break;
            case 2: f0r -= fact * a; f1r -= fact * b; f2r -= fact * d; break;
            default: f0i -= fact * a; f1i -= fact * b; f2i -= fact * d; break;
        }
    }
    M3<R> o;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const R qr = Q.v[2 * i], pr = Q2.v[2 * i];
        if (i % 4 == 0) {  // diagonal: Q and Q^2 real there
            o.v[2 * i] = qr * f1r + pr * f2r;
            o.v[2 * i + 1] = qr * f1i + pr * f2i;
            o.v[2 * i] += f0r;
            o.v[2 * i + 1] += f0i;
        } else {
            const R qi = Q.v[2 * i + 1], pi = Q2.v[2 * i + 1];
            o.v[2 * i] = qr * f1r - qi * f1i + pr * f2r - pi * f2i;
            o.v[2 * i + 1] = qr * f1i + qi * f1r + pr * f2i + pi * f2r;
        }
    }
    for (int s = 0; s < k; ++s) o = mul_nn(o, o);
    return o;
}

// 1/sqrt(n) for n near 1 (a row norm of a nearly unitary matrix): the
// hardware reciprocal square root refined by one Newton step,
// y <- y (3 - n y^2) / 2, which squares its relative error (to ~1e-17).
LGF_DEV double rsqrt_refined(double n) {
    const double y = __builtin_amdgcn_rsq(n);
    return y * (1.5 - 0.5 * n * y * y);
}

// Project onto SU(3) in double precision: Gram-Schmidt on rows 0 and 1,
// row 2 = conj(row0 x row1).  FAST selects rsqrt_refined for the two
// normalisations (the per-step projection of the mixed-precision flow);
// otherwise 1/sqrt, as the CPU oracle (generator, hot start).
template <bool FAST = false>
LGF_DEV void reunitarize(M3<double>& m) {
    double n = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) n += m.v[i] * m.v[i];
    n = FAST ? rsqrt_refined(n) : 1.0 / sqrt(n);
#pragma unroll
    for (int i = 0; i < 6; ++i) m.v[i] *= n;
    double pr = 0, pi = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double ar = m.v[2 * c], ai = m.v[2 * c + 1], br = m.v[6 + 2 * c], bi = m.v[7 + 2 * c];
        pr += ar * br + ai * bi;
        pi += ar * bi - ai * br;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double ar = m.v[2 * c], ai = m.v[2 * c + 1];
        m.v[6 + 2 * c] -= pr * ar - pi * ai;
        m.v[7 + 2 * c] -= pr * ai + pi * ar;
    }
    n = 0;
#pragma unroll
    for (int i = 6; i < 12; ++i) n += m.v[i] * m.v[i];
    n = FAST ? rsqrt_refined(n) : 1.0 / sqrt(n);
#pragma unroll
    for (int i = 6; i < 12; ++i) m.v[i] *= n;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        double xr = (m.v[2 * c1] * m.v[6 + 2 * c2] - m.v[2 * c1 + 1] * m.v[7 + 2 * c2]) -
                    (m.v[2 * c2] * m.v[6 + 2 * c1] - m.v[2 * c2 + 1] * m.v[7 + 2 * c1]);
        double xi = (m.v[2 * c1] * m.v[7 + 2 * c2] + m.v[2 * c1 + 1] * m.v[6 + 2 * c2]) -
                    (m.v[2 * c2] * m.v[7 + 2 * c1] + m.v[2 * c2 + 1] * m.v[6 + 2 * c1]);
        m.v[12 + 2 * c] = xr;
        m.v[13 + 2 * c] = -xi;
    }
}

}  // namespace lgf
