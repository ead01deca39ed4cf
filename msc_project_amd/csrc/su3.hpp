// 3x3 complex matrix arithmetic for gfx950 vector ALUs.
//
// A matrix lives in 18 registers: element (r,c) is (v[2*(3r+c)], v[2*(3r+c)+1])
// = (Re, Im), row-major -- the element order of the reference's su3_mat
// (LGF/Math/SU3_mat.h:15-17) and of its saved-configuration records.
// Everything is templated on the real type R (double for the parity path,
// float for the mixed-precision path) and force-inlined so that the compiler
// keeps matrices in VGPRs and schedules the products as FMA chains.
#pragma once
#include <hip/hip_runtime.h>

#define LGF_DEV __device__ __forceinline__

namespace lgf {

template <typename R>
struct M3 {
    R v[18];
};

template <typename R>
LGF_DEV R& re(M3<R>& m, int r, int c) { return m.v[2 * (3 * r + c)]; }
template <typename R>
LGF_DEV R& im(M3<R>& m, int r, int c) { return m.v[2 * (3 * r + c) + 1]; }

template <typename R>
LGF_DEV void set_zero(M3<R>& m) {
#pragma unroll
    for (int i = 0; i < 18; ++i) m.v[i] = R(0);
}

template <typename R>
LGF_DEV void set_unit(M3<R>& m) {
    set_zero(m);
    m.v[0] = m.v[8] = m.v[16] = R(1);
}

// Products.  Suffix n = operand as is, d = operand daggered.  Every complex
// multiply-add is written as four separate accumulations so that each
// compiles to one FMA (x += a*b - c*d would leave a multiply and an add).
// The _acc forms add the product into an existing matrix.
#define LGF_CMAC(X, Y, AR, AI, BR, BI) \
    X += (AR) * (BR);                  \
    X -= (AI) * (BI);                  \
    Y += (AR) * (BI);                  \
    Y += (AI) * (BR);

template <typename R>
LGF_DEV void mul_nn_acc(M3<R>& o, const M3<R>& a, const M3<R>& b) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            R x = o.v[2 * (3 * r + c)], y = o.v[2 * (3 * r + c) + 1];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                LGF_CMAC(x, y, a.v[2 * (3 * r + k)], a.v[2 * (3 * r + k) + 1], b.v[2 * (3 * k + c)],
                         b.v[2 * (3 * k + c) + 1])
            }
            o.v[2 * (3 * r + c)] = x;
            o.v[2 * (3 * r + c) + 1] = y;
        }
}

// o += a * b^dag : (a b^dag)_{rc} = sum_k a_{rk} conj(b_{ck})
template <typename R>
LGF_DEV void mul_nd_acc(M3<R>& o, const M3<R>& a, const M3<R>& b) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            R x = o.v[2 * (3 * r + c)], y = o.v[2 * (3 * r + c) + 1];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                LGF_CMAC(x, y, a.v[2 * (3 * r + k)], a.v[2 * (3 * r + k) + 1], b.v[2 * (3 * c + k)],
                         -b.v[2 * (3 * c + k) + 1])
            }
            o.v[2 * (3 * r + c)] = x;
            o.v[2 * (3 * r + c) + 1] = y;
        }
}

// o += a^dag * b : sum_k conj(a_{kr}) b_{kc}
template <typename R>
LGF_DEV void mul_dn_acc(M3<R>& o, const M3<R>& a, const M3<R>& b) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            R x = o.v[2 * (3 * r + c)], y = o.v[2 * (3 * r + c) + 1];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                LGF_CMAC(x, y, a.v[2 * (3 * k + r)], -a.v[2 * (3 * k + r) + 1], b.v[2 * (3 * k + c)],
                         b.v[2 * (3 * k + c) + 1])
            }
            o.v[2 * (3 * r + c)] = x;
            o.v[2 * (3 * r + c) + 1] = y;
        }
}

template <typename R>
LGF_DEV M3<R> mul_nn(const M3<R>& a, const M3<R>& b) {
    M3<R> o;
#pragma unroll
    for (int i = 0; i < 18; ++i) o.v[i] = R(0);
    mul_nn_acc(o, a, b);
    return o;
}

template <typename R>
LGF_DEV M3<R> mul_nd(const M3<R>& a, const M3<R>& b) {
    M3<R> o;
#pragma unroll
    for (int i = 0; i < 18; ++i) o.v[i] = R(0);
    mul_nd_acc(o, a, b);
    return o;
}

template <typename R>
LGF_DEV M3<R> mul_dn(const M3<R>& a, const M3<R>& b) {
    M3<R> o;
#pragma unroll
    for (int i = 0; i < 18; ++i) o.v[i] = R(0);
    mul_dn_acc(o, a, b);
    return o;
}

template <typename R>
LGF_DEV void add_to(M3<R>& acc, const M3<R>& b) {
#pragma unroll
    for (int i = 0; i < 18; ++i) acc.v[i] += b.v[i];
}

template <typename R>
LGF_DEV R retr(const M3<R>& a) { return a.v[0] + a.v[8] + a.v[16]; }

// Re Tr(a b) without forming the product
template <typename R>
LGF_DEV R retr_mul(const M3<R>& a, const M3<R>& b) {
    R s = R(0);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            s += a.v[2 * (3 * r + k)] * b.v[2 * (3 * k + r)];
            s -= a.v[2 * (3 * r + k) + 1] * b.v[2 * (3 * k + r) + 1];
        }
    return s;
}

// Traceless anti-Hermitian part scaled: out = s * [ (W - W^dag) - 1/3 Tr(W - W^dag) ]
template <typename R>
LGF_DEV M3<R> ta_scaled(const M3<R>& W, R s) {
    M3<R> o;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            o.v[2 * (3 * r + c)] = W.v[2 * (3 * r + c)] - W.v[2 * (3 * c + r)];
            o.v[2 * (3 * r + c) + 1] = W.v[2 * (3 * r + c) + 1] + W.v[2 * (3 * c + r) + 1];
        }
    // the real part of the diagonal of W - W^dag is exactly zero
    R tr_re = (o.v[0] + o.v[8] + o.v[16]) * (R(1) / R(3));
    R tr_im = (o.v[1] + o.v[9] + o.v[17]) * (R(1) / R(3));
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        o.v[8 * d] -= tr_re;
        o.v[8 * d + 1] -= tr_im;
    }
#pragma unroll
    for (int i = 0; i < 18; ++i) o.v[i] *= s;
    return o;
}

// o += a b, o += a b^dag, o += a^dag b without the real parts of the
// diagonal, for sums whose only use is their traceless anti-Hermitian part
// (ta_scaled), where those parts cancel exactly whatever their values.
#define LGF_TA_ACC_BODY(AIDX, BIDX, CONJA, CONJB)                                              \
    _Pragma("unroll") for (int r = 0; r < 3; ++r) _Pragma("unroll") for (int c = 0; c < 3; ++c) { \
        R x = o.v[2 * (3 * r + c)], y = o.v[2 * (3 * r + c) + 1];                               \
        _Pragma("unroll") for (int k = 0; k < 3; ++k) {                                          \
            const R ar = a.v[AIDX], ai = CONJA a.v[AIDX + 1];                                     \
            const R br = b.v[BIDX], bi = CONJB b.v[BIDX + 1];                                     \
            if (r != c) {                                                                         \
                x += ar * br;                                                                     \
                x -= ai * bi;                                                                     \
            }                                                                                     \
            y += ar * bi;                                                                         \
            y += ai * br;                                                                         \
        }                                                                                         \
        o.v[2 * (3 * r + c)] = x;                                                                 \
        o.v[2 * (3 * r + c) + 1] = y;                                                             \
    }

template <typename R>
LGF_DEV void mul_nn_acc_ta(M3<R>& o, const M3<R>& a, const M3<R>& b) {
    LGF_TA_ACC_BODY(2 * (3 * r + k), 2 * (3 * k + c), , )
}
template <typename R>
LGF_DEV void mul_nd_acc_ta(M3<R>& o, const M3<R>& a, const M3<R>& b) {
    LGF_TA_ACC_BODY(2 * (3 * r + k), 2 * (3 * c + k), , -)
}
template <typename R>
LGF_DEV void mul_dn_acc_ta(M3<R>& o, const M3<R>& a, const M3<R>& b) {
    LGF_TA_ACC_BODY(2 * (3 * k + r), 2 * (3 * k + c), -, )
}
#undef LGF_TA_ACC_BODY

// A sum whose only use is its traceless anti-Hermitian part, folded to nine
// reals as it is accumulated: v[d] = sum Im W_dd (d = 0..2) and, for the
// pairs (r,c) = (0,1), (0,2), (1,2), v[3 + 2p], v[4 + 2p] = Re, Im of
// sum (W_rc - conj(W_cr)) -- the upper triangle of W - W^dag.  Each product
// term goes into one running sum, so this costs the FMAs of the fifteen
// entries and holds nine registers instead of fifteen.
template <typename R>
struct Ta9 {
    R v[9];
};

template <typename R>
LGF_DEV void set_zero(Ta9<R>& t) {
#pragma unroll
    for (int i = 0; i < 9; ++i) t.v[i] = R(0);
}

template <typename R>
LGF_DEV void scale(Ta9<R>& t, R w) {
#pragma unroll
    for (int i = 0; i < 9; ++i) t.v[i] *= w;
}

#define LGF_TA9_TERMS(AIDX, BIDX, CONJA, CONJB, SX)      \
    _Pragma("unroll") for (int k = 0; k < 3; ++k) {       \
        const R ar = a.v[AIDX], ai = CONJA a.v[AIDX + 1]; \
        const R br = b.v[BIDX], bi = CONJB b.v[BIDX + 1]; \
        x SX ar * br;                                     \
        x SX - ai * bi;                                   \
        y += ar * bi;                                     \
        y += ai * br;                                     \
    }
#define LGF_TA9_BODY(AIDX, BIDX, CONJA, CONJB)                                     \
    _Pragma("unroll") for (int d = 0; d < 3; ++d) {                                \
        const int r = d, c = d;                                                    \
        R y = o.v[d];                                                              \
        _Pragma("unroll") for (int k = 0; k < 3; ++k) {                            \
            y += a.v[AIDX] * (CONJB b.v[BIDX + 1]);                                \
            y += (CONJA a.v[AIDX + 1]) * b.v[BIDX];                                \
        }                                                                          \
        o.v[d] = y;                                                                \
    }                                                                              \
    _Pragma("unroll") for (int p = 0; p < 3; ++p) {                                \
        const int r0 = p == 2 ? 1 : 0, c0 = p == 0 ? 1 : 2;                        \
        R x = o.v[3 + 2 * p], y = o.v[4 + 2 * p];                                  \
        {                                                                          \
            const int r = r0, c = c0;                                              \
            LGF_TA9_TERMS(AIDX, BIDX, CONJA, CONJB, +=)                            \
        }                                                                          \
        {                                                                          \
            const int r = c0, c = r0;                                              \
            LGF_TA9_TERMS(AIDX, BIDX, CONJA, CONJB, -=)                            \
        }                                                                          \
        o.v[3 + 2 * p] = x;                                                        \
        o.v[4 + 2 * p] = y;                                                        \
    }

// The folded sum made traceless: v[d] -= (v[0] + v[1] + v[2]) / 3.  The
// nine reals of -1/2 TA(W) (ta_pack of ta_scaled(W, -1/2)) are then
// (-v[0], -v[1], -v[2], -v[3]/2, ..., -v[8]/2), so sums of such
// projections can be added in this form and scaled once.
template <typename R>
LGF_DEV void make_traceless(Ta9<R>& t) {
    const R m = (t.v[0] + t.v[1] + t.v[2]) * (R(1) / R(3));
#pragma unroll
    for (int d = 0; d < 3; ++d) t.v[d] -= m;
}

// o += a b, a b^dag, a^dag b in the folded form
template <typename R>
LGF_DEV void mul_nn_acc_ta(Ta9<R>& o, const M3<R>& a, const M3<R>& b) {
    LGF_TA9_BODY(2 * (3 * r + k), 2 * (3 * k + c), , )
}
template <typename R>
LGF_DEV void mul_nd_acc_ta(Ta9<R>& o, const M3<R>& a, const M3<R>& b) {
    LGF_TA9_BODY(2 * (3 * r + k), 2 * (3 * c + k), , -)
}
template <typename R>
LGF_DEV void mul_dn_acc_ta(Ta9<R>& o, const M3<R>& a, const M3<R>& b) {
    LGF_TA9_BODY(2 * (3 * k + r), 2 * (3 * k + c), -, )
}
#undef LGF_TA9_BODY
#undef LGF_TA9_TERMS

// Products of two SU(3) matrices whose third row is rebuilt as the
// conjugated cross product of the first two (exact for SU(3)): 96 FMAs
// instead of 108.
template <typename R>
LGF_DEV void su3_third_row(M3<R>& o) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        R x = R(0), y = R(0);
        LGF_CMAC(x, y, o.v[2 * c1], o.v[2 * c1 + 1], o.v[6 + 2 * c2], o.v[7 + 2 * c2])
        LGF_CMAC(x, y, -o.v[2 * c2], -o.v[2 * c2 + 1], o.v[6 + 2 * c1], o.v[7 + 2 * c1])
        o.v[12 + 2 * c] = x;
        o.v[13 + 2 * c] = -y;
    }
}

// a b from rows 0, 1 and their cross product
template <typename R>
LGF_DEV M3<R> mul_nn_su3(const M3<R>& a, const M3<R>& b) {
    M3<R> o;
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            R x = R(0), y = R(0);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                LGF_CMAC(x, y, a.v[2 * (3 * r + k)], a.v[2 * (3 * r + k) + 1], b.v[2 * (3 * k + c)],
                         b.v[2 * (3 * k + c) + 1])
            }
            o.v[2 * (3 * r + c)] = x;
            o.v[2 * (3 * r + c) + 1] = y;
        }
    su3_third_row(o);
    return o;
}

// a b^dag from rows 0, 1 and their cross product
template <typename R>
LGF_DEV M3<R> mul_nd_su3(const M3<R>& a, const M3<R>& b) {
    M3<R> o;
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            R x = R(0), y = R(0);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                LGF_CMAC(x, y, a.v[2 * (3 * r + k)], a.v[2 * (3 * r + k) + 1], b.v[2 * (3 * c + k)],
                         -b.v[2 * (3 * c + k) + 1])
            }
            o.v[2 * (3 * r + c)] = x;
            o.v[2 * (3 * r + c) + 1] = y;
        }
    su3_third_row(o);
    return o;
}

// a^dag b from rows 0, 1 and their cross product
template <typename R>
LGF_DEV M3<R> mul_dn_su3(const M3<R>& a, const M3<R>& b) {
    M3<R> o;
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            R x = R(0), y = R(0);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                LGF_CMAC(x, y, a.v[2 * (3 * k + r)], -a.v[2 * (3 * k + r) + 1], b.v[2 * (3 * k + c)],
                         b.v[2 * (3 * k + c) + 1])
            }
            o.v[2 * (3 * r + c)] = x;
            o.v[2 * (3 * r + c) + 1] = y;
        }
    su3_third_row(o);
    return o;
}

// Partial products whose result is only ever the left factor of one of the
// SU(3) products above (which read rows 0, 1 of a left factor of mul_nn_su3
// and mul_nd_su3, and columns 0, 1 of a left factor of mul_dn_su3): 72 FMAs.
// The entries they do not compute are left unset.
//   rows 0, 1 of a b        (a: rows 0, 1; b: whole)
//   rows 0, 1 of a b^dag    (a: rows 0, 1; b: whole)
//   rows 0, 1 of a^dag b^dag (a: columns 0, 1; b: whole)
//   columns 0, 1 of a b     (a: whole; b: columns 0, 1)
#define LGF_PART_BODY(RMAX, CMAX, AR, AI, BR, BI)                   \
    M3<R> o;                                                         \
    _Pragma("unroll") for (int r = 0; r < RMAX; ++r)                 \
    _Pragma("unroll") for (int c = 0; c < CMAX; ++c) {               \
        R x = R(0), y = R(0);                                        \
        _Pragma("unroll") for (int k = 0; k < 3; ++k) {              \
            LGF_CMAC(x, y, AR, AI, BR, BI)                           \
        }                                                            \
        o.v[2 * (3 * r + c)] = x;                                    \
        o.v[2 * (3 * r + c) + 1] = y;                                \
    }                                                                \
    return o;

template <typename R>
LGF_DEV M3<R> mul_nn_r01(const M3<R>& a, const M3<R>& b) {
    LGF_PART_BODY(2, 3, a.v[2 * (3 * r + k)], a.v[2 * (3 * r + k) + 1], b.v[2 * (3 * k + c)], b.v[2 * (3 * k + c) + 1])
}
template <typename R>
LGF_DEV M3<R> mul_nd_r01(const M3<R>& a, const M3<R>& b) {
    LGF_PART_BODY(2, 3, a.v[2 * (3 * r + k)], a.v[2 * (3 * r + k) + 1], b.v[2 * (3 * c + k)], -b.v[2 * (3 * c + k) + 1])
}
template <typename R>
LGF_DEV M3<R> mul_dd_r01(const M3<R>& a, const M3<R>& b) {
    // conj(a_kr) conj(b_ck) = (ar br - ai bi) - i (ar bi + ai br)
    LGF_PART_BODY(2, 3, a.v[2 * (3 * k + r)], -a.v[2 * (3 * k + r) + 1], b.v[2 * (3 * c + k)], -b.v[2 * (3 * c + k) + 1])
}
template <typename R>
LGF_DEV M3<R> mul_nn_c01(const M3<R>& a, const M3<R>& b) {
    LGF_PART_BODY(3, 2, a.v[2 * (3 * r + k)], a.v[2 * (3 * r + k) + 1], b.v[2 * (3 * k + c)], b.v[2 * (3 * k + c) + 1])
}
#undef LGF_PART_BODY

// y = w p, y += w p, y = w p^dag, y += w p^dag
template <typename R>
LGF_DEV void set_scaled(M3<R>& y, R w, const M3<R>& p) {
#pragma unroll
    for (int i = 0; i < 18; ++i) y.v[i] = w * p.v[i];
}
template <typename R>
LGF_DEV void axpy(M3<R>& y, R w, const M3<R>& p) {
#pragma unroll
    for (int i = 0; i < 18; ++i) y.v[i] += w * p.v[i];
}
template <typename R>
LGF_DEV void set_scaled_dag(M3<R>& y, R w, const M3<R>& p) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            y.v[2 * (3 * r + c)] = w * p.v[2 * (3 * c + r)];
            y.v[2 * (3 * r + c) + 1] = -w * p.v[2 * (3 * c + r) + 1];
        }
}
template <typename R>
LGF_DEV void axpy_dag(M3<R>& y, R w, const M3<R>& p) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            y.v[2 * (3 * r + c)] += w * p.v[2 * (3 * c + r)];
            y.v[2 * (3 * r + c) + 1] -= w * p.v[2 * (3 * c + r) + 1];
        }
}

// Products of SU(3) matrices in the arithmetic of the mixed-precision path
// (float: links projected onto SU(3) every flow step) use the cross-product
// row; in double, the parity path, the full product as the reference.
template <typename R>
LGF_DEV M3<R> su3_nn(const M3<R>& a, const M3<R>& b) {
    if constexpr (sizeof(R) == 4)
        return mul_nn_su3(a, b);
    else
        return mul_nn(a, b);
}

template <typename R>
LGF_DEV M3<R> su3_nd(const M3<R>& a, const M3<R>& b) {
    if constexpr (sizeof(R) == 4)
        return mul_nd_su3(a, b);
    else
        return mul_nd(a, b);
}

template <typename R>
LGF_DEV M3<R> su3_dn(const M3<R>& a, const M3<R>& b) {
    if constexpr (sizeof(R) == 4)
        return mul_dn_su3(a, b);
    else
        return mul_dn(a, b);
}

// ta_scaled(a * b, s) without the real parts of the product's diagonal,
// which the projection cancels exactly: 90 FMAs for the product instead of
// 108.
template <typename R>
LGF_DEV M3<R> ta_of_product(const M3<R>& a, const M3<R>& b, R s) {
    M3<R> W;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            R x = R(0), y = R(0);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const R ar = a.v[2 * (3 * r + k)], ai = a.v[2 * (3 * r + k) + 1];
                const R br = b.v[2 * (3 * k + c)], bi = b.v[2 * (3 * k + c) + 1];
                if (r != c) {
                    x += ar * br;
                    x -= ai * bi;
                }
                y += ar * bi;
                y += ai * br;
            }
            W.v[2 * (3 * r + c)] = x;
            W.v[2 * (3 * r + c) + 1] = y;
        }
    return ta_scaled(W, s);
}

// ta_scaled(a * b^dag, s) without the real parts of the product's diagonal
template <typename R>
LGF_DEV M3<R> ta_of_product_nd(const M3<R>& a, const M3<R>& b, R s) {
    M3<R> W;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            R x = R(0), y = R(0);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const R ar = a.v[2 * (3 * r + k)], ai = a.v[2 * (3 * r + k) + 1];
                const R br = b.v[2 * (3 * c + k)], bi = b.v[2 * (3 * c + k) + 1];
                if (r != c) {
                    x += ar * br;
                    x += ai * bi;
                }
                y += ai * br;
                y -= ar * bi;
            }
            W.v[2 * (3 * r + c)] = x;
            W.v[2 * (3 * r + c) + 1] = y;
        }
    return ta_scaled(W, s);
}

// The nine reals that hold a traceless anti-Hermitian matrix exactly:
// (Im m00, Im m11, Im m22, Re m01, Im m01, Re m02, Im m02, Re m12, Im m12)
template <typename R>
LGF_DEV void ta_pack(const M3<R>& m, R* a) {
    a[0] = m.v[1], a[1] = m.v[9], a[2] = m.v[17];
    a[3] = m.v[2], a[4] = m.v[3], a[5] = m.v[4], a[6] = m.v[5], a[7] = m.v[10], a[8] = m.v[11];
}

}  // namespace lgf
