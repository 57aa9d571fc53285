// weno_math.hpp -- per-cell arithmetic of the WENO -> flux -> divergence -> TVD-stage path.
//
// Two arithmetic forms of the same mathematics are provided, selected by a template flag:
//
//  * Exact  : the reference's own evaluation order, operation for operation (left-to-right
//             sums, (eps+beta)**2 as a product, 12 IEEE divides per cell for k=3).  The
//             library is compiled with -ffp-contract=off, so no multiply-add is fused and the
//             results are bit-identical to the reference Fortran built for baseline x86-64.
//  * Fast   : the same formulas with explicit fused multiply-adds and the nonlinear weights
//             consolidated over a common denominator (two divides per cell instead of twelve):
//                 w_r = d_r/q_r / sum_s d_s/q_s  =  d_r*prod_{s!=r} q_s / sum_s d_s*prod_{t!=s} q_t
//             with q_r = (eps+beta_r)**2.  This is an exact rewrite of the algebra; only the
//             rounding differs (a few ulp per reconstruction).  The products of two q's stay
//             finite while |beta| < ~1e77, i.e. for cell-value jumps below ~1e38.
//
// The coefficient tables follow src/hrweno_weno.f90:12-21 (Shu, ICASE 97-65, Table 2.1 and
// eqs. 2.58-2.63).  Window convention: x[K-1+m] holds v_{i+m}, m = -(K-1)..K-1.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HRW_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define HRW_HD inline
#endif

namespace hrw {

// --------------------------------------------------------------------------------------
// Tables (src/hrweno_weno.f90:12-21).  C<K>::c[r+1][j] == c(j, r).
template <int K> struct Tab;
template <> struct Tab<1> {
    static constexpr double d[1] = {1.0};
    static constexpr double c[2][1] = {{1.0}, {1.0}};
};
template <> struct Tab<2> {
    static constexpr double d[2] = {2.0 / 3, 1.0 / 3};
    static constexpr double c[3][2] = {{3.0 / 2, -1.0 / 2}, {1.0 / 2, 1.0 / 2}, {-1.0 / 2, 3.0 / 2}};
};
template <> struct Tab<3> {
    static constexpr double d[3] = {0.3, 0.6, 0.1};
    static constexpr double c[4][3] = {{11.0 / 6, -7.0 / 6, 1.0 / 3},
                                       {1.0 / 3, 5.0 / 6, -1.0 / 6},
                                       {-1.0 / 6, 5.0 / 6, 1.0 / 3},
                                       {1.0 / 3, -7.0 / 6, 11.0 / 6}};
};

HRW_HD double fma_(double a, double b, double c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fma(a, b, c);
#else
    return std::fma(a, b, c);
#endif
}

// --------------------------------------------------------------------------------------
// Smoothness indicators, src/hrweno_weno.f90:183-204 (eqs. 2.62, 2.63), reference order.
template <int K> HRW_HD void beta_exact(const double *x, double *beta)
{
    const double *v = x + (K - 1);   // v[m] = v_{i+m}
    if constexpr (K == 1) {
        beta[0] = 0.0;
    } else if constexpr (K == 2) {
        double a = v[1] - v[0];
        double b = v[0] - v[-1];
        beta[0] = a * a;
        beta[1] = b * b;
    } else {
        double a, b;
        a = (v[0] - 2 * v[1]) + v[2];
        b = (3 * v[0] - 4 * v[1]) + v[2];
        beta[0] = 13.0 / 12 * (a * a) + 1.0 / 4 * (b * b);
        a = (v[-1] - 2 * v[0]) + v[1];
        b = v[-1] - v[1];
        beta[1] = 13.0 / 12 * (a * a) + 1.0 / 4 * (b * b);
        a = (v[-2] - 2 * v[-1]) + v[0];
        b = (v[-2] - 4 * v[-1]) + 3 * v[0];
        beta[2] = 13.0 / 12 * (a * a) + 1.0 / 4 * (b * b);
    }
}

// Exact form of one cell, src/hrweno_weno.f90:176-214.  ci == nullptr: uniform tables;
// otherwise ci[(r+1)*K + j] == cnu(j, r, i) (non-uniform grid, eq. 2.20).
template <int K> HRW_HD void weno_cell_exact(const double *x, double eps, const double *ci,
                                             double &vl, double &vr)
{
    const double *v = x + (K - 1);
    double vrr[K], vlr[K], beta[K], alfa[K], alfat[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
        double s = 0.0, sl = 0.0;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const double cr = ci ? ci[(r + 1) * K + j] : Tab<K>::c[r + 1][j];
            s = s + cr * v[-r + j];
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const double cl = ci ? ci[r * K + j] : Tab<K>::c[r][j];
            sl = sl + cl * v[-r + j];
        }
        vrr[r] = s;
        vlr[r] = sl;
    }
    beta_exact<K>(x, beta);
#pragma unroll
    for (int r = 0; r < K; ++r) {
        double q = eps + beta[r];
        q = q * q;
        alfa[r] = Tab<K>::d[r] / q;
        alfat[r] = Tab<K>::d[K - 1 - r] / q;
    }
    double sa = 0.0, sat = 0.0;
#pragma unroll
    for (int r = 0; r < K; ++r) sa = sa + alfa[r];
#pragma unroll
    for (int r = 0; r < K; ++r) sat = sat + alfat[r];
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int r = 0; r < K; ++r) a = a + (alfa[r] / sa) * vrr[r];
#pragma unroll
    for (int r = 0; r < K; ++r) b = b + (alfat[r] / sat) * vlr[r];
    vr = a;
    vl = b;
}

// Fast form of one cell on a uniform grid (see the header comment).
template <int K> HRW_HD void weno_cell_fast(const double *x, double eps, double &vl, double &vr)
{
    const double *v = x + (K - 1);
    if constexpr (K == 1) {
        vr = v[0];
        vl = v[0];
    } else if constexpr (K == 2) {
        // candidates: vrr0=(v0+v1)/2, vrr1=(-v-1+3v0)/2, vlr0=(3v0-v1)/2, vlr1=(v-1+v0)/2
        const double a = v[1] - v[0], b = v[0] - v[-1];
        double q0 = eps + a * a;
        double q1 = eps + b * b;
        q0 = q0 * q0;
        q1 = q1 * q1;
        // w_r ~ d_r / q_r  ->  weights proportional to d_0*q1, d_1*q0
        const double ar0 = (2.0 / 3) * q1, ar1 = (1.0 / 3) * q0;
        const double al0 = (1.0 / 3) * q1, al1 = (2.0 / 3) * q0;
        const double vrr0 = 0.5 * (v[0] + v[1]);
        const double vrr1 = fma_(1.5, v[0], -0.5 * v[-1]);
        const double vlr0 = fma_(1.5, v[0], -0.5 * v[1]);
        const double vlr1 = 0.5 * (v[-1] + v[0]);
        vr = fma_(ar0, vrr0, ar1 * vrr1) / (ar0 + ar1);
        vl = fma_(al0, vlr0, al1 * vlr1) / (al0 + al1);
    } else {
        const double vm2 = v[-2], vm1 = v[-1], v0 = v[0], vp1 = v[1], vp2 = v[2];
        // candidate stencils, c3 rows r+1 (right faces) and r (left faces)
        const double vrr0 = fma_(-1.0 / 6, vp2, fma_(5.0 / 6, vp1, (1.0 / 3) * v0));
        const double vrr1 = fma_(1.0 / 3, vp1, fma_(5.0 / 6, v0, (-1.0 / 6) * vm1));
        const double vrr2 = fma_(11.0 / 6, v0, fma_(-7.0 / 6, vm1, (1.0 / 3) * vm2));
        const double vlr0 = fma_(1.0 / 3, vp2, fma_(-7.0 / 6, vp1, (11.0 / 6) * v0));
        const double vlr1 = fma_(-1.0 / 6, vp1, fma_(5.0 / 6, v0, (1.0 / 3) * vm1));
        const double vlr2 = fma_(1.0 / 3, v0, fma_(5.0 / 6, vm1, (-1.0 / 6) * vm2));
        // smoothness indicators (eq. 2.63)
        double a, b;
        a = fma_(-2.0, vp1, v0) + vp2;
        b = fma_(-4.0, vp1, fma_(3.0, v0, vp2));
        const double b0 = fma_(13.0 / 12 * a, a, 0.25 * (b * b));
        a = fma_(-2.0, v0, vm1) + vp1;
        b = vm1 - vp1;
        const double b1 = fma_(13.0 / 12 * a, a, 0.25 * (b * b));
        a = fma_(-2.0, vm1, vm2) + v0;
        b = fma_(3.0, v0, fma_(-4.0, vm1, vm2));
        const double b2 = fma_(13.0 / 12 * a, a, 0.25 * (b * b));
        double q0 = eps + b0, q1 = eps + b1, q2 = eps + b2;
        q0 = q0 * q0;
        q1 = q1 * q1;
        q2 = q2 * q2;
        const double m0 = q1 * q2, m1 = q0 * q2, m2 = q0 * q1;
        // right face: d = (0.3, 0.6, 0.1); left face: reversed d = (0.1, 0.6, 0.3)
        const double sr = fma_(0.3, m0, fma_(0.6, m1, 0.1 * m2));
        const double sl = fma_(0.1, m0, fma_(0.6, m1, 0.3 * m2));
        const double nr = fma_(0.3 * m0, vrr0, fma_(0.6 * m1, vrr1, (0.1 * m2) * vrr2));
        const double nl = fma_(0.1 * m0, vlr0, fma_(0.6 * m1, vlr1, (0.3 * m2) * vlr2));
        vr = nr / sr;
        vl = nl / sl;
    }
}

// --------------------------------------------------------------------------------------
// Numerical fluxes, src/hrweno_fluxes.f90:42-43 and :66-74, given fm=f(vm), fp=f(vp).
HRW_HD double lax_friedrichs(double fm, double fp, double vm, double vp, double alpha)
{
    return ((fm + fp) - alpha * (vp - vm)) / 2;
}
HRW_HD double godunov(double fm, double fp, double vm, double vp)
{
    if (vm <= vp) return fm < fp ? fm : fp;
    return fm > fp ? fm : fp;
}

// Burgers f(v) = v**2/2 (example/example1_burgers_1d_fv.f90:120)
HRW_HD double burgers_f(double v) { return (v * v) / 2; }

enum NumFlux : int { LF = 0, GODUNOV = 1 };

template <int FLUX, bool FAST>
HRW_HD double burgers_numflux(double vm, double vp, double alpha)
{
    if constexpr (FLUX == GODUNOV) {
        return godunov(burgers_f(vm), burgers_f(vp), vm, vp);
    } else if constexpr (FAST) {
        // ((vm^2 + vp^2)/2 - alpha*(vp-vm)) / 2
        return fma_(0.25, fma_(vm, vm, vp * vp), (-0.5 * alpha) * (vp - vm));
    } else {
        return lax_friedrichs(burgers_f(vm), burgers_f(vp), vm, vp, alpha);
    }
}

// Linear flux (a + b*x)*v (example/example2_pbe_2d_fv.f90:131-155 with a growth rate g(x)).
// The coefficient is formed without contraction in both forms so that g(x) is the
// reference's value exactly.
HRW_HD double linear_coef(double a, double b, double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __dadd_rn(a, __dmul_rn(b, x));
#else
    return a + b * x;
#endif
}

// Edge i of a linear grid, xmin + rx*i (src/hrweno_grids.f90:76-78), never contracted.
HRW_HD double linear_edge(double xmin, double rx, long long i)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __dadd_rn(xmin, __dmul_rn(rx, (double)i));
#else
    return xmin + rx * (double)i;
#endif
}
// The same for 32-bit indices (one conversion instruction instead of four).
HRW_HD double linear_edge32(double xmin, double rx, int i)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __dadd_rn(xmin, __dmul_rn(rx, (double)i));
#else
    return xmin + rx * (double)i;
#endif
}

// x/3 correctly rounded except in rare ties, from one multiply and two FMAs.
HRW_HD double div3_fast(double x)
{
    const double inv3 = 1.0 / 3;
    double q = x * inv3;
    double r = fma_(-3.0, q, x);
    return fma_(r, inv3, q);
}

// --------------------------------------------------------------------------------------
// TVD stage combinations, src/hrweno_tvdode.f90:139,150-152,163-167 and :257.
// u0 = state at the start of the step, us = stage input, L = rhs(us), h = dt.
enum Stage : int {
    RK1 = 0,     // u  = u + dt*L(u)
    RK2_1 = 1,   // ui = u + dt*L(u)
    RK2_2 = 2,   // u  = (u + ui + dt*L(ui))/2
    RK3_1 = 3,   // ui = u + dt*L(u)
    RK3_2 = 4,   // ui = (3u + ui + dt*L(ui))/4
    RK3_3 = 5,   // u  = (u + 2ui + 2dt*L(ui))/3
    MS3 = 6      // ui = (25u + 50dt*L(u) + 7u_{n-4} + 10dt*L(u_{n-4}))/32
};

template <bool FAST>
HRW_HD double stage_update(int stage, double u0, double us, double L, double h, double uo4,
                           double Lo4)
{
    switch (stage) {
    case RK1:
    case RK2_1:
    case RK3_1:
        return FAST ? fma_(h, L, us) : us + h * L;
    case RK2_2:
        return FAST ? 0.5 * fma_(h, L, u0 + us) : ((u0 + us) + h * L) / 2;
    case RK3_2:
        return FAST ? 0.25 * fma_(h, L, fma_(3.0, u0, us)) : ((3 * u0 + us) + h * L) / 4;
    case RK3_3:
        return FAST ? div3_fast(fma_(2.0 * h, L, fma_(2.0, us, u0)))
                    : ((u0 + 2 * us) + (2 * h) * L) / 3;
    default:  // MS3
        return FAST ? 0.03125 * fma_(10.0 * h, Lo4, fma_(7.0, uo4, fma_(50.0 * h, L, 25.0 * us)))
                    : ((((25 * us) + (50 * h) * L) + 7 * uo4) + (10 * h) * Lo4) / 32;
    }
}

}  // namespace hrw
