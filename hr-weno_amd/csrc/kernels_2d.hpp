// kernels_2d.hpp -- fused dimension-by-dimension pass for the 2D population balance
//   n_t + (g1(x1) n)_x1 + (g2(x2) n)_x2 = 0,   u((j-1)*n1 + i)   (x1 contiguous)
// composing example/example2_pbe_2d_fv.f90:73-129 (x1 and x2 WENO sweeps, Godunov fluxes,
// boundary fluxes, divergence) with the TVD stage update of src/hrweno_tvdode.f90.
//
// Work decomposition (no LDS, no barriers): one wavefront owns a strip of 64 consecutive
// columns i = X0-1 .. X0+62 and marches over TY rows.  Lane l reconstructs column X0-1+l
// along x1 (its 2K-1 window read through the caches, coalesced across lanes) and along x2
// (a window sliding down the column in registers).  Interface fluxes along x1 pass between
// neighbouring lanes by shuffles, so lanes 1..62 own output cells.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels_1d.hpp"

namespace hrw {

struct Stage2DArgs {
    const double *us, *u0;
    double *uo;
    const double *uo4, *Lo4;
    double *Lout;
    int n1, n2;
    double x1min, rx1, x2min, rx2;        // linear grids: edge i = xmin + rx*i
    double g1a, g1b, g2a, g2b;            // g1 = g1a + g1b*x1, g2 = g2a + g2b*x2
    double eps, h;
    int stage;                            // hrw::Stage, or -1 for rhs only
    int ty;                               // rows per wavefront
    int nstrips;                          // column strips (of 62 output columns)
};

constexpr int kStrip = 62;

template <int K, bool FAST>
__device__ __forceinline__ void recon(const double *x, double eps, double &vl, double &vr)
{
    if constexpr (FAST) weno_cell_fast<K>(x, eps, vl, vr);
    else weno_cell_exact<K>(x, eps, nullptr, vl, vr);
}

// Godunov flux of the linear flux g*v at one interface (src/hrweno_fluxes.f90:66-74).
__device__ __forceinline__ double godunov_lin(double g, double vm, double vp)
{
    return godunov(g * vm, g * vp, vm, vp);
}

template <int K, int BC, bool FAST>
__global__ __launch_bounds__(kBlock) void pbe2d_kernel(Stage2DArgs a)
{
    const int lane = threadIdx.x & 63;
    const long long wave = ((long long)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int sx = (int)(wave % a.nstrips);
    const long long sy = wave / a.nstrips;
    const int n1 = a.n1, n2 = a.n2;
    const long long Y0 = sy * a.ty;
    if (Y0 >= n2) return;                               // whole wavefront leaves together
    const int Y1 = (int)min((long long)n2, Y0 + a.ty);
    const int i = sx * kStrip - 1 + lane;               // column of this lane (may be -1 or >= n1)
    const int ic = i < 0 ? 0 : (i >= n1 ? n1 - 1 : i);  // clamped column for loads
    const bool own = lane >= 1 && lane <= kStrip && i < n1;

    // per-lane constants: x1 width, g1 at the right interface x_{i+1/2}
    const double e_l = linear_edge(a.x1min, a.rx1, ic), e_r = linear_edge(a.x1min, a.rx1, ic + 1);
    const double w1 = e_r - e_l;
    const double g1 = linear_coef(a.g1a, a.g1b, e_r);
    int xi[2 * K - 1];                                  // clamped x1 window columns
#pragma unroll
    for (int m = -(K - 1); m <= K - 1; ++m) {
        int c = i + m;
        xi[m + K - 1] = c < 0 ? 0 : (c >= n1 ? n1 - 1 : c);
    }
    auto rowc = [n2](long long j) -> long long { return j < 0 ? 0 : (j >= n2 ? n2 - 1 : j); };
    const double *us = a.us;

    // column window for the x2 reconstruction of row r: cw[m] = u(i, r-K+1+m)
    double cw[2 * K - 1];
    // prologue: reconstruction of row Y0-1 (for F2 at Y0-1/2) and of row Y0
    double vl2c, vr2c, f2_dn;
    {
#pragma unroll
        for (int m = 0; m < 2 * K - 1; ++m) cw[m] = us[rowc(Y0 - 1 - (K - 1) + m) * n1 + ic];
        double vlp, vrp;
        recon<K, FAST>(cw, a.eps, vlp, vrp);
#pragma unroll
        for (int m = 0; m < 2 * K - 2; ++m) cw[m] = cw[m + 1];
        cw[2 * K - 2] = us[rowc(Y0 + K - 1) * n1 + ic];
        recon<K, FAST>(cw, a.eps, vl2c, vr2c);
        if (Y0 == 0) {
            f2_dn = 0.0;                                // zero flux at x2 = x2min
        } else {
            const double g2 = linear_coef(a.g2a, a.g2b, linear_edge(a.x2min, a.rx2, Y0));
            f2_dn = godunov_lin(g2, vrp, vl2c);
        }
    }

    for (int j = (int)Y0; j < Y1; ++j) {
        // x2: reconstruction of row j+1, flux at j+1/2
#pragma unroll
        for (int m = 0; m < 2 * K - 2; ++m) cw[m] = cw[m + 1];
        cw[2 * K - 2] = us[rowc((long long)j + K) * n1 + ic];
        double vl2n, vr2n;
        recon<K, FAST>(cw, a.eps, vl2n, vr2n);
        const double e2l = linear_edge(a.x2min, a.rx2, j), e2r = linear_edge(a.x2min, a.rx2, j + 1);
        const double w2 = e2r - e2l;
        double f2_up;
        if (j == n2 - 1) {
            f2_up = BC == 0 ? 0.0 : godunov_lin(linear_coef(a.g2a, a.g2b, e2r), vr2c, vr2c);
        } else {
            f2_up = godunov_lin(linear_coef(a.g2a, a.g2b, e2r), vr2c, vl2n);
        }

        // x1: reconstruction of (i, j), flux at i+1/2 with the right neighbour's vl
        const double *row = us + (long long)j * n1;
        double xw[2 * K - 1];
#pragma unroll
        for (int m = 0; m < 2 * K - 1; ++m) xw[m] = row[xi[m]];
        double vl1, vr1;
        recon<K, FAST>(xw, a.eps, vl1, vr1);
        const double vl1_right = __shfl_down(vl1, 1, 64);
        double f1_r;
        if (i < 0) f1_r = 0.0;                           // zero flux at x1 = x1min
        else if (i >= n1 - 1) f1_r = BC == 0 ? 0.0 : godunov_lin(g1, vr1, vr1);
        else f1_r = godunov_lin(g1, vr1, vl1_right);
        const double f1_l = __shfl_up(f1_r, 1, 64);

        if (own) {
            const long long g = (long long)j * n1 + i;
            // example2:124-126
            double L;
            if constexpr (FAST) L = (f1_l - f1_r) / w1 - (f2_up - f2_dn) / w2;
            else L = -((f1_r - f1_l) / w1) - (f2_up - f2_dn) / w2;
            const double uc = xw[K - 1];
            if (a.stage < 0) {
                a.Lout[g] = L;
            } else {
                double z = uc, u4 = 0.0, l4 = 0.0;
                if (a.stage == RK2_2 || a.stage == RK3_2 || a.stage == RK3_3) z = a.u0[g];
                if (a.stage == MS3) {
                    u4 = a.uo4[g];
                    l4 = a.Lo4[g];
                }
                const double un = stage_update<FAST>(a.stage, z, uc, L, a.h, u4, l4);
                if (a.Lout) a.Lout[g] = L;
                a.uo[g] = un;
            }
        }
        f2_dn = f2_up;
        vl2c = vl2n;
        vr2c = vr2n;
    }
}

}  // namespace hrw
