// kernels_2d.hpp -- fused dimension-by-dimension pass for the 2D population balance
//   n_t + (g1(x1) n)_x1 + (g2(x2) n)_x2 = 0,   u((j-1)*n1 + i)   (x1 contiguous)
// composing example/example2_pbe_2d_fv.f90:73-129 (x1 and x2 WENO sweeps, Godunov fluxes,
// boundary fluxes, divergence) with the TVD stage update of src/hrweno_tvdode.f90.
//
// Work decomposition (no LDS, no barriers): one wavefront owns a strip of 64 consecutive
// columns i = X0-1 .. X0+62 and marches over TY rows.  Lane l reconstructs column X0-1+l
// along x2 (a window sliding down the column in registers) and along x1 (its window gathered
// from the neighbouring lanes' column values by lane shifts, plus K-1 halo values loaded by
// the edge lanes).  Interface fluxes along x1 pass between neighbouring lanes the same way, so
// lanes 1..62 own output cells.  (K = 3, FAST: columns X0-2 .. X0+61, lanes 2..61 own, and the
// x1 smoothness triples are shared between lanes -- strip_width.)  Each cell value is read
// from HBM once per row band.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels_1d_fast.hpp"

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
// Output columns of a wavefront's strip.  K = 3 in the FAST form: 60, with two halo lanes each
// side, so that every lane forms only the smoothness triple centred on its own column and takes
// the triples of its neighbours' columns by lane shifts (the outer halo lanes' centred triples
// feed lanes 1 and 62, whose reconstructions give the strip's two outer fluxes).
__host__ __device__ constexpr int strip_width(int K, bool fast) { return (K == 3 && fast) ? 60 : kStrip; }

template <int K, bool FAST>
__device__ __forceinline__ void recon(const double *x, double eps, double &vl, double &vr)
{
    if constexpr (!FAST) {
        weno_cell_exact<K>(x, eps, nullptr, vl, vr);
    } else if constexpr (K == 3) {   // the 1D kernel's form (scaled triples, reciprocals)
        recon3<false, true>(12.0 * eps, tri3(x[0], x[1], x[2]), tri3(x[1], x[2], x[3]), tri3(x[2], x[3], x[4]), vl, vr);
    } else if constexpr (K == 2) {
        recon2(eps, pair2(x[0], x[1]), pair2(x[1], x[2]), vl, vr);
    } else {
        vl = vr = x[0];
    }
}

// The x2 reconstruction marching down a column: the pieces of each triple (pair) of rows are
// formed once and reused by the three (two) rows whose windows contain it.
template <int K, bool FAST>
struct ColumnRecon {
    double cw[2 * K - 1];                // u(i, r-K+1 .. r+K-1) for the current row r
    [[maybe_unused]] Tri3 t2, t1;        // K = 3, FAST: T(r-2), T(r-1)
    [[maybe_unused]] Pair2 p1;           // K = 2, FAST: pair (r-1, r)

    __device__ __forceinline__ void start(double eps)
    {
        (void)eps;
        if constexpr (FAST && K == 3) {
            t2 = tri3(cw[0], cw[1], cw[2]);
            t1 = tri3(cw[1], cw[2], cw[3]);
        } else if constexpr (FAST && K == 2) {
            p1 = pair2(cw[0], cw[1]);
        }
    }
    // reconstruction of the window now in cw (the triples of its lower part already formed)
    __device__ __forceinline__ void recon_current(double eps, double &vl, double &vr)
    {
        if constexpr (FAST && K == 3) {
            const Tri3 t0 = tri3(cw[2], cw[3], cw[4]);
            recon3<false, true>(12.0 * eps, t2, t1, t0, vl, vr);
            t2 = t1;
            t1 = t0;
        } else if constexpr (FAST && K == 2) {
            const Pair2 p0 = pair2(cw[1], cw[2]);
            recon2(eps, p1, p0, vl, vr);
            p1 = p0;
        } else {
            recon<K, FAST>(cw, eps, vl, vr);
        }
    }
    __device__ __forceinline__ void push(double v)
    {
#pragma unroll
        for (int m = 0; m < 2 * K - 2; ++m) cw[m] = cw[m + 1];
        cw[2 * K - 2] = v;
    }
};

// Godunov flux of the linear flux g*v at one interface (src/hrweno_fluxes.f90:66-74).
__device__ __forceinline__ double godunov_lin(double g, double vm, double vp)
{
    return godunov(g * vm, g * vp, vm, vp);
}

// Whole-wavefront shifts by one lane (DPP wave_shr:1 / wave_shl:1): lane l receives src of lane
// l-1 (l+1); the lane with no source (0, resp. 63) receives its own `old`.  Every lane of the
// wavefront must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double old, double src)
{
    const long long o = __builtin_bit_cast(long long, old), v = __builtin_bit_cast(long long, src);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)v, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(v >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double from_left(double old, double v) { return dpp_shift<0x138>(old, v); }
__device__ __forceinline__ double from_right(double old, double v) { return dpp_shift<0x130>(old, v); }

// A copy of x in a register of its own.  A loaded value that becomes loop-carried state (the
// newest value of the column window) would otherwise keep the load's destination register
// alive, and the next load into that register would first copy it out, waiting for the load.
__device__ __forceinline__ double fresh(double x)
{
    double y;
    asm volatile("v_mov_b64 %0, %1" : "=v"(y) : "v"(x));
    return y;
}

// Blocks are dispatched round-robin over the 8 XCDs.  This maps block b to a logical block so
// that each XCD works on one contiguous range of logical blocks (the rows a band of
// wavefronts shares with the bands above and below stay in that XCD's L2).  A bijection on
// [0, nb).
__device__ __forceinline__ unsigned xcd_major(unsigned b, unsigned nb)
{
    constexpr unsigned kXcd = 8;
    const unsigned q = nb / kXcd, r = nb % kXcd, x = b % kXcd, loc = b / kXcd;
    return x < r ? x * (q + 1) + loc : r * (q + 1) + (x - r) * q + loc;
}

// The x1 window of row j, u(clamp(i+m), j) for m = -(K-1)..K-1, from the lanes' centre values
// c = u(clamp(i), j) and the halo values: lanes 0..K-2 hold u(clamp(i-(K-1)), j) and lanes
// 64-(K-1)..63 hold u(clamp(i+(K-1)), j).
template <int K>
__device__ __forceinline__ void x1_window(double c, double halo, double *x)
{
    x[K - 1] = c;
    if constexpr (K == 2) {
        x[0] = from_left(halo, c);
        x[2] = from_right(halo, c);
    } else if constexpr (K == 3) {
        x[1] = from_left(from_right(0.0, halo), c);    // lane 0: lane 1's halo
        x[0] = from_left(halo, x[1]);                  // lane 0: its own halo
        x[3] = from_right(from_left(0.0, halo), c);    // lane 63: lane 62's halo
        x[4] = from_right(halo, x[3]);                 // lane 63: its own halo
    }
}

// Work decomposition: one wavefront owns a strip of 64 columns and TY rows.  The loads of row
// j+1 (the new x2 window value, the x1 halo, the stage's pointwise operands) are issued before
// row j is computed, by every lane and on every row (rows and columns clamped into the grid):
// a load on only some paths would make the compiler's wait for row j's values cover them too.
// NPW = pointwise operands of the stage (0; 1: u0; 2: u_{n-4}, L(u_{n-4})).
template <int K, int BC, bool FAST, int NPW>
__device__ __forceinline__ void pbe2d_wave(const Stage2DArgs &a)
{
    const int lane = threadIdx.x & 63;
    const unsigned blk = xcd_major(blockIdx.x, gridDim.x);
    const long long wave = __builtin_amdgcn_readfirstlane((int)(blk * (kBlock / 64) + (threadIdx.x >> 6)));
    const int sx = (int)(wave % a.nstrips);
    const long long sy = wave / a.nstrips;
    const int n1 = a.n1, n2 = a.n2;
    const long long Y0 = sy * a.ty;
    if (Y0 >= n2) return;                               // whole wavefront leaves together
    const int Y1 = (int)min((long long)n2, Y0 + a.ty);
    constexpr bool SHARE = K == 3 && FAST;             // centred triples shared between lanes
    constexpr int S = strip_width(K, FAST), H0 = SHARE ? 2 : 1;   // output columns, left halo lanes
    const int i = sx * S - H0 + lane;                   // column of this lane (may be < 0 or >= n1)
    auto colc = [n1](int c) { return c < 0 ? 0 : (c >= n1 ? n1 - 1 : c); };
    const int ic = colc(i);                             // clamped column for loads
    const bool own = lane >= H0 && lane < H0 + S && i < n1;
    // halo column: lanes 0..K-2 need i-(K-1), lanes 64-(K-1)..63 need i+(K-1) (shared triples:
    // lane 0 needs i-1 and lane 63 i+1); the other lanes read the nearest halo lane's column (no
    // extra cache lines)
    const int hl = lane < 32 ? min(lane, K - 2) : max(lane, 64 - (K - 1));
    const int hc = SHARE ? colc(sx * S - H0 + (lane < 32 ? -1 : 64))
                         : colc(sx * S - H0 + hl + (lane < 32 ? -(K - 1) : K - 1));

    // per-lane constants: x1 width, g1 at the right interface x_{i+1/2}
    const double e_l = linear_edge(a.x1min, a.rx1, ic), e_r = linear_edge(a.x1min, a.rx1, ic + 1);
    const double w1 = e_r - e_l;
    const double g1 = linear_coef(a.g1a, a.g1b, e_r);
    auto rowc = [n2](long long j) -> long long { return j < 0 ? 0 : (j >= n2 ? n2 - 1 : j); };
    const double *us = a.us;
    const int stage = a.stage;

    // column window for the x2 reconstruction of row r: cw[m] = u(i, r-K+1+m)
    ColumnRecon<K, FAST> col;
    // prologue: reconstruction of row Y0-1 (for F2 at Y0-1/2) and of row Y0
    double vl2c, vr2c, f2_dn;
    {
#pragma unroll
        for (int m = 0; m < 2 * K - 1; ++m) col.cw[m] = us[rowc(Y0 - 1 - (K - 1) + m) * n1 + ic];
        col.start(a.eps);
        double vlp, vrp;
        col.recon_current(a.eps, vlp, vrp);
        col.push(us[rowc(Y0 + K - 1) * n1 + ic]);
        col.recon_current(a.eps, vl2c, vr2c);
        if (Y0 == 0) {
            f2_dn = 0.0;                                // zero flux at x2 = x2min
        } else {
            const double g2 = linear_coef(a.g2a, a.g2b, linear_edge(a.x2min, a.rx2, Y0));
            f2_dn = godunov_lin(g2, vrp, vl2c);
        }
    }

    // operands of row r (r may be Y1: clamped, loaded and not used)
    struct RowOps {
        double push, halo, p1, p2;
    };
    auto load_row = [&](long long r, RowOps &o) {
        const long long rr = rowc(r);
        o.push = us[rowc(r + K) * n1 + ic];
        o.halo = us[rr * n1 + hc];
        if constexpr (NPW == 1) o.p1 = a.u0[rr * n1 + ic];
        if constexpr (NPW == 2) {
            o.p1 = a.uo4[rr * n1 + ic];
            o.p2 = a.Lo4[rr * n1 + ic];
        }
    };

    const double iw1 = FAST ? rcp_nr(w1) : 0.0;
    auto row = [&](int j, const RowOps &o) {
        const double uc = col.cw[K - 1];                // u(i, j)

        // x2: reconstruction of row j+1, flux at j+1/2
        col.push(fresh(o.push));
        double vl2n, vr2n;
        col.recon_current(a.eps, vl2n, vr2n);
        const double e2l = linear_edge(a.x2min, a.rx2, j), e2r = linear_edge(a.x2min, a.rx2, j + 1);
        const double w2 = e2r - e2l;
        double f2_up;
        if (j == n2 - 1) {
            f2_up = BC == 0 ? 0.0 : godunov_lin(linear_coef(a.g2a, a.g2b, e2r), vr2c, vr2c);
        } else {
            f2_up = godunov_lin(linear_coef(a.g2a, a.g2b, e2r), vr2c, vl2n);
        }

        // x1: reconstruction of (i, j), flux at i+1/2 with the right neighbour's vl
        double vl1, vr1;
        if constexpr (SHARE) {
            // T(i-1) = (u(i-1), u(i), u(i+1)) here; T(i-2) and T(i) are the neighbours' own
            const double xl = from_left(o.halo, uc), xr = from_right(o.halo, uc);
            const Tri3 tm = tri3(xl, uc, xr);
            Tri3 tl{}, tr{};
            tl.bC = from_left(tm.bC, tm.bC);
            tl.P2 = from_left(tm.P2, tm.P2);
            tl.P1 = from_left(tm.P1, tm.P1);
            tr.bA = from_right(tm.bA, tm.bA);
            tr.P0 = from_right(tm.P0, tm.P0);
            tr.Pm = from_right(tm.Pm, tm.Pm);
            recon3<false, true>(12.0 * a.eps, tl, tm, tr, vl1, vr1);
        } else {
            double xw[2 * K - 1];
            x1_window<K>(uc, o.halo, xw);
            recon<K, FAST>(xw, a.eps, vl1, vr1);
        }
        const double vl1_right = from_right(vl1, vl1);
        double f1_r;
        if (i < 0) f1_r = 0.0;                           // zero flux at x1 = x1min
        else if (i >= n1 - 1) f1_r = BC == 0 ? 0.0 : godunov_lin(g1, vr1, vr1);
        else f1_r = godunov_lin(g1, vr1, vl1_right);
        const double f1_l = from_left(f1_r, f1_r);

        if (own) {
            const long long g = (long long)j * n1 + i;
            // example2:124-126
            double L;
            if constexpr (FAST) L = fma_(f1_l - f1_r, iw1, -((f2_up - f2_dn) * rcp_nr(w2)));
            else L = -((f1_r - f1_l) / w1) - (f2_up - f2_dn) / w2;
            if (stage < 0) {
                a.Lout[g] = L;
            } else {
                const double z = NPW == 1 ? o.p1 : uc;
                const double un = stage_update<FAST>(stage, z, uc, L, a.h, o.p1, o.p2);
                if (a.Lout) a.Lout[g] = L;
                a.uo[g] = un;
            }
        }
        f2_dn = f2_up;
        vl2c = vl2n;
        vr2c = vr2n;
    };

    // two operand sets in turn: a set is reloaded only after its row is done, so no register
    // copy (which would wait for the load) sits between a load and its use
    RowOps A{0.0, 0.0, 0.0, 0.0}, B{0.0, 0.0, 0.0, 0.0};
    load_row(Y0, A);
    for (int j = (int)Y0; j < Y1; j += 2) {
        load_row(j + 1, B);
        row(j, A);
        if (j + 1 >= Y1) break;
        load_row(j + 2, A);
        row(j + 1, B);
    }
}

template <int K, int BC, bool FAST>
__global__ __launch_bounds__(kBlock) void pbe2d_kernel(Stage2DArgs a)
{
    if (a.stage == MS3) pbe2d_wave<K, BC, FAST, 2>(a);
    else if (a.stage == RK2_2 || a.stage == RK3_2 || a.stage == RK3_3) pbe2d_wave<K, BC, FAST, 1>(a);
    else pbe2d_wave<K, BC, FAST, 0>(a);
}

}  // namespace hrw
