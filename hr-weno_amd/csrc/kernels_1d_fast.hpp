// kernels_1d_fast.hpp -- register-blocked fused stage kernel for the FAST arithmetic form.
//
// The stage pass is bound by FP64 instruction issue, not by HBM (its traffic is already the
// algorithmic 16/24/24 bytes per cell).  This kernel cuts the instructions per cell:
//
//  * each lane owns R consecutive cells; the K-1... halo comes from the neighbouring lanes by
//    DPP wave shifts (shfl_up1 / shfl_down1: a VALU operand modifier, no LDS round trip),
//    lanes 0 and 63 load theirs directly;
//  * adjacent cells share work (and, for K = 3 with R = 4, adjacent lanes: a lane takes its
//    left neighbour's last two triples by shuffles instead of forming them again): with
//    D = v0 - 2v1 + v2 and E = v0 - v2 on the triple
//    T(m) = (v_m, v_m+1, v_m+2), the three smoothness indicators drawing on T(m) are
//        12*beta_0(m)   = 13 D^2 + 3 (E + 2D)^2
//        12*beta_1(m+1) = 13 D^2 + 3 E^2
//        12*beta_2(m+2) = 13 D^2 + 3 (2D - E)^2
//    and the four candidate values are 6*c(:,r).T(m), r = -1..2 (used as vlr0(m), vrr0(m) =
//    vlr1(m+1), vrr1(m+1) = vlr2(m+2), vrr2(m+2)), formed from one another and the smoothness
//    terms; the factors 12 and 6 cancel in the normalised weights (eps is scaled by 12 to
//    match);
//  * one Lax-Friedrichs/Godunov flux and one grid edge per interface;
//  * weights over a common denominator, w_r ~ d_r prod_{s!=r} q_s, with d scaled to integers
//    (3, 6, 1) so they need three multiplies, and one v_rcp_f64 + two Newton steps per face
//    instead of IEEE divides; 1/width by one Newton step from 1/rx.
//
// All of this is an exact rewrite of the same algebra; results differ from the reference's
// evaluation order by a few ulp per reconstruction (the FAST contract).
#pragma once
#include <hip/hip_runtime.h>

#include "kernels_1d.hpp"

namespace hrw {

// 1/x to ~1 ulp: hardware estimate + two Newton steps.
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma_(-x, r, 1.0);
    r = fma_(r, e, r);
    e = fma_(-x, r, 1.0);
    return fma_(r, e, r);
}

// Shared pieces of the triple (a, b, c) = (v_m, v_m+1, v_m+2) for K = 3.
struct Tri3 {
    double bA, bB, bC;       // 12*beta_0(m), 12*beta_1(m+1), 12*beta_2(m+2)
    double Pm, P0, P1, P2;   // 6*c(:,-1).T, 6*c(:,0).T, 6*c(:,1).T, 6*c(:,2).T
};

// With D = a - 2b + c, E = a - c, X = 2D + E, Y = 2D - E:  12 beta = 13 D^2 + 3 E^2 (middle),
// 13 D^2 + 3 X^2, 13 D^2 + 3 Y^2; and the candidates follow from P0 = 2a + 5b - c by one term
// each: P1 = P0 - 3E, Pm = P0 + 3X, P2 = P1 + 3Y (18 operations instead of 23).
__device__ __forceinline__ Tri3 tri3(double a, double b, double c)
{
    Tri3 t;
    const double D = fma_(-2.0, b, a) + c;
    const double E = a - c;
    const double dd = (13.0 * D) * D;
    const double X = fma_(2.0, D, E), Y = fma_(2.0, D, -E);
    const double E3 = 3.0 * E, X3 = 3.0 * X, Y3 = 3.0 * Y;
    t.bB = fma_(E3, E, dd);
    t.bA = fma_(X3, X, dd);
    t.bC = fma_(Y3, Y, dd);
    t.P0 = fma_(5.0, b, fma_(2.0, a, -c));   // 2a + 5b - c
    t.P1 = t.P0 - E3;                        // -a + 5b + 2c
    t.Pm = t.P0 + X3;                        // 11a - 7b + 2c
    t.P2 = t.P1 + Y3;                        // 2a - 7b + 11c
    return t;
}

// One reconstruction from the pieces of T(j-2), T(j-1), T(j).  eps12 = 12*eps.  Each face
// takes one reciprocal of its own sum of weights, so the only products are pairs of q's: finite
// while beta < ~1e77, i.e. for cell-value jumps below ~1e38.  SHARED: the two faces share one
// reciprocal, of the product of their sums of weights (within a factor 3 of each other): one
// transcendental and one FP64 operation fewer per cell, two more values live across it.  With
// eps above machine epsilon every q is at least (12 eps)^2 ~ 1e-29, so the product stays above
// ~1e-115; it stays finite while beta < ~1e37 (cell-value jumps below ~1e18).  SCALED: vl and vr come out
// multiplied by 6 (the candidates are 6 c(:,r).T; the caller's flux constants absorb the
// factor, two multiplies fewer per cell).
template <bool SCALED = false, bool SHARED = false>
__device__ __forceinline__ void recon3(double eps12, const Tri3 &t2, const Tri3 &t1,
                                       const Tri3 &t0, double &vl, double &vr)
{
    double q0 = eps12 + t0.bA, q1 = eps12 + t1.bB, q2 = eps12 + t2.bC;
    q0 = q0 * q0;
    q1 = q1 * q1;
    q2 = q2 * q2;
    const double m0 = q1 * q2, m1 = q0 * q2, m2 = q0 * q1;
    // d = (0.3, 0.6, 0.1) ~ (3, 6, 1) on the right face, reversed on the left face
    const double a0 = 3.0 * m0, a1 = 6.0 * m1, l2 = 3.0 * m2;
    double sr = a0 + a1 + m2, sl = m0 + a1 + l2;
    if constexpr (!SCALED) {
        sr = 6.0 * sr;
        sl = 6.0 * sl;
    }
    const double nr = fma_(a0, t0.P0, fma_(a1, t1.P1, m2 * t2.P2));
    const double nl = fma_(m0, t0.Pm, fma_(a1, t1.P0, l2 * t2.P1));
    if constexpr (SHARED) {
        const double i2 = rcp_nr(sr * sl);
        vr = (nr * sl) * i2;
        vl = (nl * sr) * i2;
    } else {
        vr = nr * rcp_nr(sr);
        vl = nl * rcp_nr(sl);
    }
}

// K = 2: pair (a, b) = (v_m, v_m+1) gives beta_0(m) = beta_1(m+1) = (b-a)^2 and the
// candidates 2*vrr0(m) = a+b = 2*vlr1(m+1), 2*vrr1(m+1) = 3b-a, 2*vlr0(m) = 3a-b.
struct Pair2 {
    double B, S, Rr, Lr;
};
__device__ __forceinline__ Pair2 pair2(double a, double b)
{
    Pair2 p;
    const double d = b - a;
    p.B = d * d;
    p.S = a + b;
    p.Rr = fma_(3.0, b, -a);
    p.Lr = fma_(3.0, a, -b);
    return p;
}
// SCALED: vl and vr come out multiplied by 2.
template <bool SCALED = false>
__device__ __forceinline__ void recon2(double eps, const Pair2 &p1, const Pair2 &p0, double &vl,
                                       double &vr)
{
    double q0 = eps + p0.B, q1 = eps + p1.B;
    q0 = q0 * q0;
    q1 = q1 * q1;
    // right: d = (2/3, 1/3) ~ (2, 1): weights ~ 2 q1, q0; left: (1, 2): q1, 2 q0
    const double a0 = 2.0 * q1, b1 = 2.0 * q0;
    double sr = a0 + q0, sl = q1 + b1;
    if constexpr (!SCALED) {
        sr = 2.0 * sr;
        sl = 2.0 * sl;
    }
    const double nr = fma_(a0, p0.S, q0 * p1.Rr);
    const double nl = fma_(q1, p0.Lr, b1 * p1.S);
    vr = nr * rcp_nr(sr);
    vl = nl * rcp_nr(sl);
}

// --------------------------------------------------------------------------------------
// R consecutive doubles from p (16-byte aligned when R is even) into out.
template <int R> __device__ __forceinline__ void load_run(const double *p, double *out)
{
    if constexpr (R % 2 == 0) {
#pragma unroll
        for (int i = 0; i < R; i += 2) {
            const double2 t = *reinterpret_cast<const double2 *>(p + i);
            out[i] = t.x;
            out[i + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int i = 0; i < R; ++i) out[i] = p[i];
    }
}
template <int R> __device__ __forceinline__ void store_run(double *p, const double *in)
{
    if constexpr (R % 2 == 0) {
#pragma unroll
        for (int i = 0; i < R; i += 2) *reinterpret_cast<double2 *>(p + i) = make_double2(in[i], in[i + 1]);
    } else {
#pragma unroll
        for (int i = 0; i < R; ++i) p[i] = in[i];
    }
}

// Stage combination in one form for every TVD stage: with L = rhs(us),
//   un = post * (C*L + B*p1 + A*us + D*p2)
// p1/p2 are the pointwise operands (u at the start of the step; or u_{n-4}, L(u_{n-4}) for
// the multistep method).  The register-blocked kernels evaluate it with the factor post folded
// into the coefficients.  In every stage with a pointwise operand the weights of us and p1 sum
// to one (A' + B' = 1 with A' = A*post, B' = B*post), so it is taken as
// fma(D', p2, fma(C', L, fma(A', us - p1, p1))): the rounding of a constant such as 2/3 then
// scales only the O(dt) difference us - p1, and the state's total is not biased step after
// step (which near a shock would move the shock).  The first stages (A = post = 1, no
// pointwise operand) are fma(C, L, us).
struct StageCoef {
    double A, B, C, D, post;
    bool div3;
};

__host__ __device__ inline StageCoef stage_coef(int stage, double h)
{
    switch (stage) {
    case RK1:
    case RK2_1:
    case RK3_1: return {1.0, 0.0, h, 0.0, 1.0, false};          // us + h L
    case RK2_2: return {1.0, 1.0, h, 0.0, 0.5, false};          // (u0 + us + h L)/2
    case RK3_2: return {1.0, 3.0, h, 0.0, 0.25, false};         // (3u0 + us + h L)/4
    case RK3_3: return {2.0, 1.0, 2.0 * h, 0.0, 1.0, true};     // (u0 + 2us + 2h L)/3
    case MS3: return {25.0, 7.0, 50.0 * h, 10.0 * h, 0.03125, false};
    default: return {1.0, 0.0, 0.0, 0.0, 1.0, false};           // rhs only: un = us
    }
}

// One wavefront's chunk of a line: OUTPUT cells [W, W + 62R).  Lane l reconstructs the R cells
// from c0 = W + (l-1)*R, each once (window halo K-1 cells each side, from the neighbouring
// lanes by lane shuffles; lanes 0 and 63 load their outer halo).  A lane's first and last
// cells need a face value from a neighbouring lane (vr of cell c0-1, vl of cell c0+R), which
// that lane has only at the end of its own march: they are finished after the march, from
// the values kept for them.  Lanes 0 and 63 are halo lanes: their cells only supply the outer
// faces of lanes 1 and 62 and they write nothing.  Interior chunks (range and halo inside the
// line) take vector loads/stores and no boundary checks.  NPW = number of pointwise operands
// of the stage (0, 1 or 2).  Returns the lane's max|output| (the next stage's alpha
// contribution).  Requires n < 2^31 and 16-byte aligned arrays (checked on the host).  Must be
// called by all 64 lanes of the wavefront.
constexpr int kOutLanes = 62;

// What an interior chunk reads from memory, held in registers so that it can be loaded ahead
// of its use: the lane's R cells of us, the pointwise operands and, on the two wave-edge
// lanes, the K-1 halo cells beyond the chunk.  All of it is loaded together, so that a chunk
// computed from loaded data issues no load after the loads of the next one (a later load
// would make its wait cover them too).
template <int K, int R, int NPW>
struct ChunkData {
    double run[R];
    double p1[NPW >= 1 ? R : 1], p2[NPW >= 2 ? R : 1];
    double halo[K > 1 ? K - 1 : 1];
};

__device__ __forceinline__ bool chunk_interior(long long W, int R, int K, long long n)
{
    return W - R - (K - 1) >= 0 && W + 63LL * R + (K - 1) <= n;
}

template <int K, int R, int NPW>
__device__ __forceinline__ void load_chunk(const Stage1DArgs &a, int W, ChunkData<K, R, NPW> &d)
{
    const int lane = threadIdx.x & 63;
    const int c0 = W + (lane - 1) * R;
    load_run<R>(a.us + c0, d.run);
    const double *s1 = a.stage == MS3 ? a.uo4 : a.u0;
    if constexpr (NPW >= 1) load_run<R>(s1 + c0, d.p1);
    if constexpr (NPW >= 2) load_run<R>(a.Lo4 + c0, d.p2);
    // lanes 0..31 read the K-1 cells before lane 0's, lanes 32..63 the K-1 after lane 63's:
    // two addresses per wave, no divergent branch (lanes 0 and 63 use them)
    if constexpr (K > 1) {
        const double *h = a.us + (lane < 32 ? W - R - (K - 1) : W + 63 * R);
#pragma unroll
        for (int i = 0; i < K - 1; ++i) d.halo[i] = h[i];
    }
}

// The chunks [lo, hi] of a line of n cells are interior (hi < lo: none).
__device__ __forceinline__ void interior_range(long long n, int R, int K, long long &lo, long long &hi)
{
    const long long cw = (long long)kOutLanes * R;
    const long long need_lo = R + (K - 1), need_hi = 63LL * R + (K - 1);
    lo = (need_lo + cw - 1) / cw;
    hi = n >= need_hi ? (n - need_hi) / cw : -1;
}

// INTERIOR: the chunk is interior and d holds its loads (issued by the caller, so that the
// stage itself issues no load); otherwise d is not read and the cells are loaded here with
// the boundary handling.  The line is on a linear grid (lines with a width array take the tile
// kernel), so an interior chunk's stage has no load at all -- the persistent form needs that: a
// load on only some paths makes the compiler's wait after the paths join cover every load in
// flight, the next chunk's prefetch included.  NEWTON_W: 1/width by one Newton step from
// 1/rx (the host keeps the line's edges within 1e8 cell widths of the origin, where every
// width is rx to ~1e-8 and one step gives 1/w to a few ulp); otherwise by the hardware estimate
// with two steps.
template <int K, int FLUX, bool PERIODIC, int R, int NPW, bool INTERIOR, bool NEWTON_W = true,
          bool CONSTW = false>
__device__ __forceinline__ double chunk_stage(const Stage1DArgs &a, int W, double alpha,
                                              const ChunkData<K, R, NPW> &d)
{
    static_assert(R >= 4 && R % 2 == 0, "pairs of cells per lane");
    static_assert(!CONSTW || NEWTON_W, "the constant width takes the Newton form's value");
    constexpr int H = K - 1;
    constexpr int NV = R + 2 * H;
    const int lane = threadIdx.x & 63;
    const int n = (int)a.n;
    const int c0 = W + (lane - 1) * R;
    const bool owned = lane >= 1 && lane <= kOutLanes;
    const StageCoef sc = stage_coef(a.stage, a.h);
    // post (1/3 for the stages flagged div3) folded into the coefficients (NPW >= 1; the
    // NPW = 0 stages have A = post = 1)
    const double post = sc.div3 ? 1.0 / 3.0 : sc.post;
    [[maybe_unused]] const double cA = sc.A * post, cC = sc.C * post, cD = sc.D * post;
    const bool write_u = a.stage >= 0;
    // the reconstructions come out scaled by S (recon2/recon3 <SCALED>); the flux constants
    // take it back: f(v) = v^2/2 -> (v')^2 / (2 S^2), Lax-Friedrichs
    // ((vm^2 + vp^2)/2 - alpha (vp - vm))/2 -> ((vm'^2 + vp'^2)/(4 S^2) - alpha/(2S) (vp' - vm'))
    constexpr double S = K == 3 ? 6.0 : (K == 2 ? 2.0 : 1.0);
    [[maybe_unused]] const double cq = 0.25 / (S * S), ca = (-0.5 / S) * alpha;
    [[maybe_unused]] const double cg = 0.5 / (S * S);
    auto numflux = [&](double vm, double vp) {
        if constexpr (FLUX == GODUNOV) return godunov((vm * vm) * cg, (vp * vp) * cg, vm, vp);
        else return fma_(cq, fma_(vm, vm, vp * vp), ca * (vp - vm));
    };

    // ---- the lane's cells and window halo (v[i] <-> cell c0 - H + i) and pointwise operands
    double v[NV];
    [[maybe_unused]] double p1[NPW >= 1 ? R : 1], p2[NPW >= 2 ? R : 1];
    if constexpr (INTERIOR) {
#pragma unroll
        for (int i = 0; i < R; ++i) v[H + i] = d.run[i];
        if constexpr (NPW >= 1) {
#pragma unroll
            for (int i = 0; i < R; ++i) p1[i] = d.p1[i];
        }
        if constexpr (NPW >= 2) {
#pragma unroll
            for (int i = 0; i < R; ++i) p2[i] = d.p2[i];
        }
#pragma unroll
        for (int i = 0; i < H; ++i) {
            v[i] = shfl_up1(v[R + i]);            // lane-1's last H cells
            v[H + R + i] = shfl_down1(v[H + i]);  // lane+1's first H cells
        }
        if constexpr (H > 0) {
            if (lane == 0) {
#pragma unroll
                for (int i = 0; i < H; ++i) v[i] = d.halo[i];
            }
            if (lane == 63) {
#pragma unroll
                for (int i = 0; i < H; ++i) v[H + R + i] = d.halo[i];
            }
        }
    } else {
        const double *s1 = a.stage == MS3 ? a.uo4 : a.u0;
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] = a.us[cell_index((long long)c0 - H + i, n, PERIODIC)];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int g = c0 + i < 0 ? 0 : (c0 + i < n ? c0 + i : n - 1);
            if constexpr (NPW >= 1) p1[i] = s1[g];
            if constexpr (NPW >= 2) p2[i] = a.Lo4[g];
        }
    }
    // a cell is real (inside the line) -- every cell of an interior chunk is
    auto real = [&](int g) { return INTERIOR || (g >= 0 && g < n); };

    // ---- width of cell g, recomputed from the edges like the reference
    // (src/hrweno_grids.f90:76-78,247), and its reciprocal (see NEWTON_W)
    auto edge = [&](int g) {
        return linear_edge32(a.xmin, a.rx, INTERIOR ? g : (g < 0 ? 0 : (g < n ? g : n)));
    };
    auto width_of = [&](int, double e_lo, double e_hi) { return e_hi - e_lo; };
    [[maybe_unused]] const double inv_rx = NEWTON_W ? rcp_nr(a.rx) : 0.0;
    // CONSTW: every edge xmin + rx*i of the line is exact (checked on the host), so every
    // width is rx and 1/width is the value the Newton form gives for w = rx, once per chunk
    [[maybe_unused]] const double inv_w_const =
        CONSTW ? fma_(inv_rx, fma_(-a.rx, inv_rx, 1.0), inv_rx) : 0.0;
    auto inv_width = [&](double w) {
        if constexpr (CONSTW) return inv_w_const;
        else if constexpr (NEWTON_W) return fma_(inv_rx, fma_(-w, inv_rx, 1.0), inv_rx);
        else return rcp_nr(w);
    };

    // ---- stage combination and store of cell t (global g); un/L of a pair's first cell wait
    double m_out = 0.0;
    auto finish = [&](int t, double fleft, double fright, double w, double &un, double &L) {
        const int g = c0 + t;
        if constexpr (!PERIODIC) {
            if (!INTERIOR) {   // example1:103-104
                if (g == 0) fleft = fright;
                if (g == n - 1) fright = fleft;
            }
        }
        L = (fleft - fright) * inv_width(w);
        const double us = v[H + t];
        if constexpr (NPW == 0) {
            un = fma_(sc.C, L, us);
        } else {
            double tt = fma_(cA, us - p1[t], p1[t]);
            tt = fma_(cC, L, tt);
            if constexpr (NPW >= 2) tt = fma_(cD, p2[t], tt);
            un = tt;
        }
        if (real(g)) m_out = fmax(m_out, fabs(un));
    };
    auto store_pair = [&](int t, double un0, double un1, double L0, double L1) {   // cells t, t+1
        if (!owned) return;
        const int g = c0 + t;
        if constexpr (INTERIOR) {
            if (write_u) *reinterpret_cast<double2 *>(a.uo + g) = make_double2(un0, un1);
            if (a.Lout) *reinterpret_cast<double2 *>(a.Lout + g) = make_double2(L0, L1);
        } else {
            if (g >= 0 && g < n) {
                if (write_u) a.uo[g] = un0;
                if (a.Lout) a.Lout[g] = L0;
            }
            if (g + 1 >= 0 && g + 1 < n) {
                if (write_u) a.uo[g + 1] = un1;
                if (a.Lout) a.Lout[g + 1] = L1;
            }
        }
    };

    if constexpr (K == 3 && R == 4) {
        // ---- march over the triples T(c0+t), t = 0..R-1, reconstructing cells 2..R-1 as their
        // three triples are formed; T(c0-2) and T(c0-1) are the left neighbour's last two
        // triples: the parts cells 0 and 1 use come over by lane shuffles after the march (the
        // same values a lane would form from its own window, so the results are those of the
        // march below), and those two cells follow.  Two triples (36 FP64 operations) fewer per
        // lane; with R = 6 or 8 the values held across the march no longer fit the register
        // budget of the occupancy those kernels run at, so they take the march below.
        const double eps12 = 12.0 * a.eps;
        const Tri3 ta = tri3(v[2], v[3], v[4]);   // T(c0)
        const Tri3 tb = tri3(v[3], v[4], v[5]);   // T(c0+1)
        Tri3 t2 = ta, t1 = tb;
        double vl2 = 0.0, vr_prev = 0.0, f_first = 0.0, fl = 0.0;
        double un3 = 0.0, L3 = 0.0, unR2 = 0.0, LR2 = 0.0, un_pair = 0.0, L_pair = 0.0;
#pragma unroll
        for (int t = 2; t < R; ++t) {
            const Tri3 t0 = tri3(v[t + 2], v[t + 3], v[t + 4]);   // T(c0+t)
            double vl, vr;
            recon3<true, true>(eps12, t2, t1, t0, vl, vr);
            t2 = t1;
            t1 = t0;
            if (t == 2) {
                vl2 = vl;
                vr_prev = vr;
                continue;
            }
            const double f = numflux(vr_prev, vl);   // face c0+t-1/2
            vr_prev = vr;
            if (t == 3) {
                f_first = f;
                fl = f;
                continue;
            }
            // cell t-1 (3 <= t-1 <= R-2) between fl and f
            const int c = t - 1;
            double un, L;
            finish(c, fl, f, width_of(c0 + c, edge(c0 + c), edge(c0 + c + 1)), un, L);
            fl = f;
            if (c == 3) {
                un3 = un;
                L3 = L;
            } else if (c == R - 2) {
                unR2 = un;
                LR2 = L;
            } else if (c % 2 == 0) {
                un_pair = un;
                L_pair = L;
            } else {
                store_pair(c - 1, un_pair, un, L_pair, L);
            }
        }
        // ---- T(c0-2) (bC, P2, P1) and T(c0-1) (bB, P1, P0, bC, P2) from lane-1, cells 0 and 1
        Tri3 l2{}, l1{};
        l2.bC = shfl_up1(t2.bC);
        l2.P2 = shfl_up1(t2.P2);
        l2.P1 = shfl_up1(t2.P1);
        l1.bB = shfl_up1(t1.bB);
        l1.P1 = shfl_up1(t1.P1);
        l1.P0 = shfl_up1(t1.P0);
        l1.bC = shfl_up1(t1.bC);
        l1.P2 = shfl_up1(t1.P2);
        const double vr_left = shfl_up1(vr_prev);   // vr of cell c0-1 (lane-1's last)
        double vl0, vr0, vl1, vr1;
        recon3<true, true>(eps12, l2, l1, ta, vl0, vr0);
        recon3<true, true>(eps12, l1, ta, tb, vl1, vr1);
        const double vl_right = shfl_down1(vl0);    // vl of cell c0+R (lane+1's first)
        const double f0 = numflux(vr_left, vl0);    // face c0-1/2
        const double f1 = numflux(vr0, vl1);        // face c0+1/2
        const double f2 = numflux(vr1, vl2);        // face c0+3/2
        const double fR = numflux(vr_prev, vl_right);   // face c0+R-1/2
        const double e1 = edge(c0 + 1), e2 = edge(c0 + 2);
        double un0, L0, un1, L1, un2, L2;
        finish(0, f0, f1, width_of(c0, edge(c0), e1), un0, L0);
        finish(1, f1, f2, width_of(c0 + 1, e1, e2), un1, L1);
        finish(2, f2, f_first, width_of(c0 + 2, e2, edge(c0 + 3)), un2, L2);
        store_pair(0, un0, un1, L0, L1);
        double unl, Ll;
        finish(R - 1, fl, fR, width_of(c0 + R - 1, edge(c0 + R - 1), edge(c0 + R)), unl, Ll);
        if constexpr (R == 4) {
            store_pair(2, un2, unl, L2, Ll);
        } else {
            store_pair(2, un2, un3, L2, L3);
            store_pair(R - 2, unR2, unl, LR2, Ll);
        }
        return owned ? m_out : 0.0;
    }

    // ---- march over the lane's cells t = 0..R-1: reconstruction of cell c0+t; the interface
    // between cells t-1 and t; cell t-1 finished for t >= 2
    [[maybe_unused]] const double eps12 = 12.0 * a.eps;
    [[maybe_unused]] Tri3 t2, t1;
    [[maybe_unused]] Pair2 pp1;
    if constexpr (K == 3) {
        t2 = tri3(v[0], v[1], v[2]);
        t1 = tri3(v[1], v[2], v[3]);
    } else if constexpr (K == 2) {
        pp1 = pair2(v[0], v[1]);
    }
    double vl0 = 0.0, vr_prev = 0.0, f_first = 0.0, fl = 0.0;
    const double e_c0 = edge(c0), e_c1 = edge(c0 + 1);
    double e_prev = e_c1, e_last = 0.0;
    double un_pair = 0.0, L_pair = 0.0, un1 = 0.0, L1 = 0.0, unR2 = 0.0, LR2 = 0.0;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        double vl, vr;
        if constexpr (K == 3) {
            const Tri3 t0 = tri3(v[t + 2], v[t + 3], v[t + 4]);   // T(c0+t)
            recon3<true, true>(eps12, t2, t1, t0, vl, vr);
            t2 = t1;
            t1 = t0;
        } else if constexpr (K == 2) {
            const Pair2 p0 = pair2(v[t + 1], v[t + 2]);
            recon2<true>(a.eps, pp1, p0, vl, vr);
            pp1 = p0;
        } else {
            vl = vr = v[t];
        }
        if (t == 0) {
            vl0 = vl;
            vr_prev = vr;
            continue;
        }
        const double f = numflux(vr_prev, vl);   // face c0+t-1/2
        vr_prev = vr;
        if (t == 1) {
            f_first = f;
            fl = f;
            continue;
        }
        // cell t-1 (1 <= t-1 <= R-2) between fl and f
        const int c = t - 1;
        const double e = edge(c0 + c + 1);
        const double w = width_of(c0 + c, e_prev, e);
        e_prev = e;
        double un, L;
        finish(c, fl, f, w, un, L);
        fl = f;
        if (c == 1) {
            un1 = un;
            L1 = L;
        } else if (c == R - 2) {
            unR2 = un;
            LR2 = L;
            e_last = e;
        } else if (c % 2 == 0) {
            un_pair = un;
            L_pair = L;
        } else {
            store_pair(c - 1, un_pair, un, L_pair, L);
        }
    }
    // ---- the two faces shared with the neighbouring lanes, then cells 0 and R-1
    const double vr_left = shfl_up1(vr_prev);    // vr of cell c0-1 (lane-1's last)
    const double vl_right = shfl_down1(vl0);     // vl of cell c0+R (lane+1's first)
    {
        const double f0 = numflux(vr_left, vl0);
        double un, L;
        finish(0, f0, f_first, width_of(c0, e_c0, e_c1), un, L);
        store_pair(0, un, un1, L, L1);
    }
    {
        const double fR = numflux(vr_prev, vl_right);
        double un, L;
        finish(R - 1, fl, fR, width_of(c0 + R - 1, e_last, edge(c0 + R)), un, L);
        store_pair(R - 2, unR2, un, LR2, L);
    }
    return owned ? m_out : 0.0;
}

// One chunk with its own loads (the batch kernel; the host checks its lines against the
// NEWTON_W bound, so that a line gives the single-line kernels' bits).
template <int K, int FLUX, bool PERIODIC, int R, int NPW>
__device__ __forceinline__ double chunk_load_stage(const Stage1DArgs &a, int W, double alpha)
{
    ChunkData<K, R, NPW> d;
    if (chunk_interior(W, R, K, a.n)) {
        load_chunk<K, R, NPW>(a, W, d);
        return chunk_stage<K, FLUX, PERIODIC, R, NPW, true>(a, W, alpha, d);
    }
    return chunk_stage<K, FLUX, PERIODIC, R, NPW, false>(a, W, alpha, d);
}

// Pointwise operands of a stage: u0 for the second and third RK stages, u_{n-4} and
// L(u_{n-4}) for the multistep stage, none otherwise.
__host__ __device__ constexpr int stage_npw(int stage)
{
    return stage == MS3 ? 2 : ((stage == RK2_2 || stage == RK3_2 || stage == RK3_3) ? 1 : 0);
}

// Waves per SIMD each stage kernel is compiled for (a budget of 512/waves VGPRs per lane).  With
// K = 3 the kernels take 102..128 VGPRs at R = 4 (0..2 pointwise operands), 120..168 at R = 6
// and 126..168 at R = 8.  R = 6 runs at 4 waves without a pointwise operand (the
// Lax-Friedrichs kernel with alpha = max|u| on an exact grid fits in 128 VGPRs; other variants
// spill) and at 3 with one or two.  The Godunov one-chunk kernel at R = 6 runs at 3 in every
// stage: at 4, its stage without a pointwise operand spilled 36-48 registers and was 7 % slower.
__host__ __device__ constexpr int fast_waves(int R, int NPW, int K = 3, int FLUX = LF)
{
    if (K == 3 && FLUX == GODUNOV && R == 6) return 3;
    return R >= 8 ? 3 : (R == 6 ? (NPW == 0 ? 4 : 3) : 4);
}

// The start of every stage wave: the alpha partial maxima of the previous launch are loaded
// first and the wave's first chunk right after them, so that the reduction of the partials
// waits for those loads only (vector memory counters complete in issue order) and overlaps
// the chunk's loads; then this launch's alpha is recorded and the slot set two launches ahead
// cleared.
template <bool ALPHAMAX>
__device__ __forceinline__ unsigned long long alpha_partial(const Stage1DArgs &a)
{
    if constexpr (ALPHAMAX) return a.alpha_in[((threadIdx.x & 63) % kNSlot) * kSlotStride];
    else return 0ull;
}
template <bool ALPHAMAX>
__device__ __forceinline__ double alpha_reduce(const Stage1DArgs &a, unsigned long long part)
{
    double alpha = a.alpha;
    if constexpr (ALPHAMAX) {
        alpha = __longlong_as_double((long long)wave_umax(part));
        if (a.alpha_record && blockIdx.x == 0 && threadIdx.x == 0) *a.alpha_record = alpha;
    }
    clear_set(a.alpha_clear);
    return alpha;
}

// Persistent form: wave w of the grid takes the interior chunks lo + w, lo + w + N, ... (N =
// waves in the grid) and loads the next one into a second register set before it computes the
// current one, so every wave keeps a chunk of loads in flight behind its FP64 work.  The loads
// are unconditional (the last iteration reloads its own chunk, an L2 hit): a load on only some
// paths would make the compiler's wait for the current chunk cover the prefetch as well.  The
// few boundary chunks follow, one per wave.
template <int K, int FLUX, bool PERIODIC, bool ALPHAMAX, int R, int NPW, bool CONSTW>
__device__ __forceinline__ void persistent_wave(const Stage1DArgs &a, double *s_red)
{
    const unsigned long long part = alpha_partial<ALPHAMAX>(a);
    const long long nw = (long long)gridDim.x * (kBlock / 64);
    const long long cw = (long long)kOutLanes * R;
    const long long nch = (a.n + cw - 1) / cw;
    long long lo, hi;
    interior_range(a.n, R, K, lo, hi);
    const long long w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)));
    double m = 0.0, alpha;
    ChunkData<K, R, NPW> cur, nxt;
    if (lo <= hi) {
        long long c = lo + w;
        load_chunk<K, R, NPW>(a, (int)((c <= hi ? c : hi) * cw), cur);
        alpha = alpha_reduce<ALPHAMAX>(a, part);
        for (; c <= hi; c += nw) {
            const long long c1 = c + nw <= hi ? c + nw : c;
            load_chunk<K, R, NPW>(a, (int)(c1 * cw), nxt);
            m = fmax(m, chunk_stage<K, FLUX, PERIODIC, R, NPW, true, true, CONSTW>(a, (int)(c * cw), alpha, cur));
            cur = nxt;
        }
    } else {
        alpha = alpha_reduce<ALPHAMAX>(a, part);
    }
    // boundary chunks [0, lo) and (hh, nch), hh = max(hi, lo - 1)
    const long long hh = hi >= lo ? hi : lo - 1;
    const long long nb = lo + (nch - 1 - hh);
    for (long long b = w; b < nb; b += nw) {
        const long long c = b < lo ? b : hh + 1 + (b - lo);
        m = fmax(m, chunk_stage<K, FLUX, PERIODIC, R, NPW, false, true, CONSTW>(a, (int)(c * cw), alpha, cur));
    }
    if constexpr (ALPHAMAX) publish_max(m, a.alpha_out, s_red);
}

// (The second register set of the prefetch costs one wave per SIMD against the plain form.
// With the shared triples K = 3, R = 4 would fit four with at most one pointwise operand; it
// measured no faster than three.)
template <int K, int FLUX, bool PERIODIC, bool ALPHAMAX, int R, int NPW, bool CONSTW>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(fast_waves(R, NPW) > 2 ? fast_waves(R, NPW) - 1 : 2)))
void stage1d_pf_kernel(Stage1DArgs a)
{
    __shared__ double s_red[kBlock / 64 + 1];
    persistent_wave<K, FLUX, PERIODIC, ALPHAMAX, R, NPW, CONSTW>(a, s_red);
}

// One chunk per wave: wave w of block b owns the chunk starting at W = (4b + w)*62R.
// When the line has interior chunks every wave loads one unconditionally (a boundary wave
// the nearest interior chunk, unused), so that no join of paths with and without loads comes
// before the alpha reduction, whose wait would then cover the chunk's loads.
template <int K, int FLUX, bool PERIODIC, bool ALPHAMAX, int R, int NPW, bool CONSTW>
__device__ __forceinline__ void chunk_wave(const Stage1DArgs &a, double *s_red)
{
    const long long cw = (long long)kOutLanes * R;
    const long long c = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)));
    const int W = (int)(c * cw);
    long long lo, hi;
    interior_range(a.n, R, K, lo, hi);
    const unsigned long long part = alpha_partial<ALPHAMAX>(a);
    ChunkData<K, R, NPW> d;
    double m_out;
    if (lo <= hi) {
        load_chunk<K, R, NPW>(a, (int)((c < lo ? lo : (c > hi ? hi : c)) * cw), d);
        const double alpha = alpha_reduce<ALPHAMAX>(a, part);
        if (c >= lo && c <= hi) m_out = chunk_stage<K, FLUX, PERIODIC, R, NPW, true, true, CONSTW>(a, W, alpha, d);
        else m_out = chunk_stage<K, FLUX, PERIODIC, R, NPW, false, true, CONSTW>(a, W, alpha, d);
    } else {
        const double alpha = alpha_reduce<ALPHAMAX>(a, part);
        m_out = chunk_stage<K, FLUX, PERIODIC, R, NPW, false, true, CONSTW>(a, W, alpha, d);
    }
    if constexpr (ALPHAMAX) publish_max(m_out, a.alpha_out, s_red);
}

// NPW = stage_npw(a.stage), chosen by the host.
template <int K, int FLUX, bool PERIODIC, bool ALPHAMAX, int R, int NPW, bool CONSTW>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(fast_waves(R, NPW, K, FLUX))))
void stage1d_fast_kernel(Stage1DArgs a)
{
    __shared__ double s_red[kBlock / 64 + 1];
    chunk_wave<K, FLUX, PERIODIC, ALPHAMAX, R, NPW, CONSTW>(a, s_red);
}

// --------------------------------------------------------------------------------------
// Many independent periodic lines in one launch.  Line l: cells [off[l], off[l] + n[l]) of
// the flat state, WENO order k[l], step h[l], linear grid xmin[l] + rx[l]*i.  Chunk c of the
// table covers cells [start[c], start[c] + 62R) of line line[c]; in batch_stage_kernel
// wavefront c of the grid takes chunk c (batch_pf_kernel below walks the table).
// Lax-Friedrichs alpha = max|v| per line: slot set s holds one partial per line.
struct BatchArgs {
    const double *us, *u0;
    double *uo;
    const long long *off;
    const int *n, *k;
    const double *h, *xmin, *rx;
    const int *chunk_line, *chunk_start;
    long long nchunks;
    int nlines, stage, flux;
    double eps, alpha;
    const unsigned long long *alpha_in;   // nlines partials, read
    unsigned long long *alpha_out;        // nlines partials, atomicMax
    unsigned long long *alpha_clear;      // nlines partials, cleared
    long long ninterior;                  // chunks [0, ninterior) are interior for K = 3
};

// The block's wavefront maxima into their lines' slots, one atomic per line present in the
// block: the block's wavefronts hold consecutive chunks of the table, mostly of one line, and on
// a few long lines thousands of wavefronts would otherwise meet on one slot's cache line.
// line < 0: the wavefront has no chunk.  Every thread of the block calls it.
__device__ __forceinline__ void block_line_max(unsigned long long *slots, int line, double m)
{
    __shared__ unsigned long long s_m[kBlock / 64];
    __shared__ int s_line[kBlock / 64];
    m = wave_max(m);
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_m[wv] = (unsigned long long)__double_as_longlong(m);
        s_line[wv] = line;
    }
    __syncthreads();
    if (threadIdx.x < kBlock / 64) {
        const int me = threadIdx.x, l = s_line[me];
        bool first = l >= 0;
        unsigned long long mx = s_m[me];
#pragma unroll
        for (int v = 0; v < kBlock / 64; ++v) {
            if (v < me && s_line[v] == l) first = false;
            if (v > me && s_line[v] == l) mx = s_m[v] > mx ? s_m[v] : mx;
        }
        if (first) atomicMax(slots + l, mx);
    }
}

// Line `line` of the batch as single-line stage arguments.
__device__ __forceinline__ Stage1DArgs batch_line_args(const BatchArgs &b, int line)
{
    const long long off = b.off[line];
    Stage1DArgs a{};
    a.us = b.us + off;
    a.u0 = b.u0 + off;
    a.uo = b.uo + off;
    a.n = b.n[line];
    a.xmin = b.xmin[line];
    a.rx = b.rx[line];
    a.eps = b.eps;
    a.h = b.h[line];
    a.stage = b.stage;
    return a;
}

template <int FLUX, bool ALPHAMAX, int R, int NPW>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(fast_waves(R, NPW))))
void batch_stage_kernel(BatchArgs b)
{
    // wave-uniform, so the chunk table and the line's parameters come through scalar loads
    const long long wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)));
    if (ALPHAMAX && b.alpha_clear) {
        for (long long l = (long long)blockIdx.x * kBlock + threadIdx.x; l < b.nlines;
             l += (long long)gridDim.x * kBlock)
            b.alpha_clear[l] = 0ull;
    }
    // (whole wavefronts past the table do no chunk, and take part in the block's barrier)
    int line = -1;
    double m = 0.0;
    if (wave < b.nchunks) {
        line = b.chunk_line[wave];
        const Stage1DArgs a = batch_line_args(b, line);
        double alpha = b.alpha;
        if constexpr (ALPHAMAX) alpha = __longlong_as_double((long long)b.alpha_in[line]);
        const int W = b.chunk_start[wave];
        switch (b.k[line]) {
        case 1: m = chunk_load_stage<1, FLUX, true, R, NPW>(a, W, alpha); break;
        case 2: m = chunk_load_stage<2, FLUX, true, R, NPW>(a, W, alpha); break;
        default: m = chunk_load_stage<3, FLUX, true, R, NPW>(a, W, alpha); break;
        }
    }
    if constexpr (ALPHAMAX) block_line_max(b.alpha_out, line, m);
}

// Persistent form of the batch: as many wavefronts as fit on the device, wave w taking chunks
// w, w + N, ... of the table (N waves in the grid), whose first ninterior chunks are interior
// for K = 3 and so for every k.  An interior chunk's loads (with the K = 3 halo, of which a line
// of k < 3 uses part) are issued into a second register set before the previous chunk's
// arithmetic, unconditionally (the last iteration reloads its own chunk); the chunks that are
// not interior follow, one per iteration, with their own loads.  (The second register set
// costs a wave per SIMD against batch_stage_kernel; which form is faster depends on the
// ensemble, and hrw_batch_tune times both.)
template <int K, int R, int NPW>
__device__ __forceinline__ ChunkData<K, R, NPW> narrow_chunk(const ChunkData<3, R, NPW> &d)
{
    ChunkData<K, R, NPW> o;
#pragma unroll
    for (int i = 0; i < R; ++i) o.run[i] = d.run[i];
    if constexpr (NPW >= 1) {
#pragma unroll
        for (int i = 0; i < R; ++i) o.p1[i] = d.p1[i];
    }
    if constexpr (NPW >= 2) {
#pragma unroll
        for (int i = 0; i < R; ++i) o.p2[i] = d.p2[i];
    }
    // K = 3 halo: cells W-R-2, W-R-1 (lanes 0..31) and W+63R, W+63R+1 (lanes 32..63)
    if constexpr (K == 2) o.halo[0] = (threadIdx.x & 63) < 32 ? d.halo[1] : d.halo[0];
    if constexpr (K == 3) {
        o.halo[0] = d.halo[0];
        o.halo[1] = d.halo[1];
    }
    return o;
}

template <int FLUX, bool ALPHAMAX, int R, int NPW>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(fast_waves(R, NPW) > 2 ? fast_waves(R, NPW) - 1 : 2)))
void batch_pf_kernel(BatchArgs b)
{
    const long long nw = (long long)gridDim.x * (kBlock / 64);
    const long long w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)));
    const int lane = threadIdx.x & 63;
    if (ALPHAMAX && b.alpha_clear) {
        for (long long l = (long long)blockIdx.x * kBlock + threadIdx.x; l < b.nlines;
             l += (long long)gridDim.x * kBlock)
            b.alpha_clear[l] = 0ull;
    }
    auto publish = [&](int line, double m) {
        if constexpr (ALPHAMAX) {
            m = wave_max(m);
            if (lane == 0) atomicMax(b.alpha_out + line, (unsigned long long)__double_as_longlong(m));
        }
    };
    auto line_alpha = [&](int line) {
        return ALPHAMAX ? __longlong_as_double((long long)b.alpha_in[line]) : b.alpha;
    };
    if (w < b.ninterior) {
        int line = b.chunk_line[w];
        Stage1DArgs a = batch_line_args(b, line);
        int W = b.chunk_start[w];
        ChunkData<3, R, NPW> cur, nxt;
        load_chunk<3, R, NPW>(a, W, cur);
        for (long long i = w; i < b.ninterior; i += nw) {
            const long long i1 = i + nw < b.ninterior ? i + nw : i;
            const int line1 = b.chunk_line[i1];
            const Stage1DArgs a1 = batch_line_args(b, line1);
            const int W1 = b.chunk_start[i1];
            load_chunk<3, R, NPW>(a1, W1, nxt);
            const double alpha = line_alpha(line);
            double m;
            switch (b.k[line]) {
            case 1: m = chunk_stage<1, FLUX, true, R, NPW, true>(a, W, alpha, narrow_chunk<1>(cur)); break;
            case 2: m = chunk_stage<2, FLUX, true, R, NPW, true>(a, W, alpha, narrow_chunk<2>(cur)); break;
            default: m = chunk_stage<3, FLUX, true, R, NPW, true>(a, W, alpha, cur); break;
            }
            publish(line, m);
            cur = nxt;
            line = line1;
            a = a1;
            W = W1;
        }
    }
    for (long long i = b.ninterior + w; i < b.nchunks; i += nw) {
        const int line = b.chunk_line[i];
        const Stage1DArgs a = batch_line_args(b, line);
        const int W = b.chunk_start[i];
        const double alpha = line_alpha(line);
        double m;
        switch (b.k[line]) {
        case 1: m = chunk_load_stage<1, FLUX, true, R, NPW>(a, W, alpha); break;
        case 2: m = chunk_load_stage<2, FLUX, true, R, NPW>(a, W, alpha); break;
        default: m = chunk_load_stage<3, FLUX, true, R, NPW>(a, W, alpha); break;
        }
        publish(line, m);
    }
}

// max|u| of every line into slot set `set` (one partial per line).
static __global__ __launch_bounds__(kBlock) void batch_maxabs_kernel(const double *__restrict__ u,
                                                              const long long *off, const int *n,
                                                              const int *chunk_line,
                                                              const int *chunk_start,
                                                              long long nchunks, int R,
                                                              unsigned long long *set)
{
    const long long wave = ((long long)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    int line = -1;
    double m = 0.0;
    if (wave < nchunks) {   // (wavefronts past the table take part in the block's barrier)
        line = chunk_line[wave];
        const long long base = off[line] + chunk_start[wave];
        const int len = min(kOutLanes * R, n[line] - chunk_start[wave]);
        for (int i = lane; i < len; i += 64) m = fmax(m, fabs(u[base + i]));
    }
    block_line_max(set, line, m);
}

static __global__ void fill_u64_kernel(unsigned long long *p, long long count)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

}  // namespace hrw
