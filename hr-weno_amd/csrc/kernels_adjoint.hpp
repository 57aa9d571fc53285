// kernels_adjoint.hpp -- reverse (vector-Jacobian) pass of one TVD stage for periodic 1D
// Burgers lines with the Lax-Friedrichs flux: the discrete adjoint of the fused stage pass.
//
// Forward stage (src/hrweno_tvdode.f90:136-171):  u_out = A*us + B*u_n + C*h*L(us), with
// L(us)_i = -(F_{i+1/2} - F_{i-1/2})/w_i, F = LF(vr_i, vl_{i+1}) (src/hrweno_fluxes.f90:43),
// vr/vl the WENO reconstruction of us (src/hrweno_weno.f90:174-216).  The reverse pass
// computes, for a cotangent lam of u_out,
//     out = a*lam + c*h*J(us)^T lam + e1*x1 + e2*x2
// (a = A, c = C; x1, x2 are pointwise cotangents folded in by the driver, e.g. the B terms of
// the later stages of the step).  Per cell i, with phi = lam/w:
//     F-bar_{i+1/2} = phi_{i+1} - phi_i
//     vr-bar_i = F-bar_{i+1/2} (vr_i + alpha)/2,   vl-bar_i = F-bar_{i-1/2} (vl_i - alpha)/2
//     alpha-bar = sum_i (F-bar_{i+1/2} vr_i - F-bar_{i-1/2} vl_i)/2
// then through the nonlinear weights: d vr / d p_r = -2 w_r (q_r - vr) / p_r with
// p_r = 12 (eps + beta_r) and q_r the candidate values, and through beta_r and q_r (linear /
// quadratic in the window) to the window cells.  With alpha = max|us| (a subgradient at the
// first cell attaining the maximum) alpha-bar lands on that cell with the sign of us there.
//
// Layout.  Wavefront c of the grid takes chunk c of a chunk table: OUTPUT cells
// [W, W + 62R) of line `line` (W = start[c]; line < 0 marks an idle padding chunk).  Lane l (0..63) owns the R consecutive cells
// from c0 = W + (l-1)*R; lanes 0 and 63 are halo lanes whose results are only used through
// the neighbour exchange.  Each lane scatters the window cotangents of its R reconstructions
// into R + 2(K-1) accumulators and hands the K-1 cells that spill past each end to the
// neighbouring lane (DPP wave shifts, shfl_up1 / shfl_down1).  The wave-edge lanes load the window and flux-cotangent
// cells beyond them directly (no neighbour to shuffle from), so their spills into lanes 1
// and 62 are exact for any R >= 2; only their own cells (which miss the outer spills) are
// dropped.
// Positions outside [0, n) stand for their periodic images (every (cell, offset) pair of the
// periodic stencil is counted once, also for lines shorter than the stencil).
//
// alpha-bar is a reduction over a whole line, so its correction to the output (at the argmax
// cell) is known only when the launch has finished.  Each wave writes its partial sum to
// partial_out[c] (and each block the sum of its four, in order, to bpartial_out[b]) and, if it
// holds the argmax, atomicMax-es the key (tag << 32 | ~cell) into
// rec_out[line].key (tags grow with every launch, so no slot is ever reset and the smallest cell
// wins).  The launch that next reads this output (as lam, x1 or x2) gets it as a Pending and
// adds the correction  cf*h * sum(partials) * sign(s_prev[cell])  to the value it loads at
// that cell, in registers.  A launch reading it as lam sums the partials (every wave that needs
// the sum, in the same order) and records the correction in rec[line].resolved; a launch
// reading it as x1 or x2 loads that (recorded by an earlier launch, or by
// adjoint_resolve_kernel when none has read the buffer as lam).
// The host keeps four slots of these arrays and the launches write them in turn.
// No wave ever waits for another, and no memory fence is needed inside a launch.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels_1d_fast.hpp"

namespace hrw {

// Wavefront sum and minimum by DPP row shifts and row broadcasts (no LDS round trips), in one
// fixed order: after the row shifts by 1, 2, 4, 8 each lane holds its row's prefix, lane 15 of
// a row the row's total; row_bcast:15 and row_bcast:31 carry rows 0-2 up so that lane 63 holds
// the wavefront's, which every lane gets.  Every lane of the wavefront must be active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64_or_zero(double v)   // no source: 0
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xF, true);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_f64_or_zero<0x111, 0xF>(v);
    v += dpp_f64_or_zero<0x112, 0xF>(v);
    v += dpp_f64_or_zero<0x114, 0xF>(v);
    v += dpp_f64_or_zero<0x118, 0xF>(v);
    v += dpp_f64_or_zero<0x142, 0xA>(v);
    v += dpp_f64_or_zero<0x143, 0xC>(v);
    const long long b = __builtin_bit_cast(long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __builtin_bit_cast(double, ((long long)hi << 32) | lo);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned dpp_u32_or_max(unsigned v)   // no source: 0xffffffff
{
    return (unsigned)__builtin_amdgcn_update_dpp((int)0xffffffffu, (int)v, CTRL, ROW_MASK, 0xF, false);
}
__device__ __forceinline__ unsigned wave_umin(unsigned v)
{
    v = min(v, dpp_u32_or_max<0x111, 0xF>(v));
    v = min(v, dpp_u32_or_max<0x112, 0xF>(v));
    v = min(v, dpp_u32_or_max<0x114, 0xF>(v));
    v = min(v, dpp_u32_or_max<0x118, 0xF>(v));
    v = min(v, dpp_u32_or_max<0x142, 0xA>(v));
    v = min(v, dpp_u32_or_max<0x143, 0xC>(v));
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// A launch's record per line: its argmax key, and the correction it owes at that cell once a
// later launch has summed it.
struct LineRec {
    unsigned long long key;
    double resolved;
};
static_assert(sizeof(LineRec) == 16, "one 16-byte record per line");

// The not yet applied alpha correction of an earlier launch's output.  A tag of kNoPendTag
// (which no launch uses) stands for none owed.
constexpr unsigned kNoPendTag = 0xffffffffu;
struct Pending {
    const double *partial;              // that launch's per-chunk partial sums (null: none)
    const double *bpartial;             // ... and per-block sums of its four chunks
    LineRec *rec;                       // its per-line records
    const double *s;                    // its stage input (for the sign at the argmax)
    double cf;                          // its c (the correction is cf * h[line] * sum)
    unsigned tag;
    int ready;                          // rec[].resolved was written by an earlier launch
    int slot;                           // the slot it wrote
};

// One wavefront's work in one 64-byte record, so that a single scalar load gives the chunk
// and its line's parameters (instead of a load of the line index and then of each parameter).
struct AdjChunk {
    long long off;              // the line's first cell in the flat state
    int line, start, n, k;      // line (-1: an idle padding chunk), first output cell, n, WENO k
    int first, nc;              // the line's chunks [first, first + nc)
    double h, xmin, rx;
    double constw;              // 1: every edge xmin + rx*i is exact, so every width is rx
};
static_assert(sizeof(AdjChunk) == 64, "one 64-byte record per chunk");

struct AdjArgs {
    const double *s;            // stage input state (flat, line l at chunk.off)
    const double *lam;          // cotangent of the stage output
    const double *x1, *x2;      // pointwise cotangents folded into out (or null)
    double *out;                // may alias x1 or x2 (read and written by the same lane)
    const AdjChunk *chunks;     // the chunk table
    long long nchunks;
    double a, c, e1, e2, eps, alpha;
    const unsigned long long *aset;   // alpha partials: line l, partial p at
    int aline_stride, anpart, apart_stride;   //   aset[l*aline_stride + p*apart_stride]
    double *partial_out;        // this launch: one partial alpha-bar per chunk
    double *bpartial_out;       // this launch: their sums per block of four chunks
    LineRec *rec_out;           // this launch: its records (argmax key per line)
    unsigned tag;
    // the corrections owed by lam, x1 and x2: the slot each was written to (bits 4i, 4i+1; the
    // records of slot q at recs[q * nlines]) and its tag (kNoPendTag: none owed)
    const LineRec *recs;
    int nlines;
    unsigned pend_slots;
    unsigned pend_tag[3];
    Pending lam_pend;           // lam's, to sum it here
};

// The cell a pending correction applies to on `line` (wave-uniform), or -1 if none.
__device__ __forceinline__ int pending_cell(const Pending &p, int line)
{
    if (!p.partial) return -1;
    const unsigned long long key = p.rec[line].key;
    if ((unsigned)(key >> 32) != p.tag) return -1;
    return (int)(0xffffffffu - (unsigned)(key & 0xffffffffull));
}
// The line's alpha-bar from an earlier launch: the whole blocks of four chunks [b0, b1) by
// their sums, and the chunks outside them ([first, 4 b0) and [4 b1, end), or all of the line's
// chunks when it holds no whole block) one by one.  Lane l adds the block sums l, l + 64,
// l + 128, ... in that order, the wavefront adds the 64 lane sums (wave_sum), and then the
// single chunks (at most six, one per lane) the same way.  That order does not depend on U,
// the loads a lane has in flight at once, so every wave that needs the sum gets the same bits.
// Loads past b1 stay inside the block-sum array (it holds kBlockPad more) and are not counted.
constexpr int kBlockPad = 64 * 40;
template <int U>
__device__ __forceinline__ double line_alpha_bar(const Pending &p, int first, int nc)
{
    static_assert(64 * U <= kBlockPad, "block-sum array padding");
    const int lane = threadIdx.x & 63;
    const int end = first + nc;
    const int b0 = (first + 3) >> 2, b1 = end >> 2;
    const int nb = b1 > b0 ? b1 - b0 : 0;
    const int e0 = nb ? 4 * b0 : end, e1 = nb ? 4 * b1 : end;   // singles [first, e0), [e1, end)
    const int hh = e0 - first, tt = end - e1;
    double single = 0.0;
    if (lane < hh) single = p.partial[first + lane];
    else if (lane < hh + tt) single = p.partial[e1 + (lane - hh)];
    const double *q = p.bpartial + b0;
    double acc = 0.0;
    for (int base = 0; base < nb; base += 64 * U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = q[base + 64 * u + lane];
#pragma unroll
        for (int u = 0; u < U; ++u) acc += base + 64 * u + lane < nb ? v[u] : 0.0;
    }
    return wave_sum(acc) + wave_sum(single);
}

// The correction owed at cell j of a line (whole wave), with up to UMAX block sums per lane in
// flight: about nc/4 block sums, so one round of loads for lines of up to 16 UMAX blocks of four
// chunks (1.27M cells at UMAX = 40 and 2 cells per lane).
template <int UMAX>
__device__ __forceinline__ double pending_delta(const Pending &p, int j, long long off, double h,
                                                int first, int nc)
{
    const double sj = p.s[off + j];
    const double sum = (nc >> 2) <= 512 ? line_alpha_bar<8>(p, first, nc)
                                        : line_alpha_bar<UMAX>(p, first, nc);
    return p.cf * h * sum * (double)((sj > 0.0) - (sj < 0.0));
}

// A pending correction's cell on a line (-1: none) and its value if an earlier launch resolved
// it.  Lanes 0, 1, 2 load the records of lam's, x1's and x2's corrections (one vector load,
// issued before the chunk's own loads, so that waiting for it never waits for those), and
// each is taken from its lane where it is used.
struct PendAt {
    int j;
    double d;
};
__device__ __forceinline__ LineRec pend_recs(const AdjArgs &a, int line)
{
    const int i = min((int)(threadIdx.x & 63), 2);
    const int slot = (int)(a.pend_slots >> (4 * i)) & 3;
    return a.recs[(long long)slot * a.nlines + line];
}
__device__ __forceinline__ PendAt pend_at(unsigned tag, const LineRec &r, int i)
{
    const unsigned klo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)r.key, i);
    const unsigned khi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(r.key >> 32), i);
    const long long rb = __builtin_bit_cast(long long, r.resolved);
    const unsigned rlo = (unsigned)__builtin_amdgcn_readlane((int)rb, i);
    const unsigned rhi = (unsigned)__builtin_amdgcn_readlane((int)(rb >> 32), i);
    const double d = __builtin_bit_cast(double, ((unsigned long long)rhi << 32) | rlo);
    return PendAt{khi == tag ? (int)(0xffffffffu - klo) : -1, d};
}

// Reverse of one K = 3 reconstruction of cell i from its window x[0..4] = v_{i-2..i+2}.  The
// cotangents of the two faces come from face_cot(vr, vl, ir, il, sgr, sgl), which sets
// sgr = vr-bar * ir and sgl = vl-bar * il (ir, il: the reciprocals of the two sums of weights,
// sr = 6 sum_r (d_r Q_r), vr = nr/sr).  The triple pieces are rebuilt from the window for each
// cell (cheaper in registers than carrying them between cells).  The window cotangent is added
// to acc[0..4].
template <class FaceCot>
__device__ __forceinline__ void weno5_vjp(double eps12, const double *x, FaceCot &&face_cot,
                                          double *acc)
{
    // T(i-2) = x0..x2 (beta_2, right cand P2, left cand P1), T(i-1) = x1..x3 (beta_1, P1, P0),
    // T(i) = x2..x4 (beta_0, P0, Pm); with D = a - 2b + c, E = a - c:
    //   12 beta_0 = 13 D^2 + 3 (2D + E)^2, 12 beta_1 = 13 D^2 + 3 E^2, 12 beta_2 = 13 D^2 + 3 (2D - E)^2
    const double D2 = fma_(-2.0, x[1], x[0]) + x[2], E2 = x[0] - x[2];
    const double D1 = fma_(-2.0, x[2], x[1]) + x[3], E1 = x[1] - x[3];
    const double D0 = fma_(-2.0, x[3], x[2]) + x[4], E0 = x[2] - x[4];
    const double Y2 = fma_(2.0, D2, -E2), X0 = fma_(2.0, D0, E0);
    // (13 D and 3 Y, 3 E, 3 X are kept for the reverse: 26 D pb = 13 D (2 pb) and 6 Y pb =
    // 3 Y (2 pb) exactly, with 2 pb formed directly below)
    const double t2 = 13.0 * D2, t1 = 13.0 * D1, t0 = 13.0 * D0;
    const double y2 = 3.0 * Y2, e1 = 3.0 * E1, x0 = 3.0 * X0;
    const double p2 = eps12 + fma_(y2, Y2, t2 * D2);
    const double p1 = eps12 + fma_(e1, E1, t1 * D1);
    const double p0 = eps12 + fma_(x0, X0, t0 * D0);
    // 6x candidates: right P0(T(i)), P1(T(i-1)), P2(T(i-2)); left Pm(T(i)), P0(T(i-1)), P1(T(i-2))
    const double cr0 = fma_(5.0, x[3], fma_(2.0, x[2], -x[4]));
    const double cr1 = fma_(5.0, x[2], fma_(2.0, x[3], -x[1]));
    const double cr2 = fma_(11.0, x[2], fma_(-7.0, x[1], x[0] + x[0]));
    const double cl0 = fma_(11.0, x[2], fma_(-7.0, x[3], x[4] + x[4]));
    const double cl1 = fma_(5.0, x[2], fma_(2.0, x[1], -x[3]));
    const double cl2 = fma_(5.0, x[1], fma_(2.0, x[2], -x[0]));
    const double Q0 = p0 * p0, Q1 = p1 * p1, Q2 = p2 * p2;
    const double m0 = Q1 * Q2, m1 = Q0 * Q2, m2 = Q0 * Q1;
    const double a0 = 3.0 * m0, a1 = 6.0 * m1, l2 = 3.0 * m2;
    const double sr = 6.0 * (a0 + a1 + m2), sl = 6.0 * (m0 + a1 + l2);
    // one reciprocal for both faces (sr and sl lie within a factor 3 of each other; see recon3)
    const double i2 = rcp_nr(sr * sl);
    const double ir = sl * i2, il = sr * i2;
    const double vr = fma_(a0, cr0, fma_(a1, cr1, m2 * cr2)) * ir;
    const double vl = fma_(m0, cl0, fma_(a1, cl1, l2 * cl2)) * il;
    double sgr, sgl;
    face_cot(vr, vl, ir, il, sgr, sgl);
    const double vr6 = 6.0 * vr, vl6 = 6.0 * vl;
    // twice the cotangent of p_r = 12 (eps + beta_r):  -4/p_r [ w-numerator_r sgr (cr_r - 6vr) + ... ]
    const double ipi = -4.0 * rcp_nr(p0 * p1 * p2);
    const double pb0 = (p1 * p2 * ipi) * fma_(a0 * sgr, cr0 - vr6, (m0 * sgl) * (cl0 - vl6));
    const double pb1 = (p0 * p2 * ipi) * fma_(a1 * sgr, cr1 - vr6, (a1 * sgl) * (cl1 - vl6));
    const double pb2 = (p0 * p1 * ipi) * fma_(m2 * sgr, cr2 - vr6, (l2 * sgl) * (cl2 - vl6));
    // T(i-2): beta_2 (D, Y = 2D - E), candidates P2 (right, m2) and P1 (left, l2)
    {
        const double Db = t2 * pb2, Yb = y2 * pb2;
        const double Pr = m2 * sgr, Pl = l2 * sgl;
        acc[0] += fma_(2.0, Pr, Db + Yb - Pl);
        acc[1] += fma_(-7.0, Pr, fma_(5.0, Pl, fma_(-4.0, Yb, -2.0 * Db)));
        acc[2] += fma_(11.0, Pr, fma_(2.0, Pl, fma_(3.0, Yb, Db)));
    }
    // T(i-1): beta_1 (D, E), candidates P1 (right, a1) and P0 (left, a1)
    {
        const double Db = t1 * pb1, Eb = e1 * pb1;
        const double Pr = a1 * sgr, Pl = a1 * sgl;
        acc[1] += fma_(2.0, Pl, Db + Eb - Pr);
        acc[2] += fma_(5.0, Pr + Pl, -2.0 * Db);
        acc[3] += fma_(2.0, Pr, Db - Eb - Pl);
    }
    // T(i): beta_0 (D, X = 2D + E), candidates P0 (right, a0) and Pm (left, m0)
    {
        const double Db = t0 * pb0, Xb = x0 * pb0;
        const double Pr = a0 * sgr, Pl = m0 * sgl;
        acc[2] += fma_(11.0, Pl, fma_(2.0, Pr, fma_(3.0, Xb, Db)));
        acc[3] += fma_(-7.0, Pl, fma_(5.0, Pr, fma_(-4.0, Xb, -2.0 * Db)));
        acc[4] += fma_(2.0, Pl, Db + Xb - Pr);
    }
}

// The Lax-Friedrichs form, for the halved flux cotangents phi3 = (phi_{i-1}, phi_i, phi_{i+1}),
// phi = lam/(2w) (the 1/2 of the flux folded in: an exact scaling): vr-bar = (phi_{i+1} - phi_i)
// (vr + alpha), vl-bar = (phi_i - phi_{i-1}) (vl - alpha).  Returns the cell's share of alpha-bar.
__device__ __forceinline__ double recon3_vjp(double eps12, const double *x, const double *phi3,
                                             double alpha, double *acc)
{
    double abar;
    weno5_vjp(eps12, x,
              [&](double vr, double vl, double ir, double il, double &sgr, double &sgl) {
                  const double Gp = phi3[2] - phi3[1], Gm = phi3[1] - phi3[0];
                  sgr = (Gp * (vr + alpha)) * ir;
                  sgl = (Gm * (vl - alpha)) * il;
                  abar = fma_(Gp, vr, -(Gm * vl));
              },
              acc);
    return abar;
}

// Reverse of one K = 2 reconstruction from the pairs (v_{i-1}, v_i) and (v_i, v_{i+1});
// acc[0..2] = cells i-1..i+1.  The cotangents of the faces come from face_cot as for weno5_vjp.
template <class FaceCot>
__device__ __forceinline__ void weno3_vjp(double eps, const Pair2 &p1, const Pair2 &p0,
                                          FaceCot &&face_cot, double *acc)
{
    const double q0 = eps + p0.B, q1 = eps + p1.B;
    const double Q0 = q0 * q0, Q1 = q1 * q1;
    const double a0 = 2.0 * Q1, b1 = 2.0 * Q0;
    const double sr = 2.0 * (a0 + Q0), sl = 2.0 * (Q1 + b1);
    const double nr = fma_(a0, p0.S, Q0 * p1.Rr);
    const double nl = fma_(Q1, p0.Lr, b1 * p1.S);
    const double ir = rcp_nr(sr), il = rcp_nr(sl);
    const double vr = nr * ir;
    const double vl = nl * il;
    double sgr, sgl;
    face_cot(vr, vl, ir, il, sgr, sgl);
    const double vr2 = 2.0 * vr, vl2 = 2.0 * vl;
    const double iq = rcp_nr(q0 * q1);
    // q-bar_r = -2/q_r [ (2x right weight numerator) sgr (cand - 2vr) + ... ]
    const double qb0 = -2.0 * q1 * iq * fma_(a0 * sgr, p0.S - vr2, (Q1 * sgl) * (p0.Lr - vl2));
    const double qb1 = -2.0 * q0 * iq * fma_(Q0 * sgr, p1.Rr - vr2, (b1 * sgl) * (p1.S - vl2));
    // beta = d^2 with d = b - a = (Rr - S)/2, so d-bar = 2 d q-bar moves -1 to a, +1 to b
    // pair (v_i, v_{i+1}): S = a+b, Lr = 3a-b
    {
        const double Sb = a0 * sgr, Lb = Q1 * sgl;
        const double dd = (p0.Rr - p0.S) * qb0;
        acc[1] += fma_(3.0, Lb, Sb - dd);
        acc[2] += Sb - Lb + dd;
    }
    // pair (v_{i-1}, v_i): Rr = 3b-a, S = a+b
    {
        const double Rb = Q0 * sgr, Sb = b1 * sgl;
        const double dd = (p1.Rr - p1.S) * qb1;
        acc[0] += Sb - Rb - dd;
        acc[1] += fma_(3.0, Rb, Sb + dd);
    }
}

// The Lax-Friedrichs form (halved phi3 as for recon3_vjp); abar = the cell's share of alpha-bar.
__device__ __forceinline__ void recon2_vjp(double eps, const Pair2 &p1, const Pair2 &p0,
                                           double *phi3, double alpha, double *acc, double &abar)
{
    weno3_vjp(eps, p1, p0,
              [&](double vr, double vl, double ir, double il, double &sgr, double &sgl) {
                  const double Gp = phi3[2] - phi3[1], Gm = phi3[1] - phi3[0];
                  const double gr = Gp * (vr + alpha), gl = Gm * (vl - alpha);
                  abar = fma_(Gp, vr, -(Gm * vl));
                  sgr = gr * ir;
                  sgl = gl * il;
              },
              acc);
}

// One wavefront's chunk of a line (see the header comment).  Returns the lane's share of
// alpha-bar over its owned cells; *cand = the smallest owned cell with |us| == alpha.
template <int K, int R, bool ALPHAMAX, bool INTERIOR>
__device__ __forceinline__ double chunk_adjoint(const AdjArgs &a, const double *s,
                                                const double *lam, const double *x1,
                                                const double *x2, double *out, int n, int W,
                                                double h, double xmin, double rx,
                                                bool constw, unsigned long long alpha_part,
                                                LineRec prec, int line, long long off,
                                                int first, int nc, unsigned *cand)
{
    constexpr int H = K - 1;
    constexpr int NV = R + 2 * H;
    static_assert(R >= 2, "two or more cells per lane");
    const int lane = threadIdx.x & 63;
    const long long c0 = (long long)W + (long long)(lane - 1) * R;
    const bool owned = lane >= 1 && lane <= 62;
    constexpr bool interior = INTERIOR;

    // ---- every load of the chunk first, so that their latencies overlap: the lane's cells of
    // s and lam, the window halo and the lam just outside the two wave-edge lanes (both read
    // branch-free: lanes 0..31 at lane 0's positions, lanes 32..63 at lane 63's), x1 and x2
    double v[NV], mu[R], phi[R + 2], acc[NV];
    double p1v[R], p2v[R];
    [[maybe_unused]] double hv_in[H > 0 ? H : 1];
    double lam_edge;
    if constexpr (interior) {
        load_run<R>(s + c0, v + H);
        if constexpr (H > 0) {
            const double *hp = s + (lane < 32 ? (long long)W - R - H : (long long)W + 63LL * R);
#pragma unroll
            for (int i = 0; i < H; ++i) hv_in[i] = hp[i];
        }
        load_run<R>(lam + c0, mu);
        // (in range for K >= 2; for K = 1 the value is not used and the index is clamped)
        lam_edge = lam[lane < 32 ? max((long long)W - R - 1, 0LL) : min((long long)W + 63LL * R, (long long)n - 1)];
        if (x1) load_run<R>(x1 + c0, p1v);
        if (x2) load_run<R>(x2 + c0, p2v);
#pragma unroll
        for (int i = 0; i < H; ++i) {
            v[i] = shfl_up1(v[R + i]);
            v[H + R + i] = shfl_down1(v[H + i]);
        }
        // the wave-edge lanes have no neighbour to shuffle from
        if constexpr (H > 0) {
            if (lane == 0) {
#pragma unroll
                for (int i = 0; i < H; ++i) v[i] = hv_in[i];
            }
            if (lane == 63) {
#pragma unroll
                for (int i = 0; i < H; ++i) v[H + R + i] = hv_in[i];
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] = s[cell_index(c0 - H + i, n, true)];
#pragma unroll
        for (int i = 0; i < R; ++i) mu[i] = lam[cell_index(c0 + i, n, true)];
        lam_edge = lam[cell_index(lane < 32 ? c0 - 1 : c0 + R, n, true)];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const bool in = owned && c0 + i < n;
            p1v[i] = (x1 && in) ? x1[c0 + i] : 0.0;
            p2v[i] = (x2 && in) ? x2[c0 + i] : 0.0;
        }
    }
    // alpha: the partial(s) were loaded before these loads, so this waits for them only
    const double alpha = ALPHAMAX ? __longlong_as_double((long long)wave_umax(alpha_part)) : a.alpha;
    long long cell[R];
#pragma unroll
    for (int i = 0; i < R; ++i) cell[i] = interior ? c0 + i : cell_index(c0 + i, n, true);
    // lanes 0 and 63 also form the phi just outside their cells themselves: with R < 2K-1
    // their reconstructions reach the owned cells of lanes 1 and 62
    const bool edge = lane == 0 || lane == 63;
    const long long hc = interior ? (lane == 0 ? c0 - 1 : c0 + R) : cell_index(lane == 0 ? c0 - 1 : c0 + R, n, true);
    // the pending corrections: lam's at any loaded cell, x1's and x2's (resolved by an earlier
    // launch) at the owned cells; only the alpha-max form owes any
    const PendAt none{-1, 0.0};
    const PendAt q0 = ALPHAMAX ? pend_at(a.pend_tag[0], prec, 0) : none;
    const PendAt q1 = ALPHAMAX ? pend_at(a.pend_tag[1], prec, 1) : none;
    const PendAt q2 = ALPHAMAX ? pend_at(a.pend_tag[2], prec, 2) : none;
    bool h0 = edge && hc == q0.j, h1 = false, h2 = false;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        h0 = h0 || cell[i] == q0.j;
        h1 = h1 || (owned && cell[i] == q1.j);
        h2 = h2 || (owned && cell[i] == q2.j);
    }
    const bool hit0 = ALPHAMAX && q0.j >= 0 && __any(h0);
    const bool hit1 = ALPHAMAX && q1.j >= 0 && __any(h1);
    const bool hit2 = ALPHAMAX && q2.j >= 0 && __any(h2);
    if (hit0) {
        // summed here (every wave that sums it writes the same bits) unless resolved earlier
        double d = q0.d;
        if (!a.lam_pend.ready) {
            // lam_pend read from the kernel arguments here, where it is needed, rather than
            // held in registers from the kernel's start
            const auto *kp = (const __attribute__((address_space(4))) AdjArgs *)__builtin_amdgcn_kernarg_segment_ptr();
            Pending lp;
            lp.partial = kp->lam_pend.partial;
            lp.bpartial = kp->lam_pend.bpartial;
            lp.rec = kp->lam_pend.rec;
            lp.s = kp->lam_pend.s;
            lp.cf = kp->lam_pend.cf;
            // (fewer block sums in flight with 4 cells per lane, which need more registers)
            d = pending_delta<R <= 2 ? 40 : 16>(lp, q0.j, off, h, first, nc);
            if (lane == 0) lp.rec[line].resolved = d;
        }
#pragma unroll
        for (int i = 0; i < R; ++i)
            if (cell[i] == q0.j) mu[i] += d;
        if (edge && hc == q0.j) lam_edge += d;
    }
    // phi = lam / (2 width) for the lane's cells, neighbours' from the adjacent lanes; on a grid
    // whose edges are all exact every width is rx, and 1/(2rx) is formed once (the same value)
    const double iw_const = constw ? 0.5 * rcp_nr(rx) : 0.0;
    auto inv_w = [&](long long g) {
        if (constw) return iw_const;
        const double w = linear_edge(xmin, rx, g + 1) - linear_edge(xmin, rx, g);   // src/hrweno_grids.f90:76-78,247
        return 0.5 * rcp_nr(w);
    };
#pragma unroll
    for (int i = 0; i < R; ++i) phi[i + 1] = mu[i] * inv_w(cell[i]);
    phi[0] = shfl_up1(phi[R]);
    phi[R + 1] = shfl_down1(phi[1]);
    if (edge) {
        const double ph = lam_edge * inv_w(hc);
        if (lane == 0) phi[0] = ph;
        else phi[R + 1] = ph;
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;

    double abar = 0.0;
    unsigned cnd = 0xffffffffu;
    if constexpr (K == 3) {
        const double eps12 = 12.0 * a.eps;
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const double ab = recon3_vjp(eps12, v + t, phi + t, alpha, acc + t);
            if (owned && c0 + t < n) {
                abar += ab;
                if (ALPHAMAX && fabs(v[t + 2]) == alpha && cnd == 0xffffffffu) cnd = (unsigned)(c0 + t);
            }
            __builtin_amdgcn_sched_barrier(0);   // one cell at a time: bounds the live values
        }
    } else if constexpr (K == 2) {
        Pair2 p1 = pair2(v[0], v[1]);
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const Pair2 p0 = pair2(v[t + 1], v[t + 2]);
            double ab;
            recon2_vjp(a.eps, p1, p0, phi + t, alpha, acc + t, ab);
            if (owned && c0 + t < n) {
                abar += ab;
                if (ALPHAMAX && fabs(v[t + 1]) == alpha && cnd == 0xffffffffu) cnd = (unsigned)(c0 + t);
            }
            p1 = p0;
        }
    } else {
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const double vv = v[t];
            const double Gp = phi[t + 2] - phi[t + 1], Gm = phi[t + 1] - phi[t];
            acc[t] += Gp * (vv + alpha) + Gm * (vv - alpha);
            if (owned && c0 + t < n) {
                abar += (Gp - Gm) * vv;
                if (ALPHAMAX && fabs(vv) == alpha && cnd == 0xffffffffu) cnd = (unsigned)(c0 + t);
            }
        }
    }
    // hand the spill-over cells to the neighbours: acc[j] <-> cell c0 - H + j
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const double fromR = shfl_down1(acc[i]);           // right lane's cells c0+R+i-H...
        const double fromL = shfl_up1(acc[R + H + i]);
        acc[R + i] += fromR;
        acc[H + i] += fromL;
    }
    // out = a*lam + c*h*acc + e1*x1 + e2*x2 on the owned cells (x1, x2 with their own
    // pending corrections)
    if (hit1) {
#pragma unroll
        for (int i = 0; i < R; ++i)
            if (owned && cell[i] == q1.j) p1v[i] += q1.d;
    }
    if (hit2) {
#pragma unroll
        for (int i = 0; i < R; ++i)
            if (owned && cell[i] == q2.j) p2v[i] += q2.d;
    }
    if (owned) {
        const double ch = a.c * h;
        double o[R];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            double t = fma_(ch, acc[H + i], a.a * mu[i]);
            if (x1) t = fma_(a.e1, p1v[i], t);
            if (x2) t = fma_(a.e2, p2v[i], t);
            o[i] = t;
        }
        if (interior) {
            store_run<R>(out + c0, o);
        } else {
#pragma unroll
            for (int i = 0; i < R; ++i)
                if (c0 + i < n) out[c0 + i] = o[i];
        }
    }
    *cand = cnd;
    return abar;
}

// Occupancy targets that keep the K = 3 reverse free of spills: at R = 2 it fits 168 VGPRs
// (3 waves per SIMD), at R = 4 it needs ~250 (2 waves).
template <int R, bool ALPHAMAX>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(R == 2 ? 3 : 2)))
void adjoint_stage_kernel(AdjArgs a)
{
    // wave-uniform: the chunk table and the line's parameters come through scalar loads
    const long long wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)));
    const int lane = threadIdx.x & 63;
    __shared__ double s_part[kBlock / 64];
    const AdjChunk &ch = a.chunks[wave < a.nchunks ? wave : 0];
    if (wave >= a.nchunks || ch.line < 0) {   // idle chunks only join the block sum
        if constexpr (ALPHAMAX) {
            if (lane == 0) s_part[threadIdx.x >> 6] = 0.0;
            __syncthreads();
        }
        return;
    }
    const int line = ch.line;
    const long long off = ch.off;
    const int n = ch.n;
    const double h = ch.h;
    // the alpha partials (reduced inside the chunk, after its loads are issued)
    unsigned long long apart = 0ull;
    if constexpr (ALPHAMAX) {
        const unsigned long long *set = a.aset + (long long)line * a.aline_stride;
        // lanes past the partials repeat the last one (the maximum is unchanged)
        apart = set[min(lane, a.anpart - 1) * a.apart_stride];
    }
    // the corrections owed by lam, x1 and x2 (no x1, x2: none owed)
    LineRec prec{0ull, 0.0};
    if constexpr (ALPHAMAX) prec = pend_recs(a, line);
    const double *x1 = a.x1 ? a.x1 + off : nullptr;
    const double *x2 = a.x2 ? a.x2 + off : nullptr;
    const int first = ch.first, nc = ch.nc;
    const int W = ch.start;
    const int K = ch.k;
    unsigned cand;
    double ab;
#define HRW_ADJ_CHUNK(KK, IN)                                                                    \
    chunk_adjoint<KK, R, ALPHAMAX, IN>(a, a.s + off, a.lam + off, x1, x2, a.out + off, n, W, h,  \
                                        ch.xmin, ch.rx, ch.constw != 0.0, apart, prec, line,     \
                                        off, first, nc,                                          \
                                        &cand)
    if ((long long)W - R - (K - 1) >= 0 && (long long)W + 63LL * R + (K - 1) <= n) {
        if (K == 1) ab = HRW_ADJ_CHUNK(1, true);
        else if (K == 2) ab = HRW_ADJ_CHUNK(2, true);
        else ab = HRW_ADJ_CHUNK(3, true);
    } else {
        if (K == 1) ab = HRW_ADJ_CHUNK(1, false);
        else if (K == 2) ab = HRW_ADJ_CHUNK(2, false);
        else ab = HRW_ADJ_CHUNK(3, false);
    }
#undef HRW_ADJ_CHUNK
    if constexpr (ALPHAMAX) {
        const double wsum = wave_sum(ab);
        const unsigned wmin = wave_umin(cand);
        if (lane == 0) {
            a.partial_out[wave] = wsum;
            s_part[threadIdx.x >> 6] = wsum;
            if (wmin != 0xffffffffu)
                atomicMax(&a.rec_out[line].key,
                          ((unsigned long long)a.tag << 32) | (unsigned long long)(0xffffffffu - wmin));
        }
        __syncthreads();
        // used only for blocks whose four chunks lie on one line
        if (threadIdx.x == 0) a.bpartial_out[blockIdx.x] = ((s_part[0] + s_part[1]) + s_part[2]) + s_part[3];
    }
}

// Record a pending correction per line (one wave per line), for a launch that reads the buffer
// as x1 or x2 before any launch has read it as lam.
static __global__ __launch_bounds__(64) void adjoint_resolve_kernel(Pending p, const long long *off,
                                                                   const double *h,
                                                                   const int *line_chunk0,
                                                                   const int *line_nchunks,
                                                                   int nlines)
{
    const int line = blockIdx.x;
    if (line >= nlines) return;
    const int j = pending_cell(p, line);
    if (j < 0) return;
    const double d = pending_delta<40>(p, j, off[line], h[line], line_chunk0[line], line_nchunks[line]);
    if ((threadIdx.x & 63) == 0) p.rec[line].resolved = d;
}

// Apply a pending correction to a final output: one wave per line.
static __global__ __launch_bounds__(64) void adjoint_fixup_kernel(Pending p, double *out,
                                                           const long long *off, const double *h,
                                                           const int *line_chunk0,
                                                           const int *line_nchunks, int nlines)
{
    const int line = blockIdx.x;
    if (line >= nlines) return;
    const int j = pending_cell(p, line);
    if (j < 0) return;
    const double d = p.ready ? p.rec[line].resolved
                             : pending_delta<40>(p, j, off[line], h[line], line_chunk0[line], line_nchunks[line]);
    if ((threadIdx.x & 63) == 0) out[off[line] + j] += d;
}

}  // namespace hrw
