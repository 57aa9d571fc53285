// kernels_adjoint_general.hpp -- reverse (vector-Jacobian) pass of the 1D Burgers right-hand
// side for every composition the forward runs on a linear grid: the Lax-Friedrichs flux
// (alpha fixed, or alpha = max|s|) or the endpoint Godunov flux, with periodic or zero-gradient
// ends (example/example1_burgers_1d_fv.f90:72-109; src/hrweno_weno.f90:174-216 per cell;
// src/hrweno_fluxes.f90:42-43, 66-74).  One launch computes
//     out = sum_i e_i x_i + ch * J(s)^T (b1 y1 + b2 y2)
// so a stage reverse of either integrator (rktvd: one cotangent and up to three pointwise
// terms; mstvd: the cotangents of the next state and of the state four steps on) is one pass.
//
// Per cell p (phi = (b1 y1 + b2 y2)/w, the cotangent of F_{p+1/2} is phi_{p+1} - phi_p):
//   gr_p = (phi_{p+1} - phi_p) dF/dvm (vr_p, vl_{p+1}),  gl_p = (phi_p - phi_{p-1}) dF/dvp (vr_{p-1}, vl_p)
// then through the nonlinear weights, d vr / d beta_r = -2 w_r (q_r - vr) / (eps + beta_r), and
// through beta_r (quadratic) and the candidates q_r, l_r (linear) to the 2K-1 window cells.
// With zero-gradient ends the reference copies the boundary fluxes inward (example1:103-104),
// so the end cells' RHS is identically zero (phi_0 = phi_{n-1} = 0) and only the n-1 interior
// interfaces exist; window positions past an end are the end cell (constant extrapolation,
// src/hrweno_weno.f90:171-173), so their cotangents land on it.
//
// Layout: block b takes the output cells [Tb, Tb + T) (T = kGenTile).  It loads the stage input over
// the window halo into LDS, reconstructs the T + 2K positions whose faces its interfaces need,
// forms the window cotangent of the T + 2K - 2 positions whose windows reach its cells, and
// gathers each output cell's share in a fixed order (so the result does not depend on timing).
// alpha-bar (Lax-Friedrichs, alpha = max|s|) is a sum over the whole line that lands on the
// first cell attaining the maximum: each block writes its partial sum and atomicMin-s the first
// such cell it holds; rhs_vjp_alpha_fixup adds the correction once the pass has finished.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>

#include "kernels_1d.hpp"
#include "kernels_1d_fast.hpp"
#include "weno_math.hpp"

namespace hrw {

// Output cells per block (one thread each).  The block also reconstructs the 2K positions of
// its window halo, which takes one of its waves a second round: a quarter of the work on top
// with 4 waves, an eighth with 8.  The pass is bound by the latency of its barrier-separated
// phases, and 4-wave blocks (five resident per CU at 92-94 VGPRs, against two of 8 waves)
// overlap one block's waits with another's work: 4 % faster at 2^20 cells than 8-wave blocks
// despite the halo, and 2-wave blocks lose it again to the halo (11 % slower).
constexpr int kGenTile = 256;

struct GenAdjArgs {
    const double *s;                 // the stage input the Jacobian is taken at
    const double *y1, *y2;           // J^T acts on b1*y1 + b2*y2 (y2 may be null)
    double b1, b2;
    const double *x[4];              // pointwise terms e[i]*x[i] (null: absent)
    double e[4];
    double *out;                     // may alias an x[i] (read and written by one thread)
    double ch;                       // the scale of J^T (the stage's c*h)
    long long n;
    double xmin, rx, eps, alpha;     // linear grid; alpha: Lax-Friedrichs with a fixed alpha
    unsigned long long *aset;        // alpha = max|s|: a slot set holding max|s| (the fixup clears it)
    double *partial;                 // alpha = max|s|: one alpha-bar partial sum per block
    int *argmax;                     // alpha = max|s|: first cell with |s| == alpha (INT_MAX: none)
};

// One cell's reconstruction with what its reverse needs: candidates q_r (right face), l_r (left
// face), weights w_r, wt_r and 1/(eps + beta_r).
template <int K>
struct GCell {
    double vr, vl, q[K], l[K], w[K], wt[K], ib[K];
};

// x[0 .. 2K-2] = v_{i-K+1 .. i+K-1}.  The weights over a common denominator,
//   w_r = d_r Q_r / sum_s d_s Q_s,  Q_r = prod_{s != r} q_s,  q_s = (eps + beta_s)^2,
// with every q scaled by the largest first (the products stay in (0, 1]): 6 reciprocals per
// cell (hardware estimate + two Newton steps, ~1 ulp; every operand is at least eps or d_min
// times such a product, far from zero) instead of the reference order's 15 divides (the
// reverse needs the derivatives at the point to ~1e-15, not the reference's rounding).
template <int K>
__device__ __forceinline__ void gcell_fwd(const double *x, double eps, GCell<K> &c)
{
    const double *v = x + (K - 1);
    double beta[K], e[K], q[K];
    beta_exact<K>(x, beta);
    double qm = 0.0;
#pragma unroll
    for (int r = 0; r < K; ++r) {
        double qr = 0.0, lr = 0.0;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            qr = fma_(Tab<K>::c[r + 1][j], v[-r + j], qr);
            lr = fma_(Tab<K>::c[r][j], v[-r + j], lr);
        }
        c.q[r] = qr;
        c.l[r] = lr;
        e[r] = eps + beta[r];
        c.ib[r] = rcp_nr(e[r]);
        q[r] = e[r] * e[r];
        qm = fmax(qm, q[r]);
    }
    if constexpr (K == 1) {
        c.w[0] = 1.0;
        c.wt[0] = 1.0;
    } else {
        const double iq = rcp_nr(qm);
        double qn[K], Q[K];
#pragma unroll
        for (int r = 0; r < K; ++r) qn[r] = q[r] * iq;
        if constexpr (K == 2) {
            Q[0] = qn[1];
            Q[1] = qn[0];
        } else {
            Q[0] = qn[1] * qn[2];
            Q[1] = qn[0] * qn[2];
            Q[2] = qn[0] * qn[1];
        }
        double S = 0.0, St = 0.0;
#pragma unroll
        for (int r = 0; r < K; ++r) {
            S = fma_(Tab<K>::d[r], Q[r], S);
            St = fma_(Tab<K>::d[K - 1 - r], Q[r], St);
        }
        const double iS = rcp_nr(S), iSt = rcp_nr(St);
#pragma unroll
        for (int r = 0; r < K; ++r) {
            c.w[r] = (Tab<K>::d[r] * Q[r]) * iS;
            c.wt[r] = (Tab<K>::d[K - 1 - r] * Q[r]) * iSt;
        }
    }
    double vr = 0.0, vl = 0.0;
#pragma unroll
    for (int r = 0; r < K; ++r) {
        vr = fma_(c.w[r], c.q[r], vr);
        vl = fma_(c.wt[r], c.l[r], vl);
    }
    c.vr = vr;
    c.vl = vl;
}

// Window cotangents xb[0 .. 2K-2] of one cell for the cotangents gr, gl of its faces.
template <int K>
__device__ __forceinline__ void gcell_bwd(const double *x, const GCell<K> &c, double gr, double gl,
                                          double *xb)
{
    constexpr int H = K - 1;
    const double *v = x + H;
#pragma unroll
    for (int m = 0; m < 2 * H + 1; ++m) xb[m] = 0.0;
#pragma unroll
    for (int r = 0; r < K; ++r) {
        const double P = gr * c.w[r], Q = gl * c.wt[r];
#pragma unroll
        for (int j = 0; j < K; ++j) xb[H - r + j] += P * Tab<K>::c[r + 1][j] + Q * Tab<K>::c[r][j];
        if constexpr (K > 1) {
            const double bb = -2.0 * c.ib[r] * (P * (c.q[r] - c.vr) + Q * (c.l[r] - c.vl));
            if constexpr (K == 2) {
                // beta_0 = (v1 - v0)^2, beta_1 = (v0 - v-1)^2
                if (r == 0) {
                    const double a2 = 2.0 * (v[1] - v[0]) * bb;
                    xb[H + 1] += a2;
                    xb[H] -= a2;
                } else {
                    const double b2 = 2.0 * (v[0] - v[-1]) * bb;
                    xb[H] += b2;
                    xb[H - 1] -= b2;
                }
            } else {
                // beta_r = 13/12 a^2 + 1/4 b^2 (src/hrweno_weno.f90:193-202): d beta = 13/6 a da + 1/2 b db
                if (r == 0) {
                    const double A = 13.0 / 6 * ((v[0] - 2 * v[1]) + v[2]) * bb;
                    const double B = 0.5 * ((3 * v[0] - 4 * v[1]) + v[2]) * bb;
                    xb[H] += A + 3.0 * B;
                    xb[H + 1] += -2.0 * A - 4.0 * B;
                    xb[H + 2] += A + B;
                } else if (r == 1) {
                    const double A = 13.0 / 6 * ((v[-1] - 2 * v[0]) + v[1]) * bb;
                    const double B = 0.5 * (v[-1] - v[1]) * bb;
                    xb[H - 1] += A + B;
                    xb[H] += -2.0 * A;
                    xb[H + 1] += A - B;
                } else {
                    const double A = 13.0 / 6 * ((v[-2] - 2 * v[-1]) + v[0]) * bb;
                    const double B = 0.5 * ((v[-2] - 4 * v[-1]) + 3 * v[0]) * bb;
                    xb[H - 2] += A + B;
                    xb[H - 1] += -2.0 * A - 4.0 * B;
                    xb[H] += A + 3.0 * B;
                }
            }
        }
    }
}

// Partial derivatives of the interface flux for f(v) = v^2/2.  Lax-Friedrichs: (vm + alpha)/2,
// (vp - alpha)/2 and -(vp - vm)/2 for alpha.  Godunov over the two end states: the branch the
// flux takes; at vm == vp both branches give f(v) and F is f of the upwind state nearby (a
// smooth point), so the derivative is that state's.
template <int FLUX>
__device__ __forceinline__ void gflux_partials(double alpha, double vm, double vp, double &dm,
                                               double &dp, double &da)
{
    if constexpr (FLUX == LF) {
        dm = 0.5 * (vm + alpha);
        dp = 0.5 * (vp - alpha);
        da = -0.5 * (vp - vm);
    } else {
        const double fm = (vm * vm) / 2, fp = (vp * vp) / 2;
        const bool takem = vm == vp ? vm > 0.0 : (vm <= vp ? fm < fp : fm > fp);
        dm = takem ? vm : 0.0;
        dp = takem ? 0.0 : vp;
        da = 0.0;
    }
}

template <int K, int FLUX, bool PERIODIC, bool AMAX>
__global__ __launch_bounds__(kGenTile) void rhs_vjp_kernel(GenAdjArgs a)
{
    constexpr int H = K - 1;
    constexpr int T = kGenTile;
    constexpr int NR = T + 2 * H + 2;   // reconstructed positions [B0-H-1, B0+T+H+1)
    constexpr int NS = NR + 2 * H;      // loaded positions [B0-2H-1, B0+T+2H+1)
    constexpr int NX = T + 2 * H;       // positions with a window cotangent [B0-H, B0+T+H)
    constexpr int NW = 2 * H + 1;
    __shared__ double s_v[NS], s_phi[NR], s_vr[NR], s_vl[NR], s_xb[NX * NW], s_sum[T];
    __shared__ int s_arg;
    const long long n = a.n;
    const long long B0 = (long long)blockIdx.x * T;
    const int t = threadIdx.x;

    double alpha = a.alpha;
    if constexpr (AMAX) {
        unsigned long long m = 0ull;
#pragma unroll
        for (int i = 0; i < kNSlot; ++i) {
            const unsigned long long x = a.aset[i * kSlotStride];
            m = x > m ? x : m;
        }
        alpha = __longlong_as_double((long long)m);
        if (t == 0) s_arg = INT_MAX;
    }
    // the stage input over the window halo; phi on the positions whose interfaces are needed
    for (int i = t; i < NS; i += T) s_v[i] = a.s[cell_index(B0 - 2 * H - 1 + i, n, PERIODIC)];
    for (int i = t; i < NR; i += T) {
        const long long p = B0 - H - 1 + i;
        double ph = 0.0;
        if (PERIODIC || (p >= 1 && p <= n - 2)) {
            const long long c = PERIODIC ? cell_index(p, n, true) : p;
            double lj = a.b1 * a.y1[c];
            if (a.y2) lj = lj + a.b2 * a.y2[c];
            const double w = linear_edge(a.xmin, a.rx, c + 1) - linear_edge(a.xmin, a.rx, c);
            ph = lj / w;
        }
        s_phi[i] = ph;
    }
    __syncthreads();
    // faces of every position an interface of the tile's window positions touches; reconstructed
    // position r (p = B0-H-1+r) has its window at s_v[r .. r+2H].  Thread t reconstructs r = t+1
    // and keeps the cell state for that position's reverse below; threads t < 2H+2 also take one
    // of r = 0 and r = T+1 .. T+2H+1 (the states of r = T+1 .. T+2H kept as well)
    GCell<K> cm, cx;
    const int rx_ = t == 0 ? 0 : T + t;   // the extra position of thread t < 2H+2
    const bool has_x = t < 2 * H + 2;
    gcell_fwd<K>(s_v + t + 1, a.eps, cm);
    s_vr[t + 1] = cm.vr;
    s_vl[t + 1] = cm.vl;
    if (has_x) {
        gcell_fwd<K>(s_v + rx_, a.eps, cx);
        s_vr[rx_] = cx.vr;
        s_vl[rx_] = cx.vl;
    }
    __syncthreads();
    // window cotangents of the positions [B0-H, B0+T+H) (position p = B0-H+j is reconstructed
    // position r = j+1); alpha-bar over the interfaces p+1/2 of the block's own cells
    double ab = 0.0;
    auto reverse_at = [&](int r, const GCell<K> &c) {
        const int j = r - 1;
        const long long p = B0 - H + j;
        const bool right_if = PERIODIC || (p >= 0 && p <= n - 2);
        const bool left_if = PERIODIC || (p >= 1 && p <= n - 1);
        double gr = 0.0, gl = 0.0;
        if (right_if) {
            const double G = s_phi[r + 1] - s_phi[r];
            double dm, dp, da;
            gflux_partials<FLUX>(alpha, s_vr[r], s_vl[r + 1], dm, dp, da);
            gr = G * dm;
            if (AMAX && p >= B0 && p < B0 + T && p < n) ab += G * da;
        }
        if (left_if) {
            const double G = s_phi[r] - s_phi[r - 1];
            double dm, dp, da;
            gflux_partials<FLUX>(alpha, s_vr[r - 1], s_vl[r], dm, dp, da);
            gl = G * dp;
        }
        double xb[NW];
        gcell_bwd<K>(s_v + r, c, gr, gl, xb);
#pragma unroll
        for (int m = 0; m < NW; ++m) s_xb[j * NW + m] = xb[m];
    };
    reverse_at(t + 1, cm);
    if (has_x && rx_ >= T + 1 && rx_ <= T + 2 * H) reverse_at(rx_, cx);
    if constexpr (AMAX) {
        const long long p = B0 + t;
        if (p < n && fabs(a.s[p]) == alpha) atomicMin(&s_arg, (int)p);
        s_sum[t] = ab;
    }
    __syncthreads();
    // gather: output cell j receives offset m of the window of position j - m
    const long long j = B0 + t;
    if (j < n) {
        double acc = 0.0;
#pragma unroll
        for (int m = -H; m <= H; ++m) {
            const long long p = j - m;
            if (!PERIODIC && (p < 0 || p >= n)) continue;
            acc += s_xb[(p - (B0 - H)) * NW + (m + H)];
        }
        if constexpr (!PERIODIC && H > 0) {
            // window positions past an end are the end cell
            if (j == 0) {
                for (long long i = 0; i < H && i < n; ++i)
                    for (long long m = -H; m <= -i - 1; ++m) acc += s_xb[(i - (B0 - H)) * NW + (m + H)];
            }
            if (j == n - 1) {
                for (long long i = n - H > 0 ? n - H : 0; i < n; ++i)
                    for (long long m = n - i; m <= H; ++m) acc += s_xb[(i - (B0 - H)) * NW + (m + H)];
            }
        }
        double o = 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (a.x[i]) o = o + a.e[i] * a.x[i][j];
        a.out[j] = o + a.ch * acc;
    }
    if constexpr (AMAX) {
        // the block's alpha-bar in a fixed order
        for (int w = T / 2; w > 0; w >>= 1) {
            if (t < w) s_sum[t] = s_sum[t] + s_sum[t + w];
            __syncthreads();
        }
        if (t == 0) {
            a.partial[blockIdx.x] = s_sum[0];
            if (s_arg != INT_MAX) atomicMin(a.argmax, s_arg);
        }
    }
}

// The alpha-bar correction of a finished rhs_vjp_kernel pass: out[j] += ch * (sum of the block
// partials) * sign(s[j]) at the first cell j attaining max|s|; the argmax slot is reset and the
// max|s| slot set cleared for the next pass (which reads them only after this launch).
static __global__ __launch_bounds__(kGenTile) void rhs_vjp_alpha_fixup(const double *partial,
                                                                     int nblocks, int *argmax,
                                                                     unsigned long long *aset,
                                                                     const double *s, double *out,
                                                                     double ch)
{
    __shared__ double sm[kGenTile];
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int b = t; b < nblocks; b += kGenTile) acc += partial[b];
    sm[t] = acc;
    __syncthreads();
    for (int w = kGenTile / 2; w > 0; w >>= 1) {
        if (t < w) sm[t] = sm[t] + sm[t + w];
        __syncthreads();
    }
    if (t == 0) {
        const int j = *argmax;
        if (j != INT_MAX) {
            const double sj = s[j];
            out[j] += ch * sm[0] * (double)((sj > 0.0) - (sj < 0.0));
        }
        *argmax = INT_MAX;
    }
    if (t < kNSlot) aset[t * kSlotStride] = 0ull;
}

}  // namespace hrw
