// kernels_1d_fast.hpp -- register-blocked fused stage kernel for the FAST arithmetic form.
//
// The stage pass is bound by FP64 instruction issue, not by HBM (its traffic is already the
// algorithmic 16/24/24 bytes per cell).  This kernel cuts the instructions per cell:
//
//  * each lane owns R consecutive cells; the K-1... halo comes from the neighbouring lanes by
//    ds_bpermute shuffles (LDS pipe, no VALU), lanes 0 and 63 load theirs directly;
//  * adjacent cells share work: with D = v0 - 2v1 + v2 and E = v0 - v2 on the triple
//    T(m) = (v_m, v_m+1, v_m+2), the three smoothness indicators drawing on T(m) are
//        12*beta_0(m)   = 13 D^2 + 3 (E + 2D)^2
//        12*beta_1(m+1) = 13 D^2 + 3 E^2
//        12*beta_2(m+2) = 13 D^2 + 3 (2D - E)^2
//    and the four candidate values are 6*c(:,r).T(m), r = -1..2 (used as vlr0(m), vrr0(m) =
//    vlr1(m+1), vrr1(m+1) = vlr2(m+2), vrr2(m+2)); the factors 12 and 6 cancel in the
//    normalised weights (eps is scaled by 12 to match);
//  * one Lax-Friedrichs/Godunov flux and one grid edge per interface;
//  * weights over a common denominator, w_r ~ d_r prod_{s!=r} q_s, with d scaled to integers
//    (3, 6, 1) so they need three multiplies, and one v_rcp_f64 + two Newton steps per face
//    instead of IEEE divides (also for 1/width).
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

__device__ __forceinline__ Tri3 tri3(double a, double b, double c)
{
    Tri3 t;
    const double D = fma_(-2.0, b, a) + c;
    const double E = a - c;
    const double dd = (13.0 * D) * D;
    t.bB = fma_(3.0 * E, E, dd);
    const double X = fma_(2.0, D, E);
    t.bA = fma_(3.0 * X, X, dd);
    const double Y = fma_(2.0, D, -E);
    t.bC = fma_(3.0 * Y, Y, dd);
    t.P0 = fma_(5.0, b, fma_(2.0, a, -c));          // 2a + 5b - c
    t.P1 = fma_(5.0, b, fma_(2.0, c, -a));          // -a + 5b + 2c
    t.Pm = fma_(11.0, a, fma_(-7.0, b, c + c));     // 11a - 7b + 2c
    t.P2 = fma_(11.0, c, fma_(-7.0, b, a + a));     // 2a - 7b + 11c
    return t;
}

// One reconstruction from the pieces of T(j-2), T(j-1), T(j).  eps12 = 12*eps.  Each face
// takes one reciprocal of its own sum of weights, so the only products are pairs of q's: finite
// while beta < ~1e77, i.e. for cell-value jumps below ~1e38.
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
    const double sr = 6.0 * (a0 + a1 + m2), sl = 6.0 * (m0 + a1 + l2);
    const double nr = fma_(a0, t0.P0, fma_(a1, t1.P1, m2 * t2.P2));
    const double nl = fma_(m0, t0.Pm, fma_(a1, t1.P0, l2 * t2.P1));
    vr = nr * rcp_nr(sr);
    vl = nl * rcp_nr(sl);
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
__device__ __forceinline__ void recon2(double eps, const Pair2 &p1, const Pair2 &p0, double &vl,
                                       double &vr)
{
    double q0 = eps + p0.B, q1 = eps + p1.B;
    q0 = q0 * q0;
    q1 = q1 * q1;
    // right: d = (2/3, 1/3) ~ (2, 1): weights ~ 2 q1, q0; left: (1, 2): q1, 2 q0
    const double a0 = 2.0 * q1, b1 = 2.0 * q0;
    const double sr = 2.0 * (a0 + q0), sl = 2.0 * (q1 + b1);
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
//   un = post * (C*L + B*p1 + A*us + D*p2)      (post = 1/3 uses div3_fast)
// evaluated as fma(D, p2, fma(C, L, fma(B, p1, A*us))).  p1/p2 are the pointwise operands
// (u at the start of the step; or u_{n-4}, L(u_{n-4}) for the multistep method).
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

// One wavefront's chunk [W, W + 64R) of a line: lane l owns the R cells from c0 = W + l*R.
// Halo: H = K cells on each side (K-1 for the reconstruction window, one more for the
// neighbouring reconstruction that the chunk-edge interface needs).  Interior chunks (range and
// halo inside the line) take vector loads/stores and no boundary checks.  NPW = number of
// pointwise operands of the stage (0, 1 or 2).  Returns the lane's max|output| (the next
// stage's alpha contribution).  Requires n < 2^31 and 16-byte aligned arrays (checked on the
// host).  Must be called by all 64 lanes of the wavefront.
// What an interior chunk reads from memory, held in registers so that it can be loaded one
// chunk ahead (the persistent kernel below): the lane's R cells of us and the pointwise
// operands.  (The K-cell halo that only the two wave-edge lanes load is read when used.)
template <int K, int R, int NPW>
struct ChunkData {
    double run[R];
    double p1[NPW >= 1 ? R : 1], p2[NPW >= 2 ? R : 1];
};

__device__ __forceinline__ bool chunk_interior(long long W, int R, int H, long long n)
{
    return W >= H && W + 64LL * R + H <= n;
}

template <int K, int R, int NPW>
__device__ __forceinline__ void load_chunk(const Stage1DArgs &a, int W, ChunkData<K, R, NPW> &d)
{
    const int c0 = W + (int)(threadIdx.x & 63) * R;
    load_run<R>(a.us + c0, d.run);
    const double *s1 = a.stage == MS3 ? a.uo4 : a.u0;
    if constexpr (NPW >= 1) load_run<R>(s1 + c0, d.p1);
    if constexpr (NPW >= 2) load_run<R>(a.Lo4 + c0, d.p2);
}

template <int K, int FLUX, bool PERIODIC, int R, int NPW>
__device__ __forceinline__ double chunk_stage(const Stage1DArgs &a, int W, double alpha,
                                              const ChunkData<K, R, NPW> *pre = nullptr)
{
    constexpr int H = K;
    constexpr int NV = R + 2 * H;
    const int lane = threadIdx.x & 63;
    const int n = (int)a.n;
    const int c0 = W + lane * R;
    const StageCoef sc = stage_coef(a.stage, a.h);
    const bool write_u = a.stage >= 0;

    // ---- the lane's cells and halo (v[i] <-> cell c0 - H + i) and pointwise operands
    double v[NV];
    [[maybe_unused]] double p1[NPW >= 1 ? R : 1], p2[NPW >= 2 ? R : 1];
    const double *s1 = a.stage == MS3 ? a.uo4 : a.u0;
    const bool interior = chunk_interior(W, R, H, n);
    if (interior) {
        ChunkData<K, R, NPW> d;
        if (pre) d = *pre;
        else load_chunk<K, R, NPW>(a, W, d);
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
            v[i] = __shfl_up(v[R + i], 1, 64);            // lane-1's last H cells
            v[H + R + i] = __shfl_down(v[H + i], 1, 64);  // lane+1's first H cells
        }
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < H; ++i) v[i] = a.us[c0 - H + i];
        }
        if (lane == 63) {
#pragma unroll
            for (int i = 0; i < H; ++i) v[H + R + i] = a.us[c0 + R + i];
        }
    } else {
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] = a.us[cell_index(c0 - H + i, n, PERIODIC)];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int g = c0 + i < n ? c0 + i : n - 1;
            if constexpr (NPW >= 1) p1[i] = s1[g];
            if constexpr (NPW >= 2) p2[i] = a.Lo4[g];
        }
    }

    // ---- march over the cells: reconstruction of cell c0-1+s for s = 0..R+1
    double m_out = 0.0;
    double fl = 0.0, vr_prev = 0.0;
    double e_prev = linear_edge32(a.xmin, a.rx, c0 < n ? c0 : n);
    const double eps12 = 12.0 * a.eps;
    double un_pair[2], L_pair[2];
    [[maybe_unused]] Tri3 t2, t1;
    [[maybe_unused]] Pair2 pp1;
    if constexpr (K == 3) {
        t2 = tri3(v[0], v[1], v[2]);
        t1 = tri3(v[1], v[2], v[3]);
    } else if constexpr (K == 2) {
        pp1 = pair2(v[0], v[1]);
    }
#pragma unroll
    for (int s = 0; s <= R + 1; ++s) {
        double vl, vr;
        if constexpr (K == 3) {
            const Tri3 t0 = tri3(v[s + 2], v[s + 3], v[s + 4]);   // T(c0-1+s)
            recon3(eps12, t2, t1, t0, vl, vr);
            t2 = t1;
            t1 = t0;
        } else if constexpr (K == 2) {
            const Pair2 p0 = pair2(v[s + 1], v[s + 2]);
            recon2(a.eps, pp1, p0, vl, vr);
            pp1 = p0;
        } else {
            vl = vr = v[s];
        }
        if (s == 0) {
            vr_prev = vr;
            continue;
        }
        // interface (c0-1+s) - 1/2 between cells c0-2+s and c0-1+s
        const double f = burgers_numflux<FLUX, true>(vr_prev, vl, alpha);
        vr_prev = vr;
        if (s == 1) {
            fl = f;
            continue;
        }
        // cell g = c0 + i, i = s - 2, between fluxes fleft and fr
        const int i = s - 2;
        const int g = c0 + i;
        double fr = f, fleft = fl;
        fl = f;
        double w;
        if (a.width) {
            w = a.width[g < n ? g : n - 1];
        } else {
            const double e = linear_edge32(a.xmin, a.rx, g + 1);
            w = e - e_prev;
            e_prev = e;
        }
        if constexpr (!PERIODIC) {
            if (!interior) {   // example1:103-104
                if (g == 0) fleft = fr;
                if (g == n - 1) fr = fleft;
            }
        }
        const double L = (fleft - fr) * rcp_nr(w);
        const double us = v[H + i];
        double t = sc.A * us;
        if constexpr (NPW >= 1) t = fma_(sc.B, p1[i], t);
        t = fma_(sc.C, L, t);
        if constexpr (NPW >= 2) t = fma_(sc.D, p2[i], t);
        const double un = sc.div3 ? div3_fast(t) : sc.post * t;
        // write back in pairs (16-byte stores) inside the line, one by one at its ends
        un_pair[i & 1] = un;
        L_pair[i & 1] = L;
        if (interior) {
            m_out = fmax(m_out, fabs(un));
            if (i & 1) {
                if (write_u) *reinterpret_cast<double2 *>(a.uo + g - 1) = make_double2(un_pair[0], un_pair[1]);
                if (a.Lout) *reinterpret_cast<double2 *>(a.Lout + g - 1) = make_double2(L_pair[0], L_pair[1]);
            }
        } else if (g < n) {
            m_out = fmax(m_out, fabs(un));
            if (write_u) a.uo[g] = un;
            if (a.Lout) a.Lout[g] = L;
        }
    }
    return m_out;
}

template <int K, int FLUX, bool PERIODIC, int R>
__device__ __forceinline__ double chunk_stage_any(const Stage1DArgs &a, int W, double alpha)
{
    if (a.stage == MS3) return chunk_stage<K, FLUX, PERIODIC, R, 2>(a, W, alpha);
    if (a.stage == RK2_2 || a.stage == RK3_2 || a.stage == RK3_3)
        return chunk_stage<K, FLUX, PERIODIC, R, 1>(a, W, alpha);
    return chunk_stage<K, FLUX, PERIODIC, R, 0>(a, W, alpha);
}

#ifndef HRW_FAST_WAVES_PER_EU
#define HRW_FAST_WAVES_PER_EU 4
#endif

// Persistent form: wave w of the grid takes chunks w, w + N, w + 2N, ... (N = waves in the
// grid) and loads the next interior chunk into a second register set before it computes the
// current one, so every wave keeps a chunk of loads in flight behind its FP64 work.
template <int K, int FLUX, bool PERIODIC, int R, int NPW>
__device__ __forceinline__ double persistent_stage(const Stage1DArgs &a, double alpha)
{
    const long long nw = (long long)gridDim.x * (kBlock / 64);
    const long long nch = (a.n + 64LL * R - 1) / (64LL * R);
    long long c = (long long)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    ChunkData<K, R, NPW> cur, nxt;
    double m = 0.0;
    if (c < nch && chunk_interior(c * 64 * R, R, K, a.n)) load_chunk<K, R, NPW>(a, (int)(c * 64 * R), cur);
    for (; c < nch; c += nw) {
        const long long c1 = c + nw;
        if (c1 < nch && chunk_interior(c1 * 64 * R, R, K, a.n)) load_chunk<K, R, NPW>(a, (int)(c1 * 64 * R), nxt);
        m = fmax(m, chunk_stage<K, FLUX, PERIODIC, R, NPW>(a, (int)(c * 64 * R), alpha, &cur));
        cur = nxt;
    }
    return m;
}

template <int K, int FLUX, bool PERIODIC, int R>
__device__ __forceinline__ double persistent_stage_any(const Stage1DArgs &a, double alpha)
{
    if (a.stage == MS3) return persistent_stage<K, FLUX, PERIODIC, R, 2>(a, alpha);
    if (a.stage == RK2_2 || a.stage == RK3_2 || a.stage == RK3_3)
        return persistent_stage<K, FLUX, PERIODIC, R, 1>(a, alpha);
    return persistent_stage<K, FLUX, PERIODIC, R, 0>(a, alpha);
}

#ifndef HRW_PF_WAVES_PER_EU
#define HRW_PF_WAVES_PER_EU 3
#endif

template <int K, int FLUX, bool PERIODIC, bool ALPHAMAX, int R>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(HRW_PF_WAVES_PER_EU)))
void stage1d_pf_kernel(Stage1DArgs a)
{
    __shared__ double s_red[kBlock / 64 + 1];
    const int lane = threadIdx.x & 63;
    double alpha = a.alpha;
    if constexpr (ALPHAMAX) {
        alpha = wave_max(__longlong_as_double((long long)a.alpha_in[lane * kSlotStride]));
        if (a.alpha_record && blockIdx.x == 0 && threadIdx.x == 0) *a.alpha_record = alpha;
    }
    clear_set(a.alpha_clear);
    const double m_out = persistent_stage_any<K, FLUX, PERIODIC, R>(a, alpha);
    if constexpr (ALPHAMAX) publish_max(m_out, a.alpha_out, s_red);
}

// One line: wave w of block b owns the chunk starting at W = (4b + w)*64R.
template <int K, int FLUX, bool PERIODIC, bool ALPHAMAX, int R>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(HRW_FAST_WAVES_PER_EU)))
void stage1d_fast_kernel(Stage1DArgs a)
{
    __shared__ double s_red[kBlock / 64 + 1];
    const int lane = threadIdx.x & 63;
    const int W = (int)(((long long)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * (64LL * R));
    double alpha = a.alpha;
    if constexpr (ALPHAMAX) {
        // every wave reduces the 64 partial maxima itself (no barrier)
        alpha = wave_max(__longlong_as_double((long long)a.alpha_in[lane * kSlotStride]));
        if (a.alpha_record && blockIdx.x == 0 && threadIdx.x == 0) *a.alpha_record = alpha;
    }
    clear_set(a.alpha_clear);
    const double m_out = chunk_stage_any<K, FLUX, PERIODIC, R>(a, W, alpha);
    if constexpr (ALPHAMAX) publish_max(m_out, a.alpha_out, s_red);
}

// --------------------------------------------------------------------------------------
// Many independent periodic lines in one launch.  Line l: cells [off[l], off[l] + n[l]) of
// the flat state, WENO order k[l], step h[l], linear grid xmin[l] + rx[l]*i.  Chunk c of the
// table covers cells [start[c], start[c] + 64R) of line line[c]; wavefront c of the grid takes
// chunk c.  Lax-Friedrichs alpha = max|v| per line: slot set s holds one partial per line.
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
};

template <int FLUX, bool ALPHAMAX, int R>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(HRW_FAST_WAVES_PER_EU)))
void batch_stage_kernel(BatchArgs b)
{
    const long long wave = ((long long)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (ALPHAMAX && b.alpha_clear) {
        for (long long l = (long long)blockIdx.x * kBlock + threadIdx.x; l < b.nlines;
             l += (long long)gridDim.x * kBlock)
            b.alpha_clear[l] = 0ull;
    }
    if (wave >= b.nchunks) return;   // whole wavefronts leave together
    const int line = b.chunk_line[wave];
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
    double alpha = b.alpha;
    if constexpr (ALPHAMAX) alpha = __longlong_as_double((long long)b.alpha_in[line]);
    const int W = b.chunk_start[wave];
    double m;
    switch (b.k[line]) {
    case 1: m = chunk_stage_any<1, FLUX, true, R>(a, W, alpha); break;
    case 2: m = chunk_stage_any<2, FLUX, true, R>(a, W, alpha); break;
    default: m = chunk_stage_any<3, FLUX, true, R>(a, W, alpha); break;
    }
    if constexpr (ALPHAMAX) {
        m = wave_max(m);
        if (lane == 0) atomicMax(b.alpha_out + line, (unsigned long long)__double_as_longlong(m));
    }
}

// max|u| of every line into slot set `set` (one partial per line).
__global__ __launch_bounds__(kBlock) void batch_maxabs_kernel(const double *__restrict__ u,
                                                              const long long *off, const int *n,
                                                              const int *chunk_line,
                                                              const int *chunk_start,
                                                              long long nchunks, int R,
                                                              unsigned long long *set)
{
    const long long wave = ((long long)blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= nchunks) return;
    const int line = chunk_line[wave];
    const long long base = off[line] + chunk_start[wave];
    const int len = min(64 * R, n[line] - chunk_start[wave]);
    double m = 0.0;
    for (int i = lane; i < len; i += 64) m = fmax(m, fabs(u[base + i]));
    m = wave_max(m);
    if (lane == 0) atomicMax(set + line, (unsigned long long)__double_as_longlong(m));
}

__global__ void fill_u64_kernel(unsigned long long *p, long long count)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

}  // namespace hrw
