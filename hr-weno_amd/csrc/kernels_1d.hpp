// kernels_1d.hpp -- HIP kernels for gfx950 on 1D lines.
//
//  reconstruct_kernel : weno%reconstruct (src/hrweno_weno.f90:129-219) on many strided lines.
//  stage1d_kernel     : one fused pass of  reconstruct -> numerical flux -> FV divergence ->
//                       TVD stage combination  for 1D Burgers (the user RHS of
//                       example/example1_burgers_1d_fv.f90:72-109 composed with
//                       src/hrweno_tvdode.f90:136-171 / :257).
//  maxabs_kernel      : max|u| seeding the Lax-Friedrichs alpha.
//
// Lax-Friedrichs alpha = max|v| of each stage input is carried between launches in "slot
// sets": NSLOT partial maxima (one per 128-byte line, fed by atomicMax on the bit pattern of
// |u|, which orders like the value for non-negative doubles).  Launch j of a sequence reads
// set j%3, accumulates the maximum of its own output into set (j+1)%3 and clears set (j+2)%3,
// so no set is read, written and cleared by the same launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "weno_math.hpp"

namespace hrw {

constexpr int kBlock = 256;
// Partial maxima per slot set: every block atomicMax-es into one of them, every wave of the
// next launch reads them all (16 cache lines), so fewer means less L2 traffic per wave and
// more atomics per slot.
constexpr int kNSlot = 16;
constexpr int kSlotStride = 16;  // 16 x 8 B = one 128-byte line per partial

__device__ __forceinline__ long long cell_index(long long i, long long n, bool periodic)
{
    if (i >= 0 && i < n) return i;
    if (periodic) {
        long long r = i % n;
        return r < 0 ? r + n : r;
    }
    return i < 0 ? 0 : n - 1;   // constant extrapolation, src/hrweno_weno.f90:171-173
}

// Maximum over the wavefront of an unsigned 64-bit value (the bit patterns of non-negative
// doubles order like the doubles) by DPP, without LDS round trips: after the row shifts by 1,
// 2, 4, 8 (a lane with no source keeps its own value) every lane holds the maximum of its row
// up to itself, so lane 15 of each row holds the row's; row_bcast:15 carries rows 0 and 2 into
// rows 1 and 3, row_bcast:31 carries lane 31 into rows 2 and 3, leaving the wavefront's
// maximum in lane 63, which is broadcast.  Every lane of the wavefront must be active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v)
{
    const int lo = __builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(v >> 32), (int)(v >> 32), CTRL, ROW_MASK, 0xF, false);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
__device__ __forceinline__ unsigned long long wave_umax(unsigned long long v)
{
    auto mx = [](unsigned long long x, unsigned long long y) { return x > y ? x : y; };
    v = mx(v, dpp_u64<0x111, 0xF>(v));   // row_shr:1
    v = mx(v, dpp_u64<0x112, 0xF>(v));   // row_shr:2
    v = mx(v, dpp_u64<0x114, 0xF>(v));   // row_shr:4
    v = mx(v, dpp_u64<0x118, 0xF>(v));   // row_shr:8
    v = mx(v, dpp_u64<0x142, 0xA>(v));   // row_bcast:15
    v = mx(v, dpp_u64<0x143, 0xC>(v));   // row_bcast:31
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)v, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}

// Whole-wavefront shifts by one lane by DPP (wave_shr:1 / wave_shl:1: a VALU operand modifier,
// no LDS round trip): lane l gets v of lane l-1 (shfl_up1) or of lane l+1 (shfl_down1); the lane
// with no source (0, resp. 63) keeps its own v -- what __shfl_up / __shfl_down by one over 64
// lanes return.  Every lane of the wavefront must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_wave_shift1(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double shfl_up1(double v) { return dpp_wave_shift1<0x138>(v); }
__device__ __forceinline__ double shfl_down1(double v) { return dpp_wave_shift1<0x130>(v); }

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Maximum over the slot set, broadcast to the whole block through LDS.  Must be called by
// every thread of the block (contains a barrier).
__device__ __forceinline__ double read_alpha(const unsigned long long *set, double *s_tmp)
{
    if (threadIdx.x < 64) {
        double a = __longlong_as_double((long long)set[(threadIdx.x % kNSlot) * kSlotStride]);
        a = wave_max(a);
        if (threadIdx.x == 0) s_tmp[0] = a;
    }
    __syncthreads();
    return s_tmp[0];
}

// Block-wide max of m, then one atomicMax into slot (blockIdx.x % kNSlot) of set.
__device__ __forceinline__ void publish_max(double m, unsigned long long *set, double *s_red)
{
    m = wave_max(m);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) s_red[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = s_red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = fmax(b, s_red[w]);
        atomicMax(set + (blockIdx.x % kNSlot) * kSlotStride,
                  (unsigned long long)__double_as_longlong(b));
    }
}

__device__ __forceinline__ void clear_set(unsigned long long *set)
{
    if (set && blockIdx.x == 0 && threadIdx.x < kNSlot) set[threadIdx.x * kSlotStride] = 0ull;
}

// --------------------------------------------------------------------------------------
static __global__ void clear_slots_kernel(unsigned long long *slots, int nsets)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nsets * kNSlot) slots[t * kSlotStride] = 0ull;
}

static __global__ __launch_bounds__(kBlock) void maxabs_kernel(const double *__restrict__ u, long long n,
                                                        unsigned long long *set)
{
    __shared__ double s_red[kBlock / 64];
    double m = 0.0;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock)
        m = fmax(m, fabs(u[i]));
    publish_max(m, set, s_red);
}

// --------------------------------------------------------------------------------------
// weno%reconstruct over nlines lines of n cells: element i of line l is v[l*ldv + i*incv].
// One thread per cell; the (2K-1)-wide window is read through the caches.
template <int K, bool NONUNIFORM, bool FAST>
__global__ __launch_bounds__(kBlock) void reconstruct_kernel(
    const double *__restrict__ v, long long incv, long long ldv, long long n, double eps,
    const double *__restrict__ cnu, double *__restrict__ vl, long long incl, long long ldl,
    double *__restrict__ vr, long long incr, long long ldr)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const long long l = blockIdx.y;
    const double *vline = v + l * ldv;
    double x[2 * K - 1];
#pragma unroll
    for (int m = -(K - 1); m <= K - 1; ++m) x[m + K - 1] = vline[cell_index(i + m, n, false) * incv];
    double a, b;
    if constexpr (NONUNIFORM) {
        weno_cell_exact<K>(x, eps, cnu + i * (K * (K + 1)), a, b);
    } else if constexpr (FAST) {
        weno_cell_fast<K>(x, eps, a, b);
    } else {
        weno_cell_exact<K>(x, eps, nullptr, a, b);
    }
    vl[l * ldl + i * incl] = a;
    vr[l * ldr + i * incr] = b;
}

// --------------------------------------------------------------------------------------
struct Stage1DArgs {
    const double *us;    // stage input, read with a halo of K cells
    const double *u0;    // state at the start of the step (pointwise)
    double *uo;          // stage output (nullptr: rhs only)
    const double *uo4;   // MS3: u_{n-4}        (may alias uo)
    const double *Lo4;   // MS3: L(u_{n-4})     (may alias Lout)
    double *Lout;        // L(us) (MS3 history / rhs-only), or nullptr
    const double *width; // cell widths, or nullptr for the linear-grid formula
    const double *cnu;   // non-uniform grid: cnu(j, r, i) of weno_calc_cnu, (n, K+1, K), or
                         // nullptr for the uniform coefficients
    double xmin, rx;     // linear grid: edge i = xmin + rx*i
    long long n;
    double eps, h, alpha;
    int stage;           // hrw::Stage, or -1 for rhs only
    const unsigned long long *alpha_in;   // slot set holding max|us| (ALPHAMAX)
    unsigned long long *alpha_out;        // slot set receiving max|output|
    unsigned long long *alpha_clear;      // slot set to clear
    double *alpha_record;                 // if set: receives the alpha this launch used
};

// R = kBlock*CPT cells are reconstructed per block; the block owns the TILE = R-2 interior
// ones, so each interface flux it needs comes from cells it reconstructed itself.
template <int K, int FLUX, bool PERIODIC, bool FAST, bool ALPHAMAX, int CPT>
__global__ __launch_bounds__(kBlock) void stage1d_kernel(Stage1DArgs a)
{
    constexpr int R = kBlock * CPT;
    constexpr int TILE = R - 2;
    constexpr int NLOAD = TILE + 2 * K;
    __shared__ double s_u[NLOAD];
    __shared__ double s_vl[R];
    __shared__ double s_vr[R];
    __shared__ double s_tmp[kBlock / 64 + 1];

    const long long n = a.n;
    const long long base = (long long)blockIdx.x * TILE;
    const int tid = threadIdx.x;

    // 1. stage input tile plus halo; s_u[j] <-> cell base-K+j
    for (int j = tid; j < NLOAD; j += kBlock) s_u[j] = a.us[cell_index(base - K + j, n, PERIODIC)];
    clear_set(a.alpha_clear);
    double alpha = a.alpha;
    if constexpr (ALPHAMAX) {
        alpha = read_alpha(a.alpha_in, s_tmp);   // contains a barrier
        if (a.alpha_record && blockIdx.x == 0 && threadIdx.x == 0) *a.alpha_record = alpha;
    }
    else __syncthreads();

    // 2. reconstruct cells base-1 .. base+TILE (local c <-> cell base-1+c)
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
        const int c = tid + m * kBlock;
        const double *x = s_u + c;   // window centre at s_u[c+K-1]
        double vl, vr;
        if (a.cnu) {
            // non-uniform grid (src/hrweno_weno.f90:174-178): the cell's own coefficients, in
            // the reference's order in both arithmetic forms; cells outside the line, whose
            // faces the boundary fluxes replace, take the nearest cell's
            const long long g = base - 1 + c;
            const long long gi = g < 0 ? 0 : (g >= n ? n - 1 : g);
            weno_cell_exact<K>(x, a.eps, a.cnu + gi * (K * (K + 1)), vl, vr);
        } else if constexpr (FAST) {
            weno_cell_fast<K>(x, a.eps, vl, vr);
        } else {
            weno_cell_exact<K>(x, a.eps, nullptr, vl, vr);
        }
        s_vl[c] = vl;
        s_vr[c] = vr;
    }
    __syncthreads();

    // 3. fluxes, divergence, stage update for the TILE owned cells
    double m_out = 0.0;
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
        const int o = tid + m * kBlock;
        const long long g = base + o;
        if (o >= TILE || g >= n) continue;
        double fr = burgers_numflux<FLUX, FAST>(s_vr[o + 1], s_vl[o + 2], alpha);
        double fl = burgers_numflux<FLUX, FAST>(s_vr[o], s_vl[o + 1], alpha);
        if constexpr (!PERIODIC) {
            // example1:103-104: fedges(0)=fedges(1), fedges(n)=fedges(n-1)
            if (g == 0) fl = fr;
            if (g == n - 1) fr = fl;
        }
        double w;
        if (a.width) w = a.width[g];
        else w = linear_edge(a.xmin, a.rx, g + 1) - linear_edge(a.xmin, a.rx, g);
        const double L = (fl - fr) / w;   // == -(fr - fl)/w exactly (example1:107)
        const double us = s_u[o + K];
        if (a.stage < 0) {
            a.Lout[g] = L;
            m_out = fmax(m_out, fabs(us));
            continue;
        }
        double u0 = us, uo4 = 0.0, Lo4 = 0.0;
        if (a.stage == RK2_2 || a.stage == RK3_2 || a.stage == RK3_3) u0 = a.u0[g];
        if (a.stage == MS3) {
            uo4 = a.uo4[g];
            Lo4 = a.Lo4[g];
        }
        const double un = stage_update<FAST>(a.stage, u0, us, L, a.h, uo4, Lo4);
        if (a.Lout) a.Lout[g] = L;
        a.uo[g] = un;
        m_out = fmax(m_out, fabs(un));
    }
    if constexpr (ALPHAMAX) publish_max(m_out, a.alpha_out, s_tmp);
}

// --------------------------------------------------------------------------------------
// Elementwise stage combination for integrands given as callbacks (the user computes L).
template <bool FAST>
__global__ __launch_bounds__(kBlock) void combine_kernel(int stage, long long n, double h,
                                                         const double *u0, const double *us,
                                                         const double *L, const double *uo4,
                                                         const double *Lo4, double *uo)
{
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock) {
        const double s = us[i];
        const double z = (stage == RK2_2 || stage == RK3_2 || stage == RK3_3) ? u0[i] : s;
        const double a4 = stage == MS3 ? uo4[i] : 0.0;
        const double l4 = stage == MS3 ? Lo4[i] : 0.0;
        uo[i] = stage_update<FAST>(stage, z, s, L[i], h, a4, l4);
    }
}

// Numerical fluxes over arrays of interfaces (src/hrweno_fluxes.f90:22-76) for the built-in
// flux functions: fn 0 -> v**2/2, fn 1 -> (ca + cb*x)*v.
static __global__ __launch_bounds__(kBlock) void flux_kernel(int which, int fn, double ca, double cb,
                                                      long long n, const double *vm,
                                                      const double *vp, const double *x,
                                                      double t, double alpha, double *h)
{
    (void)t;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock) {
        const double a = vm[i], b = vp[i];
        double fm, fp;
        if (fn == 0) {
            fm = burgers_f(a);
            fp = burgers_f(b);
        } else {
            const double g = linear_coef(ca, cb, x ? x[i] : 0.0);
            fm = g * a;
            fp = g * b;
        }
        h[i] = which == 0 ? lax_friedrichs(fm, fp, a, b, alpha) : godunov(fm, fp, a, b);
    }
}

}  // namespace hrw
