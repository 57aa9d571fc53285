// kernels_adjoint_blocked.hpp -- register-blocked form of the general reverse pass for the
// third- and fifth-order reconstructions (K = 2, 3): the endpoint Godunov flux, or
// Lax-Friedrichs with a fixed alpha or alpha = max|s|; periodic or zero-gradient ends
// (example/example1_burgers_1d_fv.f90:72-109; src/hrweno_weno.f90:174-216 per cell;
// src/hrweno_fluxes.f90:42-43, 66-74).  It computes what rhs_vjp_kernel computes,
//     out = ey1 y1 + ey2 y2 + ex0 x0 + ex1 x1 + ch J(s)^T (b1 y1 + b2 y2),
// with the pointwise terms on y1 and y2 folded into their coefficients by the launcher.
//
// Layout: wavefront c takes the output cells [W, W + 60R), W = 60Rc; lane l holds the R cells
// from c0 = W + (l-2)R and their window halo from its neighbours (DPP wave shifts).  Lanes 0,
// 1, 62 and 63 are halo lanes: 1 and 62 run the reverse of the cells whose windows spill into
// the wavefront's first and last cells, 0 and 63 only supply the faces and phi those need (and
// load the H = K-1 cells beyond the wavefront), so no lane reconstructs a cell twice.  Each lane
// forms the faces of its cells, the interface cotangents
// (phi = (b1 y1 + b2 y2)/w, the cotangent of F_{p+1/2} is phi_{p+1} - phi_p):
//     vr-bar_p = (phi_{p+1} - phi_p) dF/dvm (vr_p, vl_{p+1}),
//     vl-bar_p = (phi_p - phi_{p-1}) dF/dvp (vr_{p-1}, vl_p),
// runs the reverse of each reconstruction into R + 2H accumulators and hands the H that spill
// past each end to the neighbouring lane.  No LDS tile and no barrier (the alpha = max|s| form
// has one, for its block sum at the end).
// Ends: with periodic ends every position stands for its periodic image (every (cell, offset)
// pair of the stencil is counted once, also for lines shorter than the stencil).  With
// zero-gradient ends the end cells' RHS is zero, so phi vanishes at cells 0 and n-1 and past
// the ends (no interface cotangent outside [1/2, n-3/2]); window positions past an end are the
// end cell (src/hrweno_weno.f90:171-173), so the lanes holding cells 0 and n-1 fold the
// accumulators past the end into them.  Only the first and last wavefronts of a line touch an
// end; the others take the unchecked path.
// alpha = max|s| (Lax-Friedrichs): alpha-bar is a sum over the whole line that lands on the
// first cell attaining the maximum.  Each block writes its share (its waves' sums in a fixed
// order) and atomicMin-s the first such cell it holds; rhs_vjp_alpha_fixup adds the correction
// once the pass has finished, as for the LDS-tile pass.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels_adjoint.hpp"
#include "kernels_adjoint_general.hpp"

namespace hrw {

constexpr int kBvjpR = 2;   // cells per lane

struct BvjpArgs {
    const double *s;             // the stage input the Jacobian is taken at
    const double *y1, *y2;       // J^T acts on b1*y1 + b2*y2 (y2 may be null)
    double b1, b2;
    double ey1, ey2;             // pointwise terms on y1 and y2
    const double *x[2];          // other pointwise terms (null: absent)
    double ex[2];
    double *out;                 // may alias x[i], not y1 or y2
    double ch;                   // the scale of J^T (the stage's c*h)
    double eps, alpha;           // alpha: Lax-Friedrichs with a fixed alpha
    double xmin, rx;             // linear grid
    int constw;                  // 1: every edge xmin + rx*i is exact, so every width is rx
    long long n, nchunks;
    unsigned long long *aset;        // alpha = max|s|: a slot set holding max|s| (the fixup clears it)
    const double *alpha_ptr;         // alpha = max|s|: if set, max|s| itself (aset is not read)
    double *partial;                 // alpha = max|s|: one alpha-bar share per block
    int *argmax;                     // alpha = max|s|: first cell with |s| == alpha (INT_MAX: none)
};

// Returns the lane's share of alpha-bar over its owned cells (AMAX); *cand = the first owned cell
// with |s| == alpha (0xffffffff: none).
template <int K, int FLUX, bool PERIODIC, bool INTERIOR, bool AMAX>
__device__ __forceinline__ double bvjp_chunk(const BvjpArgs &a, long long W, int lane, double alpha,
                                             unsigned *cand)
{
    static_assert(K == 2 || K == 3, "third- and fifth-order reconstructions");
    constexpr int R = kBvjpR, H = K - 1, NV = R + 2 * H;
    static_assert(H <= R, "the spills reach the neighbouring lanes only");
    const long long n = a.n;
    const long long c0 = W + (long long)(lane - 2) * R;
    // the edge lanes' window halo beyond the wavefront: the H cells left of lane 0's cells
    // (lanes < 32 load them) or right of lane 63's
    const long long hc = lane < 32 ? W - 2 * R - H : W + 62LL * R;
    auto at = [&](long long c) { return INTERIOR ? c : cell_index(c, n, PERIODIC); };
    // 1/w; zero-gradient ends: phi is zero at the end cells and past them
    const double iwc = a.constw ? rcp_nr(a.rx) : 0.0;
    auto phi_of = [&](long long c, double lj) {
        if (!INTERIOR && !PERIODIC && (c < 1 || c > n - 2)) return 0.0;
        const long long g = at(c);
        if (a.constw) return lj * iwc;
        return lj * rcp_nr(linear_edge(a.xmin, a.rx, g + 1) - linear_edge(a.xmin, a.rx, g));
    };

    double v[NV], y1[R], y2[R], hv[H];
    double o[R];   // the pointwise terms
    if constexpr (INTERIOR) {
        load_run<R>(a.s + c0, v + H);
    } else {
#pragma unroll
        for (int i = 0; i < R; ++i) v[H + i] = a.s[at(c0 + i)];
    }
#pragma unroll
    for (int i = 0; i < H; ++i) hv[i] = a.s[at(hc + i)];
    if constexpr (INTERIOR) {
        load_run<R>(a.y1 + c0, y1);
    } else {
#pragma unroll
        for (int i = 0; i < R; ++i) y1[i] = a.y1[at(c0 + i)];
    }
    if (a.y2) {
        if constexpr (INTERIOR) {
            load_run<R>(a.y2 + c0, y2);
        } else {
#pragma unroll
            for (int i = 0; i < R; ++i) y2[i] = a.y2[at(c0 + i)];
        }
    } else {
#pragma unroll
        for (int i = 0; i < R; ++i) y2[i] = 0.0;
    }
#pragma unroll
    for (int i = 0; i < R; ++i) o[i] = fma_(a.ey2, y2[i], a.ey1 * y1[i]);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (!a.x[k]) continue;
        double xr[R];
        if constexpr (INTERIOR) {
            load_run<R>(a.x[k] + c0, xr);
        } else {
#pragma unroll
            for (int i = 0; i < R; ++i) xr[i] = a.x[k][at(c0 + i)];
        }
#pragma unroll
        for (int i = 0; i < R; ++i) o[i] = fma_(a.ex[k], xr[i], o[i]);
    }

    // the window halo from the neighbouring lanes; the edge lanes' from hv
#pragma unroll
    for (int i = 0; i < H; ++i) {
        v[i] = shfl_up1(v[R + i]);
        v[H + R + i] = shfl_down1(v[H + i]);
    }
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < H; ++i) v[i] = hv[i];
    }
    if (lane == 63) {
#pragma unroll
        for (int i = 0; i < H; ++i) v[H + R + i] = hv[i];
    }

    // faces of the lane's cells (triples or pairs shared between neighbouring cells) and the
    // neighbours' (lanes 0 and 63 get their own back: unused)
    [[maybe_unused]] const double eps12 = 12.0 * a.eps;
    double vr[R], vl[R];
    if constexpr (K == 3) {
        Tri3 T[R + 2];
#pragma unroll
        for (int j = 0; j < R + 2; ++j) T[j] = tri3(v[j], v[j + 1], v[j + 2]);
#pragma unroll
        for (int t = 0; t < R; ++t) recon3<false, true>(eps12, T[t], T[t + 1], T[t + 2], vl[t], vr[t]);
    } else {
        Pair2 P[R + 1];
#pragma unroll
        for (int j = 0; j < R + 1; ++j) P[j] = pair2(v[j], v[j + 1]);
#pragma unroll
        for (int t = 0; t < R; ++t) recon2(a.eps, P[t], P[t + 1], vl[t], vr[t]);
    }
    const double vr_prev = shfl_up1(vr[R - 1]), vl_next = shfl_down1(vl[0]);
    // flux derivatives at faces f = 0..R (between local cells f-1 and f)
    double dm[R + 1], dp[R + 1], da[R + 1];
#pragma unroll
    for (int f = 0; f <= R; ++f)
        gflux_partials<FLUX>(alpha, f == 0 ? vr_prev : vr[f - 1], f == R ? vl_next : vl[f], dm[f],
                             dp[f], da[f]);
    // phi of cells c0-1 .. c0+R
    double phi[R + 2];
#pragma unroll
    for (int i = 0; i < R; ++i) phi[i + 1] = phi_of(c0 + i, fma_(a.b2, y2[i], a.b1 * y1[i]));
    phi[0] = shfl_up1(phi[R]);
    phi[R + 1] = shfl_down1(phi[1]);

    const bool owned = lane >= 2 && lane <= 61;
    double abar = 0.0;
    unsigned cnd = 0xffffffffu;
    if constexpr (AMAX) {
        // the right interface of each owned cell (phi vanishes where an interface does not exist)
#pragma unroll
        for (int t = 0; t < R; ++t)
            if (owned && c0 + t < n) {
                abar = fma_(phi[t + 2] - phi[t + 1], da[t + 1], abar);
                if (fabs(v[H + t]) == alpha && cnd == 0xffffffffu) cnd = (unsigned)(c0 + t);
            }
    }
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const double gr = (phi[t + 2] - phi[t + 1]) * dm[t + 1];
        const double gl = (phi[t + 1] - phi[t]) * dp[t];
        auto face_cot = [&](double, double, double ir, double il, double &sgr, double &sgl) {
            sgr = gr * ir;
            sgl = gl * il;
        };
        if constexpr (K == 3)
            weno5_vjp(eps12, v + t, face_cot, acc + t);
        else
            weno3_vjp(a.eps, pair2(v[t], v[t + 1]), pair2(v[t + 1], v[t + 2]), face_cot, acc + t);
    }
    // the spills into the neighbours' cells
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const double fromR = shfl_down1(acc[i]);
        const double fromL = shfl_up1(acc[R + H + i]);
        acc[R + i] += fromR;
        acc[H + i] += fromL;
    }
    if constexpr (!INTERIOR && !PERIODIC) {
        // window positions past an end are the end cell: positions c0-H .. c0-1 of the lane
        // holding cell 0 (only its own cells' windows reach them) and positions n .. of the
        // lane holding cell n-1 (its own cells' and, through the exchange, its left neighbour's;
        // the cells right of n-1 are images with phi = 0 on both sides, which add nothing)
        if (c0 == 0) {
            double t = 0.0;
#pragma unroll
            for (int m = 0; m < H; ++m) t += acc[m];
            acc[H] += t;
        }
        const long long jl = n - 1 - c0;
        if (jl >= 0 && jl < R) {
            double t = 0.0;
#pragma unroll
            for (int m = 0; m < NV; ++m)
                if (m > H + jl) t += acc[m];
#pragma unroll
            for (int i = 0; i < R; ++i)
                if (i == jl) acc[H + i] += t;
        }
    }
    if (owned) {
        double r[R];
#pragma unroll
        for (int i = 0; i < R; ++i) r[i] = fma_(a.ch, acc[H + i], o[i]);
        if constexpr (INTERIOR) {
            store_run<R>(a.out + c0, r);
        } else {
#pragma unroll
            for (int i = 0; i < R; ++i)
                if (c0 + i < n) a.out[c0 + i] = r[i];
        }
    }
    *cand = cnd;
    return abar;
}

// One wavefront per chunk of 60R output cells, 4 waves per SIMD (up to 128 VGPRs: 118-128
// without spills; left at 3, the alpha = max|s| form took 130).  Asking for 5 (96 VGPRs) spills
// 33-48 registers.
template <int K, int FLUX, bool PERIODIC, bool AMAX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void rhs_vjp5_kernel(
    BvjpArgs a)
{
    constexpr int R = kBvjpR, H = K - 1;
    __shared__ double s_ab[4];
    __shared__ unsigned s_cnd[4];
    const long long c = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    const int lane = threadIdx.x & 63;
    double alpha = a.alpha;
    if constexpr (AMAX) {
        if (a.alpha_ptr) {
            alpha = *a.alpha_ptr;
        } else {
            unsigned long long m = 0ull;
#pragma unroll
            for (int i = 0; i < kNSlot; ++i) {
                const unsigned long long x = a.aset[i * kSlotStride];
                m = x > m ? x : m;
            }
            alpha = __longlong_as_double((long long)m);
        }
    }
    double ab = 0.0;
    unsigned cnd = 0xffffffffu;
    if (c < a.nchunks) {   // (whole wavefronts past the last chunk only join the block's sums)
        const long long W = c * 60LL * R;
        // every cell the wavefront touches lies in [W - 2R - H, W + 62R + H - 1]
        if (W - 2 * R - H >= 0 && W + 62LL * R + H - 1 <= a.n - 1)
            ab = bvjp_chunk<K, FLUX, PERIODIC, true, AMAX>(a, W, lane, alpha, &cnd);
        else
            ab = bvjp_chunk<K, FLUX, PERIODIC, false, AMAX>(a, W, lane, alpha, &cnd);
    }
    if constexpr (AMAX) {
        // the block's share of alpha-bar in a fixed order, and its first cell at the maximum
        const double ws = wave_sum(ab);
        const unsigned wm = wave_umin(cnd);
        if (lane == 0) {
            s_ab[threadIdx.x >> 6] = ws;
            s_cnd[threadIdx.x >> 6] = wm;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            a.partial[blockIdx.x] = ((s_ab[0] + s_ab[1]) + s_ab[2]) + s_ab[3];
            const unsigned m = min(min(s_cnd[0], s_cnd[1]), min(s_cnd[2], s_cnd[3]));
            if (m != 0xffffffffu) atomicMin(a.argmax, (int)m);
        }
    }
}

}  // namespace hrw
