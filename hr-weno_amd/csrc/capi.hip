// capi.hip -- extern "C" entry points of libhrweno_hip.so (declared in include/hrweno_capi.h).
//
// Host side of the MI355X path: objects mirroring the reference's derived types (weno,
// rktvd/mstvd, grid1), the reference's loop control and error texts, and the dispatch of the
// gfx950 kernels in kernels_1d.hpp / kernels_2d.hpp.  All arithmetic on cell data runs in
// those kernels; the host only walks time, counts evaluations and owns buffers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <climits>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "../../include/hrweno_capi.h"
#include "kernels_1d.hpp"
#include "kernels_1d_fast.hpp"
#include "kernels_2d.hpp"
#include "kernels_adjoint.hpp"
#include "launch.hpp"

using namespace hrw;

// ------------------------------------------------------------------ errors ------------
static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(HRW_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

#define LAUNCH_CHECK()                                                                    \
    do {                                                                                  \
        hipError_t e_ = hipGetLastError();                                                \
        if (e_ != hipSuccess)                                                             \
            return fail(HRW_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e_));  \
    } while (0)

#define CALL(expr)                                                                        \
    do {                                                                                  \
        int rc_ = (expr);                                                                 \
        if (rc_ != HRW_OK) return rc_;                                                    \
    } while (0)

static inline hipStream_t S(hrw_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

static bool is_done(double t, double tout, double dt)   // src/hrweno_tvdode.f90:273-284
{
    return (t - tout) * std::copysign(1.0, dt) > 0.0;
}

template <class T> static int dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) return HRW_OK;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
    if (e != hipSuccess)
        return fail(HRW_ERR_ALLOC, "hipMalloc(%zu bytes) failed: %s", count * sizeof(T),
                    hipGetErrorString(e));
    return HRW_OK;
}
template <class T> static void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

static unsigned grid_for(long long n, int per_block, unsigned cap = 1u << 20)
{
    long long g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    return (unsigned)(g > cap ? cap : g);
}

// ------------------------------------------------------------------ runtime -----------
extern "C" const char *hrw_version(void) { return "hrweno-mi355x 0.1.0 (gfx950)"; }
extern "C" const char *hrw_last_error(void) { return g_err.c_str(); }
extern "C" int hrw_last_error_copy(char *buf, int64_t len)
{
    if (!buf || len < 1) return HRW_ERR_ARG;
    const size_t n = std::min((size_t)(len - 1), g_err.size());
    std::memcpy(buf, g_err.data(), n);
    buf[n] = '\0';
    return HRW_OK;
}

extern "C" int hrw_device_count(int *count)
{
    HIP_TRY(hipGetDeviceCount(count));
    return HRW_OK;
}
extern "C" int hrw_device_alloc(size_t bytes, void **ptr)
{
    if (!ptr) return fail(HRW_ERR_ARG, "hrw_device_alloc: null output pointer");
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess)
        return fail(HRW_ERR_ALLOC, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return HRW_OK;
}
extern "C" int hrw_device_mem_info(size_t *free_bytes, size_t *total_bytes)
{
    size_t f = 0, t = 0;
    HIP_TRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return HRW_OK;
}
extern "C" int hrw_device_free(void *ptr)
{
    HIP_TRY(hipFree(ptr));
    return HRW_OK;
}
extern "C" int hrw_memcpy_htod(void *dst, const void *src, size_t bytes, hrw_stream_t s)
{
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(s)));
    HIP_TRY(hipStreamSynchronize(S(s)));
    return HRW_OK;
}
extern "C" int hrw_memcpy_dtoh(void *dst, const void *src, size_t bytes, hrw_stream_t s)
{
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(s)));
    HIP_TRY(hipStreamSynchronize(S(s)));
    return HRW_OK;
}
extern "C" int hrw_memcpy_dtod(void *dst, const void *src, size_t bytes, hrw_stream_t s)
{
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(s)));
    return HRW_OK;
}
extern "C" int hrw_stream_create(hrw_stream_t *stream)
{
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = reinterpret_cast<hrw_stream_t>(s);
    return HRW_OK;
}
extern "C" int hrw_stream_destroy(hrw_stream_t stream)
{
    HIP_TRY(hipStreamDestroy(S(stream)));
    return HRW_OK;
}
extern "C" int hrw_stream_synchronize(hrw_stream_t stream)
{
    HIP_TRY(hipStreamSynchronize(S(stream)));
    return HRW_OK;
}

static thread_local hipEvent_t g_ev[2] = {nullptr, nullptr};
extern "C" int hrw_timer_start(hrw_stream_t stream)
{
    for (auto &e : g_ev)
        if (!e) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipEventRecord(g_ev[0], S(stream)));
    return HRW_OK;
}
extern "C" int hrw_timer_stop(hrw_stream_t stream, double *ms)
{
    if (!g_ev[0]) return fail(HRW_ERR_STATE, "hrw_timer_stop without hrw_timer_start");
    HIP_TRY(hipEventRecord(g_ev[1], S(stream)));
    HIP_TRY(hipEventSynchronize(g_ev[1]));
    float f = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, g_ev[0], g_ev[1]));
    *ms = f;
    return HRW_OK;
}

// ------------------------------------------------------------------ grids -------------
// src/hrweno_grids.f90:41-230 (setup arithmetic, evaluated once on the host).
extern "C" int hrw_grid_linear(double xmin, double xmax, int n, double *edges)
{
    if (xmax <= xmin) return fail(HRW_ERR_ARG, "Invalid input 'xmin', 'xmax'. Valid range: xmax > xmin.");
    if (n < 1) return fail(HRW_ERR_ARG, "Invalid input 'ncells'. Valid range: ncells > 1.");
    const double rx = (xmax - xmin) / n;
    for (int i = 0; i <= n; ++i) edges[i] = linear_edge(xmin, rx, i);
    return HRW_OK;
}
extern "C" int hrw_grid_bilinear(double xmin, double xcross, double xmax, int n1, int n2,
                                 double *edges)
{
    if (xcross <= xmin)
        return fail(HRW_ERR_ARG, "Invalid input 'xmin', 'xcross'. Valid range: xcross > xmin.");
    if (xmax <= xcross)
        return fail(HRW_ERR_ARG, "Invalid input 'xcross', 'xmax'. Valid range: xmax > xcross.");
    if (n1 < 1 || n2 < 1)
        return fail(HRW_ERR_ARG, "Invalid input 'ncells'. Valid range: ncells(i) >= 1.");
    double rx = (xcross - xmin) / n1;
    for (int i = 0; i <= n1; ++i) edges[i] = linear_edge(xmin, rx, i);
    rx = (xmax - xcross) / n2;
    for (int i = 1; i <= n2; ++i) edges[n1 + i] = linear_edge(xcross, rx, i);
    return HRW_OK;
}
extern "C" int hrw_grid_log(double xmin, double xmax, int n, double *edges)
{
    if (xmin <= 0.0) return fail(HRW_ERR_ARG, "Invalid input 'xmin'. Valid range: xmin > 0.");
    if (!(xmax > xmin)) return fail(HRW_ERR_ARG, "Invalid input 'xmin', 'xmax'. Valid range: xmax > xmin.");
    if (n < 1) return fail(HRW_ERR_ARG, "Invalid input 'ncells'. Valid range: ncells > 1.");
    const double rx = (std::log(xmax / xmin)) / n;
    const double l0 = std::log(xmin);
    for (int i = 0; i <= n; ++i) edges[i] = linear_edge(l0, rx, i);
    for (int i = 0; i <= n; ++i) edges[i] = std::exp(edges[i]);
    return HRW_OK;
}
static double powi(double x, int m)   // real**integer by binary powering
{
    unsigned e = m < 0 ? -(unsigned)m : (unsigned)m;
    double y = (e % 2) ? x : 1.0;
    while (e >>= 1) {
        x = x * x;
        if (e % 2) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}
extern "C" int hrw_grid_geometric(double xmin, double xmax, double ratio, int n, double *edges)
{
    if (xmax <= xmin) return fail(HRW_ERR_ARG, "Invalid input 'xmin', 'xmax'. Valid range: xmax > xmin");
    if (ratio <= 0.0) return fail(HRW_ERR_ARG, "Invalid input 'ratio'. Valid range: ratio > 0");
    if (n < 1) return fail(HRW_ERR_ARG, "Invalid input 'ncells'. Valid range: ncells > 1");
    const double a = (xmax - xmin) / (powi(ratio, n) - 1.0);
    for (int i = 0; i <= n; ++i) {
        const double p = powi(ratio, i) - 1;
        const double ap = a * p;
        edges[i] = xmin + ap;
    }
    return HRW_OK;
}
extern "C" int hrw_grid_compute(int n, const double *edges, double *center, double *width)
{
    if (n < 1) return fail(HRW_ERR_ARG, "Invalid input 'ncells'. Valid range: ncells > 1.");
    for (int i = 0; i < n; ++i) {
        if (center) center[i] = (edges[i] + edges[i + 1]) / 2;
        if (width) width[i] = edges[i + 1] - edges[i];
    }
    return HRW_OK;
}

// ------------------------------------------------------------------ WENO --------------
struct hrw_weno_s {
    int n = 0, k = 3, arith = HRW_ARITH_FAST;
    double eps = 1e-6;
    bool uniform = true;
    double *d_cnu = nullptr;                 // (n, k+1, k) on the device
    ~hrw_weno_s() { dfree(d_cnu); }
};

extern "C" int hrw_weno_tables(int k, double *d, double *c)
{
    if (k < 1 || k > 3) return fail(HRW_ERR_ARG, "Invalid input 'k'. Valid range: 1 <= k <= 3.");
    auto put = [&](auto tab) {
        for (int r = 0; r < k; ++r) d[r] = tab.d[r];
        for (int r = 0; r <= k; ++r)
            for (int j = 0; j < k; ++j) c[r * k + j] = tab.c[r][j];
    };
    if (k == 1) put(Tab<1>{});
    else if (k == 2) put(Tab<2>{});
    else put(Tab<3>{});
    return HRW_OK;
}

// cnu(j, r, i), src/hrweno_weno.f90:261-295, one thread per (i, r, j); xl(m)=xext(m-1),
// xr(m)=xext(m) with xext already extended by k+1 ghost edges per side (offset k+1).
__global__ void calc_cnu_kernel(int n, int k, const double *xext, double *cnu)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int per = k * (k + 1);
    if (t >= (long long)n * per) return;
    const int i = (int)(t / per) + 1;
    const int rj = (int)(t % per);
    const int r = rj / k - 1, j = rj % k;
    const int ng = k + 1;
    auto XL = [&](int m) { return xext[m - 1 + ng]; };
    auto XR = [&](int m) { return xext[m + ng]; };
    double sum2 = 0.0;
    for (int m = j + 1; m <= k; ++m) {
        double prod2 = 1.0;
        for (int l = 0; l <= k; ++l) {
            if (l == m) continue;
            prod2 = prod2 * (XL(i - r + m) - XL(i - r + l));
        }
        double sum1 = 0.0;
        for (int l = 0; l <= k; ++l) {
            if (l == m) continue;
            double prod1 = 1.0;
            for (int q = 0; q <= k; ++q) {
                if (q == m || q == l) continue;
                prod1 = prod1 * (XR(i) - XL(i - r + q));
            }
            sum1 = sum1 + prod1;
        }
        sum2 = sum2 + sum1 / prod2;
    }
    cnu[t] = sum2 * (XR(i - r + j) - XL(i - r + j));
}

// Thread-local device staging for host-array calls (grows, never shrinks).
struct Staging {
    double *buf = nullptr;
    size_t cap = 0;
    ~Staging() { dfree(buf); }
    int reserve(size_t count)
    {
        if (count <= cap) return HRW_OK;
        dfree(buf);
        cap = 0;
        CALL(dalloc(&buf, count));
        cap = count;
        return HRW_OK;
    }
};
static thread_local Staging g_stage;

static int upload_strided(double *dst, const double *src, int64_t n, int64_t inc)
{
    if (inc == 1) {
        HIP_TRY(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        HIP_TRY(hipMemcpy2D(dst, sizeof(double), src, inc * sizeof(double), sizeof(double), n,
                            hipMemcpyHostToDevice));
    }
    return HRW_OK;
}
static int download_strided(double *dst, int64_t inc, const double *src, int64_t n)
{
    if (inc == 1) {
        HIP_TRY(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        HIP_TRY(hipMemcpy2D(dst, inc * sizeof(double), src, sizeof(double), sizeof(double), n,
                            hipMemcpyDeviceToHost));
    }
    return HRW_OK;
}

static int check_weno_args(int n, int k, double eps, int arith)
{
    if (!(n > 0)) return fail(HRW_ERR_ARG, "Invalid input 'ncells'. Valid range: ncells > 0.");
    if (!(k >= 1 && k <= 3)) return fail(HRW_ERR_ARG, "Invalid input 'k'. Valid range: 1 <= k <= 3.");
    if (!(eps > 2.220446049250313e-16))
        return fail(HRW_ERR_ARG, "Invalid input 'eps'. Valid range: eps > epsilon.");
    if (arith != HRW_ARITH_EXACT && arith != HRW_ARITH_FAST)
        return fail(HRW_ERR_ARG, "Invalid arithmetic form %d", arith);
    return HRW_OK;
}

// cnu on the device from host edges xedges(0:n) (src/hrweno_weno.f90:241-255 extend the
// grid by k+1 ghost edges on each side; the extension is a handful of host subtractions).
static int compute_cnu(int n, int k, const double *xedges, double *d_cnu)
{
    const int ng = k + 1;
    std::vector<double> xext(n + 1 + 2 * ng);
    for (int i = 0; i <= n; ++i) xext[i + ng] = xedges[i];
    double dx = xext[1 + ng] - xext[0 + ng];               // :246-249
    for (int i = -1; i >= -ng; --i) xext[i + ng] = xext[i + 1 + ng] - dx;
    dx = xext[n + ng] - xext[n - 1 + ng];                  // :252-255
    for (int i = n + 1; i <= n + ng; ++i) xext[i + ng] = xext[i - 1 + ng] + dx;
    double *d_xext = nullptr;
    CALL(dalloc(&d_xext, xext.size()));
    hipError_t e = hipMemcpy(d_xext, xext.data(), xext.size() * sizeof(double),
                             hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const long long tot = (long long)n * k * (k + 1);
        calc_cnu_kernel<<<grid_for(tot, 256, 1u << 30), 256>>>(n, k, d_xext, d_cnu);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    dfree(d_xext);
    if (e != hipSuccess) return fail(HRW_ERR_HIP, "cnu setup failed: %s", hipGetErrorString(e));
    return HRW_OK;
}

extern "C" int hrw_weno_calc_cnu(int ncells, int k, const double *xedges, double *cnu_host)
{
    CALL(check_weno_args(ncells, k, 1.0, HRW_ARITH_EXACT));
    if (!xedges || !cnu_host) return fail(HRW_ERR_ARG, "null array");
    const size_t count = (size_t)ncells * k * (k + 1);
    CALL(g_stage.reserve(count));
    CALL(compute_cnu(ncells, k, xedges, g_stage.buf));
    HIP_TRY(hipMemcpy(cnu_host, g_stage.buf, count * sizeof(double), hipMemcpyDeviceToHost));
    return HRW_OK;
}

extern "C" int hrw_weno_create(int ncells, int k, double eps, const double *xedges, int arith,
                               hrw_weno_t *out)
{
    if (!out) return fail(HRW_ERR_ARG, "hrw_weno_create: null output pointer");
    *out = nullptr;
    CALL(check_weno_args(ncells, k, eps, arith));          // src/hrweno_weno.f90:72-98
    auto *w = new (std::nothrow) hrw_weno_s;
    if (!w) return fail(HRW_ERR_ALLOC, "out of host memory");
    w->n = ncells;
    w->k = k;
    w->eps = eps;
    w->arith = arith;
    if (xedges) {   // :100-109
        w->uniform = false;
        int rc = dalloc(&w->d_cnu, (size_t)ncells * k * (k + 1));
        if (rc == HRW_OK) rc = compute_cnu(ncells, k, xedges, w->d_cnu);
        if (rc != HRW_OK) {
            delete w;
            return rc;
        }
    }
    *out = w;
    return HRW_OK;
}

extern "C" int hrw_weno_destroy(hrw_weno_t w)
{
    delete w;
    return HRW_OK;
}

extern "C" int hrw_weno_info(hrw_weno_t w, int *ncells, int *k, double *eps, int *uniform)
{
    if (!w) return fail(HRW_ERR_ARG, "null weno object");
    if (ncells) *ncells = w->n;
    if (k) *k = w->k;
    if (eps) *eps = w->eps;
    if (uniform) *uniform = w->uniform ? 1 : 0;
    return HRW_OK;
}

extern "C" int hrw_weno_get_cnu(hrw_weno_t w, double *cnu_host)
{
    if (!w) return fail(HRW_ERR_ARG, "null weno object");
    if (w->uniform) return fail(HRW_ERR_STATE, "uniform grid: no cnu table");
    HIP_TRY(hipMemcpy(cnu_host, w->d_cnu, (size_t)w->n * w->k * (w->k + 1) * sizeof(double),
                      hipMemcpyDeviceToHost));
    return HRW_OK;
}

// Reconstruction of nlines lines on the device; cnu (device, or nullptr) selects the
// non-uniform coefficients.
struct ReconArgs {
    long long n;
    int k, arith;
    double eps;
    const double *cnu;
    long long nlines;
    const double *v;
    long long incv, ldv;
    double *vl;
    long long incl, ldl;
    double *vr;
    long long incr, ldr;
};

template <int K> static void launch_reconstruct(const ReconArgs &a, hipStream_t s)
{
    dim3 grid(grid_for(a.n, kBlock, 1u << 30), (unsigned)a.nlines);
    if (a.cnu)
        reconstruct_kernel<K, true, false><<<grid, kBlock, 0, s>>>(
            a.v, a.incv, a.ldv, a.n, a.eps, a.cnu, a.vl, a.incl, a.ldl, a.vr, a.incr, a.ldr);
    else if (a.arith == HRW_ARITH_FAST)
        reconstruct_kernel<K, false, true><<<grid, kBlock, 0, s>>>(
            a.v, a.incv, a.ldv, a.n, a.eps, nullptr, a.vl, a.incl, a.ldl, a.vr, a.incr, a.ldr);
    else
        reconstruct_kernel<K, false, false><<<grid, kBlock, 0, s>>>(
            a.v, a.incv, a.ldv, a.n, a.eps, nullptr, a.vl, a.incl, a.ldl, a.vr, a.incr, a.ldr);
}

static int reconstruct_core(const ReconArgs &a, hipStream_t s)
{
    if (a.nlines < 1 || a.nlines > 65535) return fail(HRW_ERR_ARG, "nlines must be in [1, 65535]");
    if (!a.v || !a.vl || !a.vr) return fail(HRW_ERR_ARG, "null array");
    if (a.k == 1) launch_reconstruct<1>(a, s);
    else if (a.k == 2) launch_reconstruct<2>(a, s);
    else launch_reconstruct<3>(a, s);
    LAUNCH_CHECK();
    return HRW_OK;
}

extern "C" int hrw_weno_reconstruct_lines(hrw_weno_t w, int64_t nlines, const double *v,
                                          int64_t incv, int64_t ldv, double *vl, double *vr,
                                          int64_t inco, int64_t ldo, hrw_stream_t stream)
{
    if (!w) return fail(HRW_ERR_ARG, "null weno object");
    ReconArgs a{w->n, w->k, w->arith, w->eps, w->d_cnu, nlines, v, incv, ldv,
                vl, inco, ldo, vr, inco, ldo};
    return reconstruct_core(a, S(stream));
}

extern "C" int hrw_weno_reconstruct(hrw_weno_t w, const double *v, int64_t incv, double *vl,
                                    int64_t incl, double *vr, int64_t incr, hrw_stream_t stream)
{
    if (!w) return fail(HRW_ERR_ARG, "null weno object");
    ReconArgs a{w->n, w->k, w->arith, w->eps, w->d_cnu, 1, v, incv, 0, vl, incl, 0, vr, incr, 0};
    return reconstruct_core(a, S(stream));
}

extern "C" int hrw_reconstruct_host(int n, int k, double eps, int arith, const double *cnu,
                                    const double *v, int64_t incv, double *vl, int64_t incl,
                                    double *vr, int64_t incr)
{
    CALL(check_weno_args(n, k, eps, arith));
    const size_t ncnu = cnu ? (size_t)n * k * (k + 1) : 0;
    CALL(g_stage.reserve(3 * (size_t)n + ncnu));
    double *dv = g_stage.buf, *dl = dv + n, *dr = dl + n, *dc = cnu ? dr + n : nullptr;
    CALL(upload_strided(dv, v, n, incv));
    if (cnu) HIP_TRY(hipMemcpy(dc, cnu, ncnu * sizeof(double), hipMemcpyHostToDevice));
    ReconArgs a{n, k, arith, eps, dc, 1, dv, 1, 0, dl, 1, 0, dr, 1, 0};
    CALL(reconstruct_core(a, nullptr));
    CALL(download_strided(vl, incl, dl, n));
    CALL(download_strided(vr, incr, dr, n));
    return HRW_OK;
}

extern "C" int hrw_reconstruct_dev(int n, int k, double eps, int arith, const double *v,
                                   int64_t incv, double *vl, int64_t incl, double *vr,
                                   int64_t incr, hrw_stream_t stream)
{
    CALL(check_weno_args(n, k, eps, arith));
    ReconArgs a{n, k, arith, eps, nullptr, 1, v, incv, 0, vl, incl, 0, vr, incr, 0};
    return reconstruct_core(a, S(stream));
}

// Elements addressed by nlines lines of n elements (element i of line l at l*ld + i*inc).
static int64_t lines_span(int64_t n, int64_t nlines, int64_t inc, int64_t ld)
{
    return (nlines - 1) * ld + (n - 1) * inc + 1;
}

static int check_lines(int64_t n, int64_t nlines, int64_t incv, int64_t ldv, int64_t inco,
                       int64_t ldo)
{
    if (nlines < 1 || nlines > 65535) return fail(HRW_ERR_ARG, "nlines must be in [1, 65535]");
    if (incv < 1 || inco < 1 || ldv < 0 || ldo < 0)
        return fail(HRW_ERR_ARG, "line strides must be positive (inc >= 1, ld >= 0)");
    (void)n;
    return HRW_OK;
}

extern "C" int hrw_reconstruct_lines_dev(int ncells, int k, double eps, int arith, int64_t nlines,
                                         const double *v, int64_t incv, int64_t ldv, double *vl,
                                         double *vr, int64_t inco, int64_t ldo,
                                         hrw_stream_t stream)
{
    CALL(check_weno_args(ncells, k, eps, arith));
    CALL(check_lines(ncells, nlines, incv, ldv, inco, ldo));
    ReconArgs a{ncells, k, arith, eps, nullptr, nlines, v, incv, ldv, vl, inco, ldo, vr, inco, ldo};
    return reconstruct_core(a, S(stream));
}

extern "C" int hrw_reconstruct_lines_host(int ncells, int k, double eps, int arith,
                                          const double *cnu, int64_t nlines, const double *v,
                                          int64_t incv, int64_t ldv, double *vl, double *vr,
                                          int64_t inco, int64_t ldo)
{
    CALL(check_weno_args(ncells, k, eps, arith));
    CALL(check_lines(ncells, nlines, incv, ldv, inco, ldo));
    if (!v || !vl || !vr) return fail(HRW_ERR_ARG, "null array");
    const int64_t sv = lines_span(ncells, nlines, incv, ldv);
    const int64_t so = lines_span(ncells, nlines, inco, ldo);
    const size_t ncnu = cnu ? (size_t)ncells * k * (k + 1) : 0;
    CALL(g_stage.reserve((size_t)sv + 2 * (size_t)so + ncnu));
    double *dv = g_stage.buf, *dl = dv + sv, *dr = dl + so, *dc = cnu ? dr + so : nullptr;
    HIP_TRY(hipMemcpy(dv, v, (size_t)sv * sizeof(double), hipMemcpyHostToDevice));
    if (cnu) HIP_TRY(hipMemcpy(dc, cnu, ncnu * sizeof(double), hipMemcpyHostToDevice));
    // the outputs' untouched elements (gaps between strided lines) keep the caller's values
    HIP_TRY(hipMemcpy(dl, vl, (size_t)so * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dr, vr, (size_t)so * sizeof(double), hipMemcpyHostToDevice));
    ReconArgs a{ncells, k, arith, eps, dc, nlines, dv, incv, ldv, dl, inco, ldo, dr, inco, ldo};
    CALL(reconstruct_core(a, nullptr));
    HIP_TRY(hipMemcpy(vl, dl, (size_t)so * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(vr, dr, (size_t)so * sizeof(double), hipMemcpyDeviceToHost));
    return HRW_OK;
}

extern "C" int hrw_weno_reconstruct_host(hrw_weno_t w, const double *v, int64_t incv,
                                         double *vl, double *vr)
{
    if (!w) return fail(HRW_ERR_ARG, "null weno object");
    CALL(g_stage.reserve(3 * (size_t)w->n));
    const int64_t n = w->n;
    double *dv = g_stage.buf, *dl = dv + n, *dr = dl + n;
    CALL(upload_strided(dv, v, n, incv));
    ReconArgs a{w->n, w->k, w->arith, w->eps, w->d_cnu, 1, dv, 1, 0, dl, 1, 0, dr, 1, 0};
    CALL(reconstruct_core(a, nullptr));
    CALL(download_strided(vl, 1, dl, n));
    CALL(download_strided(vr, 1, dr, n));
    return HRW_OK;
}

// ------------------------------------------------------------------ fluxes ------------
static int flux_common(int which, int fn, double a, double b, int64_t n, const double *vm,
                       const double *vp, const double *x, double t, double alpha, double *h,
                       hrw_stream_t stream)
{
    if (fn != HRW_FN_BURGERS && fn != HRW_FN_LINEAR) return fail(HRW_ERR_ARG, "unknown flux function %d", fn);
    if (n < 0) return fail(HRW_ERR_ARG, "negative size");
    if (n == 0) return HRW_OK;
    flux_kernel<<<grid_for(n, kBlock, 65536), kBlock, 0, S(stream)>>>(which, fn, a, b, n, vm, vp,
                                                                       x, t, alpha, h);
    LAUNCH_CHECK();
    return HRW_OK;
}
extern "C" int hrw_flux_lax_friedrichs(int fn, double a, double b, int64_t n, const double *vm,
                                       const double *vp, const double *x, double t, double alpha,
                                       double *h, hrw_stream_t stream)
{
    return flux_common(0, fn, a, b, n, vm, vp, x, t, alpha, h, stream);
}
extern "C" int hrw_flux_godunov(int fn, double a, double b, int64_t n, const double *vm,
                                const double *vp, const double *x, double t, double *h,
                                hrw_stream_t stream)
{
    return flux_common(1, fn, a, b, n, vm, vp, x, t, 0.0, h, stream);
}

// ------------------------------------------------------------------ problems ----------
enum { PROB_BURGERS1D = 1, PROB_PBE2D = 2 };

struct hrw_problem_s {
    int kind = 0;
    int64_t neq = 0;
    int k = 3, arith = HRW_ARITH_FAST, numflux = HRW_FLUX_LAX_FRIEDRICHS, bc = HRW_BC_PERIODIC;
    double eps = 1e-6, alpha = -1.0;
    // 1D
    double xmin = 0, rx = 0;
    double *d_width = nullptr;
    double *d_cnu = nullptr;          // non-uniform grid: (n, k+1, k)
    // 2D
    int n1 = 0, n2 = 0, ty = 0;      // rows per wavefront, 0 = chosen at launch
    double x1min = 0, rx1 = 0, x2min = 0, rx2 = 0, g1a = 0, g1b = 0, g2a = 0, g2b = 0;
    // kernel choice for the 1D FAST form: 0 = register-blocked (cells_per_lane, 0 = by stage),
    // 1 = LDS tile, 2 = register-blocked persistent
    int variant = 0, cells_per_lane = 0;
    // alpha slot sets for rhs-only calls
    unsigned long long *slots = nullptr;
    bool alphamax() const { return kind == PROB_BURGERS1D && numflux == HRW_FLUX_LAX_FRIEDRICHS && alpha < 0; }
    ~hrw_problem_s()
    {
        dfree(d_width);
        dfree(d_cnu);
        dfree(slots);
    }
};

extern "C" int hrw_problem_burgers1d_edges(int n, int k, double eps, int numflux, int bc,
                                           double alpha, const double *xedges, int arith,
                                           hrw_problem_t *out)
{
    if (!out) return fail(HRW_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!xedges) return fail(HRW_ERR_ARG, "null xedges");
    if (bc != HRW_BC_ZERO_GRADIENT)
        return fail(HRW_ERR_ARG, "a non-uniform grid (xedges) needs HRW_BC_ZERO_GRADIENT");
    if (!(n > 0)) return fail(HRW_ERR_ARG, "Invalid input 'ncells'. Valid range: ncells > 0.");
    std::vector<double> width((size_t)n);
    for (int i = 0; i < n; ++i) width[(size_t)i] = xedges[i + 1] - xedges[i];   // src/hrweno_grids.f90:247
    hrw_problem_t p = nullptr;
    CALL(hrw_problem_burgers1d(n, k, eps, numflux, bc, alpha, xedges[0], xedges[n], width.data(),
                               arith, &p));
    int rc = dalloc(&p->d_cnu, (size_t)n * k * (k + 1));
    if (rc == HRW_OK) rc = compute_cnu(n, k, xedges, p->d_cnu);
    if (rc != HRW_OK) {
        delete p;
        return rc;
    }
    *out = p;
    return HRW_OK;
}

extern "C" int hrw_problem_burgers1d(int n, int k, double eps, int numflux, int bc, double alpha,
                                     double xmin, double xmax, const double *width, int arith,
                                     hrw_problem_t *out)
{
    if (!out) return fail(HRW_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!(n > 0)) return fail(HRW_ERR_ARG, "Invalid input 'ncells'. Valid range: ncells > 0.");
    if (!(k >= 1 && k <= 3)) return fail(HRW_ERR_ARG, "Invalid input 'k'. Valid range: 1 <= k <= 3.");
    if (!(eps > 2.220446049250313e-16))
        return fail(HRW_ERR_ARG, "Invalid input 'eps'. Valid range: eps > epsilon.");
    if (numflux != HRW_FLUX_LAX_FRIEDRICHS && numflux != HRW_FLUX_GODUNOV)
        return fail(HRW_ERR_ARG, "unknown numerical flux %d", numflux);
    if (bc != HRW_BC_PERIODIC && bc != HRW_BC_ZERO_GRADIENT)
        return fail(HRW_ERR_ARG, "1D Burgers supports HRW_BC_PERIODIC and HRW_BC_ZERO_GRADIENT");
    if (arith != HRW_ARITH_EXACT && arith != HRW_ARITH_FAST)
        return fail(HRW_ERR_ARG, "Invalid arithmetic form %d", arith);
    if (!width && !(xmax > xmin))
        return fail(HRW_ERR_ARG, "Invalid input 'xmin', 'xmax'. Valid range: xmax > xmin.");
    auto *p = new (std::nothrow) hrw_problem_s;
    if (!p) return fail(HRW_ERR_ALLOC, "out of host memory");
    p->kind = PROB_BURGERS1D;
    p->neq = n;
    p->k = k;
    p->eps = eps;
    p->numflux = numflux;
    p->bc = bc;
    p->alpha = alpha;
    p->arith = arith;
    p->xmin = xmin;
    p->rx = (xmax - xmin) / n;   // src/hrweno_grids.f90:76
    int rc = HRW_OK;
    if (width) {
        rc = dalloc(&p->d_width, (size_t)n);
        if (rc == HRW_OK && hipMemcpy(p->d_width, width, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(HRW_ERR_HIP, "width upload failed");
    }
    if (rc == HRW_OK) rc = dalloc(&p->slots, (size_t)3 * kNSlot * kSlotStride);
    if (rc != HRW_OK) {
        delete p;
        return rc;
    }
    *out = p;
    return HRW_OK;
}

extern "C" int hrw_problem_pbe2d(int n1, int n2, int k, double eps, int bc, double x1min,
                                 double x1max, double x2min, double x2max, double g1a, double g1b,
                                 double g2a, double g2b, int arith, hrw_problem_t *out)
{
    if (!out) return fail(HRW_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!(n1 > 0 && n2 > 0)) return fail(HRW_ERR_ARG, "Invalid input 'ncells'. Valid range: ncells > 0.");
    if ((long long)n1 * n2 > (1ll << 31) - 1) return fail(HRW_ERR_ARG, "n1*n2 must fit a 32-bit integer");
    if (!(k >= 1 && k <= 3)) return fail(HRW_ERR_ARG, "Invalid input 'k'. Valid range: 1 <= k <= 3.");
    if (!(eps > 2.220446049250313e-16))
        return fail(HRW_ERR_ARG, "Invalid input 'eps'. Valid range: eps > epsilon.");
    if (bc != HRW_BC_ZERO_FLUX && bc != HRW_BC_INFLOW0_OUTFLOW)
        return fail(HRW_ERR_ARG, "2D PBE supports HRW_BC_ZERO_FLUX and HRW_BC_INFLOW0_OUTFLOW");
    if (!(x1max > x1min) || !(x2max > x2min))
        return fail(HRW_ERR_ARG, "Invalid input 'xmin', 'xmax'. Valid range: xmax > xmin.");
    if (arith != HRW_ARITH_EXACT && arith != HRW_ARITH_FAST)
        return fail(HRW_ERR_ARG, "Invalid arithmetic form %d", arith);
    auto *p = new (std::nothrow) hrw_problem_s;
    if (!p) return fail(HRW_ERR_ALLOC, "out of host memory");
    p->kind = PROB_PBE2D;
    p->neq = (int64_t)n1 * n2;
    p->n1 = n1;
    p->n2 = n2;
    p->k = k;
    p->eps = eps;
    p->bc = bc;
    p->numflux = HRW_FLUX_GODUNOV;
    p->arith = arith;
    p->x1min = x1min;
    p->rx1 = (x1max - x1min) / n1;
    p->x2min = x2min;
    p->rx2 = (x2max - x2min) / n2;
    p->g1a = g1a;
    p->g1b = g1b;
    p->g2a = g2a;
    p->g2b = g2b;
    *out = p;
    return HRW_OK;
}

extern "C" int hrw_problem_tune(hrw_problem_t p, int variant, int cells_per_lane)
{
    if (!p) return fail(HRW_ERR_ARG, "null problem");
    if (p->kind == PROB_PBE2D) {
        if (variant != 0) return fail(HRW_ERR_ARG, "variant must be 0 for a 2D problem");
        if (cells_per_lane != 0 && (cells_per_lane < 1 || cells_per_lane > 4096))
            return fail(HRW_ERR_ARG, "rows per wavefront must be 0 (automatic) or 1..4096");
        p->ty = cells_per_lane;
        return HRW_OK;
    }
    if (variant < 0 || variant > 2) return fail(HRW_ERR_ARG, "variant must be 0, 1 or 2");
    if (cells_per_lane != 0 && cells_per_lane != 4 && cells_per_lane != 6 && cells_per_lane != 8)
        return fail(HRW_ERR_ARG, "cells_per_lane must be 0 (automatic), 4, 6 or 8");
    p->variant = variant;
    p->cells_per_lane = cells_per_lane;
    return HRW_OK;
}

extern "C" int hrw_problem_destroy(hrw_problem_t p)
{
    delete p;
    return HRW_OK;
}
extern "C" int hrw_problem_neq(hrw_problem_t p, int64_t *neq)
{
    if (!p) return fail(HRW_ERR_ARG, "null problem");
    *neq = p->neq;
    return HRW_OK;
}

// ---- kernel dispatch ----------------------------------------------------------------
static bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

// Every edge xmin + rx*i (i = 0..n) of a linear grid is computed without rounding, so every
// width e(i+1) - e(i) (src/hrweno_grids.f90:247) is exactly rx: with rx = Mr*2^er and
// xmin = Mx*2^ex (Mr, Mx odd integers), rx*i is exact while Mr*n < 2^53, and xmin + rx*i, a
// multiple of 2^min(er, ex), is exact while its magnitude stays below 2^(53 + min(er, ex)).
// (Grids of 2^m cells on [-1, 1] or [0, 1], for instance.)
static bool exact_linear_grid(double xmin, double rx, long long n)
{
    if (!(rx > 0.0) || !std::isfinite(rx) || !std::isfinite(xmin) || n < 1) return false;
    auto odd_significand = [](double x, int &e) {   // |x| = M * 2^e, M odd
        int ex = 0;
        const double m = std::frexp(std::fabs(x), &ex);
        unsigned long long M = (unsigned long long)std::ldexp(m, 53);
        e = ex - 53;
        const int tz = __builtin_ctzll(M);
        M >>= tz;
        e += tz;
        return M;
    };
    int er = 0;
    const unsigned long long Mr = odd_significand(rx, er);
    if ((long double)Mr * (long double)n >= 9007199254740992.0L) return false;
    int e = er;
    if (xmin != 0.0) {
        int ex = 0;
        (void)odd_significand(xmin, ex);
        e = std::min(e, ex);
    }
    const long double big = std::max(std::fabs((long double)xmin),
                                     std::fabs((long double)xmin + (long double)rx * (long double)n));
    return big < std::ldexp(1.0L, 53 + e);
}

static void launch_stage1d(const hrw_problem_s *p, const Stage1DArgs &a, hipStream_t s)
{
    const bool amax0 = p->alphamax();
    // the register-blocked kernels take linear grids whose edges stay within 1e8 cell widths
    // of the origin (their 1/width is one Newton step from 1/rx); a width array, a non-uniform
    // grid's coefficients or a grid far from the origin go to the tile kernel
    const bool linear_near = !a.cnu && !a.width &&
                             std::max(std::fabs(a.xmin), std::fabs(a.xmin + a.rx * (double)a.n)) <=
                                 1e8 * a.rx;
    if (p->arith == HRW_ARITH_FAST && p->variant != 1 && linear_near &&
        a.n < (1LL << 31) - (1LL << 20) &&
        aligned16(a.us) && aligned16(a.u0) && aligned16(a.uo) && aligned16(a.uo4) &&
        aligned16(a.Lo4) && aligned16(a.Lout)) {
        const bool per = p->bc == HRW_BC_PERIODIC;
        // automatic (variant 0, cells_per_lane 0), measured best on 2^16..2^24 cells: lines of
        // 2^21 cells or more (several chunks per resident wavefront) the persistent form with 4
        // cells per lane; 2^20..2^21 one chunk per wavefront with 6; shorter lines one chunk per
        // wavefront with 4 (more wavefronts to hide the load latency; with K = 3 the 4-cell
        // kernel also shares its neighbours' triples).
        const bool automatic = p->variant == 0 && p->cells_per_lane == 0;
        const bool persistent = p->variant == 2 || (automatic && a.n >= (1LL << 21));
        const int cpl = p->cells_per_lane ? p->cells_per_lane
                                          : (!persistent && a.n >= (1LL << 20) ? 6 : 4);
        const int form = persistent ? FORM_PERSISTENT : FORM_CHUNK;
        // every width exactly rx: the kernels take 1/width once per chunk (the same value)
        const bool constw = exact_linear_grid(a.xmin, a.rx, a.n);
        if (p->k == 1) launch_fast_1d_k1(p->numflux, per, amax0, constw, form, cpl, a, s);
        else if (p->k == 2) launch_fast_1d_k2(p->numflux, per, amax0, constw, form, cpl, a, s);
        else launch_fast_1d_k3(p->numflux, per, amax0, constw, form, cpl, a, s);
        return;
    }
    const bool per = p->bc == HRW_BC_PERIODIC, fast = p->arith == HRW_ARITH_FAST;
    const bool amax = p->alphamax();
    launch_tile_1d(p->k, p->numflux, per, fast, amax, a, s);
}

static void launch_stage2d(const hrw_problem_s *p, const Stage2DArgs &a, hipStream_t s)
{
    launch_pbe2d(p->k, p->bc, p->arith == HRW_ARITH_FAST, a, s);
}

// One fused pass: out = stage(us, u0, L(us)) (and/or Lout = L(us)).  `seq` indexes the
// alpha slot sets when alpha = max|v|.
struct StageIO {
    const double *us = nullptr, *u0 = nullptr, *uo4 = nullptr, *Lo4 = nullptr;
    double *uo = nullptr, *Lout = nullptr;
    double *alpha_record = nullptr;   // 1D: receives the Lax-Friedrichs alpha of us
};
// The same with the three alpha slot sets given explicitly (read / accumulate / clear).
static int run_stage_sets(const hrw_problem_s *p, const unsigned long long *ain,
                          unsigned long long *aout, unsigned long long *aclear, int stage,
                          double h, const StageIO &io, hipStream_t s);
static int run_stage(const hrw_problem_s *p, unsigned long long *slots, int seq, int stage,
                     double h, const StageIO &io, hipStream_t s)
{
    const size_t set = (size_t)kNSlot * kSlotStride;
    if (p->alphamax())
        return run_stage_sets(p, slots + (seq % 3) * set, slots + ((seq + 1) % 3) * set,
                              slots + ((seq + 2) % 3) * set, stage, h, io, s);
    return run_stage_sets(p, nullptr, nullptr, nullptr, stage, h, io, s);
}
static int run_stage_sets(const hrw_problem_s *p, const unsigned long long *ain,
                          unsigned long long *aout, unsigned long long *aclear, int stage,
                          double h, const StageIO &io, hipStream_t s)
{
    if (p->kind == PROB_BURGERS1D) {
        Stage1DArgs a{};
        a.us = io.us;
        a.u0 = io.u0;
        a.uo = io.uo;
        a.uo4 = io.uo4;
        a.Lo4 = io.Lo4;
        a.Lout = io.Lout;
        a.width = p->d_width;
        a.cnu = p->d_cnu;
        a.xmin = p->xmin;
        a.rx = p->rx;
        a.n = p->neq;
        a.eps = p->eps;
        a.h = h;
        a.alpha = p->alpha;
        a.stage = stage;
        if (p->alphamax()) {
            a.alpha_in = ain;
            a.alpha_out = aout;
            a.alpha_clear = aclear;
            a.alpha_record = io.alpha_record;
        }
        launch_stage1d(p, a, s);
    } else {
        Stage2DArgs a{};
        a.us = io.us;
        a.u0 = io.u0;
        a.uo = io.uo;
        a.uo4 = io.uo4;
        a.Lo4 = io.Lo4;
        a.Lout = io.Lout;
        a.n1 = p->n1;
        a.n2 = p->n2;
        a.x1min = p->x1min;
        a.rx1 = p->rx1;
        a.x2min = p->x2min;
        a.rx2 = p->rx2;
        a.g1a = p->g1a;
        a.g1b = p->g1b;
        a.g2a = p->g2a;
        a.g2b = p->g2b;
        a.eps = p->eps;
        a.h = h;
        a.stage = stage;
        a.ty = p->ty;
        const int sw = strip_width(p->k, p->arith == HRW_ARITH_FAST);
        a.nstrips = (p->n1 + sw - 1) / sw;
        launch_stage2d(p, a, s);
    }
    LAUNCH_CHECK();
    return HRW_OK;
}

// Seed alpha = max|u| into slot set 0 (all sets cleared first).
static int seed_alpha(const hrw_problem_s *p, unsigned long long *slots, const double *u,
                      hipStream_t s)
{
    if (!p->alphamax()) return HRW_OK;
    clear_slots_kernel<<<grid_for(3 * kNSlot, 256), 256, 0, s>>>(slots, 3);
    maxabs_kernel<<<grid_for(p->neq, kBlock * 8, 2048), kBlock, 0, s>>>(u, p->neq, slots);
    LAUNCH_CHECK();
    return HRW_OK;
}

extern "C" int hrw_problem_rhs(hrw_problem_t p, double t, const double *u, double *udot,
                               hrw_stream_t stream)
{
    (void)t;
    if (!p) return fail(HRW_ERR_ARG, "null problem");
    if (!u || !udot) return fail(HRW_ERR_ARG, "null array");
    CALL(seed_alpha(p, p->slots, u, S(stream)));
    StageIO io;
    io.us = u;
    io.Lout = udot;
    return run_stage(p, p->slots, 0, -1, 0.0, io, S(stream));
}

// ------------------------------------------------------------------ integrators -------
// ---- the general reverse pass of a 1D Burgers RHS (either flux, either kind of ends) ------
// Scratch of rhs_vjp_kernel: a slot set for max|s|, one alpha-bar partial per block and the
// argmax word (zero and INT_MAX between passes: every fixup clears and resets them).
struct GenScratch {
    unsigned long long *aset = nullptr;
    double *partial = nullptr;
    int *argmax = nullptr;
    GenScratch() = default;
    GenScratch(const GenScratch &) = delete;
    GenScratch &operator=(const GenScratch &) = delete;
    ~GenScratch()
    {
        dfree(aset);
        dfree(partial);
        dfree(argmax);
    }
    int alloc(int64_t n)
    {
        if (aset) return HRW_OK;
        CALL(dalloc(&aset, (size_t)kNSlot * kSlotStride));
        CALL(dalloc(&partial, (size_t)((n + kGenTile - 1) / kGenTile)));
        CALL(dalloc(&argmax, 1));
        const int none = INT_MAX;
        HIP_TRY(hipMemcpy(argmax, &none, sizeof none, hipMemcpyHostToDevice));
        const std::vector<unsigned long long> zero((size_t)kNSlot * kSlotStride, 0ull);
        HIP_TRY(hipMemcpy(aset, zero.data(), zero.size() * sizeof(unsigned long long),
                          hipMemcpyHostToDevice));
        return HRW_OK;
    }
};
struct GenTerm {
    const double *x;
    double e;
};
// out = sum e_i x_i + ch J(s)^T (b1 y1 + b2 y2) for problem p (the Lax-Friedrichs alpha of s
// taken as max|s| when the problem says so); out must not alias y1 or y2.  K = 2, 3 take the
// register-blocked pass (rhs_vjp5_kernel) unless the problem is tuned to the LDS-tile kernels
// (variant 1); K = 1 the LDS-tile pass (rhs_vjp_kernel).
static int gen_vjp(const hrw_problem_s *p, GenScratch &g, hipStream_t st, const double *s,
                   const double *y1, double b1, const double *y2, double b2,
                   std::initializer_list<GenTerm> terms, double *out, double ch,
                   const double *alpha_dev = nullptr)
{
    const bool amax = p->alphamax();
    if ((p->k == 2 || p->k == 3) && p->variant != 1 && p->neq > 0) {
        // the register-blocked pass (pairs of cells as 16-byte loads and stores): the terms on
        // y1 or y2 become their coefficients
        BvjpArgs b{};
        b.s = s;
        b.y1 = y1;
        b.b1 = b1;
        b.y2 = y2;
        b.b2 = y2 ? b2 : 0.0;
        int nx = 0;
        bool fits = true;
        for (const GenTerm &q : terms) {
            if (!q.x) continue;
            if (q.x == y1) b.ey1 += q.e;
            else if (y2 && q.x == y2) b.ey2 += q.e;
            else if (nx < 2) {
                b.x[nx] = q.x;
                b.ex[nx] = q.e;
                ++nx;
            } else {
                fits = false;
            }
        }
        fits = fits && aligned16(s) && aligned16(y1) && aligned16(y2) && aligned16(b.x[0]) &&
               aligned16(b.x[1]) && aligned16(out);
        if (fits) {
            b.out = out;
            b.ch = ch;
            b.eps = p->eps;
            b.alpha = p->alpha;
            b.xmin = p->xmin;
            b.rx = p->rx;
            b.constw = exact_linear_grid(p->xmin, p->rx, p->neq) ? 1 : 0;
            b.n = p->neq;
            b.nchunks = (p->neq + 60LL * kBvjpR - 1) / (60LL * kBvjpR);
            if (amax) {
                if (alpha_dev) b.alpha_ptr = alpha_dev;   // max|s| recorded by the forward
                else maxabs_kernel<<<grid_for(p->neq, kBlock * 8, 2048), kBlock, 0, st>>>(s, p->neq, g.aset);
                b.aset = g.aset;
                b.partial = g.partial;
                b.argmax = g.argmax;
            }
            launch_rhs_vjp5(p->k, p->numflux, p->bc == HRW_BC_PERIODIC, amax, b, st);
            if (amax) launch_rhs_vjp5_fixup(b, st);
            LAUNCH_CHECK();
            return HRW_OK;
        }
    }
    GenAdjArgs a{};
    a.s = s;
    a.y1 = y1;
    a.b1 = b1;
    a.y2 = y2;
    a.b2 = b2;
    int i = 0;
    for (const GenTerm &q : terms)
        if (q.x && i < 4) {
            a.x[i] = q.x;
            a.e[i] = q.e;
            ++i;
        }
    a.out = out;
    a.ch = ch;
    a.n = p->neq;
    a.xmin = p->xmin;
    a.rx = p->rx;
    a.eps = p->eps;
    a.alpha = p->alpha;
    if (amax) {
        maxabs_kernel<<<grid_for(p->neq, kBlock * 8, 2048), kBlock, 0, st>>>(s, p->neq, g.aset);
        a.aset = g.aset;
        a.partial = g.partial;
        a.argmax = g.argmax;
    }
    launch_rhs_vjp(p->k, p->numflux, p->bc == HRW_BC_PERIODIC, amax, a, st);
    if (amax) launch_rhs_vjp_fixup(a, st);
    LAUNCH_CHECK();
    return HRW_OK;
}

enum { METHOD_RK = 1, METHOD_MS = 2 };

struct hrw_ode_s {
    int method = METHOD_RK, order = 3, fevals = 0, istate = 0, arith = HRW_ARITH_FAST;
    int64_t neq = 0;
    hrw_problem_s *p = nullptr;           // fused device problem, or
    hrw_integrand_fn fu = nullptr;        // a callback integrand
    void *ctx = nullptr;
    hipStream_t stream = nullptr;
    double *ui = nullptr, *ui2 = nullptr, *udot = nullptr;
    double *S = nullptr;                   // mstvd: current state
    double *ring_u[4] = {}, *ring_L[4] = {};
    int head = 0;                          // ring index of column 4 (oldest)
    unsigned long long *slots = nullptr;
    int tape_every = 0;                    // > 0: integrate records for hrw_ode_adjoint
    hrw_adjoint_s *adj = nullptr;
    // mstvd recording: every state since the method started (S_0 .. S_N), the two inner stage
    // inputs of each of the four RK3 starting steps, each step's dt; the reverse sweep's
    // cotangent ring (6) and stage cotangents (2), and the general pass's scratch
    std::vector<double *> ms_tape;
    // the tape's storage, one block (pointer, states) per recorded call; the blocks of a
    // recording that ended are kept for the next one (ms_pool)
    std::vector<std::pair<double *, int64_t>> ms_blocks, ms_pool;
    double *ms_next = nullptr;             // next free state of the last block
    int64_t ms_free = 0;                   // states left in it
    double *ms_start[8] = {};
    std::vector<double> ms_dts;
    double *ms_lam[8] = {};
    GenScratch gen;
    bool check_finite = true;              // integrate ends with a scan of u for NaN/Inf
    unsigned long long *d_bad = nullptr;   // the scan's result word
    ~hrw_ode_s();
};
// Drop the recording; keep_storage: its blocks go to the pool for the next recording.
static void ms_tape_clear(hrw_ode_s *o, bool keep_storage = false)
{
    for (auto &b : o->ms_blocks) {
        if (keep_storage) o->ms_pool.push_back(b);
        else dfree(b.first);
    }
    if (!keep_storage) {
        for (auto &b : o->ms_pool) dfree(b.first);
        o->ms_pool.clear();
    }
    o->ms_blocks.clear();
    o->ms_tape.clear();
    o->ms_dts.clear();
    o->ms_next = nullptr;
    o->ms_free = 0;
}
static int ode_free_adjoint(hrw_ode_s *o);
hrw_ode_s::~hrw_ode_s()
{
    dfree(ui);
    dfree(ui2);
    dfree(udot);
    dfree(S);
    for (auto &b : ring_u) dfree(b);
    for (auto &b : ring_L) dfree(b);
    dfree(slots);
    dfree(d_bad);
    ms_tape_clear(this);
    for (auto &b : ms_start) dfree(b);
    for (auto &b : ms_lam) dfree(b);
    ode_free_adjoint(this);
}

static int ode_alloc(hrw_ode_s *o)
{
    CALL(dalloc(&o->ui, o->neq));
    CALL(dalloc(&o->ui2, o->neq));
    CALL(dalloc(&o->udot, o->neq));
    if (o->method == METHOD_MS) {
        CALL(dalloc(&o->S, o->neq));
        for (auto &b : o->ring_u) CALL(dalloc(&b, o->neq));
        for (auto &b : o->ring_L) CALL(dalloc(&b, o->neq));
    }
    CALL(dalloc(&o->slots, (size_t)3 * kNSlot * kSlotStride));
    CALL(dalloc(&o->d_bad, 1));
    return HRW_OK;
}

static int ode_new(int method, int order, hrw_problem_s *p, hrw_integrand_fn fu, void *ctx,
                   int64_t neq, int arith, hrw_ode_t *out)
{
    if (!out) return fail(HRW_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!(neq > 0)) return fail(HRW_ERR_ARG, "Invalid input 'neq'. Valid range: neq >= 1.");
    if (method == METHOD_RK && !(order >= 1 && order <= 3))
        return fail(HRW_ERR_ARG, "Invalid input 'order' in 'rktvd'. Valid range: 1 <= k <= 3.");
    auto *o = new (std::nothrow) hrw_ode_s;
    if (!o) return fail(HRW_ERR_ALLOC, "out of host memory");
    o->method = method;
    o->order = method == METHOD_MS ? 3 : order;
    o->p = p;
    o->fu = fu;
    o->ctx = ctx;
    o->neq = neq;
    o->arith = arith;
    int rc = ode_alloc(o);
    if (rc != HRW_OK) {
        delete o;
        return rc;
    }
    o->istate = 1;   // :93, :199
    *out = o;
    return HRW_OK;
}

extern "C" int hrw_rktvd_create(hrw_problem_t p, int order, hrw_ode_t *out)
{
    if (!p) return fail(HRW_ERR_ARG, "null problem");
    return ode_new(METHOD_RK, order, p, nullptr, nullptr, p->neq, p->arith, out);
}
extern "C" int hrw_rktvd_create_cb(hrw_integrand_fn fu, void *ctx, int64_t neq, int order,
                                   int arith, hrw_ode_t *out)
{
    if (!fu) return fail(HRW_ERR_ARG, "null integrand");
    return ode_new(METHOD_RK, order, nullptr, fu, ctx, neq, arith, out);
}
extern "C" int hrw_mstvd_create(hrw_problem_t p, hrw_ode_t *out)
{
    if (!p) return fail(HRW_ERR_ARG, "null problem");
    return ode_new(METHOD_MS, 3, p, nullptr, nullptr, p->neq, p->arith, out);
}
extern "C" int hrw_mstvd_create_cb(hrw_integrand_fn fu, void *ctx, int64_t neq, int arith,
                                   hrw_ode_t *out)
{
    if (!fu) return fail(HRW_ERR_ARG, "null integrand");
    return ode_new(METHOD_MS, 3, nullptr, fu, ctx, neq, arith, out);
}
extern "C" int hrw_ode_destroy(hrw_ode_t ode)
{
    delete ode;
    return HRW_OK;
}
extern "C" int hrw_ode_set_stream(hrw_ode_t ode, hrw_stream_t stream)
{
    if (!ode) return fail(HRW_ERR_ARG, "null ode");
    ode->stream = S(stream);
    return HRW_OK;
}
extern "C" int hrw_ode_info(hrw_ode_t ode, int64_t *neq, int *order, int *fevals, int *istate)
{
    if (!ode) return fail(HRW_ERR_ARG, "null ode");
    if (neq) *neq = ode->neq;
    if (order) *order = ode->order;
    if (fevals) *fevals = ode->fevals;
    if (istate) *istate = ode->istate;
    return HRW_OK;
}

// L(us) -> Lout through the callback, or a fused rhs-only pass.
static int ode_rhs(hrw_ode_s *o, int &seq, double t, const double *us, double *Lout)
{
    if (o->fu) {
        int rc = o->fu(o->ctx, t, us, Lout, reinterpret_cast<hrw_stream_t>(o->stream));
        if (rc != 0) return fail(HRW_ERR_STATE, "integrand callback returned %d", rc);
        return HRW_OK;
    }
    StageIO io;
    io.us = us;
    io.Lout = Lout;
    CALL(run_stage(o->p, o->slots, seq, -1, 0.0, io, o->stream));
    ++seq;
    return HRW_OK;
}

// stage(us, u0, L) -> uo; fused (L computed in the same pass) or callback + combine.
static int ode_stage(hrw_ode_s *o, int &seq, int stage, double t, double h, const double *us,
                     const double *u0, double *uo, const double *uo4 = nullptr,
                     const double *Lo4 = nullptr, double *Lout = nullptr,
                     double *alpha_record = nullptr)
{
    if (o->fu) {
        double *L = Lout ? Lout : o->udot;
        CALL(ode_rhs(o, seq, t, us, L));
        const unsigned g = grid_for(o->neq, kBlock * 4, 8192);
        if (o->arith == HRW_ARITH_FAST)
            combine_kernel<true><<<g, kBlock, 0, o->stream>>>(stage, o->neq, h, u0, us, L, uo4, Lo4, uo);
        else
            combine_kernel<false><<<g, kBlock, 0, o->stream>>>(stage, o->neq, h, u0, us, L, uo4, Lo4, uo);
        LAUNCH_CHECK();
        return HRW_OK;
    }
    StageIO io;
    io.us = us;
    io.u0 = u0;
    io.uo = uo;
    io.uo4 = uo4;
    io.Lo4 = Lo4;
    io.Lout = Lout;
    io.alpha_record = alpha_record;
    CALL(run_stage(o->p, o->slots, seq, stage, h, io, o->stream));
    ++seq;
    return HRW_OK;
}

// One RK step of order o->order on state u (in place for orders 2, 3).  For order 1 the new
// state is written to `other` and the caller swaps.
static int rk_step(hrw_ode_s *o, int &seq, double t, double dt, double *u, double *other)
{
    switch (o->order) {
    case 1:   // :138-139
        CALL(ode_stage(o, seq, RK1, t, dt, u, u, other));
        break;
    case 2:   // :149-152
        CALL(ode_stage(o, seq, RK2_1, t, dt, u, u, o->ui));
        CALL(ode_stage(o, seq, RK2_2, t + dt, dt, o->ui, u, u));
        break;
    default:  // :162-167
        CALL(ode_stage(o, seq, RK3_1, t, dt, u, u, o->ui));
        CALL(ode_stage(o, seq, RK3_2, t + dt, dt, o->ui, u, o->ui2));
        CALL(ode_stage(o, seq, RK3_3, t + dt / 2, dt, o->ui2, u, u));
        break;
    }
    return HRW_OK;
}

static int rk_integrate_taped(hrw_ode_s *o, double *u, double *t, double tout, double dt,
                              int itask);
static int rk_integrate(hrw_ode_s *o, double *u, double *t, double tout, double dt, int itask)
{
    if (o->istate < 1) return HRW_OK;                  // :126
    if (is_done(*t, tout, dt)) return HRW_OK;          // :127
    if (o->tape_every > 0) return rk_integrate_taped(o, u, t, tout, dt, itask);
    int seq = 0;
    if (o->p) CALL(seed_alpha(o->p, o->slots, u, o->stream));
    double *cur = u, *other = o->ui2;
    for (;;) {                                         // :137-171
        if (o->order == 1) {
            CALL(rk_step(o, seq, *t, dt, cur, other));
            std::swap(cur, other);
        } else {
            CALL(rk_step(o, seq, *t, dt, cur, nullptr));
        }
        *t = *t + dt;
        o->fevals += o->order;
        if (is_done(*t, tout, dt) || itask == 2) break;
    }
    if (cur != u) HIP_TRY(hipMemcpyAsync(u, cur, o->neq * sizeof(double), hipMemcpyDeviceToDevice, o->stream));
    if (o->istate == 1) o->istate = 2;                 // :176
    return HRW_OK;
}

// Recording of mstvd%integrate for the adjoint: storage for the states a call adds (one
// allocation per call, its step count known up front), a copy of a new state on the tape, or a
// new slot for a stage to write into.  Slots are 16-byte aligned (a stride of an even number of
// doubles), as the stage kernels' vector stores need; the double after a slot's state receives
// the Lax-Friedrichs alpha (max|u|) of the multistep stage that reads it, which its reverse
// reads back instead of taking the maximum again.
static size_t ms_stride(const hrw_ode_s *o) { return ((size_t)o->neq + 2) & ~(size_t)1; }
static int ms_reserve(hrw_ode_s *o, int64_t nstates)
{
    if (nstates <= 0) return HRW_OK;
    // the smallest pooled block that holds them, else a new one
    size_t best = o->ms_pool.size();
    for (size_t i = 0; i < o->ms_pool.size(); ++i)
        if (o->ms_pool[i].second >= nstates &&
            (best == o->ms_pool.size() || o->ms_pool[i].second < o->ms_pool[best].second))
            best = i;
    std::pair<double *, int64_t> blk{nullptr, nstates};
    if (best < o->ms_pool.size()) {
        blk = o->ms_pool[best];
        o->ms_pool.erase(o->ms_pool.begin() + (std::ptrdiff_t)best);
    } else {
        CALL(dalloc(&blk.first, (size_t)nstates * ms_stride(o)));
    }
    o->ms_blocks.push_back(blk);
    o->ms_next = blk.first;
    o->ms_free = blk.second;
    return HRW_OK;
}
static int ms_take(hrw_ode_s *o, double **slot)
{
    if (o->ms_free <= 0) return fail(HRW_ERR_STATE, "multistep recording: tape storage exhausted");
    *slot = o->ms_next;
    o->ms_next += ms_stride(o);
    --o->ms_free;
    o->ms_tape.push_back(*slot);
    return HRW_OK;
}
static int ms_record(hrw_ode_s *o, const double *src)
{
    double *b = nullptr;
    CALL(ms_take(o, &b));
    HIP_TRY(hipMemcpyAsync(b, src, o->neq * sizeof(double), hipMemcpyDeviceToDevice, o->stream));
    return HRW_OK;
}

static int ms_integrate(hrw_ode_s *o, double *u, double *t, double tout, double dt)
{
    if (o->istate < 1) return HRW_OK;                  // :228
    if (is_done(*t, tout, dt)) return HRW_OK;          // :229
    // recording (hrw_ode_tape): the tape holds every state since the method started
    const bool rec = o->tape_every > 0;
    if (rec) {
        if (o->istate == 1) ms_tape_clear(o, true);
        else if (o->ms_tape.empty())
            return fail(HRW_ERR_STATE, "recording a multistep integration starts with the method: "
                                       "enable tape before the first integrate call (istate = 1)");
        if (!o->ms_start[0])
            for (auto &b : o->ms_start) CALL(dalloc(&b, (size_t)o->neq));
        // the states this call adds: S_0 and the four starting steps' results when it starts
        // the method, then one per step of the loop below (the same t accumulation)
        int64_t add = 0;
        double tt = *t;
        if (o->istate == 1) {
            add = 5;
            for (int c = 0; c < 4; ++c) tt = tt + dt;
        }
        while (!is_done(tt, tout, dt)) {
            tt = tt + dt;
            ++add;
        }
        CALL(ms_reserve(o, add));
    }
    const size_t bytes = o->neq * sizeof(double);
    HIP_TRY(hipMemcpyAsync(o->S, u, bytes, hipMemcpyDeviceToDevice, o->stream));
    int seq = 0;
    if (o->p) CALL(seed_alpha(o->p, o->slots, o->S, o->stream));
    auto col = [o](int c) { return (o->head + 4 - c) % 4; };   // column c (1 newest .. 4 oldest)
    if (o->istate == 1) {                              // :236-249, 4 RK3 starting steps
        if (rec) CALL(ms_record(o, o->S));
        int fe = 0;
        for (int c = 4; c >= 1; --c) {
            HIP_TRY(hipMemcpyAsync(o->ring_u[col(c)], o->S, bytes, hipMemcpyDeviceToDevice, o->stream));
            CALL(ode_rhs(o, seq, *t, o->S, o->ring_L[col(c)]));
            CALL(rk_step(o, seq, *t, dt, o->S, nullptr));   // order 3
            if (rec) {   // the step's inner stage inputs (still in ui, ui2) and its result
                const int q = 2 * (4 - c);
                HIP_TRY(hipMemcpyAsync(o->ms_start[q], o->ui, bytes, hipMemcpyDeviceToDevice, o->stream));
                HIP_TRY(hipMemcpyAsync(o->ms_start[q + 1], o->ui2, bytes, hipMemcpyDeviceToDevice, o->stream));
                CALL(ms_record(o, o->S));
                o->ms_dts.push_back(dt);
            }
            *t = *t + dt;
            fe += 3;
        }
        o->fevals = fe;
        o->istate = 2;
    }
    if (rec && !o->fu) {
        // recorded fused steps write u_{n+1} straight onto the tape: S and the columns of u
        // become views, on the tape or on the object's own five buffers (each of those under
        // at most one view); at the end (or on an error) the views left on the tape get the
        // free buffers back, with copies of their states
        double *own[5] = {o->S, o->ring_u[0], o->ring_u[1], o->ring_u[2], o->ring_u[3]};
        auto settle = [&]() {
            double **views[5] = {&o->S, &o->ring_u[0], &o->ring_u[1], &o->ring_u[2], &o->ring_u[3]};
            bool used[5] = {};
            for (auto v : views)
                for (int i = 0; i < 5; ++i)
                    if (*v == own[i]) used[i] = true;
            int f = 0;
            hipError_t e = hipSuccess;
            for (auto v : views) {
                bool mine = false;
                for (int i = 0; i < 5; ++i) mine = mine || *v == own[i];
                if (mine) continue;
                while (used[f]) ++f;                   // (as many free buffers as tape views)
                const hipError_t ec = hipMemcpyAsync(own[f], *v, bytes, hipMemcpyDeviceToDevice, o->stream);
                if (e == hipSuccess) e = ec;
                *v = own[f];
                used[f] = true;
            }
            return e;
        };
        while (!is_done(*t, tout, dt)) {               // :252-268
            const int h4 = o->head;
            double *next = nullptr;
            int rc = ms_take(o, &next);
            if (rc == HRW_OK)   // (S_m is tape entry m, m = the steps recorded so far)
                rc = ode_stage(o, seq, MS3, *t, dt, o->S, o->S, next, o->ring_u[h4], o->ring_L[h4],
                               o->ring_L[h4], o->ms_tape[o->ms_dts.size()] + o->neq);
            if (rc != HRW_OK) {
                (void)settle();
                return rc;
            }
            *t = *t + dt;
            o->fevals += 1;
            o->ring_u[h4] = o->S;                      // column 1 <- u_n, state <- u_{n+1}
            o->S = next;
            o->head = (h4 + 1) % 4;
            o->ms_dts.push_back(dt);
        }
        HIP_TRY(settle());
    }
    for (;;) {                                         // :252-268
        if (is_done(*t, tout, dt)) break;
        const int h4 = o->head;
        if (o->fu) {
            // callback: L(u_n) lands in udot, then takes the place of L(u_{n-4})
            CALL(ode_stage(o, seq, MS3, *t, dt, o->S, o->S, o->ring_u[h4], o->ring_u[h4],
                           o->ring_L[h4], nullptr));
            std::swap(o->ring_L[h4], o->udot);
        } else {
            // fused: u_{n+1} overwrites u_{n-4} and L(u_n) overwrites L(u_{n-4}) pointwise
            CALL(ode_stage(o, seq, MS3, *t, dt, o->S, o->S, o->ring_u[h4], o->ring_u[h4],
                           o->ring_L[h4], o->ring_L[h4]));
        }
        *t = *t + dt;
        o->fevals += 1;
        std::swap(o->S, o->ring_u[h4]);                // column 1 <- u_n, state <- u_{n+1}
        o->head = (h4 + 1) % 4;
        if (rec) {
            CALL(ms_record(o, o->S));
            o->ms_dts.push_back(dt);
        }
    }
    HIP_TRY(hipMemcpyAsync(u, o->S, bytes, hipMemcpyDeviceToDevice, o->stream));
    return HRW_OK;
}

// NaN or Inf anywhere in u: one read of the state, one word back to the host.  The
// reference has no such test (its fatal-error path, src/hrweno_tvdode.f90:286-297, covers bad
// arguments); a blown-up state would otherwise integrate on silently.
static __global__ __launch_bounds__(kBlock) void nonfinite_kernel(const double *__restrict__ u,
                                                                 long long n,
                                                                 unsigned long long *bad)
{
    bool b = false;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (long long)gridDim.x * kBlock)
        b |= !isfinite(u[i]);
    if (__any(b) && (threadIdx.x & 63) == 0) atomicOr(bad, 1ull);
}

static int ode_check_finite(hrw_ode_s *o, const double *u)
{
    fill_u64_kernel<<<1, 64, 0, o->stream>>>(o->d_bad, 1);
    nonfinite_kernel<<<grid_for(o->neq, kBlock * 8, 4096), kBlock, 0, o->stream>>>(u, o->neq, o->d_bad);
    LAUNCH_CHECK();
    unsigned long long bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, o->d_bad, sizeof bad, hipMemcpyDeviceToHost, o->stream));
    HIP_TRY(hipStreamSynchronize(o->stream));
    if (bad) return fail(HRW_ERR_STATE, "non-finite state: u holds NaN or Inf after integrate");
    return HRW_OK;
}

extern "C" int hrw_ode_integrate(hrw_ode_t ode, double *u, double *t, double tout, double dt,
                                 int itask)
{
    if (!ode) return fail(HRW_ERR_ARG, "null ode");
    if (!u || !t) return fail(HRW_ERR_ARG, "null argument");
    if (itask != 1 && itask != 2) return fail(HRW_ERR_ARG, "Invalid input 'itask'. Valid values: 1, 2.");
    // the reference loops forever on dt == 0 (is_done never turns true); refuse it instead
    if (!(dt != 0.0) || !std::isfinite(dt) || !std::isfinite(*t) || std::isnan(tout))
        return fail(HRW_ERR_ARG, "Invalid input 'dt', 't' or 'tout': dt must be finite and nonzero");
    const double t_in = *t;
    int rc = ode->method == METHOD_MS ? ms_integrate(ode, u, t, tout, dt)
                                      : rk_integrate(ode, u, t, tout, dt, itask);
    if (rc == HRW_OK && ode->check_finite && *t != t_in) rc = ode_check_finite(ode, u);
    if (rc != HRW_OK) ode->istate = -1;
    return rc;
}

extern "C" int hrw_ode_set_check(hrw_ode_t ode, int check_finite)
{
    if (!ode) return fail(HRW_ERR_ARG, "null ode");
    ode->check_finite = check_finite != 0;
    return HRW_OK;
}

// ---- serialised state (resume in another object or process) ---------------------------
// Layout: StateHeader, then u (neq doubles), then for the multistep method the history
// columns 1..4 (newest first) of u and of L(u), each neq doubles.
namespace {
struct StateHeader {
    char magic[8];             // "HRWODE1\0"
    int32_t method, order;
    int64_t neq;
    int32_t fevals, istate;
    double t;
    int64_t history;           // 1: the multistep history follows u
};
constexpr char kStateMagic[8] = {'H', 'R', 'W', 'O', 'D', 'E', '1', '\0'};
}

static int64_t ode_state_bytes(const hrw_ode_s *o)
{
    const int64_t cols = 1 + (o->method == METHOD_MS ? 8 : 0);
    return (int64_t)sizeof(StateHeader) + cols * o->neq * (int64_t)sizeof(double);
}

extern "C" int hrw_ode_state_size(hrw_ode_t ode, int64_t *bytes)
{
    if (!ode || !bytes) return fail(HRW_ERR_ARG, "null argument");
    *bytes = ode_state_bytes(ode);
    return HRW_OK;
}

extern "C" int hrw_ode_save(hrw_ode_t ode, const double *u, double t, void *buf, int64_t bytes)
{
    if (!ode || !u || !buf) return fail(HRW_ERR_ARG, "null argument");
    if (bytes < ode_state_bytes(ode))
        return fail(HRW_ERR_ARG, "state buffer holds %lld bytes, %lld needed", (long long)bytes,
                    (long long)ode_state_bytes(ode));
    StateHeader h{};
    std::memcpy(h.magic, kStateMagic, sizeof h.magic);
    h.method = ode->method;
    h.order = ode->order;
    h.neq = ode->neq;
    h.fevals = ode->fevals;
    h.istate = ode->istate;
    h.t = t;
    h.history = ode->method == METHOD_MS && ode->istate >= 2 ? 1 : 0;
    char *p = static_cast<char *>(buf);
    std::memcpy(p, &h, sizeof h);
    const size_t col = (size_t)ode->neq * sizeof(double);
    double *dst = reinterpret_cast<double *>(p + sizeof h);
    HIP_TRY(hipMemcpyAsync(dst, u, col, hipMemcpyDeviceToHost, ode->stream));
    if (ode->method == METHOD_MS) {
        for (int c = 1; c <= 4; ++c) {   // column c (1 newest .. 4 oldest) is ring index head+4-c
            const int r = (ode->head + 4 - c) % 4;
            if (h.history) {
                HIP_TRY(hipMemcpyAsync(dst + (size_t)c * ode->neq, ode->ring_u[r], col,
                                       hipMemcpyDeviceToHost, ode->stream));
                HIP_TRY(hipMemcpyAsync(dst + (size_t)(4 + c) * ode->neq, ode->ring_L[r], col,
                                       hipMemcpyDeviceToHost, ode->stream));
            } else {
                std::memset(dst + (size_t)c * ode->neq, 0, col);
                std::memset(dst + (size_t)(4 + c) * ode->neq, 0, col);
            }
        }
    }
    HIP_TRY(hipStreamSynchronize(ode->stream));
    return HRW_OK;
}

extern "C" int hrw_ode_load(hrw_ode_t ode, double *u, double *t, const void *buf, int64_t bytes)
{
    if (!ode || !u || !t || !buf) return fail(HRW_ERR_ARG, "null argument");
    if (bytes < (int64_t)sizeof(StateHeader)) return fail(HRW_ERR_ARG, "state buffer too short");
    StateHeader h;
    const char *p = static_cast<const char *>(buf);
    std::memcpy(&h, p, sizeof h);
    if (std::memcmp(h.magic, kStateMagic, sizeof h.magic) != 0)
        return fail(HRW_ERR_ARG, "not a saved ODE state");
    if (h.method != ode->method || h.order != ode->order || h.neq != ode->neq)
        return fail(HRW_ERR_ARG, "saved state of a different ODE (method %d order %d neq %lld)",
                    h.method, h.order, (long long)h.neq);
    if (bytes < ode_state_bytes(ode)) return fail(HRW_ERR_ARG, "state buffer too short");
    const size_t col = (size_t)ode->neq * sizeof(double);
    const double *src = reinterpret_cast<const double *>(p + sizeof h);
    HIP_TRY(hipMemcpyAsync(u, src, col, hipMemcpyHostToDevice, ode->stream));
    if (ode->method == METHOD_MS && h.history) {
        ode->head = 0;   // column c at ring index 4 - c
        for (int c = 1; c <= 4; ++c) {
            HIP_TRY(hipMemcpyAsync(ode->ring_u[(4 - c) % 4], src + (size_t)c * ode->neq, col,
                                   hipMemcpyHostToDevice, ode->stream));
            HIP_TRY(hipMemcpyAsync(ode->ring_L[(4 - c) % 4], src + (size_t)(4 + c) * ode->neq, col,
                                   hipMemcpyHostToDevice, ode->stream));
        }
    }
    HIP_TRY(hipStreamSynchronize(ode->stream));
    ode->fevals = h.fevals;
    ms_tape_clear(ode, true);   // recorded states of another trajectory
    // a multistep state saved before its history existed starts over (istate 1)
    ode->istate = ode->method == METHOD_MS && !h.history && h.istate >= 2 ? 1 : h.istate;
    *t = h.t;
    return HRW_OK;
}

// ------------------------------------------------------------------ no-allocation steps -
// hrw_rktvd_steps_noalloc: the stage buffers and the alpha slot sets of its steps live in a
// file-scope device array instead of allocations.  Each device holds its own copy of the
// array, at its own address, until the process ends; host code reaches it only through the
// pointer hipGetSymbolAddress gives on that device (the array's name in host code is a
// stand-in, not a device address).
constexpr long long kNoAllocMaxCells = 1LL << 20;
constexpr long long kNoAllocSlotWords = 3LL * kNSlot * kSlotStride;
static __device__ __align__(256) double scratch[2 * kNoAllocMaxCells + kNoAllocSlotWords];
static std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices] = {};

// The scratch of the current device, looked up once per device with that device current.
static int noalloc_scratch(double **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(HRW_ERR_STATE, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(HRW_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return HRW_OK;
}

extern "C" int hrw_rktvd_steps_noalloc(hrw_problem_t p, int order, double *u, int64_t nsteps,
                                       double dt, hrw_stream_t stream)
{
    if (!p) return fail(HRW_ERR_ARG, "null problem");
    if (!u) return fail(HRW_ERR_ARG, "null state");
    if (p->kind != PROB_BURGERS1D)
        return fail(HRW_ERR_ARG, "hrw_rktvd_steps_noalloc takes a 1D Burgers problem");
    if (!(order >= 1 && order <= 3))
        return fail(HRW_ERR_ARG, "Invalid input 'order' in 'rktvd'. Valid range: 1 <= k <= 3.");
    if (nsteps < 0) return fail(HRW_ERR_ARG, "nsteps must be >= 0");
    if (!(dt != 0.0) || !std::isfinite(dt)) return fail(HRW_ERR_ARG, "dt must be finite and nonzero");
    if (p->neq > kNoAllocMaxCells)
        return fail(HRW_ERR_ARG,
                    "hrw_rktvd_steps_noalloc: %lld cells exceed its scratch of %lld cells per line",
                    (long long)p->neq, kNoAllocMaxCells);
    double *scr = nullptr;
    CALL(noalloc_scratch(&scr));
    if (nsteps == 0) return HRW_OK;
    // an integrator on the stack whose buffers point into the scratch (detached before it goes)
    hrw_ode_s o;
    o.method = METHOD_RK;
    o.order = order;
    o.neq = p->neq;
    o.p = p;
    o.arith = p->arith;
    o.stream = S(stream);
    o.ui = scr;
    o.ui2 = scr + kNoAllocMaxCells;
    o.slots = reinterpret_cast<unsigned long long *>(scr + 2 * kNoAllocMaxCells);
    int seq = 0;
    int rc = seed_alpha(p, o.slots, u, o.stream);   // clears all three slot sets first
    double *cur = u, *other = o.ui2;
    for (int64_t st = 0; rc == HRW_OK && st < nsteps; ++st) {   // the steps of rk_integrate
        if (order == 1) {
            rc = rk_step(&o, seq, 0.0, dt, cur, other);
            std::swap(cur, other);
        } else {
            rc = rk_step(&o, seq, 0.0, dt, cur, nullptr);
        }
    }
    if (rc == HRW_OK && cur != u) {
        const hipError_t e = hipMemcpyAsync(u, cur, (size_t)o.neq * sizeof(double),
                                            hipMemcpyDeviceToDevice, o.stream);
        if (e != hipSuccess) rc = fail(HRW_ERR_HIP, "copy of the last stage failed: %s", hipGetErrorString(e));
    }
    o.ui = o.ui2 = nullptr;
    o.slots = nullptr;
    o.p = nullptr;
    return rc;
}

// ------------------------------------------------------------------ batched lines -----
// An ensemble of independent periodic 1D Burgers lines advanced together: one launch per RK3
// stage covers every line (chunk table, per-line k/dt/grid/alpha), optionally replayed from a
// captured hipGraph.  The plain path (one set of launches per line) is kept for comparison.
struct hrw_batch_s {
    // cells per lane and kernel form until tuned: the fastest pair on the 4096-line ensemble of
    // 256-65536 cells
    int nlines = 0, numflux = HRW_FLUX_LAX_FRIEDRICHS, arith = HRW_ARITH_FAST, R = 6;
    int form = HRW_BATCH_CHUNK;
    double eps = 1e-6, alpha = -1.0;
    int64_t total = 0;
    std::vector<int64_t> h_off;
    std::vector<int> h_n, h_k;
    std::vector<double> h_dt, h_xmin, h_rx;
    long long nchunks = 0, ninterior = 0;
    int chunk_R = 0;
    long long *d_off = nullptr;
    int *d_n = nullptr, *d_k = nullptr, *d_cline = nullptr, *d_cstart = nullptr;
    double *d_h = nullptr, *d_xmin = nullptr, *d_rx = nullptr;
    double *ui = nullptr, *ui2 = nullptr;
    unsigned long long *slots = nullptr;       // 3 sets x nlines
    unsigned long long *line_slots = nullptr;  // per-line path: nlines x 3 x kNSlot x kSlotStride
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;   // ordering with the NULL stream
    hipGraphExec_t graph = nullptr;
    const double *graph_u = nullptr;
    int graph_R = 0, graph_form = 0;
    bool alphamax() const { return numflux == HRW_FLUX_LAX_FRIEDRICHS && alpha < 0; }
    ~hrw_batch_s()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        dfree(d_off); dfree(d_n); dfree(d_k); dfree(d_cline); dfree(d_cstart);
        dfree(d_h); dfree(d_xmin); dfree(d_rx); dfree(ui); dfree(ui2); dfree(slots);
        dfree(line_slots);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }
};

// A batch runs on a non-default stream (graph capture needs one).  When the caller gave none
// (hrw_batch_set_stream), each call is ordered after the work already queued on the NULL
// stream and the NULL stream after the call's work, so C and Fortran callers that use the
// NULL stream (copies, their own kernels) see the batch's results in stream order.
struct BatchCall {
    hrw_batch_s *b;
    explicit BatchCall(hrw_batch_s *bb) : b(bb)
    {
        if (b->own_stream && hipEventRecord(b->ev_in, nullptr) == hipSuccess)
            (void)hipStreamWaitEvent(b->stream, b->ev_in, 0);
    }
    ~BatchCall()
    {
        if (b->own_stream && hipEventRecord(b->ev_out, b->stream) == hipSuccess)
            (void)hipStreamWaitEvent(nullptr, b->ev_out, 0);
    }
};

template <class T> static int upload(T **d, const std::vector<T> &h)
{
    CALL(dalloc(d, h.size()));
    HIP_TRY(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return HRW_OK;
}

// Chunk table for chunks of 62*R cells (the output cells of one wavefront): first every chunk
// whose wavefront reads only cells of its line for K = 3 (the interior chunks, which need no
// boundary handling for any k), line by line, then the others.
static int batch_chunks(hrw_batch_s *b, int R)
{
    if (b->chunk_R == R) return HRW_OK;
    std::vector<int> cl, cs, bl, bs;
    for (int l = 0; l < b->nlines; ++l)
        for (int st = 0; st < b->h_n[l]; st += kOutLanes * R) {
            const bool interior = (long long)st - R - 2 >= 0 && (long long)st + 63LL * R + 2 <= b->h_n[l];
            (interior ? cl : bl).push_back(l);
            (interior ? cs : bs).push_back(st);
        }
    b->ninterior = (long long)cl.size();
    cl.insert(cl.end(), bl.begin(), bl.end());
    cs.insert(cs.end(), bs.begin(), bs.end());
    dfree(b->d_cline);
    dfree(b->d_cstart);
    CALL(upload(&b->d_cline, cl));
    CALL(upload(&b->d_cstart, cs));
    b->nchunks = (long long)cl.size();
    b->chunk_R = R;
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return HRW_OK;
}

extern "C" int hrw_batch_burgers_create(int64_t nlines, const int *ncells, const int *k,
                                        const double *dt, const double *xmin, const double *xmax,
                                        double eps, int numflux, double alpha, int arith,
                                        hrw_batch_t *out)
{
    if (!out) return fail(HRW_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (nlines < 1 || nlines > (1 << 30)) return fail(HRW_ERR_ARG, "nlines must be in [1, 2^30]");
    if (!ncells || !k || !dt) return fail(HRW_ERR_ARG, "null per-line array");
    if (!(eps > 2.220446049250313e-16))
        return fail(HRW_ERR_ARG, "Invalid input 'eps'. Valid range: eps > epsilon.");
    if (numflux != HRW_FLUX_LAX_FRIEDRICHS && numflux != HRW_FLUX_GODUNOV)
        return fail(HRW_ERR_ARG, "unknown numerical flux %d", numflux);
    if (arith != HRW_ARITH_FAST)
        return fail(HRW_ERR_ARG, "the batched path implements the FAST arithmetic form only");
    auto *b = new (std::nothrow) hrw_batch_s;
    if (!b) return fail(HRW_ERR_ALLOC, "out of host memory");
    std::unique_ptr<hrw_batch_s> guard(b);
    b->nlines = (int)nlines;
    b->eps = eps;
    b->numflux = numflux;
    b->alpha = alpha;
    b->arith = arith;
    b->h_off.resize(nlines + 1);
    int64_t off = 0;
    for (int64_t l = 0; l < nlines; ++l) {
        if (!(ncells[l] > 0)) return fail(HRW_ERR_ARG, "Invalid input 'ncells'. Valid range: ncells > 0.");
        if (!(k[l] >= 1 && k[l] <= 3)) return fail(HRW_ERR_ARG, "Invalid input 'k'. Valid range: 1 <= k <= 3.");
        if (!(dt[l] != 0.0) || !std::isfinite(dt[l])) return fail(HRW_ERR_ARG, "dt must be finite and nonzero");
        const double x0 = xmin ? xmin[l] : -1.0, x1 = xmax ? xmax[l] : 1.0;
        if (!(x1 > x0)) return fail(HRW_ERR_ARG, "Invalid input 'xmin', 'xmax'. Valid range: xmax > xmin.");
        // the stage kernel's 1/width starts from 1/rx (accurate while the edges stay within
        // 1e8 cell widths of the origin), the same as the single-line path's
        if (!(std::max(std::fabs(x0), std::fabs(x1)) <= 1e8 * ((x1 - x0) / ncells[l])))
            return fail(HRW_ERR_ARG, "line %lld: the batched path needs max(|xmin|, |xmax|) <= 1e8 cell widths", (long long)l);
        b->h_off[l] = off;
        off += ((int64_t)ncells[l] + 31) / 32 * 32;   // 256-byte aligned line starts
        b->h_n.push_back(ncells[l]);
        b->h_k.push_back(k[l]);
        b->h_dt.push_back(dt[l]);
        b->h_xmin.push_back(x0);
        b->h_rx.push_back((x1 - x0) / ncells[l]);
    }
    b->h_off[nlines] = off;
    b->total = off;
    std::vector<long long> offs(b->h_off.begin(), b->h_off.end());
    CALL(upload(&b->d_off, offs));
    CALL(upload(&b->d_n, b->h_n));
    CALL(upload(&b->d_k, b->h_k));
    CALL(upload(&b->d_h, b->h_dt));
    CALL(upload(&b->d_xmin, b->h_xmin));
    CALL(upload(&b->d_rx, b->h_rx));
    CALL(dalloc(&b->ui, (size_t)off));
    CALL(dalloc(&b->ui2, (size_t)off));
    CALL(dalloc(&b->slots, (size_t)3 * nlines));
    CALL(batch_chunks(b, b->R));
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->own_stream = true;
    HIP_TRY(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming));
    *out = guard.release();
    return HRW_OK;
}

extern "C" int hrw_batch_destroy(hrw_batch_t b)
{
    delete b;
    return HRW_OK;
}

extern "C" int hrw_batch_size(hrw_batch_t b, int64_t *total, int64_t *offsets)
{
    if (!b) return fail(HRW_ERR_ARG, "null batch");
    if (total) *total = b->total;
    if (offsets) std::memcpy(offsets, b->h_off.data(), b->h_off.size() * sizeof(int64_t));
    return HRW_OK;
}

extern "C" int hrw_batch_synchronize(hrw_batch_t b)
{
    if (!b) return fail(HRW_ERR_ARG, "null batch");
    HIP_TRY(hipStreamSynchronize(b->stream));
    return HRW_OK;
}

extern "C" int hrw_batch_set_stream(hrw_batch_t b, hrw_stream_t stream)
{
    if (!b) return fail(HRW_ERR_ARG, "null batch");
    if (!stream) return fail(HRW_ERR_ARG, "the batch needs a non-default stream (graph capture)");
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    b->own_stream = false;
    b->stream = S(stream);
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return HRW_OK;
}

static int batch_stage_sets(hrw_batch_s *b, const unsigned long long *ain,
                            unsigned long long *aout, unsigned long long *aclear, int stage,
                            const double *us, const double *u0, double *uo);
static int batch_stage(hrw_batch_s *b, int seq, int stage, const double *us, const double *u0,
                       double *uo)
{
    unsigned long long *sl = b->slots;
    const size_t L = (size_t)b->nlines;
    return batch_stage_sets(b, sl + (seq % 3) * L, sl + ((seq + 1) % 3) * L,
                            sl + ((seq + 2) % 3) * L, stage, us, u0, uo);
}
static int batch_stage_sets(hrw_batch_s *b, const unsigned long long *ain,
                            unsigned long long *aout, unsigned long long *aclear, int stage,
                            const double *us, const double *u0, double *uo)
{
    BatchArgs a{};
    a.us = us;
    a.u0 = u0;
    a.uo = uo;
    a.off = b->d_off;
    a.n = b->d_n;
    a.k = b->d_k;
    a.h = b->d_h;
    a.xmin = b->d_xmin;
    a.rx = b->d_rx;
    a.chunk_line = b->d_cline;
    a.chunk_start = b->d_cstart;
    a.nchunks = b->nchunks;
    a.ninterior = b->ninterior;
    a.nlines = b->nlines;
    a.stage = stage;
    a.eps = b->eps;
    a.alpha = b->alpha;
    const bool amax = b->alphamax();
    if (amax) {
        a.alpha_in = ain;
        a.alpha_out = aout;
        a.alpha_clear = aclear;
    }
    const bool lf = b->numflux == HRW_FLUX_LAX_FRIEDRICHS;
    launch_batch_stage(lf ? LF : GODUNOV, lf && amax, b->chunk_R, b->form, a, b->stream);
    LAUNCH_CHECK();
    return HRW_OK;
}

static int batch_seed(hrw_batch_s *b, const double *u)
{
    if (!b->alphamax()) return HRW_OK;
    fill_u64_kernel<<<grid_for(3LL * b->nlines, 256, 4096), 256, 0, b->stream>>>(b->slots, 3LL * b->nlines);
    batch_maxabs_kernel<<<(unsigned)((b->nchunks * 64 + kBlock - 1) / kBlock), kBlock, 0, b->stream>>>(
        u, b->d_off, b->d_n, b->d_cline, b->d_cstart, b->nchunks, b->chunk_R, b->slots);
    LAUNCH_CHECK();
    return HRW_OK;
}

// The three RK3 stages of one step for every line (src/hrweno_tvdode.f90:162-167).
static int batch_rk3_step(hrw_batch_s *b, double *u)
{
    CALL(batch_stage(b, 0, RK3_1, u, u, b->ui));
    CALL(batch_stage(b, 1, RK3_2, b->ui, u, b->ui2));
    CALL(batch_stage(b, 2, RK3_3, b->ui2, u, u));
    return HRW_OK;
}

extern "C" int hrw_batch_step(hrw_batch_t b, double *u, int64_t nsteps, double *t, int use_graph)
{
    if (!b) return fail(HRW_ERR_ARG, "null batch");
    if (!u) return fail(HRW_ERR_ARG, "null state");
    if (nsteps < 0) return fail(HRW_ERR_ARG, "nsteps must be >= 0");
    if (((uintptr_t)u & 15u) != 0) return fail(HRW_ERR_ARG, "the batch state must be 16-byte aligned");
    if (nsteps == 0) return HRW_OK;
    BatchCall order(b);
    CALL(batch_seed(b, u));
    if (use_graph) {
        if (!b->graph || b->graph_u != u || b->graph_R != b->chunk_R || b->graph_form != b->form) {
            if (b->graph) {
                (void)hipGraphExecDestroy(b->graph);
                b->graph = nullptr;
            }
            hipGraph_t g;
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int rc = batch_rk3_step(b, u);
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc != HRW_OK) return rc;
            if (e != hipSuccess) return fail(HRW_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return fail(HRW_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
            b->graph_u = u;
            b->graph_R = b->chunk_R;
            b->graph_form = b->form;
        }
        for (int64_t s = 0; s < nsteps; ++s) HIP_TRY(hipGraphLaunch(b->graph, b->stream));
    } else {
        for (int64_t s = 0; s < nsteps; ++s) CALL(batch_rk3_step(b, u));
    }
    if (t)
        for (int l = 0; l < b->nlines; ++l)
            for (int64_t s = 0; s < nsteps; ++s) t[l] = t[l] + b->h_dt[l];
    return HRW_OK;
}

// The plain path: every line advanced by its own launches of the single-line stage kernel.
extern "C" int hrw_batch_step_per_line(hrw_batch_t b, double *u, int64_t nsteps, double *t)
{
    if (!b) return fail(HRW_ERR_ARG, "null batch");
    if (!u) return fail(HRW_ERR_ARG, "null state");
    if (nsteps < 0) return fail(HRW_ERR_ARG, "nsteps must be >= 0");
    const size_t per = (size_t)3 * kNSlot * kSlotStride;
    if (!b->line_slots) CALL(dalloc(&b->line_slots, per * b->nlines));
    BatchCall order(b);
    hrw_problem_s p;
    p.kind = PROB_BURGERS1D;
    p.eps = b->eps;
    p.numflux = b->numflux;
    p.alpha = b->alpha;
    p.arith = HRW_ARITH_FAST;
    p.bc = HRW_BC_PERIODIC;
    p.cells_per_lane = b->chunk_R;
    for (int l = 0; l < b->nlines; ++l) {
        p.neq = b->h_n[l];
        p.k = b->h_k[l];
        p.xmin = b->h_xmin[l];
        p.rx = b->h_rx[l];
        unsigned long long *sl = b->line_slots + per * l;
        double *ul = u + b->h_off[l], *a1 = b->ui + b->h_off[l], *a2 = b->ui2 + b->h_off[l];
        CALL(seed_alpha(&p, sl, ul, b->stream));
        int seq = 0;
        for (int64_t s = 0; s < nsteps; ++s) {
            StageIO io;
            io.us = ul; io.u0 = ul; io.uo = a1;
            CALL(run_stage(&p, sl, seq++, RK3_1, b->h_dt[l], io, b->stream));
            io.us = a1; io.u0 = ul; io.uo = a2;
            CALL(run_stage(&p, sl, seq++, RK3_2, b->h_dt[l], io, b->stream));
            io.us = a2; io.u0 = ul; io.uo = ul;
            CALL(run_stage(&p, sl, seq++, RK3_3, b->h_dt[l], io, b->stream));
        }
        if (t)
            for (int64_t s = 0; s < nsteps; ++s) t[l] = t[l] + b->h_dt[l];
    }
    return HRW_OK;
}

// Autotuner: times the batched RK3 step for each kernel form and chunk width on a scratch copy
// of the state and keeps the fastest (the result is the same for every candidate).  Each
// candidate is timed in two rounds, the second in reverse order, and keeps its faster one: the
// clock drifts under the FP64 load, and a single pass in a fixed order once kept a candidate
// 6 % slower than the best.
extern "C" int hrw_batch_tune(hrw_batch_t b, const double *u, int *chosen)
{
    if (!b) return fail(HRW_ERR_ARG, "null batch");
    if (!u) return fail(HRW_ERR_ARG, "null state");
    BatchCall order(b);
    double *work = nullptr;
    CALL(dalloc(&work, (size_t)b->total));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    constexpr int kCand = 6;
    float cand_ms[kCand];
    for (int c = 0; c < kCand; ++c) cand_ms[c] = 1e30f;
    for (int pass = 0; pass < 2 * kCand; ++pass) {
        const int cand = pass < kCand ? pass : 2 * kCand - 1 - pass;
        const int R = 4 + 2 * (cand % 3);
        b->form = cand / 3;
        int rc = batch_chunks(b, R);
        if (rc == HRW_OK) {
            (void)hipMemcpyAsync(work, u, b->total * sizeof(double), hipMemcpyDeviceToDevice, b->stream);
            rc = batch_seed(b, work);
            if (rc == HRW_OK) rc = batch_rk3_step(b, work);   // warm-up
            (void)hipEventRecord(e0, b->stream);
            for (int i = 0; rc == HRW_OK && i < 3; ++i) rc = batch_rk3_step(b, work);
            (void)hipEventRecord(e1, b->stream);
            (void)hipEventSynchronize(e1);
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rc != HRW_OK) {
            dfree(work);
            return rc;
        }
        cand_ms[cand] = std::min(cand_ms[cand], ms);
    }
    int best = 6, best_form = HRW_BATCH_CHUNK;
    float best_ms = 1e30f;
    for (int cand = 0; cand < kCand; ++cand)
        if (cand_ms[cand] < best_ms) {
            best_ms = cand_ms[cand];
            best = 4 + 2 * (cand % 3);
            best_form = cand / 3;
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dfree(work);
    b->R = best;
    b->form = best_form;
    CALL(batch_chunks(b, best));
    if (chosen) *chosen = best;
    return HRW_OK;
}

extern "C" int hrw_batch_set_cells_per_lane(hrw_batch_t b, int cells_per_lane)
{
    if (!b) return fail(HRW_ERR_ARG, "null batch");
    if (cells_per_lane != 4 && cells_per_lane != 6 && cells_per_lane != 8)
        return fail(HRW_ERR_ARG, "cells_per_lane must be 4, 6 or 8");
    b->R = cells_per_lane;
    return batch_chunks(b, cells_per_lane);
}

extern "C" int hrw_batch_set_form(hrw_batch_t b, int form)
{
    if (!b) return fail(HRW_ERR_ARG, "null batch");
    if (form != HRW_BATCH_CHUNK && form != HRW_BATCH_PERSISTENT)
        return fail(HRW_ERR_ARG, "form must be HRW_BATCH_CHUNK (0) or HRW_BATCH_PERSISTENT (1)");
    b->form = form;
    return HRW_OK;
}

extern "C" int hrw_batch_config(hrw_batch_t b, int *cells_per_lane, int *form)
{
    if (!b) return fail(HRW_ERR_ARG, "null batch");
    if (cells_per_lane) *cells_per_lane = b->chunk_R;
    if (form) *form = b->form;
    return HRW_OK;
}

// ------------------------------------------------------------------ adjoint -----------
// Discrete adjoint of nsteps fixed rktvd steps (src/hrweno_tvdode.f90:136-171) of periodic
// Lax-Friedrichs Burgers: one line (from a problem) or every line of a batch.
//   forward : u0 -> u_N, keeping a checkpoint every m steps; the stage inputs of the last
//             segment and their alpha slot sets stay in the segment buffers;
//   backward: dJ/du_N -> dJ/du_0, segment by segment from the last: the stage inputs of a
//             segment are recomputed from its checkpoint (the same kernels, so the same
//             values), then its steps are reversed stage by stage with adjoint_stage_kernel.
// Segment s holds the stage inputs X[q], q = 0..order*len: X[0] = ckpt[s], X[q] = Y[q-1],
// X[order*len] = ckpt[s+1]; launch q reads alpha set q, accumulates set q+1, clears set q+2,
// so every set keeps max|X[q]| until the segment is recomputed.
struct hrw_adjoint_s {
    hrw_problem_s *p = nullptr;   // one line, or
    hrw_batch_s *b = nullptr;     // every line of a batch (which must outlive this object)
    int order = 3, m = 20, R = 2;   // R: cells per lane of the reverse kernel (2 or 4)
    int64_t nsteps = 0, nseg = 0, flat = 0;
    double dt = 0.0;
    std::vector<double *> ckpt, Y;
    double *lamA = nullptr, *lamB = nullptr;
    unsigned long long *sets = nullptr;
    size_t set_words = 0;
    double *alpha_rec = nullptr;  // one line: alpha of X[q], recorded by forward launch q
    // reverse-pass tables (the batch's own line tables are borrowed)
    long long *d_off = nullptr;
    int *d_n = nullptr, *d_k = nullptr;
    double *d_h = nullptr, *d_xmin = nullptr, *d_rx = nullptr;
    bool own_lines = false;
    AdjChunk *d_chunks = nullptr;
    int *d_lchunk0 = nullptr, *d_lnch = nullptr;
    long long nchunks = 0;
    // alpha-bar bookkeeping of the reverse launches: launch r writes slot r % 4 (per-chunk
    // partial sums, per-line records with the argmax key tagged with r's tag); the correction
    // it owes to its output is applied by the launches that read that buffer (at most 3
    // launches later; the first of them records it in the line records for the others), or by
    // a fixup at the end of the backward.
    static constexpr int kSlots = 4;
    double *partial[kSlots] = {}, *bpartial[kSlots] = {};
    LineRec *recs = nullptr;      // kSlots x nlines
    unsigned tag = 0;
    int64_t launches = 0;
    std::vector<std::pair<const double *, Pending>> owed;   // buffer -> its pending correction
    // Stage-input buffers: 0 = Y, sets, alpha_rec; 1 = Y1, sets1, alpha_rec1 (one line with more
    // than one segment, allocated by the first backward that can use it): the backward then
    // recomputes the next segment into one buffer on the second stream `side` while it reverses
    // the current segment from the other.  buf_seg[b]: the segment buffer b holds (-1: none).
    std::vector<double *> Y1;
    unsigned long long *sets1 = nullptr;
    double *alpha_rec1 = nullptr;
    double *sink = nullptr;       // a recompute's last output (the checkpoint, whose copy is read)
    hipStream_t side = nullptr;
    hipEvent_t ev_a = nullptr, ev_fill[2] = {nullptr, nullptr};
    bool two_buffers = false, no_two_buffers = false;
    int64_t buf_seg[2] = {-1, -1};
    bool recorded = false;
    bool general = false;         // one line that takes the general reverse pass
    GenScratch gen;
    hipStream_t stream = nullptr;
    int nlines() const { return b ? b->nlines : 1; }
    std::vector<double *> &ybuf(int k) { return k ? Y1 : Y; }
    unsigned long long *setsbuf(int k) const { return k ? sets1 : sets; }
    double *arec(int k) const { return k ? alpha_rec1 : alpha_rec; }
    hipStream_t strm() const { return b ? b->stream : stream; }
    bool alphamax() const { return b ? b->alphamax() : p->alphamax(); }
    int64_t seg_len(int64_t s) const { return std::min<int64_t>(m, nsteps - s * m); }
    ~hrw_adjoint_s()
    {
        if (side) (void)hipStreamSynchronize(side);
        for (auto &c : ckpt) dfree(c);
        for (auto &y : Y) dfree(y);
        for (auto &y : Y1) dfree(y);
        dfree(lamA); dfree(lamB); dfree(sets); dfree(alpha_rec);
        dfree(sets1); dfree(alpha_rec1); dfree(sink);
        if (ev_a) (void)hipEventDestroy(ev_a);
        for (auto &e : ev_fill)
            if (e) (void)hipEventDestroy(e);
        if (side) (void)hipStreamDestroy(side);
        if (own_lines) {
            dfree(d_off); dfree(d_n); dfree(d_k); dfree(d_h); dfree(d_xmin); dfree(d_rx);
        }
        dfree(d_chunks); dfree(d_lchunk0); dfree(d_lnch);
        for (auto &q : partial) dfree(q);
        for (auto &q : bpartial) dfree(q);
        dfree(recs);
    }
    Pending owed_by(const double *buf) const
    {
        for (auto &e : owed)
            if (e.first == buf) return e.second;
        return Pending{};
    }
    // after a launch that read buf: its correction is recorded per line
    void mark_resolved(const double *buf)
    {
        for (auto &e : owed)
            if (e.first == buf && e.second.partial) e.second.ready = 1;
    }
    // buf's pending correction, recorded per line (by a resolve launch if no launch has read
    // buf as lam yet)
    Pending resolved_owed(const double *buf)
    {
        Pending p = owed_by(buf);
        if (p.partial && !p.ready) {
            adjoint_resolve_kernel<<<nlines(), 64, 0, strm()>>>(p, d_off, d_h, d_lchunk0, d_lnch,
                                                                nlines());
            // (a launch error surfaces at the stage launch's check that follows)
            mark_resolved(buf);
            p.ready = 1;
        }
        return p;
    }
    void set_owed(const double *buf, const Pending &p)
    {
        for (auto &e : owed)
            if (e.first == buf) {
                e.second = p;
                return;
            }
        owed.emplace_back(buf, p);
    }
};

static int adjoint_chunks(hrw_adjoint_s *a, const std::vector<int> &ncells);
static int adjoint_alloc(hrw_adjoint_s *a, const std::vector<int> &ncells)
{
    a->nseg = a->nsteps > 0 ? (a->nsteps + a->m - 1) / a->m : 0;
    a->ckpt.assign((size_t)a->nseg + 1, nullptr);
    for (auto &c : a->ckpt) CALL(dalloc(&c, (size_t)a->flat));
    const int ny = a->nsteps > 0 ? a->order * a->m - 1 : 0;
    a->Y.assign((size_t)ny, nullptr);
    for (auto &y : a->Y) CALL(dalloc(&y, (size_t)a->flat));
    CALL(dalloc(&a->lamA, (size_t)a->flat));
    CALL(dalloc(&a->lamB, (size_t)a->flat));
    a->set_words = a->b ? (size_t)a->b->nlines : (size_t)kNSlot * kSlotStride;
    CALL(dalloc(&a->sets, a->set_words * (size_t)(a->order * a->m + 2)));
    if (!a->b) CALL(dalloc(&a->alpha_rec, (size_t)(a->order * a->m + 1)));
    CALL(adjoint_chunks(a, ncells));
    return HRW_OK;
}

// Chunk table of the reverse pass (output chunks of 62R cells) and the per-chunk slots.
static int adjoint_chunks(hrw_adjoint_s *a, const std::vector<int> &ncells)
{
    dfree(a->d_chunks); dfree(a->d_lchunk0); dfree(a->d_lnch);
    for (auto &q : a->partial) dfree(q);
    for (auto &q : a->bpartial) dfree(q);
    dfree(a->recs);
    a->owed.clear();
    std::vector<AdjChunk> ch;
    std::vector<int> c0, nc;
    for (int l = 0; l < (int)ncells.size(); ++l) {
        // every line starts on a block of four chunks (idle chunks pad the gap), so its
        // alpha-bar is summed in the same order as for a line on its own
        while (ch.size() % 4) {
            AdjChunk idle{};
            idle.line = -1;
            ch.push_back(idle);
        }
        c0.push_back((int)ch.size());
        AdjChunk c{};
        c.line = l;
        c.n = ncells[l];
        if (a->b) {
            c.off = a->b->h_off[l];
            c.k = a->b->h_k[l];
            c.h = a->b->h_dt[l];
            c.xmin = a->b->h_xmin[l];
            c.rx = a->b->h_rx[l];
        } else {
            c.off = 0;
            c.k = a->p->k;
            c.h = a->dt;
            c.xmin = a->p->xmin;
            c.rx = a->p->rx;
        }
        c.constw = exact_linear_grid(c.xmin, c.rx, c.n) ? 1.0 : 0.0;
        for (long long st = 0; st < ncells[l]; st += 62LL * a->R) {
            c.start = (int)st;
            ch.push_back(c);
        }
        nc.push_back((int)ch.size() - c0.back());
    }
    for (auto &c : ch) {
        if (c.line >= 0) {
            c.first = c0[(size_t)c.line];
            c.nc = nc[(size_t)c.line];
        }
    }
    a->nchunks = (long long)ch.size();
    CALL(upload(&a->d_chunks, ch));
    CALL(upload(&a->d_lchunk0, c0));
    CALL(upload(&a->d_lnch, nc));
    const size_t L = ncells.size();
    for (int q = 0; q < hrw_adjoint_s::kSlots; ++q) {
        CALL(dalloc(&a->partial[q], (size_t)a->nchunks));
        // (padded: the alpha-bar sums load whole rounds past a line's last block)
        CALL(dalloc(&a->bpartial[q], (size_t)(a->nchunks + 3) / 4 + kBlockPad));
    }
    CALL(dalloc(&a->recs, (size_t)hrw_adjoint_s::kSlots * L));
    const long long words = 2LL * hrw_adjoint_s::kSlots * (long long)L;
    fill_u64_kernel<<<grid_for(words, 256, 4096), 256, 0, a->strm()>>>(
        reinterpret_cast<unsigned long long *>(a->recs), words);
    LAUNCH_CHECK();
    HIP_TRY(hipStreamSynchronize(a->strm()));
    return HRW_OK;
}

// Checkpoint interval for ckpt_every = 0: the largest m whose states (nsteps/m + 1 checkpoints,
// order*m - 1 stage inputs of one segment, two cotangents) fit in half of the free device
// memory.  m = nsteps keeps every stage input of the run, so the backward recomputes nothing.
static int auto_ckpt(int64_t nsteps, int order, int64_t flat)
{
    if (nsteps <= 0) return 1;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    const double budget = 0.5 * (double)free_b;
    auto bytes = [&](int64_t m) {
        return 8.0 * (double)flat * (double)((nsteps + m - 1) / m + 1 + order * m - 1 + 2);
    };
    // the states are convex in m with the minimum near sqrt(nsteps/order): search above it
    int64_t lo = std::max<int64_t>(1, (int64_t)std::sqrt((double)nsteps / order));
    lo = std::min(lo, nsteps);
    if (bytes(lo) > budget) return (int)std::min<int64_t>(lo, INT32_MAX);
    int64_t hi = std::min<int64_t>(nsteps, INT32_MAX / 4);
    while (lo < hi) {   // largest m in [lo, hi] with bytes(m) <= budget
        const int64_t mid = lo + (hi - lo + 1) / 2;
        if (bytes(mid) <= budget) lo = mid;
        else hi = mid - 1;
    }
    return (int)lo;
}

// The adjoint covers the fused 1D Burgers problems on a linear grid, with either flux and either
// kind of ends; periodic Lax-Friedrichs lines take the register-blocked reverse kernel
// (adjoint_stage_kernel), the others the general pass (rhs_vjp_kernel).
static bool adjoint_problem_ok(const hrw_problem_s *p)
{
    return p->kind == PROB_BURGERS1D && !p->d_width && !p->d_cnu;
}
static bool adjoint_fast_form(const hrw_problem_s *p)
{
    return p->bc == HRW_BC_PERIODIC && p->numflux == HRW_FLUX_LAX_FRIEDRICHS;
}
static const char *kAdjointCovers = "the adjoint covers 1D Burgers problems on a linear grid";

extern "C" int hrw_adjoint_create(hrw_problem_t p, int order, double dt, int64_t nsteps,
                                  int ckpt_every, hrw_adjoint_t *out)
{
    if (!out) return fail(HRW_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!p) return fail(HRW_ERR_ARG, "null problem");
    if (!adjoint_problem_ok(p)) return fail(HRW_ERR_ARG, "%s", kAdjointCovers);
    if (!(order >= 1 && order <= 3))
        return fail(HRW_ERR_ARG, "Invalid input 'order' in 'rktvd'. Valid range: 1 <= k <= 3.");
    if (!(dt != 0.0) || !std::isfinite(dt)) return fail(HRW_ERR_ARG, "dt must be finite and nonzero");
    if (nsteps < 0) return fail(HRW_ERR_ARG, "nsteps must be >= 0");
    if (ckpt_every < 0) return fail(HRW_ERR_ARG, "ckpt_every must be >= 0 (0 = automatic)");
    if (p->neq >= (1LL << 31) - 64LL * 8) return fail(HRW_ERR_ARG, "the adjoint needs ncells < 2^31 - 512");
    if (ckpt_every == 0) ckpt_every = auto_ckpt(nsteps, order, p->neq);
    auto *a = new (std::nothrow) hrw_adjoint_s;
    if (!a) return fail(HRW_ERR_ALLOC, "out of host memory");
    std::unique_ptr<hrw_adjoint_s> guard(a);
    a->p = p;
    a->order = order;
    a->dt = dt;
    a->nsteps = nsteps;
    a->m = (int)std::max<int64_t>(1, std::min<int64_t>(ckpt_every, nsteps));
    a->flat = p->neq;
    a->own_lines = true;
    CALL(upload(&a->d_off, std::vector<long long>{0}));
    CALL(upload(&a->d_n, std::vector<int>{(int)p->neq}));
    CALL(upload(&a->d_k, std::vector<int>{p->k}));
    CALL(upload(&a->d_h, std::vector<double>{dt}));
    CALL(upload(&a->d_xmin, std::vector<double>{p->xmin}));
    CALL(upload(&a->d_rx, std::vector<double>{p->rx}));
    CALL(adjoint_alloc(a, std::vector<int>{(int)p->neq}));
    a->general = !adjoint_fast_form(p);
    if (a->general) CALL(a->gen.alloc(p->neq));
    *out = guard.release();
    return HRW_OK;
}

extern "C" int hrw_batch_adjoint_create(hrw_batch_t b, int64_t nsteps, int ckpt_every,
                                        hrw_adjoint_t *out)
{
    if (!out) return fail(HRW_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (!b) return fail(HRW_ERR_ARG, "null batch");
    if (b->numflux != HRW_FLUX_LAX_FRIEDRICHS)
        return fail(HRW_ERR_ARG, "the batched adjoint covers the Lax-Friedrichs flux");
    if (nsteps < 0) return fail(HRW_ERR_ARG, "nsteps must be >= 0");
    if (ckpt_every < 0) return fail(HRW_ERR_ARG, "ckpt_every must be >= 0 (0 = automatic)");
    if (ckpt_every == 0) ckpt_every = auto_ckpt(nsteps, 3, b->total);
    auto *a = new (std::nothrow) hrw_adjoint_s;
    if (!a) return fail(HRW_ERR_ALLOC, "out of host memory");
    std::unique_ptr<hrw_adjoint_s> guard(a);
    a->b = b;
    a->order = 3;
    a->nsteps = nsteps;
    a->m = (int)std::max<int64_t>(1, std::min<int64_t>(ckpt_every, nsteps));
    a->flat = b->total;
    a->d_off = b->d_off;
    a->d_n = b->d_n;
    a->d_k = b->d_k;
    a->d_h = b->d_h;
    a->d_xmin = b->d_xmin;
    a->d_rx = b->d_rx;
    CALL(adjoint_alloc(a, b->h_n));
    *out = guard.release();
    return HRW_OK;
}

extern "C" int hrw_adjoint_set_cells_per_lane(hrw_adjoint_t a, int cells_per_lane)
{
    if (!a) return fail(HRW_ERR_ARG, "null adjoint");
    if (cells_per_lane != 2 && cells_per_lane != 4)
        return fail(HRW_ERR_ARG, "cells_per_lane must be 2 or 4");
    a->R = cells_per_lane;
    return adjoint_chunks(a, a->b ? a->b->h_n : std::vector<int>{(int)a->p->neq});
}

extern "C" int hrw_adjoint_destroy(hrw_adjoint_t a)
{
    delete a;
    return HRW_OK;
}

extern "C" int hrw_adjoint_set_stream(hrw_adjoint_t a, hrw_stream_t stream)
{
    if (!a) return fail(HRW_ERR_ARG, "null adjoint");
    if (a->b) return fail(HRW_ERR_ARG, "a batch adjoint runs on the batch's stream (hrw_batch_set_stream)");
    a->stream = S(stream);
    return HRW_OK;
}

extern "C" int hrw_adjoint_info(hrw_adjoint_t a, int64_t *neq, int64_t *nsteps, int *ckpt_every,
                                int64_t *device_bytes)
{
    if (!a) return fail(HRW_ERR_ARG, "null adjoint");
    if (neq) *neq = a->flat;
    if (nsteps) *nsteps = a->nsteps;
    if (ckpt_every) *ckpt_every = a->m;
    if (device_bytes)
        *device_bytes = (int64_t)((a->ckpt.size() + a->Y.size() + 2 + (a->two_buffers ? a->Y1.size() + 1 : 0)) *
                                      (size_t)a->flat * sizeof(double) +
                                  a->set_words * (size_t)(a->order * a->m + 2) * 8 * (a->two_buffers ? 2 : 1));
    return HRW_OK;
}

static int adjoint_seed(hrw_adjoint_s *a, const double *u, int k, hipStream_t s)
{
    if (!a->alphamax()) return HRW_OK;
    unsigned long long *sets = a->setsbuf(k);
    if (a->b) {
        hrw_batch_s *b = a->b;
        fill_u64_kernel<<<grid_for(2LL * b->nlines, 256, 4096), 256, 0, s>>>(sets, 2LL * b->nlines);
        batch_maxabs_kernel<<<(unsigned)((b->nchunks * 64 + kBlock - 1) / kBlock), kBlock, 0, s>>>(
            u, b->d_off, b->d_n, b->d_cline, b->d_cstart, b->nchunks, b->chunk_R, sets);
    } else {
        clear_slots_kernel<<<grid_for(2 * kNSlot, 256), 256, 0, s>>>(sets, 2);
        maxabs_kernel<<<grid_for(a->p->neq, kBlock * 8, 2048), kBlock, 0, s>>>(u, a->p->neq, sets);
    }
    LAUNCH_CHECK();
    return HRW_OK;
}

// Stage input q of segment s, its inner ones from buffer k.
static double *seg_state(hrw_adjoint_s *a, int64_t s, int64_t q, int k = 0)
{
    const int64_t last = a->order * a->seg_len(s);
    if (q == 0) return a->ckpt[(size_t)s];
    if (q == last) return a->ckpt[(size_t)s + 1];
    return a->ybuf(k)[(size_t)q - 1];
}

// Forward over segment s from its checkpoint into buffer k on stream st; `full` = also the last
// launch (whose output is the next checkpoint, or last_out when given: a recompute running beside
// the reverse of the next segment, which reads that checkpoint, writes its copy elsewhere).
static int adjoint_segment(hrw_adjoint_s *a, int64_t s, bool full, int k, hipStream_t st,
                           double *last_out = nullptr)
{
    static const int stage_id[3][3] = {{RK1, 0, 0}, {RK2_1, RK2_2, 0}, {RK3_1, RK3_2, RK3_3}};
    const int64_t nq = a->order * a->seg_len(s) - (full ? 0 : 1);
    CALL(adjoint_seed(a, seg_state(a, s, 0, k), k, st));
    const size_t W = a->set_words;
    for (int64_t q = 0; q < nq; ++q) {
        const int st_i = (int)(q % a->order);
        const double *us = seg_state(a, s, q, k), *u0 = seg_state(a, s, q - st_i, k);
        double *uo = q + 1 == a->order * a->seg_len(s) && last_out ? last_out : seg_state(a, s, q + 1, k);
        unsigned long long *ain = a->setsbuf(k) + (size_t)q * W, *aout = ain + W, *aclr = aout + W;
        const int stage = stage_id[a->order - 1][st_i];
        if (a->b) {
            CALL(batch_stage_sets(a->b, ain, aout, aclr, stage, us, u0, uo));
        } else {
            StageIO io;
            io.us = us;
            io.u0 = u0;
            io.uo = uo;
            io.alpha_record = a->arec(k) + q;
            CALL(run_stage_sets(a->p, ain, aout, aclr, stage, a->dt, io, st));
        }
    }
    return HRW_OK;
}

// The second stage-input buffer, its stream and events (one line, more than one segment); false
// when the device has no room for it (the backward then recomputes and reverses in turn).
static bool adjoint_two_buffers(hrw_adjoint_s *a)
{
    if (a->two_buffers) return true;
    if (a->b || a->nseg < 2 || a->no_two_buffers) return false;
    bool ok = true;
    a->Y1.assign(a->Y.size(), nullptr);
    for (auto &y : a->Y1) ok = ok && dalloc(&y, (size_t)a->flat) == HRW_OK;
    ok = ok && dalloc(&a->sets1, a->set_words * (size_t)(a->order * a->m + 2)) == HRW_OK;
    ok = ok && dalloc(&a->alpha_rec1, (size_t)(a->order * a->m + 1)) == HRW_OK;
    ok = ok && dalloc(&a->sink, (size_t)a->flat) == HRW_OK;
    ok = ok && hipStreamCreateWithFlags(&a->side, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&a->ev_a, hipEventDisableTiming) == hipSuccess;
    for (auto &e : a->ev_fill) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    if (!ok) {   // no room: free what was taken, and do not try again
        (void)hipGetLastError();
        for (auto &y : a->Y1) dfree(y);
        a->Y1.clear();
        dfree(a->sets1); dfree(a->alpha_rec1); dfree(a->sink);
        if (a->ev_a) (void)hipEventDestroy(a->ev_a);
        for (auto &e : a->ev_fill)
            if (e) (void)hipEventDestroy(e);
        if (a->side) (void)hipStreamDestroy(a->side);
        a->ev_a = nullptr;
        a->ev_fill[0] = a->ev_fill[1] = nullptr;
        a->side = nullptr;
        a->no_two_buffers = true;
        return false;
    }
    a->two_buffers = true;
    a->buf_seg[1] = -1;
    return true;
}

extern "C" int hrw_adjoint_forward(hrw_adjoint_t a, const double *u0, double *uN)
{
    if (!a) return fail(HRW_ERR_ARG, "null adjoint");
    if (!u0 || !uN) return fail(HRW_ERR_ARG, "null state");
    const size_t bytes = (size_t)a->flat * sizeof(double);
    hipStream_t s = a->strm();
    std::unique_ptr<BatchCall> order(a->b ? new BatchCall(a->b) : nullptr);
    if (a->side) HIP_TRY(hipStreamSynchronize(a->side));   // (idle unless a call failed midway)
    HIP_TRY(hipMemcpyAsync(a->ckpt[0], u0, bytes, hipMemcpyDeviceToDevice, s));
    a->recorded = false;
    for (int64_t g = 0; g < a->nseg; ++g) CALL(adjoint_segment(a, g, true, 0, s));
    a->buf_seg[0] = a->nseg - 1;
    a->buf_seg[1] = -1;
    HIP_TRY(hipMemcpyAsync(uN, a->ckpt[(size_t)a->nseg], bytes, hipMemcpyDeviceToDevice, s));
    a->recorded = true;
    return HRW_OK;
}

// out = ca*lam + cc*h*J(X[q])^T lam + e1*x1 + e2*x2 over every line (X[q] of segment s from
// buffer k).
static int adjoint_stage(hrw_adjoint_s *a, int64_t s, int64_t q, int k, const double *lam,
                         double *out, double ca, double cc, const double *x1, double e1,
                         const double *x2, double e2)
{
    if (a->general)   // (its alpha-bar correction is applied within the pass: nothing owed;
                      // the alpha of X[q], max|X[q]|, was recorded by the stage that read it)
        return gen_vjp(a->p, a->gen, a->strm(), seg_state(a, s, q, k), lam, 1.0, nullptr, 0.0,
                       {{lam, ca}, {x1, e1}, {x2, e2}}, out, cc * a->dt, a->arec(k) + q);
    AdjArgs g{};
    g.s = seg_state(a, s, q, k);
    g.lam = lam;
    g.x1 = x1;
    g.x2 = x2;
    g.out = out;
    g.chunks = a->d_chunks;
    g.nchunks = a->nchunks;
    g.a = ca;
    g.c = cc;
    g.e1 = e1;
    g.e2 = e2;
    g.eps = a->b ? a->b->eps : a->p->eps;
    g.alpha = a->b ? a->b->alpha : a->p->alpha;
    g.aline_stride = a->b ? 1 : 0;
    g.anpart = 1;
    g.apart_stride = 0;
    g.aset = a->b ? a->sets + (size_t)q * a->set_words
                  : reinterpret_cast<const unsigned long long *>(a->arec(k) + q);
    const bool amax = a->alphamax();
    if (amax) {
        if (a->tag >= kNoPendTag - 16) {   // tags wrap: start over from clean slots
            const long long words = 2LL * hrw_adjoint_s::kSlots * a->nlines();
            fill_u64_kernel<<<grid_for(words, 256, 4096), 256, 0, a->strm()>>>(
                reinterpret_cast<unsigned long long *>(a->recs), words);
            a->tag = 0;
        }
        const int slot = (int)(a->launches % hrw_adjoint_s::kSlots);
        g.partial_out = a->partial[slot];
        g.bpartial_out = a->bpartial[slot];
        g.rec_out = a->recs + (size_t)slot * a->nlines();
        g.tag = ++a->tag;
        const Pending owed[3] = {a->owed_by(lam), x1 ? a->resolved_owed(x1) : Pending{},
                                 x2 ? a->resolved_owed(x2) : Pending{}};
        a->mark_resolved(lam);
        g.recs = a->recs;
        g.nlines = a->nlines();
        g.pend_slots = 0;
        for (int i = 0; i < 3; ++i) {
            // none owed: slot 0's records (the kernel reads them) with a tag never used
            g.pend_slots |= (unsigned)(owed[i].partial ? owed[i].slot : 0) << (4 * i);
            g.pend_tag[i] = owed[i].partial ? owed[i].tag : kNoPendTag;
        }
        g.lam_pend = owed[0];
        Pending mine{};
        mine.partial = g.partial_out;
        mine.bpartial = g.bpartial_out;
        mine.rec = g.rec_out;
        mine.s = g.s;
        mine.cf = cc;
        mine.tag = g.tag;
        mine.slot = slot;
        a->set_owed(out, mine);   // out's earlier debt (if out aliases x1/x2) is settled here
        ++a->launches;
    }
    launch_adjoint_stage(a->R, amax, g, a->strm());
    LAUNCH_CHECK();
    return HRW_OK;
}

extern "C" int hrw_adjoint_backward(hrw_adjoint_t a, const double *gN, double *g0)
{
    if (!a) return fail(HRW_ERR_ARG, "null adjoint");
    if (!gN || !g0) return fail(HRW_ERR_ARG, "null cotangent");
    if (!a->recorded) return fail(HRW_ERR_STATE, "hrw_adjoint_backward needs a recorded hrw_adjoint_forward");
    if (((uintptr_t)g0 & 15u) != 0) return fail(HRW_ERR_ARG, "the cotangent must be 16-byte aligned");
    const size_t bytes = (size_t)a->flat * sizeof(double);
    hipStream_t st = a->strm();
    std::unique_ptr<BatchCall> order(a->b ? new BatchCall(a->b) : nullptr);
    double *lam = g0, *A = a->lamA, *B = a->lamB;
    a->owed.clear();
    if (a->side) HIP_TRY(hipStreamSynchronize(a->side));   // (idle unless a call failed midway)
    if (gN != g0) HIP_TRY(hipMemcpyAsync(g0, gN, bytes, hipMemcpyDeviceToDevice, st));
    // With two stage-input buffers the segments' recomputation overlaps the reverse: while
    // segment s is reversed from buffer k on st, segment s-1 is recomputed into the other on the
    // side stream, after st's launches that read that buffer (those of segment s+1).
    const bool two = adjoint_two_buffers(a);
    bool filling[2] = {false, false};   // buffer filled on the side stream, not yet waited for
    for (int64_t s = a->nseg - 1; s >= 0; --s) {
        int k = a->buf_seg[0] == s ? 0 : (a->buf_seg[1] == s ? 1 : -1);
        if (k < 0) {
            // the whole segment again, in turn (its last launch rewrites ckpt[s+1] with the
            // same values, and records the alpha of the last stage input)
            k = 0;
            CALL(adjoint_segment(a, s, true, k, st));
            a->buf_seg[k] = s;
        }
        if (filling[k]) {
            HIP_TRY(hipStreamWaitEvent(st, a->ev_fill[k], 0));
            filling[k] = false;
        }
        if (two && s >= 1 && a->buf_seg[1 - k] != s - 1) {
            const int o = 1 - k;
            HIP_TRY(hipEventRecord(a->ev_a, st));
            HIP_TRY(hipStreamWaitEvent(a->side, a->ev_a, 0));
            // (its last output, ckpt[s] again, goes to the sink: this segment's reverse reads ckpt[s])
            CALL(adjoint_segment(a, s - 1, true, o, a->side, a->sink));
            HIP_TRY(hipEventRecord(a->ev_fill[o], a->side));
            a->buf_seg[o] = s - 1;
            filling[o] = true;
        }
        for (int64_t j = a->seg_len(s) - 1; j >= 0; --j) {
            const int64_t q = a->order * j;
            switch (a->order) {
            case 3:   // u' = (u + 2 u2 + 2h L(u2))/3, u2 = (3u + u1 + h L(u1))/4, u1 = u + h L(u)
                CALL(adjoint_stage(a, s, q + 2, k, lam, A, 2.0 / 3.0, 2.0 / 3.0, nullptr, 0, nullptr, 0));
                CALL(adjoint_stage(a, s, q + 1, k, A, B, 0.25, 0.25, nullptr, 0, nullptr, 0));
                CALL(adjoint_stage(a, s, q, k, B, lam, 1.0, 1.0, A, 0.75, lam, 1.0 / 3.0));
                break;
            case 2:   // u' = (u + u1 + h L(u1))/2, u1 = u + h L(u)
                CALL(adjoint_stage(a, s, q + 1, k, lam, A, 0.5, 0.5, nullptr, 0, nullptr, 0));
                CALL(adjoint_stage(a, s, q, k, A, lam, 1.0, 1.0, lam, 0.5, nullptr, 0));
                break;
            default:  // u' = u + h L(u)
                CALL(adjoint_stage(a, s, q, k, lam, A, 1.0, 1.0, nullptr, 0, nullptr, 0));
                std::swap(lam, A);
                break;
            }
        }
    }
    // the correction owed by the last launch's output
    const Pending last = a->owed_by(lam);
    if (last.partial) {
        adjoint_fixup_kernel<<<a->nlines(), 64, 0, st>>>(last, lam, a->d_off, a->d_h, a->d_lchunk0,
                                                        a->d_lnch, a->nlines());
        LAUNCH_CHECK();
    }
    a->owed.clear();
    // order 1 ping-pongs between g0 and the scratch buffer: the result may sit in either
    if (lam != g0) HIP_TRY(hipMemcpyAsync(g0, lam, bytes, hipMemcpyDeviceToDevice, st));
    a->lamA = lam == g0 ? A : lam;
    return HRW_OK;
}

// ---- recording rktvd%integrate for the adjoint (fused periodic Burgers problems) ----------
static int ode_free_adjoint(hrw_ode_s *o)
{
    delete o->adj;
    o->adj = nullptr;
    return HRW_OK;
}

// The loop of rk_integrate with its step count known up front (the same t accumulation and
// is_done test), run through a recording adjoint object.
static int rk_integrate_taped(hrw_ode_s *o, double *u, double *t, double tout, double dt,
                              int itask)
{
    int64_t nsteps = 0;
    double tt = *t;
    for (;;) {                                         // :137-171
        tt = tt + dt;
        ++nsteps;
        if (is_done(tt, tout, dt) || itask == 2) break;
    }
    hrw_adjoint_s *a = o->adj;
    if (!a || a->nsteps != nsteps || a->dt != dt || a->order != o->order ||
        a->m != (int)std::max<int64_t>(1, std::min<int64_t>(o->tape_every, nsteps))) {
        ode_free_adjoint(o);
        CALL(hrw_adjoint_create(o->p, o->order, dt, nsteps, o->tape_every, &o->adj));
        a = o->adj;
    }
    a->stream = o->stream;
    CALL(hrw_adjoint_forward(a, u, u));
    *t = tt;
    o->fevals += o->order * (int)nsteps;
    if (o->istate == 1) o->istate = 2;                 // :176
    return HRW_OK;
}

extern "C" int hrw_ode_tape(hrw_ode_t ode, int ckpt_every)
{
    if (!ode) return fail(HRW_ERR_ARG, "null ode");
    if (ckpt_every < 0) return fail(HRW_ERR_ARG, "ckpt_every must be >= 0");
    if (ckpt_every > 0) {
        if (!ode->p) return fail(HRW_ERR_ARG, "recording needs an integrator on a device problem");
        if (!adjoint_problem_ok(ode->p)) return fail(HRW_ERR_ARG, "%s", kAdjointCovers);
    } else {
        ode_free_adjoint(ode);
        ms_tape_clear(ode);
    }
    ode->tape_every = ckpt_every;
    return HRW_OK;
}

// Reverse sweep of a recorded mstvd integration (src/hrweno_tvdode.f90:236-266): S_0 .. S_N on
// the tape, step m (dt_m) maps S_m to S_{m+1} -- the RK3 method for m < 4, else
// S_{m+1} = (25 S_m + 50 dt_m L(S_m) + 7 S_{m-4} + 10 dt_m L(S_{m-4}))/32 -- and S_m is the
// oldest history column of step m + 4.  With l_m = dJ/dS_m (zero past N):
//   m >= 4: l_m = 25/32 l_{m+1} + 7/32 l_{m+5} + J_m^T (50 dt_m/32 l_{m+1} + 10 dt_{m+4}/32 l_{m+5})
//   m <  4: l_m = RK3^T(S_m) l_{m+1} + 7/32 l_{m+5} + 10 dt_{m+4}/32 J_m^T l_{m+5}, the last two
//           terms joining the first stage's reverse (the same Jacobian)
// one general pass per multistep step, three per starting step; l_m in ring slot m % 6.
static int ms_adjoint(hrw_ode_s *o, double *lambda)
{
    const int64_t N = (int64_t)o->ms_tape.size() - 1;
    if (N <= 0) return HRW_OK;
    const hrw_problem_s *p = o->p;
    hipStream_t st = o->stream;
    const size_t bytes = (size_t)o->neq * sizeof(double);
    if (!o->ms_lam[0])
        for (auto &b : o->ms_lam) CALL(dalloc(&b, (size_t)o->neq));
    CALL(o->gen.alloc(o->neq));
    auto L = [o](int64_t m) { return o->ms_lam[m % 6]; };
    double *L2 = o->ms_lam[6], *L1 = o->ms_lam[7];
    const std::vector<double> &dts = o->ms_dts;
    HIP_TRY(hipMemcpyAsync(L(N), lambda, bytes, hipMemcpyDeviceToDevice, st));
    for (int64_t m = N - 1; m >= 0; --m) {
        const double *lp = L(m + 1);
        const double *l5 = m + 5 <= N ? L(m + 5) : nullptr;
        const double d5 = l5 ? dts[(size_t)m + 4] : 0.0;
        const double dm = dts[(size_t)m];
        const double *Sm = o->ms_tape[(size_t)m];
        if (m >= 4) {
            CALL(gen_vjp(p, o->gen, st, Sm, lp, 50.0 * dm / 32.0, l5, 10.0 * d5 / 32.0,
                         {{lp, 25.0 / 32.0}, {l5, 7.0 / 32.0}}, L(m), 1.0, Sm + o->neq));
        } else {
            const double *Y1 = o->ms_start[2 * m], *Y2 = o->ms_start[2 * m + 1];
            // u' = (u + 2 Y2 + 2 dt L(Y2))/3, Y2 = (3u + Y1 + dt L(Y1))/4, Y1 = u + dt L(u)
            CALL(gen_vjp(p, o->gen, st, Y2, lp, 1.0, nullptr, 0.0, {{lp, 2.0 / 3.0}}, L2,
                         2.0 / 3.0 * dm));
            CALL(gen_vjp(p, o->gen, st, Y1, L2, 1.0, nullptr, 0.0, {{L2, 0.25}}, L1, 0.25 * dm));
            CALL(gen_vjp(p, o->gen, st, Sm, L1, dm, l5, 10.0 * d5 / 32.0,
                         {{lp, 1.0 / 3.0}, {L2, 0.75}, {L1, 1.0}, {l5, 7.0 / 32.0}}, L(m), 1.0));
        }
    }
    HIP_TRY(hipMemcpyAsync(lambda, L(0), bytes, hipMemcpyDeviceToDevice, st));
    return HRW_OK;
}

extern "C" int hrw_ode_adjoint(hrw_ode_t ode, double *lambda)
{
    if (!ode) return fail(HRW_ERR_ARG, "null ode");
    if (!lambda) return fail(HRW_ERR_ARG, "null lambda");
    if (ode->method == METHOD_MS) {
        if (ode->ms_tape.empty())
            return fail(HRW_ERR_STATE, "adjoint: no recorded integrate call (enable it with tape)");
        return ms_adjoint(ode, lambda);
    }
    if (!ode->adj || !ode->adj->recorded)
        return fail(HRW_ERR_STATE, "adjoint: no recorded integrate call (enable it with tape)");
    ode->adj->stream = ode->stream;
    return hrw_adjoint_backward(ode->adj, lambda, lambda);
}
