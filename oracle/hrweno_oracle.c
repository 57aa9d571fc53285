/*
 * hrweno_oracle.c -- CPU restatement of the HR-WENO hot path (test infrastructure only;
 * the shipped library never links or calls it).  Build with -ffp-contract=off: the
 * reference, built by gfortran for baseline x86-64, fuses no multiply-add, and every
 * expression below keeps the reference's left-to-right evaluation order.
 * Each function cites the reference lines it follows (paths relative to the
 * HugoMVale/HR-WENO repository root).
 */
#include "hrweno_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 1;

void hro_set_threads(int nthreads) { g_threads = nthreads > 0 ? nthreads : 1; }
int hro_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------------------ */
/* src/hrweno_weno.f90:12-21.  c is stored as c[(r+1)*k + j] for c(j, r), r=-1..k-1.   */
void hro_weno_tables(int k, double *d, double *c)
{
    if (k == 1) {
        d[0] = 1.0;
        c[0] = 1.0; c[1] = 1.0;
    } else if (k == 2) {
        d[0] = 2.0 / 3; d[1] = 1.0 / 3;
        const double t[6] = {3.0 / 2, -1.0 / 2, 1.0 / 2, 1.0 / 2, -1.0 / 2, 3.0 / 2};
        memcpy(c, t, sizeof t);
    } else {
        d[0] = 0.3; d[1] = 0.6; d[2] = 0.1;
        const double t[12] = {11.0 / 6, -7.0 / 6, 1.0 / 3,  1.0 / 3, 5.0 / 6, -1.0 / 6,
                              -1.0 / 6, 5.0 / 6,  1.0 / 3,  1.0 / 3, -7.0 / 6, 11.0 / 6};
        memcpy(c, t, sizeof t);
    }
}

/* ------------------------------------------------------------------------------------ */
/* src/hrweno_weno.f90:221-297 (Shu, ICASE 97-65, eq. 2.20).                            */
int hro_weno_calc_cnu(int n, int k, const double *xedges, double *cnu)
{
    if (n < 1 || k < 1 || k > 3) return -1;
    const int ng = k + 1;
    /* xext(-(k+1) : n+k+1), stored with offset ng */
    double *xext = (double *)malloc((size_t)(n + 1 + 2 * ng) * sizeof(double));
    if (!xext) return -2;
#define XEXT(m) xext[(m) + ng]
    for (int i = 0; i <= n; ++i) XEXT(i) = xedges[i];
    double dx = XEXT(1) - XEXT(0);                 /* :246-249 */
    for (int i = -1; i >= -ng; --i) XEXT(i) = XEXT(i + 1) - dx;
    dx = XEXT(n) - XEXT(n - 1);                    /* :252-255 */
    for (int i = n + 1; i <= n + ng; ++i) XEXT(i) = XEXT(i - 1) + dx;
    /* :258-259  xl(m) = xext(m-1), xr(m) = xext(m), m >= 1-ng */
#define XL(m) XEXT((m) - 1)
#define XR(m) XEXT(m)
    for (int i = 1; i <= n; ++i) {
        for (int r = -1; r <= k - 1; ++r) {
            for (int j = 0; j <= k - 1; ++j) {
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
                cnu[(size_t)(i - 1) * k * (k + 1) + (size_t)(r + 1) * k + j] =
                    sum2 * (XR(i - r + j) - XL(i - r + j));
            }
        }
    }
#undef XL
#undef XR
#undef XEXT
    free(xext);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* One cell of src/hrweno_weno.f90:174-216.  x[m] = vext(i+m), m = -(k-1)..k-1.        */
static inline void weno_cell(int k, double eps, const double *d, const double *ci,
                             const double *x, double *vl_out, double *vr_out)
{
    double vrr[3], vlr[3], beta[3], alfa[3], alfat[3];
    for (int r = 0; r < k; ++r) {                    /* :178-181 */
        double s = 0.0, sl = 0.0;
        for (int j = 0; j < k; ++j) s = s + ci[(r + 1) * k + j] * x[-r + j];
        for (int j = 0; j < k; ++j) sl = sl + ci[r * k + j] * x[-r + j];
        vrr[r] = s;
        vlr[r] = sl;
    }
    if (k == 1) {                                    /* :185-186 */
        beta[0] = 0.0;
    } else if (k == 2) {                             /* :189-191 */
        double a = x[1] - x[0];
        double b = x[0] - x[-1];
        beta[0] = a * a;
        beta[1] = b * b;
    } else {                                         /* :194-202 */
        double a, b;
        a = (x[0] - 2 * x[1]) + x[2];
        b = (3 * x[0] - 4 * x[1]) + x[2];
        beta[0] = 13.0 / 12 * (a * a) + 1.0 / 4 * (b * b);
        a = (x[-1] - 2 * x[0]) + x[1];
        b = x[-1] - x[1];
        beta[1] = 13.0 / 12 * (a * a) + 1.0 / 4 * (b * b);
        a = (x[-2] - 2 * x[-1]) + x[0];
        b = (x[-2] - 4 * x[-1]) + 3 * x[0];
        beta[2] = 13.0 / 12 * (a * a) + 1.0 / 4 * (b * b);
    }
    double sa = 0.0, sat = 0.0;                      /* :207-210 */
    for (int r = 0; r < k; ++r) {
        double q = eps + beta[r];
        q = q * q;
        alfa[r] = d[r] / q;
        alfat[r] = d[k - 1 - r] / q;
    }
    for (int r = 0; r < k; ++r) sa = sa + alfa[r];
    for (int r = 0; r < k; ++r) sat = sat + alfat[r];
    double vr = 0.0, vl = 0.0;                       /* :213-214 */
    for (int r = 0; r < k; ++r) vr = vr + (alfa[r] / sa) * vrr[r];
    for (int r = 0; r < k; ++r) vl = vl + (alfat[r] / sat) * vlr[r];
    *vr_out = vr;
    *vl_out = vl;
}

/* src/hrweno_weno.f90:129-219 */
int hro_weno_reconstruct(int n, int k, double eps, const double *cnu,
                         const double *v, int64_t incv,
                         double *vl, int64_t incl, double *vr, int64_t incr)
{
    if (n < 1 || k < 1 || k > 3) return -1;
    double d[3], c[12];
    hro_weno_tables(k, d, c);
    const int km = k - 1;
    /* vext(1-km : n+km), :171-173 (constant extrapolation) */
    double *vext = (double *)malloc((size_t)(n + 2 * km) * sizeof(double));
    if (!vext) return -2;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (n > 4096)
    for (int64_t i = 0; i < n; ++i) vext[i + km] = v[i * incv];
    for (int i = 0; i < km; ++i) {
        vext[i] = v[0];
        vext[n + km + i] = v[(int64_t)(n - 1) * incv];
    }
    const int kk = k * (k + 1);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (n > 4096)
    for (int64_t i = 0; i < n; ++i) {
        const double *ci = cnu ? cnu + (size_t)i * kk : c;
        weno_cell(k, eps, d, ci, vext + km + i, &vl[i * incl], &vr[i * incr]);
    }
    free(vext);
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* src/hrweno_fluxes.f90:42-43 */
double hro_lax_friedrichs(double fm, double fp, double vm, double vp, double alpha)
{
    return ((fm + fp) - alpha * (vp - vm)) / 2;
}

/* src/hrweno_fluxes.f90:66-74 */
double hro_godunov(double fm, double fp, double vm, double vp)
{
    if (vm <= vp) return fm < fp ? fm : fp;
    return fm > fp ? fm : fp;
}

/* ------------------------------------------------------------------------------------ */
/* src/hrweno_grids.f90:64-79 */
int hro_grid_linear(double xmin, double xmax, int n, double *edges)
{
    if (xmax <= xmin || n < 1) return -1;
    double rx = (xmax - xmin) / n;
    for (int i = 0; i <= n; ++i) edges[i] = xmin + rx * (double)i;
    return 0;
}

/* src/hrweno_grids.f90:111-131 */
int hro_grid_bilinear(double xmin, double xcross, double xmax, int n1, int n2, double *edges)
{
    if (xcross <= xmin || xmax <= xcross || n1 < 1 || n2 < 1) return -1;
    double rx = (xcross - xmin) / n1;
    for (int i = 0; i <= n1; ++i) edges[i] = xmin + rx * (double)i;
    rx = (xmax - xcross) / n2;
    for (int i = 1; i <= n2; ++i) edges[n1 + i] = xcross + rx * (double)i;
    return 0;
}

/* src/hrweno_grids.f90:156-175 */
int hro_grid_log(double xmin, double xmax, int n, double *edges)
{
    if (xmin <= 0.0 || !(xmax > xmin) || n < 1) return -1;
    double rx = (log(xmax / xmin)) / n;
    for (int i = 0; i <= n; ++i) edges[i] = log(xmin) + rx * (double)i;
    for (int i = 0; i <= n; ++i) edges[i] = exp(edges[i]);
    return 0;
}

/* real**integer by binary powering (the libgcc __powidf2 scheme gfortran emits) */
static double powi(double x, int m)
{
    unsigned int e = m < 0 ? -(unsigned int)m : (unsigned int)m;
    double y = (e % 2) ? x : 1.0;
    while (e >>= 1) {
        x = x * x;
        if (e % 2) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

/* src/hrweno_grids.f90:207-225 */
int hro_grid_geometric(double xmin, double xmax, double ratio, int n, double *edges)
{
    if (xmax <= xmin || ratio <= 0.0 || n < 1) return -1;
    double a = (xmax - xmin) / (powi(ratio, n) - 1.0);
    for (int i = 0; i <= n; ++i) edges[i] = xmin + a * (powi(ratio, i) - 1);
    return 0;
}

/* src/hrweno_grids.f90:242-247 */
void hro_grid_compute(int n, const double *edges, double *center, double *width)
{
    for (int i = 0; i < n; ++i) {
        center[i] = (edges[i] + edges[i + 1]) / 2;
        width[i] = edges[i + 1] - edges[i];
    }
}

/* ------------------------------------------------------------------------------------ */
/* src/hrweno_tvdode.f90:273-284 */
int hro_is_done(double t, double tout, double dt)
{
    return (t - tout) * copysign(1.0, dt) > 0.0;
}

/* src/hrweno_tvdode.f90:69-95 */
int hro_rktvd_init(hro_ode *ode, int neq, int order)
{
    memset(ode, 0, sizeof *ode);
    if (neq < 1) { ode->istate = -1; return -1; }
    if (order < 1 || order > 3) { ode->istate = -1; return -1; }
    ode->neq = neq;
    ode->order = order;
    ode->ui = (double *)calloc((size_t)neq, sizeof(double));
    ode->udot = (double *)calloc((size_t)neq, sizeof(double));
    ode->istate = 1;
    return 0;
}

/* src/hrweno_tvdode.f90:180-201 */
int hro_mstvd_init(hro_ode *ode, int neq)
{
    memset(ode, 0, sizeof *ode);
    if (neq < 1) { ode->istate = -1; return -1; }
    ode->neq = neq;
    ode->order = 3;
    ode->ui = (double *)calloc((size_t)neq, sizeof(double));
    ode->udot = (double *)calloc((size_t)neq, sizeof(double));
    ode->uold = (double *)calloc((size_t)neq * 4, sizeof(double));
    ode->udotold = (double *)calloc((size_t)neq * 4, sizeof(double));
    ode->istate = 1;
    return 0;
}

void hro_ode_free(hro_ode *ode)
{
    free(ode->ui); free(ode->udot); free(ode->uold); free(ode->udotold);
    memset(ode, 0, sizeof *ode);
}

/* src/hrweno_tvdode.f90:97-178 */
void hro_rktvd_integrate(hro_ode *ode, hro_rhs_fn fu, void *ctx, double *u,
                         double *t, double tout, double dt, int itask)
{
    if (ode->istate < 1) return;                      /* :126 */
    if (hro_is_done(*t, tout, dt)) return;            /* :127 */
    const int64_t n = ode->neq;
    double *ui = ode->ui, *udot = ode->udot;
    if (ode->order == 1) {                            /* :136-143 */
        for (;;) {
            fu(ctx, *t, u, udot);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (n > 4096)
            for (int64_t i = 0; i < n; ++i) u[i] = u[i] + dt * udot[i];
            *t = *t + dt;
            ode->fevals += 1;
            if (hro_is_done(*t, tout, dt) || itask == 2) break;
        }
    } else if (ode->order == 2) {                     /* :147-156 */
        for (;;) {
            fu(ctx, *t, u, udot);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (n > 4096)
            for (int64_t i = 0; i < n; ++i) ui[i] = u[i] + dt * udot[i];
            fu(ctx, *t + dt, ui, udot);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (n > 4096)
            for (int64_t i = 0; i < n; ++i) u[i] = ((u[i] + ui[i]) + dt * udot[i]) / 2;
            *t = *t + dt;
            ode->fevals += 2;
            if (hro_is_done(*t, tout, dt) || itask == 2) break;
        }
    } else {                                          /* :160-171 */
        for (;;) {
            fu(ctx, *t, u, udot);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (n > 4096)
            for (int64_t i = 0; i < n; ++i) ui[i] = u[i] + dt * udot[i];
            fu(ctx, *t + dt, ui, udot);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (n > 4096)
            for (int64_t i = 0; i < n; ++i) ui[i] = ((3 * u[i] + ui[i]) + dt * udot[i]) / 4;
            fu(ctx, *t + dt / 2, ui, udot);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (n > 4096)
            for (int64_t i = 0; i < n; ++i)
                u[i] = ((u[i] + 2 * ui[i]) + (2 * dt) * udot[i]) / 3;
            *t = *t + dt;
            ode->fevals += 3;
            if (hro_is_done(*t, tout, dt) || itask == 2) break;
        }
    }
    if (ode->istate == 1) ode->istate = 2;            /* :176 */
}

/* src/hrweno_tvdode.f90:203-271 */
void hro_mstvd_integrate(hro_ode *ode, hro_rhs_fn fu, void *ctx, double *u,
                         double *t, double tout, double dt)
{
    if (ode->istate < 1) return;                      /* :228 */
    if (hro_is_done(*t, tout, dt)) return;            /* :229 */
    const int64_t n = ode->neq;
    double *ui = ode->ui, *udot = ode->udot, *uold = ode->uold, *udotold = ode->udotold;
#define COL(a, c) ((a) + (size_t)((c) - 1) * n)
    if (ode->istate == 1) {                           /* :236-249 */
        hro_ode start;
        hro_rktvd_init(&start, ode->neq, ode->order);
        for (int c = ode->order + 1; c >= 1; --c) {
            memcpy(COL(uold, c), u, (size_t)n * sizeof(double));
            fu(ctx, *t, u, COL(udotold, c));
            hro_rktvd_integrate(&start, fu, ctx, u, t, *t + 2 * dt, dt, 2);
        }
        ode->fevals = start.fevals;
        ode->istate = 2;
        hro_ode_free(&start);
    }
    for (;;) {                                        /* :252-268 */
        if (hro_is_done(*t, tout, dt)) break;
        fu(ctx, *t, u, udot);
        const double *u4 = COL(uold, 4), *ud4 = COL(udotold, 4);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (n > 4096)
        for (int64_t i = 0; i < n; ++i)
            ui[i] = ((((25 * u[i]) + (50 * dt) * udot[i]) + 7 * u4[i]) + (10 * dt) * ud4[i]) / 32;
        *t = *t + dt;
        ode->fevals += 1;
        /* eoshift(.., shift=-1, dim=2): column c <- column c-1 (c=4..2) */
        for (int c = 4; c >= 2; --c) {
            memcpy(COL(udotold, c), COL(udotold, c - 1), (size_t)n * sizeof(double));
            memcpy(COL(uold, c), COL(uold, c - 1), (size_t)n * sizeof(double));
        }
        memcpy(COL(udotold, 1), udot, (size_t)n * sizeof(double));
        memcpy(COL(uold, 1), u, (size_t)n * sizeof(double));
        memcpy(u, ui, (size_t)n * sizeof(double));
    }
#undef COL
}

/* ------------------------------------------------------------------------------------ */
/* test/test_tvdode.f90:106-111 */
void hro_rhs_linear(void *ctx, double t, const double *u, double *udot)
{
    (void)t;
    const hro_linear_ctx *c = (const hro_linear_ctx *)ctx;
    for (int i = 0; i < c->n; ++i) udot[i] = c->a[i] * u[i];
}

static inline double burgers_f(double v) { return (v * v) / 2; }   /* example1:120 */

static inline double burgers_flux(int kind, double alpha, double vm, double vp)
{
    double fm = burgers_f(vm), fp = burgers_f(vp);
    return kind == HRO_FLUX_LF ? hro_lax_friedrichs(fm, fp, vm, vp, alpha)
                               : hro_godunov(fm, fp, vm, vp);
}

/* example/example1_burgers_1d_fv.f90:72-109, with a periodic variant */
void hro_rhs_burgers(void *ctx, double t, const double *v, double *vdot)
{
    (void)t;
    hro_burgers_ctx *c = (hro_burgers_ctx *)ctx;
    const int n = c->n, k = c->k;
    double alpha = c->alpha_fixed;
    if (c->flux == HRO_FLUX_LF && alpha < 0) {       /* alpha = maxval(abs(v)) */
        alpha = 0.0;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (n > 4096) reduction(max : alpha)
        for (int i = 0; i < n; ++i) alpha = fabs(v[i]) > alpha ? fabs(v[i]) : alpha;
    }
    double *F = c->work;                              /* fedges(0:n) */
    if (c->bc == HRO_BC_PERIODIC) {
        const int np = n + 2 * k;
        double *vp = F + (n + 1), *vl = vp + np, *vr = vl + np;
        for (int p = 0; p < k; ++p) vp[p] = v[((n - k + p) % n + n) % n];
        memcpy(vp + k, v, (size_t)n * sizeof(double));
        for (int p = 0; p < k; ++p) vp[k + n + p] = v[p % n];
        hro_weno_reconstruct(np, k, c->eps, NULL, vp, 1, vl, 1, vr, 1);
        /* interface i+1/2 (i = 0..n) sits between padded cells k-1+i and k+i */
#pragma omp parallel for schedule(static) num_threads(g_threads) if (n > 4096)
        for (int i = 0; i <= n; ++i) F[i] = burgers_flux(c->flux, alpha, vr[k - 1 + i], vl[k + i]);
    } else {
        double *vl = F + (n + 1), *vr = vl + n;
        hro_weno_reconstruct(n, k, c->eps, c->cnu, v, 1, vl, 1, vr, 1);
        for (int i = 1; i <= n - 1; ++i) F[i] = burgers_flux(c->flux, alpha, vr[i - 1], vl[i]);
        F[0] = F[1];                                  /* :103-104 */
        F[n] = F[n - 1];
    }
#pragma omp parallel for schedule(static) num_threads(g_threads) if (n > 4096)
    for (int i = 0; i < n; ++i) vdot[i] = -((F[i + 1] - F[i]) / c->width[i]);   /* :107 */
}

/* Test hook of the 2D composition (see hrweno_oracle.h). */
static int g_pbe_plant = 0;
void hro_pbe_plant(int which) { g_pbe_plant = which; }

/* Coordinate at which the growth rate of interior face i (1..n-1) is evaluated: the face
 * itself, right(i) (example2:101,110); planted mistake 1 takes the centre of the upwind cell i. */
static double pbe_face_x(const double *right, int i)
{
    if (g_pbe_plant == 1) return (right[i - 2] + right[i - 1]) / 2;   /* right[-1] = edges(0) */
    return right[i - 1];
}

/* Outflow flux at the high end of a line of n cells: f(vr(n)) with the growth rate at the
 * boundary face right(n); planted mistake 2 takes vl(n), mistake 3 the rate at the centre of
 * cell n. */
static double pbe_outflow(double ga, double gb, const double *right, int n, const double *vl,
                          const double *vr)
{
    const double x = g_pbe_plant == 3 ? (right[n - 2] + right[n - 1]) / 2 : right[n - 1];
    const double vm = g_pbe_plant == 2 ? vl[n - 1] : vr[n - 1];
    const double g = ga + gb * x;
    return hro_godunov(g * vm, g * vm, vm, vm);
}

/* example/example2_pbe_2d_fv.f90:73-129 */
void hro_rhs_pbe2d(void *ctx, double t, const double *v, double *vdot)
{
    (void)t;
    hro_pbe_ctx *c = (hro_pbe_ctx *)ctx;
    const int n1 = c->n1, n2 = c->n2;
    double *F1 = c->work;                          /* fedges1(0:n1, 0:n2) col-major */
    double *F2 = F1 + (size_t)(n1 + 1) * (n2 + 1); /* fedges2(0:n2, 0:n1)           */
    double *scratch = F2 + (size_t)(n2 + 1) * (n1 + 1);
    const int nmax = n1 > n2 ? n1 : n2;
#define F1_(i, j) F1[(size_t)(j) * (n1 + 1) + (i)]
#define F2_(j, i) F2[(size_t)(i) * (n2 + 1) + (j)]
#pragma omp parallel num_threads(g_threads)
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        double *vl = scratch + (size_t)tid * 2 * nmax, *vr = vl + nmax;
#pragma omp for schedule(static)
        for (int j = 1; j <= n2; ++j) {           /* x1 lines, example2:97-103 */
            int64_t o = (int64_t)(j - 1) * n1;
            hro_weno_reconstruct(n1, c->k, c->eps, NULL, v + o, 1, vl, 1, vr, 1);
            for (int i = 1; i <= n1 - 1; ++i) {
                double g = c->g1a + c->g1b * pbe_face_x(c->right1, i);
                F1_(i, j) = hro_godunov(g * vr[i - 1], g * vl[i], vr[i - 1], vl[i]);
            }
            F1_(0, j) = 0.0;
            F1_(n1, j) = c->bc == 0 ? 0.0 : pbe_outflow(c->g1a, c->g1b, c->right1, n1, vl, vr);
        }
#pragma omp for schedule(static)
        for (int i = 1; i <= n1; ++i) {           /* x2 lines, example2:106-112 */
            hro_weno_reconstruct(n2, c->k, c->eps, NULL, v + (i - 1), n1, vl, 1, vr, 1);
            for (int j = 1; j <= n2 - 1; ++j) {
                double g = c->g2a + c->g2b * pbe_face_x(c->right2, j);
                F2_(j, i) = hro_godunov(g * vr[j - 1], g * vl[j], vr[j - 1], vl[j]);
            }
            F2_(0, i) = 0.0;
            F2_(n2, i) = c->bc == 0 ? 0.0 : pbe_outflow(c->g2a, c->g2b, c->right2, n2, vl, vr);
        }
#pragma omp for schedule(static)
        for (int j = 1; j <= n2; ++j)             /* example2:123-127 */
            for (int i = 1; i <= n1; ++i)
                vdot[(size_t)(j - 1) * n1 + (i - 1)] =
                    -((F1_(i, j) - F1_(i - 1, j)) / c->width1[i - 1]) -
                    (F2_(j, i) - F2_(j - 1, i)) / c->width2[j - 1];
    }
#undef F1_
#undef F2_
}

/* ------------------------------------------------------------------------------------ */
int64_t hro_burgers_rk(int n, int k, double eps, int flux, int bc, double alpha_fixed,
                       const double *width, const double *cnu, int order,
                       double *u, double *t, double tout, double dt, int itask,
                       int *fevals)
{
    hro_burgers_ctx c;
    c.n = n; c.k = k; c.flux = flux; c.bc = bc; c.eps = eps; c.alpha_fixed = alpha_fixed;
    c.width = width; c.cnu = cnu;
    c.work = (double *)malloc((size_t)5 * (n + 2 * k + 2) * sizeof(double));
    if (!c.work) return -2;
    hro_ode ode;
    if (hro_rktvd_init(&ode, n, order) != 0) { free(c.work); return -1; }
    hro_rktvd_integrate(&ode, hro_rhs_burgers, &c, u, t, tout, dt, itask);
    int64_t steps = ode.fevals / order;
    if (fevals) *fevals = ode.fevals;
    hro_ode_free(&ode);
    free(c.work);
    return steps;
}

int hro_burgers_ms(int n, int k, double eps, int flux, int bc, double alpha_fixed,
                   const double *width, const double *cnu, double *u, double *t, double tout,
                   double dt)
{
    hro_burgers_ctx c;
    c.n = n; c.k = k; c.flux = flux; c.bc = bc; c.eps = eps; c.alpha_fixed = alpha_fixed;
    c.width = width; c.cnu = cnu;
    c.work = (double *)malloc((size_t)5 * (n + 2 * k + 2) * sizeof(double));
    if (!c.work) return -2;
    hro_ode ode;
    if (hro_mstvd_init(&ode, n) != 0) { free(c.work); return -1; }
    hro_mstvd_integrate(&ode, hro_rhs_burgers, &c, u, t, tout, dt);
    int fe = ode.fevals;
    hro_ode_free(&ode);
    free(c.work);
    return fe;
}

int hro_pbe_ms(int n1, int n2, int k, double eps, int bc, double g1a, double g1b,
               double g2a, double g2b, const double *edges1, const double *edges2,
               int order, double *u, double *t, double tout, double dt)
{
    hro_pbe_ctx c;
    c.n1 = n1; c.n2 = n2; c.k = k; c.bc = bc; c.eps = eps;
    c.g1a = g1a; c.g1b = g1b; c.g2a = g2a; c.g2b = g2b;
    double *w1 = (double *)malloc((size_t)n1 * sizeof(double));
    double *w2 = (double *)malloc((size_t)n2 * sizeof(double));
    double *ctr1 = (double *)malloc((size_t)n1 * sizeof(double));
    double *ctr2 = (double *)malloc((size_t)n2 * sizeof(double));
    hro_grid_compute(n1, edges1, ctr1, w1);
    hro_grid_compute(n2, edges2, ctr2, w2);
    c.right1 = edges1 + 1; c.right2 = edges2 + 1; c.width1 = w1; c.width2 = w2;
    const int nmax = n1 > n2 ? n1 : n2;
    size_t wn = (size_t)2 * (n1 + 1) * (n2 + 1) + (size_t)2 * nmax * (g_threads + 1) + 16;
    c.work = (double *)malloc(wn * sizeof(double));
    hro_ode ode;
    if (order == 0) {
        hro_mstvd_init(&ode, n1 * n2);
        hro_mstvd_integrate(&ode, hro_rhs_pbe2d, &c, u, t, tout, dt);
    } else {
        hro_rktvd_init(&ode, n1 * n2, order);
        hro_rktvd_integrate(&ode, hro_rhs_pbe2d, &c, u, t, tout, dt, 1);
    }
    int fe = ode.fevals;
    hro_ode_free(&ode);
    free(c.work); free(w1); free(w2); free(ctr1); free(ctr2);
    return fe;
}

int hro_pbe_rhs(int n1, int n2, int k, double eps, int bc, double g1a, double g1b, double g2a,
                double g2b, const double *edges1, const double *edges2, const double *v,
                double *vdot)
{
    hro_pbe_ctx c;
    c.n1 = n1; c.n2 = n2; c.k = k; c.bc = bc; c.eps = eps;
    c.g1a = g1a; c.g1b = g1b; c.g2a = g2a; c.g2b = g2b;
    double *w1 = (double *)malloc((size_t)n1 * sizeof(double));
    double *w2 = (double *)malloc((size_t)n2 * sizeof(double));
    double *ctr1 = (double *)malloc((size_t)n1 * sizeof(double));
    double *ctr2 = (double *)malloc((size_t)n2 * sizeof(double));
    const int nmax = n1 > n2 ? n1 : n2;
    size_t wn = (size_t)2 * (n1 + 1) * (n2 + 1) + (size_t)2 * nmax * (g_threads + 1) + 16;
    c.work = (double *)malloc(wn * sizeof(double));
    if (!w1 || !w2 || !ctr1 || !ctr2 || !c.work) {
        free(w1); free(w2); free(ctr1); free(ctr2); free(c.work);
        return -2;
    }
    hro_grid_compute(n1, edges1, ctr1, w1);
    hro_grid_compute(n2, edges2, ctr2, w2);
    c.right1 = edges1 + 1; c.right2 = edges2 + 1; c.width1 = w1; c.width2 = w2;
    hro_rhs_pbe2d(&c, 0.0, v, vdot);
    free(c.work); free(w1); free(w2); free(ctr1); free(ctr2);
    return 0;
}

int hro_linear_ode(int n, const double *a, int order, double *u, double *t,
                   double tout, double dt)
{
    hro_linear_ctx c = {n, a};
    hro_ode ode;
    int fe;
    if (order == 0) {
        hro_mstvd_init(&ode, n);
        hro_mstvd_integrate(&ode, hro_rhs_linear, &c, u, t, tout, dt);
    } else {
        if (hro_rktvd_init(&ode, n, order) != 0) return -1;
        hro_rktvd_integrate(&ode, hro_rhs_linear, &c, u, t, tout, dt, 1);
    }
    fe = ode.fevals;
    hro_ode_free(&ode);
    return fe;
}
