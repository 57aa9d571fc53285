/*
 * hrweno_oracle.h -- CPU restatement of the HR-WENO hot path, used ONLY as test
 * infrastructure (the checker in tests/, smoke() and the CPU-baseline leg of bench.py).
 * Nothing in the shipped library links, loads or calls this code.
 *
 * Every function restates one routine of the reference Fortran library
 * (HugoMVale/HR-WENO), keeping its evaluation order operation by operation, and is
 * compiled with -ffp-contract=off so that no multiply-add is fused.  The citation
 * next to each declaration names the reference lines it follows.
 *
 * Array conventions: all arrays are FP64, 0-based in C.  A Fortran array
 * a(lo:hi) is stored as a[0 .. hi-lo].  cnu(0:k-1,-1:k-1,1:n) keeps the Fortran
 * column-major order: cnu[(i-1)*k*(k+1) + (r+1)*k + j].
 */
#ifndef HRWENO_ORACLE_H
#define HRWENO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/hrweno_weno.f90:12-21 -- d(0:k-1) and c(0:k-1,-1:k-1) (c[(r+1)*k+j]) */
void hro_weno_tables(int k, double *d, double *c);

/* src/hrweno_weno.f90:221-297 -- non-uniform coefficients, eq. 2.20 */
int hro_weno_calc_cnu(int n, int k, const double *xedges, double *cnu);

/* src/hrweno_weno.f90:129-219 -- vl(i)=v_{i-1/2}^+, vr(i)=v_{i+1/2}^- with
 * constant extrapolation at both line ends.  cnu == NULL selects the uniform
 * tables.  incv/incl/incr are element strides (strided sections, example 2). */
int hro_weno_reconstruct(int n, int k, double eps, const double *cnu,
                         const double *v, int64_t incv,
                         double *vl, int64_t incl, double *vr, int64_t incr);

/* src/hrweno_fluxes.f90:22-45 and :47-76, with the user flux values f(vm), f(vp)
 * already evaluated (the reference evaluates f(vm) before f(vp)). */
double hro_lax_friedrichs(double fm, double fp, double vm, double vp, double alpha);
double hro_godunov(double fm, double fp, double vm, double vp);

/* src/hrweno_grids.f90:41-84, 86-136, 138-180, 182-230: edges(0:n) */
int hro_grid_linear(double xmin, double xmax, int n, double *edges);
int hro_grid_bilinear(double xmin, double xcross, double xmax, int n1, int n2,
                      double *edges);
int hro_grid_log(double xmin, double xmax, int n, double *edges);
int hro_grid_geometric(double xmin, double xmax, double ratio, int n, double *edges);
/* src/hrweno_grids.f90:232-250: center(n), width(n) from edges(0:n) */
void hro_grid_compute(int n, const double *edges, double *center, double *width);

/* ---- TVD ODE integrators, src/hrweno_tvdode.f90 ---- */
typedef void (*hro_rhs_fn)(void *ctx, double t, const double *u, double *udot);

typedef struct {
    int neq, order, fevals, istate;
    double *ui, *udot;          /* (neq)                 */
    double *uold, *udotold;     /* (neq,4) column-major, mstvd only */
} hro_ode;

/* src/hrweno_tvdode.f90:69-95 */
int hro_rktvd_init(hro_ode *ode, int neq, int order);
/* src/hrweno_tvdode.f90:180-201 */
int hro_mstvd_init(hro_ode *ode, int neq);
void hro_ode_free(hro_ode *ode);
/* src/hrweno_tvdode.f90:97-178 (itask 1: to tout, 2: single step) */
void hro_rktvd_integrate(hro_ode *ode, hro_rhs_fn fu, void *ctx, double *u,
                         double *t, double tout, double dt, int itask);
/* src/hrweno_tvdode.f90:203-271 */
void hro_mstvd_integrate(hro_ode *ode, hro_rhs_fn fu, void *ctx, double *u,
                         double *t, double tout, double dt);
/* src/hrweno_tvdode.f90:273-284 */
int hro_is_done(double t, double tout, double dt);

/* ---- user right-hand sides (compositions of the library calls) ---- */

/* Linear test ODE u' = a*u (test/test_tvdode.f90:106-111). */
typedef struct { int n; const double *a; } hro_linear_ctx;
void hro_rhs_linear(void *ctx, double t, const double *u, double *udot);

enum { HRO_FLUX_LF = 0, HRO_FLUX_GODUNOV = 1 };
enum { HRO_BC_PERIODIC = 0, HRO_BC_ZERO_GRADIENT = 1 };

/* 1D Burgers f(v)=v**2/2 (example/example1_burgers_1d_fv.f90:72-122).
 *  bc PERIODIC: the line is padded with k ghost cells per side (periodic copy), the
 *     padded line is reconstructed, and the n+1 interface fluxes between cells 0..n+1
 *     are taken from it, so every used window is interior.
 *  bc ZERO_GRADIENT: the example-1 composition: interior interfaces 1..n-1, then
 *     fedges(0)=fedges(1), fedges(n)=fedges(n-1) (example1:103-104).
 *  LF alpha: alpha_fixed if >= 0, else max|v| of the stage input (per RHS call). */
typedef struct {
    int n, k, flux, bc;
    double eps, alpha_fixed;
    const double *width;     /* (n) */
    const double *cnu;       /* NULL: uniform tables */
    double *work;            /* >= 5*(n+2k+2) doubles, owned by the caller */
} hro_burgers_ctx;
void hro_rhs_burgers(void *ctx, double t, const double *v, double *vdot);

/* 2D population balance (example/example2_pbe_2d_fv.f90:73-155), u((j-1)*n1+i).
 * Flux along x1: (g1a + g1b*x1)*v with x1 the x1-interface coordinate; along x2:
 * (g2a + g2b*x2)*v.  Both with the Godunov flux.  Boundaries:
 *  bc 0 (example 2): zero flux at both ends of both directions (example2:117-120);
 *  bc 1 (zero inflow): zero flux at the low end, outflow f(vr(n)) at the high end. */
typedef struct {
    int n1, n2, k, bc;
    double eps, g1a, g1b, g2a, g2b;
    const double *right1, *right2, *width1, *width2;   /* grid arrays */
    double *work;            /* >= 6*max(n1,n2)+2*(n1+1)*(n2+1)+16 doubles */
} hro_pbe_ctx;
void hro_rhs_pbe2d(void *ctx, double t, const double *v, double *vdot);

/* ---- drivers (one call runs a whole integration, for ctypes callers) ---- */
/* Burgers with rktvd; returns steps taken; fevals/istate reported. */
int64_t hro_burgers_rk(int n, int k, double eps, int flux, int bc, double alpha_fixed,
                       const double *width, const double *cnu, int order,
                       double *u, double *t, double tout, double dt, int itask,
                       int *fevals);
/* Burgers with mstvd from istate=1; returns fevals. */
int hro_burgers_ms(int n, int k, double eps, int flux, int bc, double alpha_fixed,
                   const double *width, const double *cnu, double *u, double *t, double tout,
                   double dt);
/* 2D PBE from istate=1 with mstvd (order 0) or rktvd (order 1..3); returns fevals. */
int hro_pbe_ms(int n1, int n2, int k, double eps, int bc, double g1a, double g1b,
               double g2a, double g2b, const double *edges1, const double *edges2,
               int order, double *u, double *t, double tout, double dt);
/* One evaluation of hro_rhs_pbe2d on the linear grids edges1(0:n1), edges2(0:n2). */
int hro_pbe_rhs(int n1, int n2, int k, double eps, int bc, double g1a, double g1b, double g2a,
                double g2b, const double *edges1, const double *edges2, const double *v,
                double *vdot);
/* Test hook of hro_rhs_pbe2d: 0 = the example-2 composition; 1, 2, 3 plant one mistake each
 * (interior growth rates at the upwind cell's centre instead of the face, the outflow flux from
 * vl(n) instead of vr(n), the outflow growth rate at the centre of cell n instead of the
 * boundary face) so tests can show the checks catch them. */
void hro_pbe_plant(int which);
/* Linear ODE with rktvd (order 1..3) or mstvd (order 0). */
int hro_linear_ode(int n, const double *a, int order, double *u, double *t,
                   double tout, double dt);

/* ---- discrete adjoint of the Burgers path (hrweno_adjoint.c) ---- */
/* One evaluation of hro_rhs_burgers (uniform tables). */
int hro_burgers_rhs(int n, int k, double eps, int flux, int bc, double alpha_fixed,
                    const double *width, const double *v, double *vdot);
/* out = (dL/dv)^T mu for L = hro_rhs_burgers with the given flux (Lax-Friedrichs with
 * alpha = alpha_fixed >= 0 or max|v|, or Godunov) and ends (periodic or zero-gradient). */
int hro_burgers_rhs_vjp_ex(int n, int k, double eps, int flux, int bc, double alpha_fixed,
                           const double *width, const double *v, const double *mu, double *out);
/* The periodic Lax-Friedrichs case of hro_burgers_rhs_vjp_ex. */
int hro_burgers_rhs_vjp(int n, int k, double eps, double alpha_fixed, const double *width,
                        const double *v, const double *mu, double *out);
/* nsteps fixed rktvd steps (order 1..3) of that RHS, u <- u_N, then the reverse sweep:
 * g in: dJ/du_N, out: dJ/du_0. */
int hro_burgers_rk_adjoint_ex(int n, int k, double eps, int flux, int bc, double alpha_fixed,
                              const double *width, int order, int64_t nsteps, double dt,
                              double *u, double *g);
int hro_burgers_rk_adjoint(int n, int k, double eps, double alpha_fixed, const double *width,
                           int order, int64_t nsteps, double dt, double *u, double *g);
/* mstvd%integrate from a fresh object (four RK3 starting steps, then the multistep method
 * until is_done), u <- u_N, *t advanced, then the reverse sweep of g; *nsteps = N, the
 * starting steps included. */
int hro_burgers_ms_adjoint(int n, int k, double eps, int flux, int bc, double alpha_fixed,
                           const double *width, double *t, double tout, double dt, double *u,
                           double *g, int64_t *nsteps);
/* Test hook: 0 = exact adjoint; each other value plants one mistake so tests can show the
 * checks catch it: 1 weights held fixed, 2 the sign of alpha in the left-face flux
 * derivative, 3 alpha held fixed, 4 Godunov differentiated on the other end state, 5 the
 * zero-gradient end cells' RHS treated as live, 6 zero-gradient window cotangents wrapped
 * around the line, 7 the multistep history state dropped, 8 the multistep history slope
 * dropped. */
void hro_adjoint_plant(int which);

/* OpenMP thread count used by the per-cell loops (results do not depend on it). */
void hro_set_threads(int nthreads);
int hro_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
