/*
 * hrweno_adjoint.c -- CPU restatement of the discrete adjoint of the periodic 1D Burgers
 * path (test infrastructure only: the checker of the device adjoint; nothing shipped links
 * it).
 *
 * The forward map being differentiated is exactly the one in hrweno_oracle.c:
 *   RHS  hro_rhs_burgers with bc PERIODIC and the Lax-Friedrichs flux
 *        (example/example1_burgers_1d_fv.f90:72-109 with a periodic line;
 *         src/hrweno_weno.f90:174-216 per cell; src/hrweno_fluxes.f90:42-43),
 *   time hro_rktvd_integrate, orders 1..3 (src/hrweno_tvdode.f90:136-171),
 * run for a fixed number of steps.  Lax-Friedrichs alpha is alpha_fixed if >= 0, else
 * max|v| of the stage input; its derivative is taken at the first index attaining the
 * maximum (the one the forward scan keeps), with sign(v) = (v > 0) - (v < 0).
 *
 * Derivatives (per cell i, window x_m = v_{i+m}, m = -(k-1)..k-1):
 *   vr = sum_r w_r q_r,  w_r = a_r / sum(a),  a_r = d_r / (eps + beta_r)^2
 *   vl = sum_r w~_r l_r, w~_r = a~_r / sum(a~), a~_r = d_{k-1-r} / (eps + beta_r)^2
 *   d vr / d beta_r = -2 w_r (q_r - vr) / (eps + beta_r)    (and likewise for vl)
 *   q_r, l_r linear in x (tables c(:,r), c(:,r-1)); beta_r quadratic in x.
 * Flux F(vm, vp) = (vm^2/2 + vp^2/2 - alpha (vp - vm)) / 2:
 *   dF/dvm = (vm + alpha)/2, dF/dvp = (vp - alpha)/2, dF/dalpha = -(vp - vm)/2.
 * Divergence L_i = -(F_{i+1/2} - F_{i-1/2}) / w_i, so the cotangent of F_{i+1/2} is
 *   mu_{i+1}/w_{i+1} - mu_i/w_i.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hrweno_oracle.h"

static int g_plant = 0;
void hro_adjoint_plant(int which) { g_plant = which; }

static inline int wrap(long long i, int n)
{
    long long r = i % n;
    return (int)(r < 0 ? r + n : r);
}

/* beta_r and d beta_r / d x_m (db[r][m + 2]) for the window x (x[m], m = -(k-1)..k-1),
 * in the order of src/hrweno_weno.f90:185-202. */
static void smoothness(int k, const double *x, double *beta, double db[3][5])
{
    memset(db, 0, sizeof(double) * 15);
    if (k == 1) {
        beta[0] = 0.0;
    } else if (k == 2) {
        double a = x[1] - x[0];
        double b = x[0] - x[-1];
        beta[0] = a * a;
        beta[1] = b * b;
        db[0][1 + 2] = 2 * a;  db[0][0 + 2] = -2 * a;
        db[1][0 + 2] = 2 * b;  db[1][-1 + 2] = -2 * b;
    } else {
        static const double A[3][5] = {{0, 0, 1, -2, 1}, {0, 1, -2, 1, 0}, {1, -2, 1, 0, 0}};
        static const double B[3][5] = {{0, 0, 3, -4, 1}, {0, 1, 0, -1, 0}, {1, -4, 3, 0, 0}};
        double a, b;
        a = (x[0] - 2 * x[1]) + x[2];
        b = (3 * x[0] - 4 * x[1]) + x[2];
        beta[0] = 13.0 / 12 * (a * a) + 1.0 / 4 * (b * b);
        for (int m = 0; m < 5; ++m) db[0][m] = 13.0 / 6 * a * A[0][m] + 0.5 * b * B[0][m];
        a = (x[-1] - 2 * x[0]) + x[1];
        b = x[-1] - x[1];
        beta[1] = 13.0 / 12 * (a * a) + 1.0 / 4 * (b * b);
        for (int m = 0; m < 5; ++m) db[1][m] = 13.0 / 6 * a * A[1][m] + 0.5 * b * B[1][m];
        a = (x[-2] - 2 * x[-1]) + x[0];
        b = (x[-2] - 4 * x[-1]) + 3 * x[0];
        beta[2] = 13.0 / 12 * (a * a) + 1.0 / 4 * (b * b);
        for (int m = 0; m < 5; ++m) db[2][m] = 13.0 / 6 * a * A[2][m] + 0.5 * b * B[2][m];
    }
}

typedef struct {
    double vr, vl, vrr[3], vlr[3], beta[3], w[3], wt[3], db[3][5];
} cell_state;

/* Forward values of one cell in the reference order (as weno_cell in hrweno_oracle.c). */
static void cell_forward(int k, double eps, const double *d, const double *c, const double *x,
                         cell_state *s)
{
    for (int r = 0; r < k; ++r) {
        double a = 0.0, b = 0.0;
        for (int j = 0; j < k; ++j) a = a + c[(r + 1) * k + j] * x[-r + j];
        for (int j = 0; j < k; ++j) b = b + c[r * k + j] * x[-r + j];
        s->vrr[r] = a;
        s->vlr[r] = b;
    }
    smoothness(k, x, s->beta, s->db);
    double al[3], alt[3], sa = 0.0, sat = 0.0;
    for (int r = 0; r < k; ++r) {
        double q = eps + s->beta[r];
        q = q * q;
        al[r] = d[r] / q;
        alt[r] = d[k - 1 - r] / q;
    }
    for (int r = 0; r < k; ++r) sa = sa + al[r];
    for (int r = 0; r < k; ++r) sat = sat + alt[r];
    double vr = 0.0, vl = 0.0;
    for (int r = 0; r < k; ++r) {
        s->w[r] = al[r] / sa;
        s->wt[r] = alt[r] / sat;
        vr = vr + s->w[r] * s->vrr[r];
        vl = vl + s->wt[r] * s->vlr[r];
    }
    s->vr = vr;
    s->vl = vl;
}

/* Cotangent of the window x_m (xb[m + 2]) for cotangents gr, gl on (vr, vl). */
static void cell_vjp(int k, double eps, const double *c, const cell_state *s, double gr,
                     double gl, double *xb)
{
    for (int m = 0; m < 5; ++m) xb[m] = 0.0;
    for (int r = 0; r < k; ++r) {
        const double P = gr * s->w[r], Q = gl * s->wt[r];
        for (int j = 0; j < k; ++j)
            xb[-r + j + 2] += P * c[(r + 1) * k + j] + Q * c[r * k + j];
        double bb = -2.0 / (eps + s->beta[r]) * (P * (s->vrr[r] - s->vr) + Q * (s->vlr[r] - s->vl));
        if (g_plant == 1) bb = 0.0;                    /* planted: weights held fixed */
        for (int m = 0; m < 5; ++m) xb[m] += bb * s->db[r][m];
    }
}

/* Index of window position i + m: wrapped on a periodic line, clamped (the constant
 * extrapolation inside src/hrweno_weno.f90:155-170) otherwise. */
static inline int win(long long i, int n, int bc)
{
    if (bc == HRO_BC_PERIODIC) return wrap(i, n);
    return (int)(i < 0 ? 0 : (i >= n ? n - 1 : i));
}

/* Partial derivatives of the interface flux F(vm, vp) for f(v) = v^2/2 (f'(v) = v):
 *   Lax-Friedrichs (src/hrweno_fluxes.f90:42-43): (vm + alpha)/2, (vp - alpha)/2, and
 *     -(vp - vm)/2 for alpha;
 *   Godunov over the two end states (src/hrweno_fluxes.f90:66-74): the derivative of the
 *     branch the flux takes.  At vm == vp (constant data reconstructs exactly) both branches
 *     give f(v), and near that point F is f of the upwind state whatever the perturbation
 *     (v > 0: min/max of an increasing f picks vm's side; v < 0: vp's), so F is smooth there
 *     with gradient (v, 0), (0, v), or (0, 0) at v = 0. */
static void flux_partials(int flux, double alpha, double vm, double vp, double *dm, double *dp,
                          double *da)
{
    if (flux == HRO_FLUX_LF) {
        *dm = 0.5 * (vm + alpha);
        *dp = 0.5 * (vp - alpha);
        if (g_plant == 2) *dp = 0.5 * (vp + alpha);   /* planted: sign of alpha */
        *da = -0.5 * (vp - vm);
        return;
    }
    const double fm = (vm * vm) / 2, fp = (vp * vp) / 2;
    int takem = vm == vp ? vm > 0 : (vm <= vp ? fm < fp : fm > fp);
    if (g_plant == 4) takem = !takem;                 /* planted: the other end state */
    *dm = takem ? vm : 0.0;
    *dp = takem ? 0.0 : vp;   /* (at vm == vp == 0 both are 0) */
    *da = 0.0;
}

int hro_burgers_rhs_vjp_ex(int n, int k, double eps, int flux, int bc, double alpha_fixed,
                           const double *width, const double *v, const double *mu, double *out)
{
    if (n < 1 || k < 1 || k > 3) return -1;
    double d[3], c[12];
    hro_weno_tables(k, d, c);
    double alpha = alpha_fixed;
    int jmax = -1;
    if (flux == HRO_FLUX_LF && alpha < 0) {           /* alpha = maxval(abs(v)) */
        alpha = 0.0;
        jmax = 0;
        for (int i = 0; i < n; ++i)
            if (fabs(v[i]) > alpha) {
                alpha = fabs(v[i]);
                jmax = i;
            }
    }
    cell_state *cs = (cell_state *)malloc((size_t)n * sizeof(cell_state));
    double *phi = (double *)malloc((size_t)n * sizeof(double));
    double *gr = (double *)calloc((size_t)n, sizeof(double));
    double *gl = (double *)calloc((size_t)n, sizeof(double));
    if (!cs || !phi || !gr || !gl) {
        free(cs);
        free(phi);
        free(gr);
        free(gl);
        return -2;
    }
    const int km = k - 1;
    /* with zero-gradient ends the reference overwrites the boundary fluxes with their
     * neighbours (example1:103-104), so L of the two end cells is identically zero */
    const int periodic = bc == HRO_BC_PERIODIC;
    for (int i = 0; i < n; ++i) {
        double xw[5];
        for (int m = -km; m <= km; ++m) xw[m + 2] = v[win((long long)i + m, n, bc)];
        cell_forward(k, eps, d, c, xw + 2, &cs[i]);
        phi[i] = mu[i] / width[i];
    }
    if (!periodic && g_plant != 5) {                  /* planted 5: end cells kept */
        phi[0] = 0.0;
        phi[n - 1] = 0.0;
    }
    for (int i = 0; i < n; ++i) out[i] = 0.0;
    /* interface i+1/2 between cell i (its vr) and cell ip (its vl): every i on a periodic
     * line, i = 0..n-2 otherwise; the cotangent of F_{i+1/2} is phi[ip] - phi[i] */
    double abar = 0.0;
    const int nif = periodic ? n : n - 1;
    for (int i = 0; i < nif; ++i) {
        const int ip = periodic ? wrap((long long)i + 1, n) : i + 1;
        const double G = phi[ip] - phi[i];
        double dm, dp, da;
        flux_partials(flux, alpha, cs[i].vr, cs[ip].vl, &dm, &dp, &da);
        gr[i] += G * dm;
        gl[ip] += G * dp;
        abar += G * da;
    }
    for (int i = 0; i < n; ++i) {
        double xb[5];
        cell_vjp(k, eps, c, &cs[i], gr[i], gl[i], xb);
        for (int m = -km; m <= km; ++m) {
            const int j = g_plant == 6 ? wrap((long long)i + m, n)   /* planted: ends wrapped */
                                       : win((long long)i + m, n, bc);
            out[j] += xb[m + 2];
        }
    }
    if (jmax >= 0 && g_plant != 3) {                   /* planted 3: alpha held fixed */
        const double sg = (v[jmax] > 0) - (v[jmax] < 0);
        out[jmax] += abar * sg;
    }
    free(cs);
    free(phi);
    free(gr);
    free(gl);
    return 0;
}

int hro_burgers_rhs_vjp(int n, int k, double eps, double alpha_fixed, const double *width,
                        const double *v, const double *mu, double *out)
{
    return hro_burgers_rhs_vjp_ex(n, k, eps, HRO_FLUX_LF, HRO_BC_PERIODIC, alpha_fixed, width,
                                  v, mu, out);
}

int hro_burgers_rhs(int n, int k, double eps, int flux, int bc, double alpha_fixed,
                    const double *width, const double *v, double *vdot)
{
    if (n < 1 || k < 1 || k > 3) return -1;
    hro_burgers_ctx ctx;
    ctx.n = n; ctx.k = k; ctx.flux = flux; ctx.bc = bc; ctx.eps = eps;
    ctx.alpha_fixed = alpha_fixed; ctx.width = width; ctx.cnu = NULL;
    ctx.work = (double *)malloc((size_t)5 * ((size_t)n + 2 * k + 2) * sizeof(double));
    if (!ctx.work) return -2;
    hro_rhs_burgers(&ctx, 0.0, v, vdot);
    free(ctx.work);
    return 0;
}

/* The forward of hro_rktvd_integrate for nsteps steps, storing the stage inputs, then the
 * reverse sweep.  u <- u_N;  g: in dJ/du_N, out dJ/du_0. */
int hro_burgers_rk_adjoint_ex(int n, int k, double eps, int flux, int bc, double alpha_fixed,
                              const double *width, int order, int64_t nsteps, double dt,
                              double *u, double *g)
{
    if (n < 1 || k < 1 || k > 3 || order < 1 || order > 3 || nsteps < 0) return -1;
    const size_t nn = (size_t)n;
    double *X = (double *)malloc((size_t)order * (size_t)(nsteps > 0 ? nsteps : 1) * nn * sizeof(double));
    double *udot = (double *)malloc(nn * sizeof(double));
    double *l1 = (double *)malloc(nn * sizeof(double));
    double *l2 = (double *)malloc(nn * sizeof(double));
    double *jt = (double *)malloc(nn * sizeof(double));
    hro_burgers_ctx ctx;
    ctx.n = n; ctx.k = k; ctx.flux = flux; ctx.bc = bc; ctx.eps = eps;
    ctx.alpha_fixed = alpha_fixed; ctx.width = width; ctx.cnu = NULL;
    ctx.work = (double *)malloc((size_t)5 * (nn + 2 * k + 2) * sizeof(double));
    int rc = 0;
    if (!X || !udot || !l1 || !l2 || !jt || !ctx.work) {
        rc = -2;
        goto done;
    }
    for (int64_t s = 0; s < nsteps; ++s) {             /* src/hrweno_tvdode.f90:136-171 */
        double *x0 = X + (size_t)(order * s) * nn;
        memcpy(x0, u, nn * sizeof(double));
        if (order == 1) {
            hro_rhs_burgers(&ctx, 0.0, u, udot);
            for (int i = 0; i < n; ++i) u[i] = u[i] + dt * udot[i];
        } else if (order == 2) {
            double *x1 = x0 + nn;
            hro_rhs_burgers(&ctx, 0.0, u, udot);
            for (int i = 0; i < n; ++i) x1[i] = u[i] + dt * udot[i];
            hro_rhs_burgers(&ctx, 0.0, x1, udot);
            for (int i = 0; i < n; ++i) u[i] = ((u[i] + x1[i]) + dt * udot[i]) / 2;
        } else {
            double *x1 = x0 + nn, *x2 = x1 + nn;
            hro_rhs_burgers(&ctx, 0.0, u, udot);
            for (int i = 0; i < n; ++i) x1[i] = u[i] + dt * udot[i];
            hro_rhs_burgers(&ctx, 0.0, x1, udot);
            for (int i = 0; i < n; ++i) x2[i] = ((3 * u[i] + x1[i]) + dt * udot[i]) / 4;
            hro_rhs_burgers(&ctx, 0.0, x2, udot);
            for (int i = 0; i < n; ++i) u[i] = ((u[i] + 2 * x2[i]) + (2 * dt) * udot[i]) / 3;
        }
    }
#define VJP(x, m, o) hro_burgers_rhs_vjp_ex(n, k, eps, flux, bc, alpha_fixed, width, x, m, o)
    for (int64_t s = nsteps - 1; s >= 0; --s) {
        const double *x0 = X + (size_t)(order * s) * nn;
        if (order == 1) {                              /* u' = u + dt L(u) */
            VJP(x0, g, jt);
            for (int i = 0; i < n; ++i) g[i] = g[i] + dt * jt[i];
        } else if (order == 2) {                       /* u' = (u + u1 + dt L(u1)) / 2 */
            VJP(x0 + nn, g, jt);
            for (int i = 0; i < n; ++i) l1[i] = 0.5 * g[i] + 0.5 * dt * jt[i];
            VJP(x0, l1, jt);
            for (int i = 0; i < n; ++i) g[i] = 0.5 * g[i] + l1[i] + dt * jt[i];
        } else {
            /* u' = (u + 2 u2 + 2 dt L(u2)) / 3 */
            VJP(x0 + 2 * nn, g, jt);
            for (int i = 0; i < n; ++i) l2[i] = (2.0 / 3) * g[i] + (2.0 / 3) * dt * jt[i];
            /* u2 = (3 u + u1 + dt L(u1)) / 4 */
            VJP(x0 + nn, l2, jt);
            for (int i = 0; i < n; ++i) l1[i] = 0.25 * l2[i] + 0.25 * dt * jt[i];
            /* u1 = u + dt L(u) */
            VJP(x0, l1, jt);
            for (int i = 0; i < n; ++i)
                g[i] = (1.0 / 3) * g[i] + 0.75 * l2[i] + l1[i] + dt * jt[i];
        }
    }
#undef VJP
done:
    free(X);
    free(udot);
    free(l1);
    free(l2);
    free(jt);
    free(ctx.work);
    return rc;
}

int hro_burgers_rk_adjoint(int n, int k, double eps, double alpha_fixed, const double *width,
                           int order, int64_t nsteps, double dt, double *u, double *g)
{
    return hro_burgers_rk_adjoint_ex(n, k, eps, HRO_FLUX_LF, HRO_BC_PERIODIC, alpha_fixed, width,
                                     order, nsteps, dt, u, g);
}

/* The multistep integrator from a fresh object (istate 1, src/hrweno_tvdode.f90:203-271),
 * S_0 = u -> S_N, storing every state, then the reverse sweep.
 *   start (:236-249): S_{c+1} = RK3(S_c) for c = 0..3, each a single step of the RK3 method
 *     of src/hrweno_tvdode.f90:160-171 (taken even past tout); L(S_c) is kept as udotold;
 *   steps (:252-268): S_{m+1} = (25 S_m + 50 dt L(S_m) + 7 S_{m-4} + 10 dt L(S_{m-4}))/32 for
 *     m = 4, 5, ... while not is_done(t, tout, dt).
 * S_n feeds S_{n+1} (RK3 for n < 4, the multistep formula otherwise) and, as the oldest
 * history column, S_{n+5}; so with lambda_m = dJ/dS_m (zero past N):
 *   n >= 4: lambda_n = 25/32 l_{n+1} + 7/32 l_{n+5} + dt/32 J_n^T (50 l_{n+1} + 10 l_{n+5})
 *   n <  4: lambda_n = RK3^T(S_n) l_{n+1} + 7/32 l_{n+5} + 10 dt/32 J_n^T l_{n+5}
 * u <- S_N, *t advanced, *nsteps = N (0 if is_done(t0) already holds: no step, g kept). */
int hro_burgers_ms_adjoint(int n, int k, double eps, int flux, int bc, double alpha_fixed,
                           const double *width, double *t, double tout, double dt, double *u,
                           double *g, int64_t *nsteps)
{
    if (n < 1 || k < 1 || k > 3 || !(dt != 0.0)) return -1;
    *nsteps = 0;
    if (hro_is_done(*t, tout, dt)) return 0;           /* :229 */
    int64_t M = 0;                                     /* multistep steps, as the loop counts */
    {
        double tt = *t;
        for (int c = 0; c < 4; ++c) tt = tt + dt;
        while (!hro_is_done(tt, tout, dt)) {
            tt = tt + dt;
            ++M;
        }
    }
    const int64_t N = 4 + M;
    const size_t nn = (size_t)n;
    double *S = (double *)malloc((size_t)(N + 1) * nn * sizeof(double));
    double *Y = (double *)malloc((size_t)8 * nn * sizeof(double));   /* RK3 stages of the start */
    double *lam = (double *)calloc((size_t)(N + 1) * nn, sizeof(double));
    double *udot = (double *)malloc(nn * sizeof(double));
    double *ud4 = (double *)malloc((size_t)(N + 1) * nn * sizeof(double));
    double *mu = (double *)malloc(nn * sizeof(double));
    double *jt = (double *)malloc(nn * sizeof(double));
    double *l1 = (double *)malloc(nn * sizeof(double));
    double *l2 = (double *)malloc(nn * sizeof(double));
    hro_burgers_ctx ctx;
    ctx.n = n; ctx.k = k; ctx.flux = flux; ctx.bc = bc; ctx.eps = eps;
    ctx.alpha_fixed = alpha_fixed; ctx.width = width; ctx.cnu = NULL;
    ctx.work = (double *)malloc((size_t)5 * (nn + 2 * k + 2) * sizeof(double));
    int rc = 0;
    if (!S || !Y || !lam || !udot || !ud4 || !mu || !jt || !l1 || !l2 || !ctx.work) {
        rc = -2;
        goto done;
    }
#define ST(m) (S + (size_t)(m) * nn)
#define LM(m) (lam + (size_t)(m) * nn)
    memcpy(ST(0), u, nn * sizeof(double));
    for (int c = 0; c < 4; ++c) {                      /* :241-245 */
        const double *x0 = ST(c);
        double *x1 = Y + (size_t)(2 * c) * nn, *x2 = x1 + nn, *xo = ST(c + 1);
        hro_rhs_burgers(&ctx, 0.0, x0, ud4 + (size_t)c * nn);   /* udotold(:, 4-c) */
        hro_rhs_burgers(&ctx, 0.0, x0, udot);
        for (int i = 0; i < n; ++i) x1[i] = x0[i] + dt * udot[i];
        hro_rhs_burgers(&ctx, 0.0, x1, udot);
        for (int i = 0; i < n; ++i) x2[i] = ((3 * x0[i] + x1[i]) + dt * udot[i]) / 4;
        hro_rhs_burgers(&ctx, 0.0, x2, udot);
        for (int i = 0; i < n; ++i) xo[i] = ((x0[i] + 2 * x2[i]) + (2 * dt) * udot[i]) / 3;
        *t = *t + dt;
    }
    for (int64_t m = 4; m < N; ++m) {                  /* :255-266 */
        const double *um = ST(m), *u4 = ST(m - 4), *d4 = ud4 + (size_t)(m - 4) * nn;
        double *dm = ud4 + (size_t)m * nn;
        hro_rhs_burgers(&ctx, 0.0, um, dm);
        double *xo = ST(m + 1);
        for (int i = 0; i < n; ++i)
            xo[i] = ((((25 * um[i]) + (50 * dt) * dm[i]) + 7 * u4[i]) + (10 * dt) * d4[i]) / 32;
        *t = *t + dt;
    }
    memcpy(u, ST(N), nn * sizeof(double));
    memcpy(LM(N), g, nn * sizeof(double));
    const double c7 = g_plant == 7 ? 0.0 : 7.0 / 32;   /* planted 7: history state dropped */
    const double c10 = g_plant == 8 ? 0.0 : 10.0 / 32; /* planted 8: history slope dropped */
#define VJP(x, m, o) hro_burgers_rhs_vjp_ex(n, k, eps, flux, bc, alpha_fixed, width, x, m, o)
    for (int64_t m = N - 1; m >= 0; --m) {
        const double *lp = LM(m + 1);
        const double *l5 = m + 5 <= N ? LM(m + 5) : NULL;
        double *lm = LM(m);
        if (m >= 4) {
            for (int i = 0; i < n; ++i) mu[i] = (50.0 / 32) * lp[i] + (l5 ? c10 * l5[i] : 0.0);
            VJP(ST(m), mu, jt);
            for (int i = 0; i < n; ++i)
                lm[i] = (25.0 / 32) * lp[i] + (l5 ? c7 * l5[i] : 0.0) + dt * jt[i];
        } else {
            const double *x0 = ST(m), *x1 = Y + (size_t)(2 * m) * nn, *x2 = x1 + nn;
            VJP(x2, lp, jt);
            for (int i = 0; i < n; ++i) l2[i] = (2.0 / 3) * lp[i] + (2.0 / 3) * dt * jt[i];
            VJP(x1, l2, jt);
            for (int i = 0; i < n; ++i) l1[i] = 0.25 * l2[i] + 0.25 * dt * jt[i];
            /* stage 1 of the RK3 step and the history term share J(S_m)^T */
            for (int i = 0; i < n; ++i) mu[i] = l1[i] + (l5 ? c10 * l5[i] : 0.0);
            VJP(x0, mu, jt);
            for (int i = 0; i < n; ++i)
                lm[i] = (1.0 / 3) * lp[i] + 0.75 * l2[i] + l1[i] + (l5 ? c7 * l5[i] : 0.0) +
                        dt * jt[i];
        }
    }
    memcpy(g, LM(0), nn * sizeof(double));
    *nsteps = N;
#undef ST
#undef LM
done:
    free(S);
    free(Y);
    free(lam);
    free(udot);
    free(ud4);
    free(mu);
    free(jt);
    free(l1);
    free(l2);
    free(ctx.work);
    return rc;
}
#undef VJP
