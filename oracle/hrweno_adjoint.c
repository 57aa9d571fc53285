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

int hro_burgers_rhs_vjp(int n, int k, double eps, double alpha_fixed, const double *width,
                        const double *v, const double *mu, double *out)
{
    if (n < 1 || k < 1 || k > 3) return -1;
    double d[3], c[12];
    hro_weno_tables(k, d, c);
    double alpha = alpha_fixed;
    int jmax = -1;
    if (alpha < 0) {                                   /* alpha = maxval(abs(v)) */
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
    if (!cs || !phi) {
        free(cs);
        free(phi);
        return -2;
    }
    const int km = k - 1;
    for (int i = 0; i < n; ++i) {
        double xw[5];
        for (int m = -km; m <= km; ++m) xw[m + 2] = v[wrap((long long)i + m, n)];
        cell_forward(k, eps, d, c, xw + 2, &cs[i]);
        phi[i] = mu[i] / width[i];
    }
    for (int i = 0; i < n; ++i) out[i] = 0.0;
    double abar = 0.0;
    for (int i = 0; i < n; ++i) {
        const int ip = wrap((long long)i + 1, n), im = wrap((long long)i - 1, n);
        const double Gp = phi[ip] - phi[i];            /* cotangent of F_{i+1/2} */
        const double Gm = phi[i] - phi[im];            /* cotangent of F_{i-1/2} */
        const double gr = Gp * (0.5 * (cs[i].vr + alpha));
        double gl = Gm * (0.5 * (cs[i].vl - alpha));
        if (g_plant == 2) gl = Gm * (0.5 * (cs[i].vl + alpha));   /* planted: sign of alpha */
        abar += Gp * (-0.5 * (cs[ip].vl - cs[i].vr));
        double xb[5];
        cell_vjp(k, eps, c, &cs[i], gr, gl, xb);
        for (int m = -km; m <= km; ++m) out[wrap((long long)i + m, n)] += xb[m + 2];
    }
    if (jmax >= 0 && g_plant != 3) {                   /* planted 3: alpha held fixed */
        const double sg = (v[jmax] > 0) - (v[jmax] < 0);
        out[jmax] += abar * sg;
    }
    free(cs);
    free(phi);
    return 0;
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

/* The forward of hro_rktvd_integrate (periodic Burgers, LF) for nsteps steps, storing the
 * stage inputs, then the reverse sweep.  u <- u_N;  g: in dJ/du_N, out dJ/du_0. */
int hro_burgers_rk_adjoint(int n, int k, double eps, double alpha_fixed, const double *width,
                           int order, int64_t nsteps, double dt, double *u, double *g)
{
    if (n < 1 || k < 1 || k > 3 || order < 1 || order > 3 || nsteps < 0) return -1;
    const size_t nn = (size_t)n;
    double *X = (double *)malloc((size_t)order * (size_t)(nsteps > 0 ? nsteps : 1) * nn * sizeof(double));
    double *udot = (double *)malloc(nn * sizeof(double));
    double *l1 = (double *)malloc(nn * sizeof(double));
    double *l2 = (double *)malloc(nn * sizeof(double));
    double *jt = (double *)malloc(nn * sizeof(double));
    hro_burgers_ctx ctx;
    ctx.n = n; ctx.k = k; ctx.flux = HRO_FLUX_LF; ctx.bc = HRO_BC_PERIODIC; ctx.eps = eps;
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
    for (int64_t s = nsteps - 1; s >= 0; --s) {
        const double *x0 = X + (size_t)(order * s) * nn;
        if (order == 1) {                              /* u' = u + dt L(u) */
            hro_burgers_rhs_vjp(n, k, eps, alpha_fixed, width, x0, g, jt);
            for (int i = 0; i < n; ++i) g[i] = g[i] + dt * jt[i];
        } else if (order == 2) {                       /* u' = (u + u1 + dt L(u1)) / 2 */
            hro_burgers_rhs_vjp(n, k, eps, alpha_fixed, width, x0 + nn, g, jt);
            for (int i = 0; i < n; ++i) l1[i] = 0.5 * g[i] + 0.5 * dt * jt[i];
            hro_burgers_rhs_vjp(n, k, eps, alpha_fixed, width, x0, l1, jt);
            for (int i = 0; i < n; ++i) g[i] = 0.5 * g[i] + l1[i] + dt * jt[i];
        } else {
            /* u' = (u + 2 u2 + 2 dt L(u2)) / 3 */
            hro_burgers_rhs_vjp(n, k, eps, alpha_fixed, width, x0 + 2 * nn, g, jt);
            for (int i = 0; i < n; ++i) l2[i] = (2.0 / 3) * g[i] + (2.0 / 3) * dt * jt[i];
            /* u2 = (3 u + u1 + dt L(u1)) / 4 */
            hro_burgers_rhs_vjp(n, k, eps, alpha_fixed, width, x0 + nn, l2, jt);
            for (int i = 0; i < n; ++i) l1[i] = 0.25 * l2[i] + 0.25 * dt * jt[i];
            /* u1 = u + dt L(u) */
            hro_burgers_rhs_vjp(n, k, eps, alpha_fixed, width, x0, l1, jt);
            for (int i = 0; i < n; ++i)
                g[i] = (1.0 / 3) * g[i] + 0.75 * l2[i] + l1[i] + dt * jt[i];
        }
    }
done:
    free(X);
    free(udot);
    free(l1);
    free(l2);
    free(jt);
    free(ctx.work);
    return rc;
}
