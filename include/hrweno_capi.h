/*
 * hrweno_capi.h -- C-ABI of the MI355X (gfx950) HR-WENO hot path.
 *
 * Plain C types only: device arrays are `double*` obtained from hrw_device_alloc (or any HIP
 * allocation, e.g. a torch tensor's data pointer), sizes are 64-bit, streams are opaque
 * hipStream_t values (NULL = the default stream).  Every function returns an hrw_status;
 * on failure hrw_last_error() holds a message for the calling thread.  Invalid arguments
 * produce the same messages as the reference's `error stop` texts.
 *
 * Citations (path:line) are relative to the reference repository HugoMVale/HR-WENO.
 */
#ifndef HRWENO_CAPI_H
#define HRWENO_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *hrw_stream_t;
typedef struct hrw_weno_s *hrw_weno_t;
typedef struct hrw_problem_s *hrw_problem_t;
typedef struct hrw_ode_s *hrw_ode_t;
typedef struct hrw_batch_s *hrw_batch_t;
typedef struct hrw_adjoint_s *hrw_adjoint_t;

enum hrw_status {
    HRW_OK = 0,
    HRW_ERR_ARG = 1,     /* invalid argument (reference: error stop)             */
    HRW_ERR_HIP = 2,     /* HIP runtime or launch failure                        */
    HRW_ERR_ALLOC = 3,   /* device allocation failed                             */
    HRW_ERR_STATE = 4    /* object in a state that forbids the call              */
};

/* Arithmetic form.  EXACT keeps the reference's evaluation order (bit-identical to the
 * reference Fortran compiled without FMA).  FAST is the same algebra with fused multiply-adds,
 * the nonlinear weights over a common denominator, work shared between neighbouring cells and
 * reciprocals by hardware estimate plus Newton steps (one shared by the two faces of a 1D
 * reconstruction): a few ulp per reconstruction.  Measured
 * agreement with the reference arithmetic: within 1e-12 of max|u| per cell (and 1e-12
 * relative where |u| >= 1e-3 max|u|) for RK1-3 at every horizon tested (100 RK3 steps at 2^24
 * cells, ~1250 to t = 0.5 at 1000 cells, 200 on the 4096-line ensemble).  The reference's own
 * 5-step multistep method amplifies rounding at CFL 0.4 (in its own arithmetic a 1-ulp change
 * of u0 grows to ~2e-9 after 100 steps of 1D Burgers, and to ~1e-11 after 200 and ~1e-1 after
 * 500 steps of the 2D PBE), so with that method FAST meets 1e-12 only over short horizons (~15
 * steps on 1D Burgers, ~20 on the 2D PBE); EXACT stays bit-identical at any horizon. */
enum hrw_arith { HRW_ARITH_EXACT = 0, HRW_ARITH_FAST = 1 };

/* Numerical flux (src/hrweno_fluxes.f90:22-76). */
enum hrw_numflux { HRW_FLUX_LAX_FRIEDRICHS = 0, HRW_FLUX_GODUNOV = 1 };

/* Built-in physical flux functions f(u, x, t) (src/hrweno_fluxes.f90:12-18). */
enum hrw_fluxfn {
    HRW_FN_BURGERS = 0,  /* f = u**2/2            (example/example1_burgers_1d_fv.f90:120) */
    HRW_FN_LINEAR = 1    /* f = (a + b*x)*u       (example/example2_pbe_2d_fv.f90:140,153) */
};

/* Boundary treatment of a problem descriptor. */
enum hrw_bc {
    HRW_BC_PERIODIC = 0,        /* periodic ghost cells                                   */
    HRW_BC_ZERO_GRADIENT = 1,   /* example1:103-104: fedges(0)=fedges(1), fedges(n)=fedges(n-1) */
    HRW_BC_ZERO_FLUX = 2,       /* example2:117-120: zero flux at both ends               */
    HRW_BC_INFLOW0_OUTFLOW = 3  /* zero flux in at the low end, upwind f(vr(n)) out at the high end */
};

/* ---------------------------------------------------------------- runtime ---------- */
const char *hrw_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char *hrw_last_error(void);
/* The same message copied into buf (NUL-terminated, truncated to len-1 characters). */
int hrw_last_error_copy(char *buf, int64_t len);
int hrw_device_count(int *count);
int hrw_device_alloc(size_t bytes, void **ptr);
int hrw_device_free(void *ptr);
/* Free and total device memory of the current device (hipMemGetInfo). */
int hrw_device_mem_info(size_t *free_bytes, size_t *total_bytes);
int hrw_memcpy_htod(void *dst, const void *src, size_t bytes, hrw_stream_t stream);
int hrw_memcpy_dtoh(void *dst, const void *src, size_t bytes, hrw_stream_t stream);
int hrw_memcpy_dtod(void *dst, const void *src, size_t bytes, hrw_stream_t stream);
int hrw_stream_create(hrw_stream_t *stream);
int hrw_stream_destroy(hrw_stream_t stream);
int hrw_stream_synchronize(hrw_stream_t stream);
/* Elapsed milliseconds between two points of a stream: record a start event, then stop. */
int hrw_timer_start(hrw_stream_t stream);
int hrw_timer_stop(hrw_stream_t stream, double *ms);

/* ---------------------------------------------------------------- WENO ------------- */
/* Public tables d(0:k-1) and c(0:k-1,-1:k-1) (src/hrweno_weno.f90:9-21); c is written
 * column-major, c[(r+1)*k + j] = c(j,r).  d and c need k and k*(k+1) doubles. */
int hrw_weno_tables(int k, double *d, double *c);

/* weno(ncells, k, eps [, xedges]) -- replaces src/hrweno_weno.f90:54-127 (weno_init).
 * xedges: host array (0:ncells) or NULL for a uniform grid; for a non-uniform grid the
 * coefficient table cnu (src/hrweno_weno.f90:221-297) is computed on the device. */
int hrw_weno_create(int ncells, int k, double eps, const double *xedges, int arith,
                    hrw_weno_t *out);
int hrw_weno_destroy(hrw_weno_t w);
int hrw_weno_info(hrw_weno_t w, int *ncells, int *k, double *eps, int *uniform);
/* Copy the cnu table, cnu(0:k-1,-1:k-1,1:ncells) column-major, to a host array. */
int hrw_weno_get_cnu(hrw_weno_t w, double *cnu_host);
/* Stateless form of the same table -- replaces src/hrweno_weno.f90:221-297 (weno_calc_cnu):
 * host xedges(0:ncells) in, host cnu out, computed on the device. */
int hrw_weno_calc_cnu(int ncells, int k, const double *xedges, double *cnu_host);

/* w%reconstruct(v, vl, vr) -- replaces src/hrweno_weno.f90:129-219 (weno_reconstruct).
 * Device arrays with element strides (strided sections as in example2:107). */
int hrw_weno_reconstruct(hrw_weno_t w, const double *v, int64_t incv, double *vl,
                         int64_t incl, double *vr, int64_t incr, hrw_stream_t stream);
/* The same on nlines lines in one launch: element i of line l is v[l*ldv + i*incv],
 * vl/vr[l*ldo + i*inco] (example2:97-112 does this line by line). */
int hrw_weno_reconstruct_lines(hrw_weno_t w, int64_t nlines, const double *v, int64_t incv,
                               int64_t ldv, double *vl, double *vr, int64_t inco,
                               int64_t ldo, hrw_stream_t stream);
/* Host arrays (copied through device staging buffers; synchronous). */
int hrw_weno_reconstruct_host(hrw_weno_t w, const double *v, int64_t incv, double *vl,
                              double *vr);
/* Stateless forms for hosts that keep the weno object themselves (the Fortran module):
 * the same src/hrweno_weno.f90:129-219 on host arrays (cnu: host table or NULL) or on
 * device arrays (uniform grid). */
int hrw_reconstruct_host(int ncells, int k, double eps, int arith, const double *cnu,
                         const double *v, int64_t incv, double *vl, int64_t incl, double *vr,
                         int64_t incr);
int hrw_reconstruct_dev(int ncells, int k, double eps, int arith, const double *v,
                        int64_t incv, double *vl, int64_t incl, double *vr, int64_t incr,
                        hrw_stream_t stream);
/* Stateless many-lines forms of src/hrweno_weno.f90:129-219, for example 2's loops over the
 * x1 lines (contiguous, ld = n1) and the x2 lines (stride n1, ld = 1) of one state
 * (example/example2_pbe_2d_fv.f90:95-112) in one launch: element i of line l is
 * v[l*ldv + i*incv], vl/vr[l*ldo + i*inco], 1 <= nlines <= 65535.  The host form copies the
 * spans the lines address (cnu: host table or NULL); the device form takes a uniform grid. */
int hrw_reconstruct_lines_host(int ncells, int k, double eps, int arith, const double *cnu,
                               int64_t nlines, const double *v, int64_t incv, int64_t ldv,
                               double *vl, double *vr, int64_t inco, int64_t ldo);
int hrw_reconstruct_lines_dev(int ncells, int k, double eps, int arith, int64_t nlines,
                              const double *v, int64_t incv, int64_t ldv, double *vl, double *vr,
                              int64_t inco, int64_t ldo, hrw_stream_t stream);

/* ---------------------------------------------------------------- fluxes ----------- */
/* h(i) = lax_friedrichs(f, vm(i), vp(i), [x(i)], t, alpha) -- src/hrweno_fluxes.f90:22-45
 * h(i) = godunov(f, vm(i), vp(i), [x(i)], t)             -- src/hrweno_fluxes.f90:47-76
 * for the built-in f = fn with coefficients (a, b); x may be NULL (taken as 0). */
int hrw_flux_lax_friedrichs(int fn, double a, double b, int64_t n, const double *vm,
                            const double *vp, const double *x, double t, double alpha,
                            double *h, hrw_stream_t stream);
int hrw_flux_godunov(int fn, double a, double b, int64_t n, const double *vm,
                     const double *vp, const double *x, double t, double *h,
                     hrw_stream_t stream);

/* ---------------------------------------------------------------- grids ------------ */
/* grid1%linear/%bilinear/%log/%geometric -- src/hrweno_grids.f90:41-230.  Host arrays:
 * edges(0:n).  hrw_grid_compute fills center(n), width(n) (src/hrweno_grids.f90:232-250). */
int hrw_grid_linear(double xmin, double xmax, int ncells, double *edges);
int hrw_grid_bilinear(double xmin, double xcross, double xmax, int ncells1, int ncells2,
                      double *edges);
int hrw_grid_log(double xmin, double xmax, int ncells, double *edges);
int hrw_grid_geometric(double xmin, double xmax, double ratio, int ncells, double *edges);
int hrw_grid_compute(int ncells, const double *edges, double *center, double *width);

/* ---------------------------------------------------------------- problems --------- */
/* A device problem descriptor stands for the user RHS `fu` of the reference examples, so
 * that the integrators can fuse RHS and stage update into one pass per stage.
 *
 * 1D Burgers u_t + (u**2/2)_x = 0 on a linear grid [xmin, xmax] of n cells, or on the cell
 * widths `width` (host, n; NULL = linear grid) -- example/example1_burgers_1d_fv.f90:72-122.
 * alpha < 0 selects alpha = max|v| of each stage input for Lax-Friedrichs. */
int hrw_problem_burgers1d(int n, int k, double eps, int numflux, int bc, double alpha,
                          double xmin, double xmax, const double *width, int arith,
                          hrw_problem_t *out);
/* The same on the non-uniform grid of edges xedges (host, n+1): the example-1 right-hand side
 * with weno(ncells, k, eps, xedges) -- reconstruction with the grid's own coefficients cnu
 * (src/hrweno_weno.f90:100-112,174-178,221-297) and widths xedges(i+1)-xedges(i)
 * (src/hrweno_grids.f90:247).  bc must be HRW_BC_ZERO_GRADIENT (the reference extends the
 * edges linearly and the cell values by a constant beyond both ends).  The reconstruction
 * takes the reference's evaluation order in both arithmetic forms. */
int hrw_problem_burgers1d_edges(int n, int k, double eps, int numflux, int bc, double alpha,
                                const double *xedges, int arith, hrw_problem_t *out);
/* 2D population balance n_t + (g1 n)_x1 + (g2 n)_x2 = 0, g1 = g1a + g1b*x1,
 * g2 = g2a + g2b*x2, Godunov flux, linear grids, u((j-1)*n1 + i) --
 * example/example2_pbe_2d_fv.f90:73-155.  bc: HRW_BC_ZERO_FLUX or HRW_BC_INFLOW0_OUTFLOW. */
int hrw_problem_pbe2d(int n1, int n2, int k, double eps, int bc, double x1min, double x1max,
                      double x2min, double x2max, double g1a, double g1b, double g2a,
                      double g2b, int arith, hrw_problem_t *out);
/* Kernel choice for the 1D FAST form (no reference counterpart; used by the autotuner):
 * variant 0 = register-blocked kernel with cells_per_lane (4, 6 or 8) consecutive cells per
 * lane, one chunk per wavefront; with cells_per_lane 0 (the default) the form and the cells per
 * lane are chosen by the line length (2^21 cells or more: variant 2 with 4; 2^20 or more: 6),
 * variant 1 = LDS-tile kernel (for the reverse passes of the adjoint too: the K = 2, 3
 * register-blocked reverse pass is then not used), variant 2 = the register-blocked kernel run persistently with
 * the next chunk's loads issued ahead of the current chunk's arithmetic (lines with a width
 * array take the variant 0 kernel).  Variants 0 and 2 give identical results; variant 1
 * agrees within the FAST tolerance.  For a 2D problem
 * variant must be 0 and cells_per_lane is the number of rows each wavefront marches over
 * (0 = chosen at launch so that every wavefront is resident at once); results are identical
 * for every value. */
int hrw_problem_tune(hrw_problem_t p, int variant, int cells_per_lane);
int hrw_problem_destroy(hrw_problem_t p);
int hrw_problem_neq(hrw_problem_t p, int64_t *neq);
/* One RHS evaluation udot = L(u) on device arrays (the user `rhs` of the examples). */
int hrw_problem_rhs(hrw_problem_t p, double t, const double *u, double *udot,
                    hrw_stream_t stream);

/* ---------------------------------------------------------------- integrators ------ */
/* Callback integrand on device arrays: udot = f(t, u); must enqueue its work on `stream`
 * and return 0 on success (src/hrweno_tvdode.f90:50-57, abstract interface integrand). */
typedef int (*hrw_integrand_fn)(void *ctx, double t, const double *u, double *udot,
                                hrw_stream_t stream);

/* rktvd(fu, neq, order) -- src/hrweno_tvdode.f90:69-95 (fused device problem or callback) */
int hrw_rktvd_create(hrw_problem_t p, int order, hrw_ode_t *out);
int hrw_rktvd_create_cb(hrw_integrand_fn fu, void *ctx, int64_t neq, int order, int arith,
                        hrw_ode_t *out);
/* mstvd(fu, neq) -- src/hrweno_tvdode.f90:180-201 */
int hrw_mstvd_create(hrw_problem_t p, hrw_ode_t *out);
int hrw_mstvd_create_cb(hrw_integrand_fn fu, void *ctx, int64_t neq, int arith,
                        hrw_ode_t *out);
int hrw_ode_destroy(hrw_ode_t ode);
int hrw_ode_set_stream(hrw_ode_t ode, hrw_stream_t stream);
/* ode%integrate(u, t, tout, dt [, itask]) -- src/hrweno_tvdode.f90:97-178 (rktvd) and
 * :203-271 (mstvd; itask ignored).  u is a device array of neq doubles; t is updated with
 * the reference's loop control (is_done, :273-284), including its overshoot of tout. */
int hrw_ode_integrate(hrw_ode_t ode, double *u, double *t, double tout, double dt, int itask);
/* Public fields neq, order, fevals, istate (src/hrweno_tvdode.f90:18-29). */
int hrw_ode_info(hrw_ode_t ode, int64_t *neq, int *order, int *fevals, int *istate);
/* Non-finite state detection (no reference counterpart; the reference integrates a blown-up
 * state on silently).  With check_finite != 0 (the default) every hrw_ode_integrate call that
 * took a step ends with one scan of u for NaN or Inf (one read of the state and a
 * synchronisation of the ODE's stream); if it finds one the call returns HRW_ERR_STATE with a
 * message and istate becomes -1, the reference's fatal-error state
 * (src/hrweno_tvdode.f90:286-297).  0 skips the scan. */
int hrw_ode_set_check(hrw_ode_t ode, int check_finite);
/* Serialised ODE state for resume in another object or process (src/hrweno_tvdode.f90:18-29:
 * fevals, istate; :44-45, :236-249: the multistep history uold/udotold).  The blob holds a
 * header (method, order, neq, fevals, istate, t), u (device, neq doubles) and, for mstvd, the
 * four history columns of u and L(u) (8 x neq doubles).  buf is host memory of at least
 * hrw_ode_state_size bytes.  hrw_ode_load needs an object of the same method, order and neq
 * (built from the same problem or integrand) and writes u (device) and t; the next
 * hrw_ode_integrate continues exactly as the saved object would have (bit for bit).  A
 * recording set by hrw_ode_tape is not part of the state. */
int hrw_ode_state_size(hrw_ode_t ode, int64_t *bytes);
int hrw_ode_save(hrw_ode_t ode, const double *u, double t, void *buf, int64_t bytes);
int hrw_ode_load(hrw_ode_t ode, double *u, double *t, const void *buf, int64_t bytes);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops that forbid hipMalloc): nsteps SSP-RK steps of order 1..3
 * with step dt of the 1D Burgers problem p on the device state u -- the steps
 * hrw_ode_integrate takes on hrw_rktvd_create(p, order), bit for bit -- with the stage buffers
 * and the Lax-Friedrichs alpha slot sets in a static device array of the library (16 MiB +
 * 6 KiB per device, held until the process ends) instead of allocations.  At most 2^20 cells
 * (larger problems are refused).  Allocates nothing, does not synchronise (no NaN/Inf scan),
 * enqueues on stream.  Every call on a device shares that device's array: calls that overlap
 * on one device (other streams or threads) race. */
int hrw_rktvd_steps_noalloc(hrw_problem_t p, int order, double *u, int64_t nsteps, double dt,
                            hrw_stream_t stream);

/* ---------------------------------------------------------------- batched lines ---- */
/* Ensemble of nlines independent periodic 1D Burgers lines (the many-lines use of
 * rktvd(fu, neq, 3)%integrate, src/hrweno_tvdode.f90:160-171, with the example-1 RHS):
 * line l has ncells[l] cells on [xmin[l], xmax[l]] (NULL: [-1, 1]), WENO order k[l] and its own
 * time step dt[l].  FAST arithmetic only.  The state is one device array of `total` doubles;
 * line l occupies [offsets[l], offsets[l] + ncells[l]) (offsets are 256-byte aligned). */
int hrw_batch_burgers_create(int64_t nlines, const int *ncells, const int *k, const double *dt,
                             const double *xmin, const double *xmax, double eps, int numflux,
                             double alpha, int arith, hrw_batch_t *out);
int hrw_batch_destroy(hrw_batch_t b);
/* total doubles of the state; offsets (nlines+1 values, may be NULL). */
int hrw_batch_size(hrw_batch_t b, int64_t *total, int64_t *offsets);
/* The batch runs on a non-default stream of its own (graph capture needs one).  Until a
 * stream is set here, every call that launches work (step, step_per_line, tune, the batch
 * adjoint's forward and backward) first waits for the work queued on the NULL stream and makes
 * the NULL stream wait for its own work, so callers that use the NULL stream see its results
 * in order.  With a stream set, ordering against other streams is the caller's
 * (hrw_batch_synchronize waits for the batch's stream). */
int hrw_batch_set_stream(hrw_batch_t b, hrw_stream_t stream);
int hrw_batch_synchronize(hrw_batch_t b);
/* nsteps SSP-RK3 steps of every line: one launch per stage for all lines; use_graph != 0 replays
 * a captured hipGraph of one step.  t (host, nlines, may be NULL) advances by nsteps*dt[l]. */
int hrw_batch_step(hrw_batch_t b, double *u, int64_t nsteps, double *t, int use_graph);
/* The same, each line advanced by its own launches (the one-launch-per-line path). */
int hrw_batch_step_per_line(hrw_batch_t b, double *u, int64_t nsteps, double *t);
/* Stage kernel of the batch: one wavefront per chunk of 62*cells_per_lane cells (CHUNK), or
 * one wavefront per SIMD slot walking the chunk table with the next chunk's loads issued ahead
 * (PERSISTENT).  Results are bit-identical for every form and width. */
enum { HRW_BATCH_CHUNK = 0, HRW_BATCH_PERSISTENT = 1 };
/* Autotuner: times each form and chunk width (cells per lane 4, 6, 8) on a scratch copy of u
 * and keeps the fastest (cells_per_lane: the width chosen, may be NULL); or set them directly. */
int hrw_batch_tune(hrw_batch_t b, const double *u, int *cells_per_lane);
int hrw_batch_set_cells_per_lane(hrw_batch_t b, int cells_per_lane);
int hrw_batch_set_form(hrw_batch_t b, int form);
int hrw_batch_config(hrw_batch_t b, int *cells_per_lane, int *form);

/* ---------------------------------------------------------------- discrete adjoint - */
/* Reverse mode of nsteps fixed rktvd steps (src/hrweno_tvdode.f90:136-171) of a fused 1D
 * Burgers problem on a linear grid: the Lax-Friedrichs or the endpoint Godunov flux, periodic
 * or zero-gradient ends (example/example1_burgers_1d_fv.f90:72-109; src/hrweno_fluxes.f90:
 * 22-76; src/hrweno_weno.f90:174-216).  The reference has no adjoint; this is the added
 * differentiation form.  Periodic Lax-Friedrichs lines take the register-blocked reverse
 * kernel, the other compositions a general reverse pass.  At vm == vp the Godunov flux is
 * differentiated as the upwind state (where it is smooth).  forward records u0 -> u_N with a
 * checkpoint every ckpt_every steps; backward maps a cotangent dJ/du_N to dJ/du_0,
 * recomputing one segment of ckpt_every steps at a time.  With alpha = max|u| the gradient
 * takes the subgradient at the first cell attaining the maximum.  Device memory:
 * (nsteps/ckpt_every + 1) + (order*ckpt_every - 1) + 2 states.  ckpt_every = 0 chooses the
 * largest interval whose states fit in half of the free device memory; when every stage input
 * of the run fits (ckpt_every = nsteps), the backward recomputes nothing.  With more than one
 * segment, a single line's first backward also takes order*ckpt_every states more when they
 * fit.  Every later backward then recomputes the next segment on a second stream of its own
 * while it reverses the current one, with the same results (its results are ordered on the
 * adjoint's stream as always). */
int hrw_adjoint_create(hrw_problem_t p, int order, double dt, int64_t nsteps, int ckpt_every,
                       hrw_adjoint_t *out);
/* Every line of a batch (RK3, each line with its own dt; the Lax-Friedrichs flux); runs on
 * the batch's stream and needs the batch to outlive it.  States are the batch's flat layout. */
int hrw_batch_adjoint_create(hrw_batch_t b, int64_t nsteps, int ckpt_every, hrw_adjoint_t *out);
int hrw_adjoint_destroy(hrw_adjoint_t a);
int hrw_adjoint_set_stream(hrw_adjoint_t a, hrw_stream_t stream);
/* Cells per lane of the reverse kernel (2 or 4, default 2); results agree to rounding. */
int hrw_adjoint_set_cells_per_lane(hrw_adjoint_t a, int cells_per_lane);
int hrw_adjoint_forward(hrw_adjoint_t a, const double *u0, double *uN);
/* gN may equal g0; needs a preceding hrw_adjoint_forward (may be called again after it). */
int hrw_adjoint_backward(hrw_adjoint_t a, const double *gN, double *g0);
int hrw_adjoint_info(hrw_adjoint_t a, int64_t *neq, int64_t *nsteps, int *ckpt_every,
                     int64_t *device_bytes);
/* ode%tape(ckpt_every) / ode%adjoint(lambda): the same reverse mode on an integrator built on
 * a fused problem of the kinds above.  rktvd (hrw_rktvd_create): with ckpt_every > 0 every
 * later hrw_ode_integrate records its steps (0 stops recording); hrw_ode_adjoint maps lambda
 * (device, neq) from dJ/du at the end of the last recorded integrate call to dJ/du at its
 * start, in place.  mstvd (hrw_mstvd_create; src/hrweno_tvdode.f90:203-271): the recording
 * keeps every state from the integrate call that starts the method (istate 1; tape must be set
 * before it, else that integrate call fails with HRW_ERR_STATE) through every later call --
 * N + 9 states for N steps -- and hrw_ode_adjoint maps dJ/du at the end of the last call to
 * dJ/du at the start of the method, the four RK3 starting steps included.  hrw_ode_load and
 * ckpt_every = 0 drop an mstvd recording. */
int hrw_ode_tape(hrw_ode_t ode, int ckpt_every);
int hrw_ode_adjoint(hrw_ode_t ode, double *lambda);

#ifdef __cplusplus
}
#endif
#endif /* HRWENO_CAPI_H */
