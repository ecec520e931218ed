/*
 * ns_api.h -- C ABI of the MI355X (gfx950) time-advance library for the
 * mathe1s/NavierStokes 2-D solver.
 *
 * The reference exposes its hot path as Fortran module procedures that share
 * state through `module global` (global.f90).  This header is the drop-in
 * boundary that replaces them: a context (`ns_ctx`) holds, on one GPU, what
 * `global` holds on the host (scalars, 1-D profiles, base flow, state) and
 * every reference routine on the path has an entry point here.  The Fortran
 * layer (navierstokes_amd/fortran/ns_capi.f90, ns_dropin.f90) keeps the reference's module
 * and procedure names and calls these through ISO_C_BINDING; the Python layer
 * (navierstokes_amd/solver.py) calls them through ctypes.
 *
 * Conventions
 *   - every function returns 0 on success or a negative NS_ERR_* code; the
 *     message of the last failure on the calling thread is ns_last_error();
 *   - host arrays use the reference's layout U(1:nk,1:imax,1:jmax):
 *     element (k,i,j), 0-based, at U[k + nk*(i + imax*j)] (AoS); they are
 *     borrowed for the duration of the call only;
 *   - device buffers belong to the context; the state stays resident in HBM
 *     between calls and is kept as structure-of-arrays planes;
 *   - one context = one device + one HIP stream; contexts are not thread
 *     safe, several contexts may coexist.
 */
#ifndef NS_API_H
#define NS_API_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NS_API_VERSION 1

/* status codes */
#define NS_OK 0
#define NS_ERR_PARAM (-1)     /* invalid argument / parameter set */
#define NS_ERR_HIP (-2)       /* HIP runtime failure */
#define NS_ERR_NOMEM (-3)     /* device allocation failed */
#define NS_ERR_STATE (-4)     /* call out of order (e.g. step before profiles) */
#define NS_ERR_NONFINITE (-5) /* NaN/Inf found by ns_check_finite */

/* precision modes */
#define NS_PREC_FP64 0      /* reference semantics, FP64 everywhere */
#define NS_PREC_FP32_PERT 1 /* FP32 perturbation about the FP64 base flow */

/* flags */
#define NS_FLAG_INVISCID 1u           /* drop the 1/Re and 1/(Re Pr) terms (Re -> inf);
                                         the species diffusion 1/Pe stays */
#define NS_FLAG_METRIC_EVERYWHERE 2u  /* apply meshx/meshy at every point (stretched
                                         grids); default = reference region tests */
#define NS_FLAG_PERIODIC_X 4u         /* periodic in x (no reference counterpart: the
                                         reference has one-sided x stencils, x buffers
                                         and inflow/outflow columns, diff.f90:18-36,
                                         rk.f90:98-104, cc.f90:23-35): wrapped interior
                                         x stencils, no x buffer or metric layers, no
                                         inflow/outflow columns, the LDDRK update on
                                         every column, and the x filter an in-place
                                         sweep around the ring from i = 1; imax >= 64 */

typedef struct ns_params {
    int32_t imax, jmax; /* grid points (global.f90:12) */
    int32_t nk;         /* 4 = (rho,u,v,p); 5 adds the passive scalar psi */
    int32_t D;          /* buffer-zone width in points (global.f90:86) */
    double dm, dn, dt;  /* grid steps and time step (global.f90:11) */
    double sigmaf;      /* selective-filter strength (global.f90:92) */
    double gamma, Re, Pr, Pe, alpha_cond, Tref;
    /* transport fits ln mu = A ln T + B/T + C/T^2 + D and reference values,
       coefficients.f90:71-123; ns_default_params fills the N2 set */
    double A_mi, B_mi, C_mi, D_mi, MI0;
    double A_k, B_k, C_k, D_k, K0;
    /* acoustic source, rk.f90:171-193 (amplitude 0 disables it) */
    double amplitude, wf, r0_acous, m_pt_acous, n_pt_acous;
    int32_t precision; /* NS_PREC_* */
    uint32_t flags;    /* NS_FLAG_* */
    int32_t device;    /* HIP device ordinal */
    int32_t reserved;
} ns_params;

typedef struct ns_ctx ns_ctx;

/* library / parameters --------------------------------------------------- */
int32_t ns_api_version(void);
const char *ns_last_error(void);
/* dados.in defaults + N2 transport (coefficients.f90:71-123, global.f90:38-39) */
int32_t ns_default_params(ns_params *p);

/* lifetime: allocates the device state (global.f90 allocatables) */
int32_t ns_create(const ns_params *p, ns_ctx **out);
int32_t ns_destroy(ns_ctx *ctx);
/* run on an external HIP stream (hipStream_t passed as void*; NULL = own) */
int32_t ns_set_stream(ns_ctx *ctx, void *hip_stream);
int32_t ns_synchronize(ns_ctx *ctx);

/* inputs produced by init.f90:56-206 / 350-404 on the host ---------------- */
/* meshx(imax), meshy(jmax), sigmax(imax), sigmay(jmax), m(imax), n(jmax),
   rhob(jmax), psib(jmax) (psib may be NULL when nk == 4) */
int32_t ns_set_profiles(ns_ctx *ctx, const double *meshx, const double *meshy,
                        const double *sigmax, const double *sigmay, const double *m,
                        const double *n, const double *rhob, const double *psib);
/* base flow Ub(nk,imax,jmax), init.f90:392-404 */
int32_t ns_set_base(ns_ctx *ctx, const double *Ub);

/* state ------------------------------------------------------------------- */
int32_t ns_put_state(ns_ctx *ctx, const double *U);  /* host AoS -> device */
int32_t ns_get_state(ns_ctx *ctx, double *U);        /* device -> host AoS */
/* gather npts points (1-based (i,j) pairs) x nk values: euler.f90:156-161 probes */
int32_t ns_probe(ns_ctx *ctx, int32_t npts, const int32_t *ij, double *out);
/* ns_probe without any device allocation (no hipMalloc, hipFree or
   allocating library call): for callers that cannot allocate, such as
   harnesses that forbid it or per-step probe loops that must not touch the
   allocator.  No counterpart in the reference.  The point list and the
   gathered values are staged in a fixed device array of NS_PROBE_NOALLOC_MAX
   points, one per device, shared by every call on that device: calls that
   overlap on one device race.  npts above NS_PROBE_NOALLOC_MAX fails with
   NS_ERR_PARAM.  Same arguments and results as ns_probe. */
#define NS_PROBE_NOALLOC_MAX 4096
int32_t ns_probe_noalloc(ns_ctx *ctx, int32_t npts, const int32_t *ij, double *out);
/* NS_ERR_NONFINITE if any state value is NaN/Inf */
int32_t ns_check_finite(ns_ctx *ctx);
/* outt.f90:49-55: the vorticity column of the snapshot files,
   omega = dv/dm - du/dn with the DRP operators of diffall.f90 (no metric
   factors), computed on the device from the current state.  omega: host
   array of jmax x imax values, row-major (omega[j*imax + i]). */
int32_t ns_vorticity(ns_ctx *ctx, double *omega);
/* outt.f90:63-110 output_energy / kineticenergy_x: ec[t] = the kinetic
   energy of column cols[t] (1-based) integrated over rows D..jmax-D,
   (2 sum_{j=D+1}^{jmax-D-1} (u^2+v^2)/2 + the two end rows) dn/2, from the
   current state, reduced on the device. */
int32_t ns_station_energy(ns_ctx *ctx, int32_t nst, const int32_t *cols, double *ec);
/* out[j] = mean over columns i0..i1 (1-based, inclusive) of component k of
   the state, j = 1..jmax (out has jmax values), reduced on the device.  For
   the statistics of the reference's mixing-layer runs (the x-averaged
   profiles behind its momentum thickness); no routine of the reference
   computes it on this path. */
int32_t ns_row_average(ns_ctx *ctx, int32_t k, int32_t i0, int32_t i1, double *out);

/* time advance ------------------------------------------------------------- */
/* one LDDRK6 step, rk.f90:12-75 (rk_euler(U,t,iter,nk)) */
int32_t ns_rk_euler(ns_ctx *ctx, int32_t iter);
/* nsteps steps with iter = iter0 .. iter0+nsteps-1 (euler.f90:123-130);
   the step is captured once into a HIP graph and replayed */
int32_t ns_rk_steps(ns_ctx *ctx, int32_t iter0, int32_t nsteps);

/* operator-level mirrors (host AoS arrays in/out; for parity checks) -------- */
/* diff.f90:6-38 deropm (+ metricsX :315-350), any nk in 1..8 */
int32_t ns_deropm(ns_ctx *ctx, int32_t nk, double *dU, const double *U);
/* diff.f90:127-151 deropn (+ metricsY :276-311) */
int32_t ns_deropn(ns_ctx *ctx, int32_t nk, double *dU, const double *U);
/* coefficients.f90:27-59 MOLECULAR_MI, k_coefficient on T(imax,jmax) */
int32_t ns_transport(ns_ctx *ctx, const double *T, double *MI, double *K1);
/* rhs.f90:10-46 rhs_euler(dUdt,U,nk) with the source of (ct, iter) */
int32_t ns_rhs(ns_ctx *ctx, double *dUdt, const double *U, double ct, int32_t iter);
/* rk.f90:153-169 rk_steep(alphat,betat,ct,U,dudt,w,nk) */
int32_t ns_rk_steep(ns_ctx *ctx, double alphat, double betat, double *U, const double *dudt,
                    double *w);
/* rk.f90:77-111 bufferxy(U,nk) */
int32_t ns_bufferxy(ns_ctx *ctx, double *U);
/* filtering.f90:9-96 lddfilterx(U,nk), :98-190 lddfiltery(U,nk) */
int32_t ns_filter_x(ns_ctx *ctx, double *U);
int32_t ns_filter_y(ns_ctx *ctx, double *U);
/* cc.f90:6-68 ccpml(U,nk) */
int32_t ns_ccpml(ns_ctx *ctx, double *U);
/* rk.f90:171-193 acousticfont(ct,iter): s_acou(imax,jmax) */
int32_t ns_acousticfont(ns_ctx *ctx, double ct, int32_t iter, double *s_acou);

/* rhs_euler split as the reference splits it, for the drop-in modules
   derivs, eqeuler and diff (host arrays; T, K1, MI, dT*, psiDiff, d2* and
   s_acou are imax x jmax, i.e. the reference's (1,imax,jmax) / (imax,jmax);
   tau, dtau* are (3,imax,jmax); dU* are (nk,imax,jmax)) ----------------- */
/* derivs.f90:9-46 deriv(U,nk): dUdm, dUdn of U; dTdm, dTdn = K1 * d(T)/dm,
   K1 * d(T)/dn (after kdT); d2Tdm, d2Tdn their derivatives; for nk = 5
   psiDiff = U1 alpha_cond (dUdm5 + dUdn5) and d2psidm, d2psidn (may be NULL
   when nk = 4).  Metric factors as deropm / deropn. */
int32_t ns_deriv(ns_ctx *ctx, int32_t nk, const double *U, const double *T, const double *K1, double *dUdm,
                 double *dUdn, double *dTdm, double *dTdn, double *d2Tdm, double *d2Tdn, double *psiDiff,
                 double *d2psidm, double *d2psidn);
/* derivs.f90:48-56 kdT(): dTdm, dTdn multiplied by K1 in place */
int32_t ns_kdt(ns_ctx *ctx, const double *K1, double *dTdm, double *dTdn);
/* eqeuler.f90:28-47 tau_ij(): tau from MI, dUdm, dUdn (components 2, 3), then
   dtaudm, dtaudn */
int32_t ns_tau_ij(ns_ctx *ctx, int32_t nk, const double *MI, const double *dUdm, const double *dUdn, double *tau,
                  double *dtaudm, double *dtaudn);
/* eqeuler.f90:53-99 EULERNONLINEAR(DUDT,U,NK) from the fields the calls above
   leave; d2psi* may be NULL when nk = 4, s_acou NULL for no source */
int32_t ns_eulernonlinear(ns_ctx *ctx, int32_t nk, double *dUdt, const double *U, const double *dUdm,
                          const double *dUdn, const double *tau, const double *dtaudm, const double *dtaudn,
                          const double *d2Tdm, const double *d2Tdn, const double *d2psidm, const double *d2psidn,
                          const double *s_acou);
/* diff.f90:315-350 metricsX(dU,nk) / :276-311 metricsY(dU,nk), in place */
int32_t ns_metrics_x(ns_ctx *ctx, int32_t nk, double *dU);
int32_t ns_metrics_y(ns_ctx *ctx, int32_t nk, double *dU);

/* batched cases (no reference counterpart: the reference runs one case per
   process) ------------------------------------------------------------------ */
/* Many independent cases of any grid sizes advanced together.  Each kernel
   class runs as one launch over a work table that spans every case, and one
   captured HIP graph replays a whole step of the batch.  All cases share nk,
   precision and flags; the acoustic source must be off (amplitude 0). */
typedef struct ns_batch ns_batch;
int32_t ns_batch_create(const ns_params *params, int32_t ncases, ns_batch **out);
/* profiles, base flow and initial state of case c (same meaning as
   ns_set_profiles / ns_set_base / ns_put_state) */
int32_t ns_batch_set_case(ns_batch *b, int32_t c, const double *meshx, const double *meshy,
                          const double *sigmax, const double *sigmay, const double *m,
                          const double *n, const double *rhob, const double *psib,
                          const double *Ub, const double *U0);
int32_t ns_batch_steps(ns_batch *b, int32_t iter0, int32_t nsteps);
int32_t ns_batch_get_state(ns_batch *b, int32_t c, double *U);
/* Time the batch's stage variants and keep the fastest, as ns_autotune:
   0 = tiles over every case (fast tiles and general tiles as two launches on
   two streams); 1 = one marching launch over every case (4 rows per step,
   blocks in each case's fast zone on the fast code, then the edge tiles),
   2 = the same on plain general regions; the marching variants at the chunk
   height that times fastest.  ms[v] = one stage launch of variant v (-1:
   cannot run these cases).  FP64 results are bitwise the same for every
   variant. */
int32_t ns_batch_autotune(ns_batch *b, int32_t *variant, double *ms);
/* Use batch stage variant 0..2 (as ns_batch_autotune) from now on; the first
   call on a batch times every variant, which sets up their kernels. */
int32_t ns_batch_set_variant(ns_batch *b, int32_t variant);
int32_t ns_batch_destroy(ns_batch *b);

/* device-resident state ------------------------------------------------------
   U points to device memory holding nk planes of imax*jmax FP64 values
   (plane k, row j, column i at U[k*imax*jmax + j*imax + i]): the layout of a
   contiguous (nk, jmax, imax) torch tensor.  Same meaning as ns_put_state /
   ns_get_state, without the host round trip.  No reference counterpart. */
int32_t ns_put_state_dev(ns_ctx *ctx, const double *U);
int32_t ns_get_state_dev(ns_ctx *ctx, double *U);

/* reverse mode ---------------------------------------------------------------
   Discrete adjoint of the LDDRK6 step for the 2-D Euler equations (nk = 4,
   NS_FLAG_INVISCID), both precision modes.  No reference counterpart: the
   reference integrates forward only.
   Objective: J = dt * sum_{n=1..nsteps} sum_arc (p_n - pb)^2, p_n the pressure
   after step n at the observer points (1-based (i,j) pairs, each listed once).
   ns_adj_forward runs nsteps from the current state (like ns_rk_steps),
   keeping a checkpoint every `every` steps, and returns J.  ns_adj_backward
   returns dJ/dU0 (host AoS like ns_get_state; _dev: device planes as
   ns_put_state_dev) by recomputing each segment from its checkpoint; the
   state is the final state again afterwards.  When 5 * every state-sized
   buffers fit in a quarter of the free device memory, the backward keeps
   every stage input of a segment there instead of recomputing them per
   step, and frees them when the backward returns. */
int32_t ns_adj_set_arc(ns_ctx *ctx, int32_t npts, const int32_t *ij);
/* keep = 0: the backward never keeps stage inputs and recomputes each step's
   stage inputs from its step state (less memory, more time); keep = 1 (the
   default): keep them when they fit.  Both give the same bits. */
int32_t ns_adj_set_keep(ns_ctx *ctx, int32_t keep);
int32_t ns_adj_forward(ns_ctx *ctx, int32_t iter0, int32_t nsteps, int32_t every, double *J);
int32_t ns_adj_backward(ns_ctx *ctx, double *dJdU0);
int32_t ns_adj_backward_dev(ns_ctx *ctx, double *dJdU0);
/* transposed operators (host AoS, FP64), the backward counterparts of
   ns_rhs (lam_U = R'(U)^T lam_f at every point), ns_filter_x / ns_filter_y
   and ns_ccpml (in place) */
int32_t ns_vjp_rhs(ns_ctx *ctx, const double *U, const double *lam_f, double *lam_U);
int32_t ns_vjp_filter_x(ns_ctx *ctx, double *lam);
int32_t ns_vjp_filter_y(ns_ctx *ctx, double *lam);
int32_t ns_vjp_ccpml(ns_ctx *ctx, double *lam);
/* lam <- (dU_{n+1}/dU_n)^T lam for one step (iter) from state U (host AoS);
   leaves the context's state at U */
int32_t ns_vjp_step(ns_ctx *ctx, int32_t iter, const double *U, double *lam);

/* tuning ---------------------------------------------------------------------- */
/* Time the stage launch's variants on this context's workload and keep the
   fastest.  FP64 (results bitwise identical across variants): 0 = the
   interior marching kernel, 4 rows per step, over the whole region 6 points
   in from the edges (buffers and metrics point by point; blocks inside the
   fast zone, a unit grid outside the buffer/metric layers, take the fast
   code) with a thin edge band beside it on a second stream; 1 = the marching
   kernel over the fast zone only, the edge ring beside it; 2 = as 0 at 8
   rows per step.  FP32: 0 = interior tiles, 1 = as FP64's 0, 2 = as FP64's
   1 (the tile and marching kernels may round differently).
   Runs by itself before the first step; call it between steps only.
   variant (may be NULL) gets the choice 0..2, ms[3] (may be NULL) the
   average time of each variant (-1 if it could not launch).
   No reference counterpart: the reference has no tunable kernels. */
int32_t ns_autotune(ns_ctx *ctx, int32_t *variant, double *ms);
/* Use stage variant 0..2 (as ns_autotune) without timing.
   Both calls also load every kernel a step of this context launches, as
   ns_rk_steps does before it captures its step graph (a kernel must not
   be loaded inside a stream capture).
   No reference counterpart. */
int32_t ns_set_variant(ns_ctx *ctx, int32_t variant);

/* The stage's two filter sweeps (filtering.f90:9-190): mode 0 runs the x and
   the y sweep as two kernels, mode 1 as one fused sweep that reads and writes
   the state once (each 64-column strip keeps its x-filtered rows in LDS for
   the y recurrence).  Both start each chunk's recurrences W points early
   from unfiltered values (exact to 1e-19 relative), so they agree to
   rounding; lines shorter than 272 points are swept in the reference's
   order by both (bitwise).  ns_autotune picks the faster mode, each timed
   after the stage kernel as in a step, unless this call pinned one;
   ns_get_filter_mode reports the mode and (ms[2], may be NULL) the tuned
   times of a stage followed by each form (-1: not tuned).  No reference
   counterpart. */
int32_t ns_set_filter_mode(ns_ctx *ctx, int32_t mode);
int32_t ns_get_filter_mode(ns_ctx *ctx, int32_t *mode, double *ms);

/* measurement ----------------------------------------------------------------- */
/* average device time (ms) of each kernel class over the last ns_rk_steps
   call with timing enabled: names "rhs" (the whole stage launch: interior
   kernel and edge tiles, fork to join), "interior" (the stage's interior
   kernel alone, events on its own stream), "filter_x", "filter_y", "bc" */
int32_t ns_enable_timing(ns_ctx *ctx, int32_t on);
int32_t ns_kernel_time(ns_ctx *ctx, const char *name, double *avg_ms, int32_t *count);
/* output points covered by the stage's interior kernel in the last timed
   stage (the rest of the grid runs as edge tiles; the whole grid when the
   edge tiles run in the interior kernel's launch).
   No reference counterpart: measurement only. */
int32_t ns_stage_interior(ns_ctx *ctx, int64_t *points);

#ifdef __cplusplus
}
#endif
#endif
