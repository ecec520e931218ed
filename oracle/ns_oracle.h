/*
 * ns_oracle.h -- CPU FP64 restatement of the mathe1s/NavierStokes time-advance
 * hot path.  TEST INFRASTRUCTURE ONLY: this library is the checker for the HIP
 * path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it.  The product (navierstokes_amd) never links or calls it.
 *
 * Layout follows the reference's Fortran arrays U(1:nk,1:imax,1:jmax)
 * (column-major, variable index fastest): element (k,i,j) with 0-based
 * indices lives at U[k + nk*(i + imax*j)].
 *
 * Arithmetic follows the reference expression by expression (same operand
 * order, divisions kept as divisions, no contraction: build with
 * -ffp-contract=off), so that on x86-64 it reproduces the Fortran build's
 * rounding.  Each function cites the reference lines it restates.
 */
#ifndef NS_ORACLE_H
#define NS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nso_params {
    int imax, jmax, nk, D;
    double dm, dn, dt, sigmaf;
    double gamma, Re, Pr, Pe, alpha_cond, Tref;
    double A_mi, B_mi, C_mi, D_mi, MI0;
    double A_k, B_k, C_k, D_k, K0;
    double amplitude, wf, r0_acous, m_pt_acous, n_pt_acous;
    /* 1-D profiles (length imax or jmax) and the base flow Ub(nk,imax,jmax) */
    const double *meshx, *meshy, *sigmax, *sigmay, *m, *n, *rhob, *psib, *Ub;
    /* 0: the reference's metric region tests (diff.f90:276-350); 1: multiply
       every derivative by meshx(i) / meshy(j) once, at every point (an
       extension of this build for stretched grids: the reference has no
       such switch) */
    int metric_everywhere;
    /* 1: periodic in x (an extension of this build for the temporal mixing
       layer; the reference has no periodic boundary): x derivatives use the
       interior stencil with wrapped indices, there are no x buffer layers,
       no x metric layers and no inflow/outflow columns, the LDDRK update
       covers every column, and the x filter is an in-place sweep around the
       ring from i = 1 (the points left of the seam still hold their
       unfiltered values when the sweep starts, the points right of the end
       their filtered ones when it finishes) */
    int periodic_x;
} nso_params;

/* RHS callback used by nso_rk_euler_cb: dudt = F(U, stage) */
typedef void (*nso_rhs_fn)(const nso_params *p, double *dudt, const double *U,
                           int irk, double stage_time, void *user);

/* coefficients.f90:71-123 */
void nso_get_coefficients(nso_params *p);
/* coefficients.f90:27-59 */
void nso_transport(const nso_params *p, const double *T, double *MI, double *K1);
/* diff.f90:6-38 + 315-350 (x), 127-151 + 276-311 (y); nk planes, AoS */
void nso_deropm(const nso_params *p, int nk, double *dU, const double *U);
void nso_deropn(const nso_params *p, int nk, double *dU, const double *U);
/* rhs.f90:10-46 (deriv, tau_ij, eulernonlinear); s_acou may be NULL (= 0) */
void nso_rhs(const nso_params *p, double *dUdt, const double *U, const double *s_acou);
/* rk.f90:153-169 */
void nso_rk_steep(const nso_params *p, double alphat, double betat, double *U,
                  const double *dudt, double *w);
/* rk.f90:77-151 */
void nso_bufferxy(const nso_params *p, double *U);
/* filtering.f90:9-96, 98-190 */
void nso_filterx(const nso_params *p, double *U);
void nso_filtery(const nso_params *p, double *U);
/* cc.f90:6-68 */
void nso_ccpml(const nso_params *p, double *U);
/* rk.f90:171-193 */
void nso_acousticfont(const nso_params *p, double ct, int iter, double *s_acou);
/* rk.f90:12-75: one time step */
void nso_rk_euler(const nso_params *p, double *U, int iter);
/* same stage loop with a user RHS (used to pin the LDDRK coefficients) */
void nso_rk_euler_cb(const nso_params *p, double *U, int iter, nso_rhs_fn fn, void *user);
/* euler.f90:123-130: nsteps calls of rk_euler with iter = iter0 .. iter0+nsteps-1 */
void nso_run(const nso_params *p, double *U, int iter0, int nsteps);
/* LDDRK6 coefficient tables, rk.f90:27-46 */
void nso_lddrk_coeffs(double *alphat, double *betat, double *ct);
/* number of OpenMP threads the build uses (1 for a serial build) */
int nso_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
