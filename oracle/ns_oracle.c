/*
 * ns_oracle.c -- CPU FP64 restatement of the mathe1s/NavierStokes time-advance
 * hot path (rk_euler and everything it reaches).
 *
 * TEST INFRASTRUCTURE.  This file is the oracle: the checker for the HIP
 * implementation in navierstokes_amd/.  It is loaded only by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  The product
 * never calls it and has no CPU fallback to it.
 *
 * Pinning (what checks this restatement independently of itself):
 *   nso_deropm / nso_deropn   bitwise against the reference's own diffall.f90
 *                             (compiled unmodified into oracle/_ref) times the
 *                             metric factor; closed forms (linear exactness,
 *                             DRP wavenumber) in tests/test_oracle_pins.py
 *   nso_transport / nso_get_coefficients
 *                             bitwise against coefficients.f90 (oracle/_ref)
 *   nso_ccpml                 bitwise against cc.f90 (oracle/_ref)
 *   nso_filterx / nso_filtery against an independent triangular-solve
 *                             formulation of the in-place sweeps (numpy/scipy),
 *                             constant preservation, the f32 y-literals
 *   nso_rk_euler_cb           LDDRK6 amplification polynomial matches exp(z)
 *                             through z^4 (4th order), stage times c_k against
 *                             a time-dependent source
 *   nso_rhs                   manufactured-solution convergence (4th order)
 *                             and pointwise identities on uniform states
 *   nso_bufferxy              closed forms (sigma=1 identity, sigma=0 reset,
 *                             corners Ub + sx*sy*(U-Ub))
 *   nso_acousticfont          closed form at the source centre
 *   periodic_x (no reference counterpart): the ring sweep against an
 *                             independent triangular-solve formulation on the
 *                             ring and its Fourier-mode response away from the
 *                             seam; wrapped DRP against the closed-form
 *                             wavenumber of a periodic mode
 *   nso_rk_euler composed over many steps: parity unpinned against the
 *                             reference (rk.f90, filtering.f90, eqeuler.f90,
 *                             derivs.f90, rhs.f90 do not compile as shipped).
 *
 * Build with -O2 -ffp-contract=off (no FMA contraction, like the reference's
 * x86-64 build).  -fopenmp gives the all-cores CPU baseline; rows/columns are
 * independent in every loop that is parallelised, so results do not depend
 * on the thread count.
 */
#include "ns_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(p, nk, k, i, j) ((size_t)(k) + (size_t)(nk) * ((size_t)(i) + (size_t)(p)->imax * (size_t)(j)))

/* DRP coefficients, init.f90:288-338 */
static const double A_INT[3] = {0.770882380518, -0.166705904415, 0.0208431427703};
static const double A4[7] = {0.049041958000, -0.46884035700, -0.47476091400, 1.27327473700,
                             -0.51848452600, 0.16613853300, -0.026369431000};
static const double A5[7] = {-0.20933762200, -1.08487567600, 2.14777605000, -1.38892832200,
                             0.76894976600, -0.28181465000, 0.048230454000};
static const double A6[7] = {-2.19228033900, 4.74861140100, -5.10885191500, 4.46156710400,
                             -2.83349874100, 1.12832886100, -0.20387637100};
/* a_dudx4b(1..7) etc., init.f90:300-306, 316-322, 332-338 */
static const double A4B[7] = {0.026369431000, -0.16613853300, 0.51848452600, -1.27327473700,
                              0.47476091400, 0.46884035700, -0.049041958000};
static const double A5B[7] = {-0.048230454000, 0.28181465000, -0.76894976600, 1.38892832200,
                              -2.14777605000, 1.08487567600, 0.20933762200};
static const double A6B[7] = {0.20387637100, -1.12832886100, 2.83349874100, -4.46156710400,
                              5.10885191500, -4.74861140100, 2.19228033900};

/* filtering.f90:54-66 (index s+6) */
static const double DF[13] = {0.001254597714, -0.008520738659, 0.029662754736, -0.069975429105,
                              0.123632891797, -0.171503832236, 0.190899511506, -0.171503832236,
                              0.123632891797, -0.069975429105, 0.029662754736, -0.008520738659,
                              0.001254597714};
/* filtering.f90:25-31, index s+1 (FP64 literals, x direction) */
static const double DFB_X[7] = {-0.057717512738, 0.199278374994, -0.292668277650, 0.244537361546,
                                -0.134605018019, 0.056184263460, -0.015009191593};
/* filtering.f90:114-120: default-real (float32) literals widened to FP64 */
static const double DFB_Y[7] = {(double)-0.057717512738f, (double)0.199278374994f,
                                (double)-0.292668277650f, (double)0.244537361546f,
                                (double)-0.134605018019f, (double)0.056184263460f,
                                (double)-0.015009191593f};

/* rk.f90:27-46 */
static const double ALPHAT[6] = {0.0, -0.737101392796, -1.634740794341,
                                 -0.744739003780, -1.469897351522, -2.813971388035};
static const double BETAT[6] = {0.032918605146, 0.823256998200, 0.381530948900,
                                0.200092213184, 1.718581042715, 0.27};
static const double CT[6] = {0.0, 0.032918605146, 0.249351723343,
                             0.466911705055, 0.582030414044, 0.847252983783};

void nso_lddrk_coeffs(double *alphat, double *betat, double *ct)
{
    for (int s = 0; s < 6; ++s) {
        alphat[s] = ALPHAT[s];
        betat[s] = BETAT[s];
        ct[s] = CT[s];
    }
}

int nso_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* coefficients.f90:113-123 (SPECIE_N) and :71-80 (GET_COEFFICIENTS)   */
void nso_get_coefficients(nso_params *p)
{
    p->A_mi = 0.62526577;
    p->B_mi = -31.779652;
    p->C_mi = -1640.7913;
    p->D_mi = 1.7454992;
    p->A_k = 0.85372829;
    p->B_k = 105.18665;
    p->C_k = -12229.753;
    p->D_k = 0.48299104;
    double Tref = p->Tref;
    p->K0 = exp(p->A_k * log(1.0 * Tref) + p->B_k / Tref + p->C_k / Tref / Tref + p->D_k);
    p->MI0 = exp(p->A_mi * log(1.0 * Tref) + p->B_mi / Tref + p->C_mi / Tref / Tref + p->D_mi);
}

/* coefficients.f90:35-42 (MOLECULAR_MI) and :50-58 (k_coefficient) */
void nso_transport(const nso_params *p, const double *T, double *MI, double *K1)
{
    const size_t N = (size_t)p->imax * p->jmax;
    const double Tref = p->Tref;
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) {
        double t = T[q];
        MI[q] = 1.0 / p->MI0 *
                (exp(p->A_mi * log(Tref * t) + p->B_mi / Tref / t + p->C_mi / Tref / t / Tref / t + p->D_mi));
        K1[q] = 1.0 / p->K0 *
                (exp(p->A_k * log(Tref * t) + p->B_k / Tref / t + p->C_k / Tref / Tref / t / t + p->D_k));
    }
}

/* ------------------------------------------------------------------ */
/* One derivative along a line of n points with stride `st`; h = dm or dn.
 * diff.f90:18-36 (deropm) / 131-149 (deropn); dudx :40-47; dx4..dx6I :49-125 */
static void drp_line(double *out, size_t ost, const double *u, size_t st, int n, double h)
{
#define U_(q) u[(size_t)(q) * st]
    for (int q = 0; q < n; ++q) {
        double s;
        if (q == 0) {               /* dx6: u(i..i+6) */
            s = (A6[0] * U_(0)) / h;
            for (int mm = 1; mm < 7; ++mm) s = s + (A6[mm] * U_(mm)) / h;
        } else if (q == 1) {        /* dx5: u(i-1..i+5) */
            s = (A5[0] * U_(0)) / h;
            for (int mm = 1; mm < 7; ++mm) s = s + (A5[mm] * U_(mm)) / h;
        } else if (q == 2) {        /* dx4: u(i-2..i+4) */
            s = (A4[0] * U_(0)) / h;
            for (int mm = 1; mm < 7; ++mm) s = s + (A4[mm] * U_(mm)) / h;
        } else if (q == n - 3) {    /* dx4I: terms m=7..1 on u(i+m-5) */
            s = (A4B[6] * U_(q + 2)) / h;
            for (int mm = 5; mm >= 0; --mm) s = s + (A4B[mm] * U_(q + mm + 1 - 5)) / h;
        } else if (q == n - 2) {    /* dx5I: u(i+m-6) */
            s = (A5B[6] * U_(q + 1)) / h;
            for (int mm = 5; mm >= 0; --mm) s = s + (A5B[mm] * U_(q + mm + 1 - 6)) / h;
        } else if (q == n - 1) {    /* dx6I: u(i+m-7) */
            s = (A6B[6] * U_(q)) / h;
            for (int mm = 5; mm >= 0; --mm) s = s + (A6B[mm] * U_(q + mm + 1 - 7)) / h;
        } else {                    /* interior Tam-Webb DRP */
            s = (A_INT[0] * (U_(q + 1) - U_(q - 1))) / h + (A_INT[1] * (U_(q + 2) - U_(q - 2))) / h +
                (A_INT[2] * (U_(q + 3) - U_(q - 3))) / h;
        }
        out[(size_t)q * ost] = s;
    }
#undef U_
}

/* periodic line (periodic_x): the interior stencil everywhere, indices wrapped */
static void drp_line_periodic(double *out, size_t ost, const double *u, size_t st, int n, double h)
{
#define UW_(q) u[(size_t)(((q) % n + n) % n) * st]
    for (int q = 0; q < n; ++q)
        out[(size_t)q * ost] = (A_INT[0] * (UW_(q + 1) - UW_(q - 1))) / h + (A_INT[1] * (UW_(q + 2) - UW_(q - 2))) / h +
                               (A_INT[2] * (UW_(q + 3) - UW_(q - 3))) / h;
#undef UW_
}

/* metric region tests, diff.f90:315-350 (metricsX) and :276-311 (metricsY) */
static int in_xstrip(const nso_params *p, int i1) { return i1 <= p->D || i1 >= p->imax - p->D + 1; }
static int in_ystrip(const nso_params *p, int j1) { return j1 <= p->D || j1 >= p->jmax - p->D + 1; }

static double metx_factor_count(const nso_params *p, int i1, int j1, int *cnt)
{
    if (p->metric_everywhere) {
        *cnt = 1;
        return p->meshx[i1 - 1];
    }
    if (p->periodic_x) { /* no x layers */
        *cnt = 0;
        return 1.0;
    }
    const int D = p->D, imax = p->imax, jmax = p->jmax;
    int c = 0;
    if (i1 >= 1 && i1 <= D && j1 >= D + 1 && j1 <= jmax - D) c++;
    if (i1 >= imax - D + 1 && i1 <= imax && j1 >= D + 1 && j1 <= jmax - D) c++;
    if (i1 >= 1 && i1 <= D && j1 >= jmax - D + 1 && j1 <= jmax) c++;
    if (i1 >= 1 && i1 <= D && j1 >= 1 && j1 <= D) c++;
    if (i1 >= imax - D + 1 && i1 <= imax && j1 >= 1 && j1 <= D) c++;
    if (i1 >= imax - D + 1 && i1 <= imax && j1 >= jmax - D + 1 && j1 <= jmax) c++;
    *cnt = c;
    return p->meshx[i1 - 1];
}

static double mety_factor_count(const nso_params *p, int i1, int j1, int *cnt)
{
    if (p->metric_everywhere) {
        *cnt = 1;
        return p->meshy[j1 - 1];
    }
    const int D = p->D, imax = p->imax, jmax = p->jmax;
    int c = 0;
    if (i1 >= D + 1 && i1 <= imax - D && j1 >= 1 && j1 <= D) c++;
    if (i1 >= D + 1 && i1 <= imax - D && j1 >= jmax - D + 1 && j1 <= jmax) c++;
    if (i1 >= 1 && i1 <= D && j1 >= jmax - D + 1 && j1 <= jmax) c++;
    if (i1 >= 1 && i1 <= D && j1 >= 1 && j1 <= D) c++;
    if (i1 >= imax - D + 1 && i1 <= imax && j1 >= 1 && j1 <= D) c++;
    if (i1 >= imax - D + 1 && i1 <= imax && j1 >= jmax - D + 1 && j1 <= jmax) c++;
    *cnt = c;
    return p->meshy[j1 - 1];
}

void nso_deropm(const nso_params *p, int nk, double *dU, const double *U)
{
    const int imax = p->imax, jmax = p->jmax;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < jmax; ++j)
        for (int k = 0; k < nk; ++k)
            (p->periodic_x ? drp_line_periodic : drp_line)(dU + IDX(p, nk, k, 0, j), (size_t)nk,
                                                           U + IDX(p, nk, k, 0, j), (size_t)nk, imax, p->dm);
    /* metricsX: du = du*meshx(i) once per satisfied region test */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < jmax; ++j)
        for (int i = 0; i < imax; ++i) {
            int c;
            double f = metx_factor_count(p, i + 1, j + 1, &c);
            for (int r = 0; r < c; ++r)
                for (int k = 0; k < nk; ++k) dU[IDX(p, nk, k, i, j)] = dU[IDX(p, nk, k, i, j)] * f;
        }
}

void nso_deropn(const nso_params *p, int nk, double *dU, const double *U)
{
    const int imax = p->imax, jmax = p->jmax;
    const size_t st = (size_t)nk * imax;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < imax; ++i)
        for (int k = 0; k < nk; ++k)
            drp_line(dU + IDX(p, nk, k, i, 0), st, U + IDX(p, nk, k, i, 0), st, jmax, p->dn);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < jmax; ++j)
        for (int i = 0; i < imax; ++i) {
            int c;
            double f = mety_factor_count(p, i + 1, j + 1, &c);
            for (int r = 0; r < c; ++r)
                for (int k = 0; k < nk; ++k) dU[IDX(p, nk, k, i, j)] = dU[IDX(p, nk, k, i, j)] * f;
        }
}

/* ------------------------------------------------------------------ */
/* workspace kept across calls (the oracle is single-context) */
typedef struct {
    size_t N;
    int nk;
    double *T, *MI, *K1, *dUdm, *dUdn, *dTdm, *dTdn, *d2Tdm, *d2Tdn, *tau, *dtaudm, *dtaudn,
        *psiDiff, *d2psidm, *d2psidn, *dudt, *w, *s;
} nso_ws;

static nso_ws WS;

static double *xalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

static void ws_free(void)
{
    double **ptrs[] = {&WS.T, &WS.MI, &WS.K1, &WS.dUdm, &WS.dUdn, &WS.dTdm, &WS.dTdn, &WS.d2Tdm,
                       &WS.d2Tdn, &WS.tau, &WS.dtaudm, &WS.dtaudn, &WS.psiDiff, &WS.d2psidm,
                       &WS.d2psidn, &WS.dudt, &WS.w, &WS.s};
    for (size_t q = 0; q < sizeof(ptrs) / sizeof(ptrs[0]); ++q) {
        free(*ptrs[q]);
        *ptrs[q] = NULL;
    }
    WS.N = 0;
}

static void ws_ensure(const nso_params *p)
{
    size_t N = (size_t)p->imax * p->jmax;
    if (WS.N == N && WS.nk == p->nk) return;
    ws_free();
    WS.N = N;
    WS.nk = p->nk;
    WS.T = xalloc(N); WS.MI = xalloc(N); WS.K1 = xalloc(N);
    WS.dUdm = xalloc(N * p->nk); WS.dUdn = xalloc(N * p->nk);
    WS.dTdm = xalloc(N); WS.dTdn = xalloc(N); WS.d2Tdm = xalloc(N); WS.d2Tdn = xalloc(N);
    WS.tau = xalloc(3 * N); WS.dtaudm = xalloc(3 * N); WS.dtaudn = xalloc(3 * N);
    WS.psiDiff = xalloc(N); WS.d2psidm = xalloc(N); WS.d2psidn = xalloc(N);
    WS.dudt = xalloc(N * p->nk); WS.w = xalloc(N * p->nk); WS.s = xalloc(N);
}

/* rhs.f90:10-46 -> derivs.f90:9-56 -> eqeuler.f90:28-99 */
void nso_rhs(const nso_params *p, double *dUdt, const double *U, const double *s_acou)
{
    ws_ensure(p);
    const int nk = p->nk, imax = p->imax, jmax = p->jmax;
    const size_t N = (size_t)imax * jmax;
    const double gamma = p->gamma, Re = p->Re, Pr = p->Pr;

    /* rhs.f90:20  TEMP = gamma*p/rho */
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) WS.T[q] = (gamma) * (U[3 + (size_t)nk * q]) / (U[(size_t)nk * q]);
    /* rhs.f90:24-25 */
    nso_transport(p, WS.T, WS.MI, WS.K1);

    /* derivs.f90:17-20 */
    nso_deropm(p, nk, WS.dUdm, U);
    nso_deropn(p, nk, WS.dUdn, U);
    /* derivs.f90:25-26 */
    nso_deropm(p, 1, WS.dTdm, WS.T);
    nso_deropn(p, 1, WS.dTdn, WS.T);
    /* derivs.f90:28 -> kdT :48-56 */
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) {
        WS.dTdn[q] = WS.dTdn[q] * WS.K1[q];
        WS.dTdm[q] = WS.dTdm[q] * WS.K1[q];
    }
    /* derivs.f90:32-33 */
    nso_deropm(p, 1, WS.d2Tdm, WS.dTdm);
    nso_deropn(p, 1, WS.d2Tdn, WS.dTdn);
    /* derivs.f90:36-43 (species, only with the passive scalar) */
    if (nk >= 5) {
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < N; ++q)
            WS.psiDiff[q] = U[(size_t)nk * q] * p->alpha_cond * (WS.dUdm[4 + (size_t)nk * q] + WS.dUdn[4 + (size_t)nk * q]);
        nso_deropm(p, 1, WS.d2psidm, WS.psiDiff);
        nso_deropn(p, 1, WS.d2psidn, WS.psiDiff);
    }

    /* eqeuler.f90:34-41 tau_ij */
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) {
        const double *dm_ = WS.dUdm + (size_t)nk * q, *dn_ = WS.dUdn + (size_t)nk * q;
        double mi = WS.MI[q];
        WS.tau[3 * q + 0] = -(2.0 / 3.0 * (mi) * (dm_[1] + dn_[2])) + 2.0 * (mi) * (dm_[1]);
        WS.tau[3 * q + 1] = (mi) * (dn_[1] + dm_[2]);
        WS.tau[3 * q + 2] = -(2.0 / 3.0 * (mi) * (dm_[1] + dn_[2])) + 2.0 * (mi) * (dn_[2]);
    }
    /* eqeuler.f90:43-45 */
    nso_deropm(p, 3, WS.dtaudm, WS.tau);
    nso_deropn(p, 3, WS.dtaudn, WS.tau);

    /* eqeuler.f90:61-97 EULERNONLINEAR */
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < N; ++q) {
        const double *u = U + (size_t)nk * q;
        const double *dm_ = WS.dUdm + (size_t)nk * q, *dn_ = WS.dUdn + (size_t)nk * q;
        const double *t = WS.tau + 3 * q, *tm = WS.dtaudm + 3 * q, *tn = WS.dtaudn + 3 * q;
        double *f = dUdt + (size_t)nk * q;
        double s = s_acou ? s_acou[q] : 0.0;
        f[0] = -(u[1] * (dm_[0]) + u[0] * (dm_[1]) + u[2] * (dn_[0]) + u[0] * (dn_[2]));
        f[1] = -(u[1] * (dm_[1]) + 1.0 / u[0] * dm_[3] + u[2] * (dn_[1]) -
                 1.0 / Re / (u[0]) * (tm[0] + tn[1]));
        f[2] = -(u[1] * (dm_[2]) + u[2] * (dn_[2]) + 1.0 / u[0] * (dn_[3]) -
                 1.0 / Re / (u[0]) * (tm[1] + tn[2]));
        f[3] = -(gamma * u[3] * (dm_[1]) + u[1] * (dm_[3]) + gamma * u[3] * (dn_[2]) + u[2] * (dn_[3]) - s -
                 (gamma - 1.0) / Re * (+(dm_[1]) * t[0] + (dm_[2]) * t[1] + (dn_[1]) * t[1] + (dn_[2]) * t[2]) -
                 1.0 / Re / Pr * (WS.d2Tdm[q] + WS.d2Tdn[q]));
        if (nk >= 5)
            f[4] = -(u[1] * dm_[4] + u[2] * dn_[4] - 1.0 / p->Pe / u[0] * (WS.d2psidm[q] + WS.d2psidn[q]));
    }
}

/* rk.f90:153-169 */
void nso_rk_steep(const nso_params *p, double alphat, double betat, double *U, const double *dudt, double *w)
{
    const int nk = p->nk, imax = p->imax, jmax = p->jmax;
    const double dt = p->dt;
    /* periodic_x: every column is interior */
    const int i0 = p->periodic_x ? 0 : 1, i1 = p->periodic_x ? imax : imax - 1;
#pragma omp parallel for schedule(static)
    for (int j = 1; j < jmax - 1; ++j)
        for (int i = i0; i < i1; ++i)
            for (int k = 0; k < nk; ++k) {
                size_t q = IDX(p, nk, k, i, j);
                w[q] = alphat * w[q] + dt * dudt[q];
                U[q] = U[q] + betat * w[q];
            }
}

/* rk.f90:77-151 */
void nso_bufferxy(const nso_params *p, double *U)
{
    const int nk = p->nk, imax = p->imax, jmax = p->jmax, D = p->D;
    const double *Ub = p->Ub;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < jmax; ++j) {
        int j1 = j + 1;
        for (int i = 0; i < imax; ++i) {
            int i1 = i + 1;
            /* buffery (:140-151) for the two y-layers, then bufferx (:113-138) */
            if (j1 >= 1 && j1 <= D)
                for (int k = 0; k < nk; ++k) {
                    size_t q = IDX(p, nk, k, i, j);
                    U[q] = Ub[q] + p->sigmay[j] * (U[q] - Ub[q]);
                }
            if (j1 >= jmax - D + 1 && j1 <= jmax)
                for (int k = 0; k < nk; ++k) {
                    size_t q = IDX(p, nk, k, i, j);
                    U[q] = Ub[q] + p->sigmay[j] * (U[q] - Ub[q]);
                }
            if (p->periodic_x) continue; /* no x layers */
            if (i1 >= 1 && i1 <= D)
                for (int k = 0; k < nk; ++k) {
                    size_t q = IDX(p, nk, k, i, j);
                    U[q] = Ub[q] + p->sigmax[i] * (U[q] - Ub[q]);
                }
            if (i1 >= imax - D + 1 && i1 <= imax)
                for (int k = 0; k < nk; ++k) {
                    size_t q = IDX(p, nk, k, i, j);
                    U[q] = Ub[q] + p->sigmax[i] * (U[q] - Ub[q]);
                }
        }
    }
}

/* In-place filter of one line of n points with stride st.
 * filtering.f90:36-49 (left block, ascending), :51-79 (interior, ascending),
 * :81-94 (right block, descending, mirrored stencil). */
static void filter_line(double *u, size_t st, int n, double sigmaf, const double *dfb)
{
#define U_(q) u[(size_t)(q) * st]
    /* left block: i = 2..7 (0-based 1..6) */
    for (int q = 1; q <= 6; ++q) {
        double Du = 0.0;
        for (int s = -1; s <= 5; ++s) Du = Du + dfb[s + 1] * U_(q + s);
        U_(q) = U_(q) - sigmaf * Du;
    }
    /* interior: i = 7..n-6 (0-based 6..n-7) */
    for (int q = 6; q <= n - 7; ++q) {
        double Du = 0.0;
        for (int s = -6; s <= 6; ++s) Du = Du + DF[s + 6] * U_(q + s);
        U_(q) = U_(q) - sigmaf * Du;
    }
    /* right block: i = n-1 down to n-6 (0-based n-2 .. n-7) */
    for (int q = n - 2; q >= n - 7; --q) {
        double Du = 0.0;
        for (int s = -1; s <= 5; ++s) Du = Du + dfb[s + 1] * U_(q - s);
        U_(q) = U_(q) - sigmaf * Du;
    }
#undef U_
}

/* periodic_x: the in-place sweep of the interior stencil around the ring,
 * i = 1..n (0-based 0..n-1) in order, every index wrapped: at q the points
 * q-6..q-1 hold values updated in this sweep except those left of the seam
 * (q-s < 0), and q+1..q+6 hold unfiltered values except those past the end
 * (q+s >= n), which are the points 0..5 updated at the start. */
static void filter_ring(double *u, size_t st, int n, double sigmaf)
{
#define UW_(q) u[(size_t)(((q) % n + n) % n) * st]
    for (int q = 0; q < n; ++q) {
        double Du = 0.0;
        for (int s = -6; s <= 6; ++s) Du = Du + DF[s + 6] * UW_(q + s);
        UW_(q) = UW_(q) - sigmaf * Du;
    }
#undef UW_
}

void nso_filterx(const nso_params *p, double *U)
{
    const int nk = p->nk, imax = p->imax, jmax = p->jmax;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < jmax; ++j)
        for (int k = 0; k < nk; ++k) {
            if (p->periodic_x)
                filter_ring(U + IDX(p, nk, k, 0, j), (size_t)nk, imax, p->sigmaf);
            else
                filter_line(U + IDX(p, nk, k, 0, j), (size_t)nk, imax, p->sigmaf, DFB_X);
        }
}

void nso_filtery(const nso_params *p, double *U)
{
    const int nk = p->nk, imax = p->imax, jmax = p->jmax;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < imax; ++i)
        for (int k = 0; k < nk; ++k)
            filter_line(U + IDX(p, nk, k, i, 0), (size_t)nk * imax, jmax, p->sigmaf, DFB_Y);
}

/* cc.f90:6-68 */
void nso_ccpml(const nso_params *p, double *U)
{
    const int nk = p->nk, imax = p->imax, jmax = p->jmax;
    const double g = p->gamma;
#define V(k, i, j) U[IDX(p, nk, k, i, j)]
    /* inflow, i = 1 (periodic_x: no inflow or outflow column) */
    for (int j = 0; j < (p->periodic_x ? 0 : jmax); ++j) {
        V(0, 0, j) = p->rhob[j];
        V(1, 0, j) = 2.0 * V(1, 1, j) - V(1, 2, j);
        V(2, 0, j) = 0.0;
        V(3, 0, j) = 1.0 / g;
        if (nk >= 5) V(4, 0, j) = p->psib[j];
    }
    /* outflow, i = imax */
    for (int j = 0; j < (p->periodic_x ? 0 : jmax); ++j) {
        V(0, imax - 1, j) = 2.0 * V(0, imax - 2, j) - V(0, imax - 3, j);
        V(1, imax - 1, j) = 2.0 * V(1, imax - 2, j) - V(1, imax - 3, j);
        V(2, imax - 1, j) = 2.0 * V(2, imax - 2, j) - V(2, imax - 3, j);
        V(3, imax - 1, j) = 1.0 / g;
        if (nk >= 5) V(4, imax - 1, j) = 2.0 * V(4, imax - 2, j) - V(4, imax - 3, j);
    }
    /* upper boundary, j = jmax */
    for (int i = 0; i < imax; ++i) {
        V(0, i, jmax - 1) = 2.0 * V(0, i, jmax - 2) - V(0, i, jmax - 3);
        V(1, i, jmax - 1) = 2.0 * V(1, i, jmax - 2) - V(1, i, jmax - 3);
        V(2, i, jmax - 1) = 2.0 * V(2, i, jmax - 2) - V(2, i, jmax - 3);
        V(3, i, jmax - 1) = 1.0 / g;
        if (nk >= 5) V(4, i, jmax - 1) = 1.0;
    }
    /* lower boundary, j = 1 */
    for (int i = 0; i < imax; ++i) {
        V(0, i, 0) = 2.0 * V(0, i, 1) - V(0, i, 2);
        V(1, i, 0) = 2.0 * V(1, i, 1) - V(1, i, 2);
        V(2, i, 0) = 2.0 * V(2, i, 1) - V(2, i, 2);
        V(3, i, 0) = 1.0 / g;
        if (nk >= 5) V(4, i, 0) = 0.0;
    }
#undef V
}

/* rk.f90:186-191 */
void nso_acousticfont(const nso_params *p, double ct, int iter, double *s_acou)
{
    const int imax = p->imax, jmax = p->jmax;
    const double amp = p->amplitude * sin(p->wf * (((double)iter + ct) * p->dt));
    const double r2 = p->r0_acous * p->r0_acous;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < jmax; ++j)
        for (int i = 0; i < imax; ++i) {
            double dx = p->m[i] - p->m_pt_acous, dy = p->n[j] - p->n_pt_acous;
            s_acou[(size_t)i + (size_t)imax * j] = amp * exp(-(log(2.0) * (dx * dx + dy * dy) / r2));
        }
}

static void rhs_default(const nso_params *p, double *dudt, const double *U, int irk, double t, void *user)
{
    (void)irk;
    (void)t;
    nso_rhs(p, dudt, U, (const double *)user);
}

/* rk.f90:12-75 */
void nso_rk_euler_cb(const nso_params *p, double *U, int iter, nso_rhs_fn fn, void *user)
{
    ws_ensure(p);
    const size_t NT = (size_t)p->imax * p->jmax * p->nk;
    memset(WS.w, 0, NT * sizeof(double)); /* rk.f90:21 */
    for (int irk = 0; irk < 6; ++irk) {
        if (fn == NULL) {
            nso_acousticfont(p, CT[irk], iter, WS.s); /* rk.f90:55 */
            rhs_default(p, WS.dudt, U, irk, 0.0, WS.s); /* rk.f90:58 */
        } else {
            fn(p, WS.dudt, U, irk, ((double)iter + CT[irk]) * p->dt, user);
        }
        nso_rk_steep(p, ALPHAT[irk], BETAT[irk], U, WS.dudt, WS.w); /* :61 */
        nso_bufferxy(p, U);                                          /* :64 */
        nso_filterx(p, U);                                           /* :67 */
        nso_filtery(p, U);                                           /* :68 */
        nso_ccpml(p, U);                                             /* :72 */
    }
}

void nso_rk_euler(const nso_params *p, double *U, int iter) { nso_rk_euler_cb(p, U, iter, NULL, NULL); }

void nso_run(const nso_params *p, double *U, int iter0, int nsteps)
{
    for (int it = 0; it < nsteps; ++it) nso_rk_euler(p, U, iter0 + it);
}
