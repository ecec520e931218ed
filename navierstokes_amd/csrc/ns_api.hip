// ns_api.hip -- C ABI (include/ns_api.h) over the gfx950 kernels.
//
// A context owns what the reference keeps in `module global` (global.f90):
// the scalar parameter block, the 1-D profiles built by init.f90, the base
// flow Ub, and the state U, all resident on one GPU.  The state lives in three
// SoA buffers S0..S2 that rotate through each LDDRK stage
//   S0 --stage--> S1 --filter_x--> S2 --filter_y+edges--> S0
// so one captured step graph replays without pointer changes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ns_api.h"
#include "ns_adjoint.h"
#include "ns_kernels.h"

namespace nsamd {
hipError_t launch_stage(const StageArgs<double> &A, int nk, bool visc, bool rhs_only, int variant, hipStream_t s,
                        hipStream_t s2, StageHooks *tm = nullptr);
hipError_t launch_stage(const StageArgs<float> &A, int nk, bool visc, bool rhs_only, int variant, hipStream_t s,
                        hipStream_t s2, StageHooks *tm = nullptr);
hipError_t prepare_step(const StageArgs<double> &A, int nk, bool visc, int variant, bool batch);
hipError_t prepare_step(const StageArgs<float> &A, int nk, bool visc, int variant, bool batch);
hipError_t launch_filter(const FilterArgs<double> &F, bool xdir, hipStream_t s);
hipError_t launch_filter(const FilterArgs<float> &F, bool xdir, hipStream_t s);
hipError_t launch_filter_xy(const FilterXYArgs<double> &F, hipStream_t s);
hipError_t launch_filter_xy(const FilterXYArgs<float> &F, hipStream_t s);
hipError_t launch_edge_i(const EdgeArgs<double> &E, hipStream_t s);
hipError_t launch_edge_i(const EdgeArgs<float> &E, hipStream_t s);
hipError_t launch_stage_batch(const StageArgs<double> *args, const int4 *ftiles, int nf, const int4 *gtiles, int ng,
                              int stage, int ncases, int nk, bool visc, int variant, hipStream_t s, hipStream_t s2);
hipError_t launch_stage_batch(const StageArgs<float> *args, const int4 *ftiles, int nf, const int4 *gtiles, int ng,
                              int stage, int ncases, int nk, bool visc, int variant, hipStream_t s, hipStream_t s2);
bool batch_march_case(const StageArgs<double> &A, int nk, bool visc, bool hybrid, int H, BatchMarch &p);
bool batch_march_case(const StageArgs<float> &A, int nk, bool visc, bool hybrid, int H, BatchMarch &p);
hipError_t launch_stage_march_batch(const StageArgs<double> *args, const BatchMarch *plans, const int4 *items,
                                    int nitems, int stage, int ncases, int nk, bool visc, bool unit, hipStream_t s);
hipError_t launch_stage_march_batch(const StageArgs<float> *args, const BatchMarch *plans, const int4 *items,
                                    int nitems, int stage, int ncases, int nk, bool visc, bool unit, hipStream_t s);
template <typename S>
hipError_t launch_vorticity(const S *u, const S *v, const double *ub1, int imax, int jmax, int pitch, double inv_dm,
                            double inv_dn, bool xper, double *out, hipStream_t s);
template <typename S>
hipError_t launch_station_energy(const S *u, const S *v, const double *ub1, const double *ub2, int pitch, int jmax,
                                 int D, double dn, int nst, const int *cols, double *out, hipStream_t s);
template <typename S>
hipError_t launch_row_mean(const S *U, int pitch, int jmax, int i0, int i1, const double *add, double *out,
                           hipStream_t s);
hipError_t launch_adj_stage(const AdjPointArgs<double> &P, const AdjGatherArgs<double> &G, hipStream_t s);
hipError_t launch_adj_stage(const AdjPointArgs<float> &P, const AdjGatherArgs<float> &G, hipStream_t s);
hipError_t launch_adj_filter(const AdjFilterArgs<double> &F, bool ydir, hipStream_t s);
hipError_t launch_adj_filter(const AdjFilterArgs<float> &F, bool ydir, hipStream_t s);
hipError_t launch_adj_edge(const AdjEdgeArgs<double> &E, hipStream_t s);
hipError_t launch_adj_edge(const AdjEdgeArgs<float> &E, hipStream_t s);
hipError_t launch_arc_reduce(const ArcArgs<double> &a, double *out, hipStream_t s);
hipError_t launch_arc_reduce(const ArcArgs<float> &a, double *out, hipStream_t s);
hipError_t launch_arc_seed(const ArcArgs<double> &a, double *lam, double weight, hipStream_t s);
hipError_t launch_arc_seed(const ArcArgs<float> &a, float *lam, double weight, hipStream_t s);
hipError_t launch_convert(const ConvArgs<double, double> &a, hipStream_t s);
hipError_t launch_convert(const ConvArgs<double, float> &a, hipStream_t s);
hipError_t launch_convert(const ConvArgs<float, double> &a, hipStream_t s);
hipError_t launch_filter_batch(const FilterArgs<double> *args, const int4 *tiles, int ntiles, int nk, bool xdir,
                               hipStream_t s);
hipError_t launch_filter_batch(const FilterArgs<float> *args, const int4 *tiles, int ntiles, int nk, bool xdir,
                               hipStream_t s);
hipError_t launch_edge_batch(const EdgeArgs<double> *args, int ncases, hipStream_t s);
hipError_t launch_edge_batch(const EdgeArgs<float> *args, int ncases, hipStream_t s);
hipError_t launch_deriv(const double *const *in, double *const *out, int nk, int imax, int jmax, int pitch, bool ydir,
                        const double *met, double inv_h, bool xper, hipStream_t s);
hipError_t launch_transport(const double *T, double *MI, double *K1, size_t n, const Transport<double> &tc,
                            hipStream_t s);
hipError_t launch_rk_steep(double *const *U, const double *const *dudt, double *const *W, int nk, int imax, int jmax,
                           int pitch, double alpha, double beta, double dt, bool xper, hipStream_t s);
hipError_t launch_buffer(double *const *U, const double *const *base, int nk, int imax, int jmax, int pitch, int D,
                         const double *sigx, const double *sigy, bool xper, hipStream_t s);
hipError_t launch_source(double *s_out, int imax, int jmax, double amp, const double *gx, const double *gy,
                         hipStream_t s);
template <typename S>
hipError_t launch_probe(const S *const *U, int nk, int pitch, int npts, const int *ij, double *out, hipStream_t s);
template <typename S>
hipError_t launch_finite(const S *const *U, int nk, int imax, int jmax, int pitch, int *flag, hipStream_t s);
hipError_t launch_fill(double *p, size_t n, double v, hipStream_t s);
hipError_t launch_kdt(const PwArgs &a, const double *k, double *dTdm, double *dTdn, hipStream_t s);
hipError_t launch_psidiff(const PwArgs &a, const double *rho, const double *pm, const double *pn, double alpha,
                          double *out, hipStream_t s);
hipError_t launch_tau(const PwArgs &a, const double *mi, const double *ux, const double *uy, const double *vx,
                      const double *vy, double *txx, double *txy, double *tyy, hipStream_t s);
hipError_t launch_euler_pw(const EulerPwArgs &e, hipStream_t s);
hipError_t launch_metric(const PwArgs &a, double *const *dU, int nk, const double *met, bool ydir, hipStream_t s);
}  // namespace nsamd

using namespace nsamd;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// A HIP failure as a status.  hipErrorStreamCaptureUnsupported comes back
// from a launch path that met a kernel not set up before a stream capture
// (the kernels' set-up must not run inside one): a call out of order.
int hip_fail(hipError_t e, const char *what)
{
    if (e == hipErrorStreamCaptureUnsupported)
        return fail(NS_ERR_STATE,
                    "%s: reached inside a stream capture before its kernels were set up (ns_set_variant, "
                    "ns_autotune or the first ns_rk_steps of a context set them up)",
                    what);
    return fail(NS_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIPCHK(expr)                                  \
    do {                                              \
        hipError_t e_ = (expr);                       \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

// "rhs": the whole stage launch (interior kernel, edge tiles beside it on
// the side stream, fork to join); "interior": the interior kernel alone,
// timed on its own stream
enum TimerClass { TC_STAGE = 0, TC_FX, TC_FY, TC_BC, TC_INT, TC_N };
const char *TC_NAMES[TC_N] = {"rhs", "filter_x", "filter_y", "bc", "interior"};

}  // namespace

struct ns_ctx {
    ns_params P{};
    int pitch = 0;
    size_t plane = 0;  // elements per plane
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipStream_t side = nullptr;  // edge-ring stage tiles run here, concurrent with the interior
    hipEvent_t fork = nullptr, join = nullptr;
    bool pert = false;        // FP32 perturbation form
    size_t esz = 8;           // bytes per state element
    void *S[3] = {nullptr, nullptr, nullptr};  // state buffers (FP64 U or FP32 q)
    void *Wr = nullptr;       // LDDRK register (same precision as the state)
    double *base = nullptr;   // Ub planes (FP64)
    double *prow = nullptr;   // perturbation form: per-row base data (PertRows)
    std::vector<double> ubh;  // perturbation form: host copy of Ub(k, j)
    RowEdge re{};             // ccpml row constants
    double *work = nullptr;   // operator-API scratch (2 x 8 planes)
    double *prof = nullptr;   // 1-D profiles
    double *metx = nullptr, *mety = nullptr, *sigx = nullptr, *sigy = nullptr;
    double *rhob = nullptr, *psib = nullptr, *gx = nullptr, *gy = nullptr;
    double **dptrs = nullptr;  // device pointer tables for operator kernels
    int *dflag = nullptr;
    bool have_profiles = false, have_base = false, have_state = false;
    int filt_L = 256, filt_W = 40;  // y sweep chunk length, look-back
    int filt_Lx = 256;              // x sweep chunk length
    double filt_ms = -1.0;  // autotuned time of both filter sweeps at filt_L (-1: not tuned)
    int filt_mode = 0;      // 0: x and y sweeps as two kernels; 1: one fused sweep (filter_xy_kernel)
    double filt_xy_ms = -1.0;  // autotuned time of the fused sweep (-1: not tuned)
    int filt_Ly = 512;      // row chunk length of the fused sweep
    bool filt_pinned = false;  // ns_set_filter_mode chose the mode: the autotuner keeps it
    hipGraphExec_t graph = nullptr;
    int stage_variant = -1;       // block-size variant of the stage kernel (-1: not tuned)
    bool same_launch = false;     // marching stages: edge tiles in the marching launch (autotuned)
    double tune_ms[3] = {0, 0, 0};
    bool timing = false;
    hipEvent_t ev[2 * 24] = {};
    double tsum[TC_N] = {};
    long tcount[TC_N] = {};
    long long interior_points = 0;  // output points of the interior kernel in the last timed stage
    // reverse mode: observer arc, checkpoints and adjoint work planes
    struct Adj {
        int2 *arc = nullptr;
        int narc = 0;
        double *jvals = nullptr;  // per-step objective values
        size_t jbytes = 0;
        void *ckpt = nullptr;     // state copies at steps 0, K, 2K, ... and the final state
        size_t ckbytes = 0;
        void *seg = nullptr;      // step states of one segment (K+1)
        size_t segbytes = 0;
        void *stg = nullptr;      // the 6 stage input states of one step
        void *stgall = nullptr;   // inputs of stages 2-6 of every step of one segment (when they fit)
        size_t stgallbytes = 0;
        void *lam = nullptr, *lamW = nullptr, *gx = nullptr, *gy = nullptr;
        int iter0 = 0, nsteps = 0, every = 0;
        bool have_fwd = false;
        bool keep = true;  // the backward may keep a segment's stage inputs (ns_adj_set_keep)
    } adj;
};

namespace {

double *plane_ptr(double *buf, const ns_ctx *c, int k) { return buf + (size_t)k * c->plane; }
template <typename S>
S *splane(void *buf, const ns_ctx *c, int k) { return static_cast<S *>(buf) + (size_t)k * c->plane; }

// ccpml row constants (cc.f90:38-59) for the FP64 form
RowEdge fp64_row_edge(const ns_params &P)
{
    RowEdge re{};
    for (int k = 0; k < NKMAX; ++k) re.bot_res[k] = re.top_res[k] = 0.0;
    re.bot_val[3] = re.top_val[3] = 1.0 / P.gamma;
    re.bot_val[4] = 0.0;
    re.top_val[4] = 1.0;
    return re;
}

// per-row base block: ub[nk], dub[nk], tb, dtb, ycorr[nk] (jmax each)
size_t prow_count(const ns_params &P) { return (size_t)(3 * P.nk + 2) * P.jmax; }

PertRows pert_rows(const ns_ctx *c)
{
    PertRows pr{};
    const size_t J = (size_t)c->P.jmax;
    const int nk = c->P.nk;
    for (int k = 0; k < NKMAX; ++k) {
        pr.ub[k] = k < nk ? c->prow + k * J : nullptr;
        pr.dub[k] = k < nk ? c->prow + (nk + k) * J : nullptr;
        pr.ycorr[k] = k < nk ? c->prow + (2 * nk + 2 + k) * J : nullptr;
    }
    pr.tb = c->prow + 2 * nk * J;
    pr.dtb = c->prow + (2 * nk + 1) * J;
    return pr;
}

// 1-D profile block: three arrays along i, five along j, one allocation
struct ProfileLayout {
    size_t metx, sigx, gx, mety, sigy, rhob, psib, gy, total;
};

ProfileLayout profile_layout(const ns_params &P)
{
    ProfileLayout L;
    const size_t I = (size_t)P.imax, J = (size_t)P.jmax;
    L.metx = 0;
    L.sigx = I;
    L.gx = 2 * I;
    L.mety = 3 * I;
    L.sigy = L.mety + J;
    L.rhob = L.sigy + J;
    L.psib = L.rhob + J;
    L.gy = L.psib + J;
    L.total = L.gy + J;
    return L;
}

size_t profile_count(const ns_params &P) { return profile_layout(P).total; }

Transport<double> make_transport(const ns_params &P)
{
    Transport<double> t;
    t.A_mi = P.A_mi;
    t.Bt_mi = P.B_mi / P.Tref;
    t.Ct_mi = P.C_mi / P.Tref / P.Tref;
    t.D_mi = P.D_mi;
    t.inv_mi0 = 1.0 / P.MI0;
    t.A_k = P.A_k;
    t.Bt_k = P.B_k / P.Tref;
    t.Ct_k = P.C_k / P.Tref / P.Tref;
    t.D_k = P.D_k;
    t.inv_k0 = 1.0 / P.K0;
    t.Tref = P.Tref;
    return t;
}

// Warm-up length of the look-back filter sweep: the homogeneous response of
// u_i + sigma*sum_{s=1..6} d_{-s} u_{i-s} = f_i must fall below 1e-19.
int filter_warmup(double sigma)
{
    double c[6];
    for (int s = 1; s <= 6; ++s) c[s - 1] = sigma * filt_d(6 - s);
    std::vector<double> h(400, 0.0);
    h[0] = 1.0;
    for (int n = 1; n < 400; ++n) {
        double acc = 0.0;
        for (int s = 1; s <= 6 && s <= n; ++s) acc -= c[s - 1] * h[n - s];
        h[n] = acc;
    }
    for (int W = 16; W < 380; ++W) {
        double tail = 0.0;
        for (int n = W; n < 400; ++n) tail += std::fabs(h[n]);
        if (tail < 1e-19) return W;
    }
    return -1;
}

template <typename S>
StageArgs<S> stage_args(ns_ctx *c, int irk, S *Uin, S *Uout, double amp, bool pert)
{
    const ns_params &P = c->P;
    StageArgs<S> A{};
    for (int k = 0; k < NKMAX; ++k) {
        const bool on = k < P.nk;
        A.U[k] = on ? Uin + (size_t)k * c->plane : nullptr;
        A.Uout[k] = on ? Uout + (size_t)k * c->plane : nullptr;
        A.W[k] = on ? splane<S>(c->Wr, c, k) : nullptr;
        A.base[k] = (on && !pert) ? reinterpret_cast<const S *>(plane_ptr(c->base, c, k)) : nullptr;
        A.dudt[k] = nullptr;
    }
    if (pert) A.pr = pert_rows(c);
    A.metx = c->metx;
    A.mety = c->mety;
    A.sigx = c->sigx;
    A.sigy = c->sigy;
    A.gx = c->gx;
    A.gy = c->gy;
    A.imax = P.imax;
    A.jmax = P.jmax;
    A.pitch = c->pitch;
    A.D = P.D;
    A.inv_dm = 1.0 / P.dm;
    A.inv_dn = 1.0 / P.dn;
    A.dt = P.dt;
    A.alpha = RK_ALPHA[irk];
    A.beta = RK_BETA[irk];
    A.gamma = P.gamma;
    A.inv_Re = 1.0 / P.Re;
    A.gm1_Re = (P.gamma - 1.0) / P.Re;
    A.inv_RePr = 1.0 / P.Re / P.Pr;
    A.inv_Pe = 1.0 / P.Pe;
    A.alpha_cond = P.alpha_cond;
    A.amp_src = amp;
    A.tc = make_transport(P);
    A.first = irk == 0;
    A.last = irk == 5;
    A.fast_ok = (P.dm == 1.0 && P.dn == 1.0 && !(P.flags & NS_FLAG_METRIC_EVERYWHERE)) ? 1 : 0;
    A.xper = (P.flags & NS_FLAG_PERIODIC_X) ? 1 : 0;
    return A;
}

template <typename S>
FilterArgs<S> filter_args(ns_ctx *c, bool xdir, S *in, S *out, bool edge_bc, bool pert)
{
    const ns_params &P = c->P;
    FilterArgs<S> F{};
    const PertRows pr = pert ? pert_rows(c) : PertRows{};
    for (int k = 0; k < NKMAX; ++k) {
        F.in[k] = k < P.nk ? in + (size_t)k * c->plane : nullptr;
        F.out[k] = k < P.nk ? out + (size_t)k * c->plane : nullptr;
        F.corr[k] = (pert && !xdir && k < P.nk) ? pr.ycorr[k] : nullptr;
    }
    F.re = pert ? c->re : fp64_row_edge(P);
    F.n = xdir ? P.imax : P.jmax;
    F.m = xdir ? P.jmax : P.imax;
    F.pitch = c->pitch;
    F.L = xdir ? c->filt_Lx : c->filt_L;
    F.W = c->filt_W;
    F.sigma = P.sigmaf;
    F.nk = P.nk;
    F.edge_bc = edge_bc;
    F.xper = (P.flags & NS_FLAG_PERIODIC_X) ? 1 : 0;
    return F;
}

template <typename S>
FilterXYArgs<S> filter_xy_args(ns_ctx *c, S *in, S *out, bool pert)
{
    const ns_params &P = c->P;
    FilterXYArgs<S> F{};
    const PertRows pr = pert ? pert_rows(c) : PertRows{};
    for (int k = 0; k < NKMAX; ++k) {
        F.in[k] = k < P.nk ? in + (size_t)k * c->plane : nullptr;
        F.out[k] = k < P.nk ? out + (size_t)k * c->plane : nullptr;
        F.corr[k] = (pert && k < P.nk) ? pr.ycorr[k] : nullptr;
    }
    F.re = pert ? c->re : fp64_row_edge(P);
    F.imax = P.imax;
    F.jmax = P.jmax;
    F.pitch = c->pitch;
    // lines shorter than 272 points are swept from their start in one
    // piece, as the two-kernel sweeps do (the reference's order, bitwise)
    F.Ly = P.jmax < 272 ? P.jmax : c->filt_Ly;
    F.x_exact = P.imax < 272;
    F.W = c->filt_W;
    F.xper = (P.flags & NS_FLAG_PERIODIC_X) ? 1 : 0;
    F.sigma = P.sigmaf;
    F.nk = P.nk;
    return F;
}

// the stage's filters: S1 -> (x) S2 -> (y + ccpml rows) out, or S1 -> out
// in one fused sweep
template <typename S>
hipError_t launch_filters(ns_ctx *c, S *S1, S *S2, S *out, bool pert, hipStream_t s)
{
    if (c->filt_mode == 1) return launch_filter_xy(filter_xy_args<S>(c, S1, out, pert), s);
    hipError_t e = launch_filter(filter_args<S>(c, true, S1, S2, false, pert), true, s);
    if (e == hipSuccess) e = launch_filter(filter_args<S>(c, false, S2, out, true, pert), false, s);
    return e;
}

template <typename S>
EdgeArgs<S> edge_args(ns_ctx *c, S *U, bool pert)
{
    EdgeArgs<S> E{};
    const PertRows pr = pert ? pert_rows(c) : PertRows{};
    for (int k = 0; k < NKMAX; ++k) {
        E.U[k] = k < c->P.nk ? U + (size_t)k * c->plane : nullptr;
        E.ub[k] = (pert && k < c->P.nk) ? pr.ub[k] : nullptr;
    }
    E.re = pert ? c->re : fp64_row_edge(c->P);
    E.rhob = c->rhob;
    E.psib = c->P.nk >= 5 ? c->psib : nullptr;
    E.imax = c->P.imax;
    E.jmax = c->P.jmax;
    E.pitch = c->pitch;
    E.nk = c->P.nk;
    E.gamma = c->P.gamma;
    E.xper = (c->P.flags & NS_FLAG_PERIODIC_X) ? 1 : 0;
    return E;
}

double stage_amp(const ns_params &P, int irk, int iter)
{
    if (P.amplitude == 0.0) return 0.0;
    return P.amplitude * std::sin(P.wf * (((double)iter + RK_C[irk]) * P.dt));
}

// one LDDRK6 stage on the device (rk.f90:52-73, stage irk of step iter):
// state S0 -> stage kernel -> S1 -> x filter -> S2 -> y filter (+ rows
// j=1,jmax) -> S0 -> columns i=1,imax.  in / out (null: S0) replace S0 as
// the stage's input and as the y filter's and ccpml columns' output.
template <typename S>
int enqueue_stage_t(ns_ctx *c, int irk, int iter, bool timed, S *in = nullptr, S *out = nullptr)
{
    const ns_params &P = c->P;
    const bool visc = !(P.flags & NS_FLAG_INVISCID);
    const bool pert = c->pert;
    S *S1 = static_cast<S *>(c->S[1]), *S2 = static_cast<S *>(c->S[2]);
    S *S0 = in ? in : static_cast<S *>(c->S[0]);
    S *S0out = out ? out : static_cast<S *>(c->S[0]);
    auto tick = [&](int slot) {
        if (timed) hipEventRecord(c->ev[slot], c->stream);
    };
    StageArgs<S> A = stage_args<S>(c, irk, S0, S1, stage_amp(P, irk, iter), pert);
    tick(0);
    // interior kernel on the context's stream; edge tiles in the same launch
    // or on the side stream (forked after the previous work, joined before
    // the filters)
    StageHooks tm;
    tm.fork = c->fork;
    tm.join = c->join;
    tm.same_launch = c->same_launch;
    if (timed) {
        tm.before = c->ev[5];
        tm.after = c->ev[6];
    }
    HIPCHK(launch_stage(A, P.nk, visc, false, c->stage_variant < 0 ? 0 : c->stage_variant, c->stream, c->side, &tm));
    tick(1);
    if (c->filt_mode == 1) {
        // one fused sweep: its time is reported under "filter_x", "filter_y" gets 0
        HIPCHK(launch_filter_xy(filter_xy_args<S>(c, S1, S0out, pert), c->stream));
        tick(2);
    } else {
        HIPCHK(launch_filter(filter_args<S>(c, true, S1, S2, false, pert), true, c->stream));
        tick(2);
        HIPCHK(launch_filter(filter_args<S>(c, false, S2, S0out, true, pert), false, c->stream));
    }
    tick(3);
    if (!(P.flags & NS_FLAG_PERIODIC_X)) HIPCHK(launch_edge_i(edge_args<S>(c, S0out, pert), c->stream));
    tick(4);
    if (timed) {
        HIPCHK(hipEventSynchronize(c->ev[4]));
        for (int t = 0; t < TC_INT; ++t) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->ev[t], c->ev[t + 1]));
            c->tsum[t] += ms;
            c->tcount[t] += 1;
        }
        if (tm.interior_points > 0) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->ev[5], c->ev[6]));
            c->tsum[TC_INT] += ms;
            c->tcount[TC_INT] += 1;
        }
        c->interior_points = tm.interior_points;
    }
    return NS_OK;
}

int enqueue_stage(ns_ctx *c, int irk, int iter, bool timed)
{
    return c->pert ? enqueue_stage_t<float>(c, irk, iter, timed) : enqueue_stage_t<double>(c, irk, iter, timed);
}

// one LDDRK6 step on the device, rk.f90:52-73
int enqueue_step_t(ns_ctx *c, int iter, bool timed)
{
    for (int irk = 0; irk < 6; ++irk) {
        int rc = enqueue_stage(c, irk, iter, timed);
        if (rc) return rc;
    }
    return NS_OK;
}

int enqueue_step(ns_ctx *c, int iter, bool timed) { return enqueue_step_t(c, iter, timed); }

// Before a step is captured into a graph: the host-side set-up of its
// launches (kernel code objects load at first use, marching plans query the
// device once), so that no such runtime call runs inside the capture.
template <typename S>
int prepare_capture_t(ns_ctx *c)
{
    const StageArgs<S> A = stage_args<S>(c, 0, static_cast<S *>(c->S[0]), static_cast<S *>(c->S[1]), 0.0, c->pert);
    HIPCHK(prepare_step(A, c->P.nk, !(c->P.flags & NS_FLAG_INVISCID), c->stage_variant < 0 ? 0 : c->stage_variant,
                        false));
    return NS_OK;
}

// Autotuner: time each variant of the stage launch on the live workload, with
// the same fork/join onto the side stream as a step, and keep the fastest.
// It runs stage 2 (which reads and writes W, like stages 2-5) from S0 into
// S1; between steps W is dead (stage 1 does not read it) and S1 is scratch,
// so the state is untouched.
template <typename S>
int autotune_stage_t(ns_ctx *c)
{
    const ns_params &P = c->P;
    const bool visc = !(P.flags & NS_FLAG_INVISCID);
    S *S0 = static_cast<S *>(c->S[0]), *S1 = static_cast<S *>(c->S[1]);
    const StageArgs<S> A = stage_args<S>(c, 1, S0, S1, 0.0, c->pert);
    auto launch = [&](int v) -> hipError_t {
        StageHooks h;
        h.fork = c->fork;
        h.join = c->join;
        h.same_launch = c->same_launch;
        return launch_stage(A, P.nk, visc, false, v, c->stream, c->side, &h);
    };
    int best = 0;
    bool best_same = false;
    double best_ms = 1e300;
    const int reps = 3;
    // each variant with its edge tiles on the second stream and in the
    // marching launch (a variant without a marching launch runs the same
    // either way); tune_ms reports each variant's faster form
    for (int v = 0; v < 3; ++v) {
        c->tune_ms[v] = -1.0;
        for (int f = 0; f < 2; ++f) {
            c->same_launch = f == 1;
            hipError_t e = launch(v);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                continue;
            }
            HIPCHK(hipEventRecord(c->ev[0], c->stream));
            for (int r = 0; r < reps; ++r) HIPCHK(launch(v));
            HIPCHK(hipEventRecord(c->ev[1], c->stream));
            HIPCHK(hipEventSynchronize(c->ev[1]));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            const double t = ms / reps;
            if (c->tune_ms[v] < 0 || t < c->tune_ms[v]) c->tune_ms[v] = t;
            // a later candidate must be more than 3 % faster to replace an
            // earlier one, so that candidates within timing noise do not flip
            // from run to run
            if (t < best_ms * (best_ms < 1e299 ? 0.97 : 1.0)) {
                best_ms = t;
                best = v;
                best_same = f == 1;
            }
        }
    }
    c->stage_variant = best;
    c->same_launch = best_same;
    // look-back chunk length of each filter sweep (more chunks: more waves in
    // flight; longer chunks: less warm-up work), timed on the scratch buffers
    // S1 -> S2 (x) and S2 -> S1 (y) so the state is untouched.  Lines
    // shorter than 272 points keep L = 256: one chunk, the exact sequential
    // sweep, so small grids stay bitwise equal to the reference order (a
    // look-back start is exact to 1e-19 relative, which can still flip a
    // rounding now and then).
    S *S2 = static_cast<S *>(c->S[2]);
    const int Ls[5] = {64, 128, 256, 512, 1024};
    auto tune_L = [&](bool xdir, int &L, double &best_ms) -> int {
        const int n = xdir ? P.imax : P.jmax;
        if (n < 272) return NS_OK;
        int bestL = L;
        best_ms = 1e300;
        for (int t = 0; t < 5; ++t) {
            if (Ls[t] >= n && t > 0) break;
            L = Ls[t];
            auto sweep = [&]() -> hipError_t {
                return xdir ? launch_filter(filter_args<S>(c, true, S1, S2, false, c->pert), true, c->stream)
                            : launch_filter(filter_args<S>(c, false, S2, S1, true, c->pert), false, c->stream);
            };
            HIPCHK(sweep());
            HIPCHK(hipEventRecord(c->ev[0], c->stream));
            for (int r = 0; r < reps; ++r) HIPCHK(sweep());
            HIPCHK(hipEventRecord(c->ev[1], c->stream));
            HIPCHK(hipEventSynchronize(c->ev[1]));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            if (ms / reps < best_ms) {
                best_ms = ms / reps;
                bestL = Ls[t];
            }
        }
        L = bestL;
        return NS_OK;
    };
    double x_ms = 0.0, y_ms = 0.0;
    int rc;
    if ((rc = tune_L(true, c->filt_Lx, x_ms))) return rc;
    if ((rc = tune_L(false, c->filt_L, y_ms))) return rc;
    const int ntune = std::max(P.imax, P.jmax) < 272 ? 0 : 3;
    const double bestL_ms = x_ms + y_ms;
    if (ntune) c->filt_ms = bestL_ms;
    // the fused sweep (filter_xy_kernel) against the two sweeps, with its
    // row chunk length; the fused one must be more than 3 % faster to be
    // chosen (a caller's ns_set_filter_mode stands)
    if (ntune && !c->filt_pinned) {
        const int Lys[3] = {256, 512, 1024};
        int bestLy = c->filt_Ly;
        double bestxy = 1e300;
        const int mode0 = c->filt_mode;
        c->filt_mode = 1;
        for (int t = 0; t < 3; ++t) {
            if (Lys[t] >= P.jmax && t > 0) break;
            c->filt_Ly = Lys[t];
            auto sweep = [&]() -> hipError_t { return launch_filters<S>(c, S1, S2, S2, c->pert, c->stream); };
            hipError_t e = sweep();
            if (e != hipSuccess) {
                c->filt_mode = mode0;
                return fail(NS_ERR_HIP, "fused filter: %s", hipGetErrorString(e));
            }
            HIPCHK(hipEventRecord(c->ev[0], c->stream));
            for (int r = 0; r < reps; ++r) HIPCHK(sweep());
            HIPCHK(hipEventRecord(c->ev[1], c->stream));
            HIPCHK(hipEventSynchronize(c->ev[1]));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            if (ms / reps < bestxy) {
                bestxy = ms / reps;
                bestLy = Lys[t];
            }
        }
        c->filt_Ly = bestLy;
        c->filt_mode = mode0;
        // the choice: each form timed after the chosen stage variant, as in
        // a step (what the stage leaves in the caches changes the filters'
        // time: in isolation the fused sweep can win where it loses in the
        // step); the fused form must be more than 3 % faster
        auto stage_then = [&](int mode) -> hipError_t {
            c->filt_mode = mode;
            hipError_t e = launch(best);
            if (e == hipSuccess) e = launch_filters<S>(c, S1, S2, S2, c->pert, c->stream);
            return e;
        };
        double form_ms[2];
        for (int mode = 0; mode < 2; ++mode) {
            HIPCHK(stage_then(mode));
            HIPCHK(hipEventRecord(c->ev[0], c->stream));
            for (int r = 0; r < reps; ++r) HIPCHK(stage_then(mode));
            HIPCHK(hipEventRecord(c->ev[1], c->stream));
            HIPCHK(hipEventSynchronize(c->ev[1]));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            form_ms[mode] = ms / reps;
        }
        c->filt_mode = form_ms[1] < 0.97 * form_ms[0] ? 1 : 0;
        c->filt_ms = form_ms[0];
        c->filt_xy_ms = form_ms[1];
    }
    return NS_OK;
}

int autotune_stage(ns_ctx *c)
{
    int rc = c->pert ? autotune_stage_t<float>(c) : autotune_stage_t<double>(c);
    if (rc) return rc;
    // the chosen variant's set-up, so that a caller may capture the steps
    return c->pert ? prepare_capture_t<float>(c) : prepare_capture_t<double>(c);
}

int check_ready(ns_ctx *c)
{
    if (!c) return fail(NS_ERR_PARAM, "null context");
    if (!c->have_profiles) return fail(NS_ERR_STATE, "ns_set_profiles has not been called");
    if (!c->have_base) return fail(NS_ERR_STATE, "ns_set_base has not been called");
    // launches and kernel set-up go to the calling thread's current device
    HIPCHK(hipSetDevice(c->P.device));
    return NS_OK;
}

// host AoS (k fastest) <-> device SoA planes
int upload_aos(ns_ctx *c, double *dst, const double *U, int nk)
{
    const ns_params &P = c->P;
    std::vector<double> tmp((size_t)nk * c->plane, 0.0);
    for (int j = 0; j < P.jmax; ++j)
        for (int i = 0; i < P.imax; ++i)
            for (int k = 0; k < nk; ++k)
                tmp[(size_t)k * c->plane + (size_t)j * c->pitch + i] = U[k + (size_t)nk * (i + (size_t)P.imax * j)];
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

int download_aos(ns_ctx *c, const double *src, double *U, int nk)
{
    const ns_params &P = c->P;
    std::vector<double> tmp((size_t)nk * c->plane);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int j = 0; j < P.jmax; ++j)
        for (int i = 0; i < P.imax; ++i)
            for (int k = 0; k < nk; ++k)
                U[k + (size_t)nk * (i + (size_t)P.imax * j)] = tmp[(size_t)k * c->plane + (size_t)j * c->pitch + i];
    return NS_OK;
}

int set_ptr_table(ns_ctx *c, int slot, void *buf, int nk, size_t esz = sizeof(double));

// One fork/join round trip between two new streams, synchronised: both
// streams are in use before any step is captured into a graph, so the
// side stream's first operation is never one recorded inside a capture.
hipError_t touch_streams(hipStream_t s, hipStream_t side, hipEvent_t fork, hipEvent_t join)
{
    hipError_t e = hipEventRecord(fork, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(side, fork, 0);
    if (e == hipSuccess) e = hipEventRecord(join, side);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, join, 0);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
}

// state <-> host AoS; the perturbation form stores q = U - Ub(k,j) in FP32
int upload_state(ns_ctx *c, const double *U)
{
    if (!c->pert) return upload_aos(c, static_cast<double *>(c->S[0]), U, c->P.nk);
    const ns_params &P = c->P;
    const int nk = P.nk;
    const size_t J = (size_t)P.jmax;
    std::vector<float> tmp((size_t)nk * c->plane, 0.f);
    for (int j = 0; j < P.jmax; ++j)
        for (int i = 0; i < P.imax; ++i)
            for (int k = 0; k < nk; ++k)
                tmp[(size_t)k * c->plane + (size_t)j * c->pitch + i] =
                    (float)(U[k + (size_t)nk * (i + (size_t)P.imax * j)] - c->ubh[k * J + j]);
    HIPCHK(hipMemcpyAsync(c->S[0], tmp.data(), tmp.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

int download_state(ns_ctx *c, double *U)
{
    if (!c->pert) return download_aos(c, static_cast<const double *>(c->S[0]), U, c->P.nk);
    const ns_params &P = c->P;
    const int nk = P.nk;
    const size_t J = (size_t)P.jmax;
    std::vector<float> tmp((size_t)nk * c->plane);
    HIPCHK(hipMemcpyAsync(tmp.data(), c->S[0], tmp.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int j = 0; j < P.jmax; ++j)
        for (int i = 0; i < P.imax; ++i)
            for (int k = 0; k < nk; ++k)
                U[k + (size_t)nk * (i + (size_t)P.imax * j)] =
                    c->ubh[k * J + j] + (double)tmp[(size_t)k * c->plane + (size_t)j * c->pitch + i];
    return NS_OK;
}

// column i=0 of nplanes device planes -> host rows (jmax each)
int extract_column(ns_ctx *c, const double *planes, int nplanes, double *rows)
{
    const ns_params &P = c->P;
    for (int k = 0; k < nplanes; ++k)
        HIPCHK(hipMemcpy2DAsync(rows + (size_t)k * P.jmax, sizeof(double), planes + (size_t)k * c->plane,
                                (size_t)c->pitch * sizeof(double), sizeof(double), P.jmax, hipMemcpyDeviceToHost,
                                c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

// Per-row base data of the perturbation form, computed with the FP64 kernels:
// d Ub/dy and d Tb/dy (DRP + metric) and Ly(Ub) - Ub (the in-place y filter
// acting on the base, which includes the float32 boundary literals' drift).
int setup_perturbation(ns_ctx *c, const double *Ub)
{
    const ns_params &P = c->P;
    const int I = P.imax, J = P.jmax, nk = P.nk;
    c->ubh.assign((size_t)nk * J, 0.0);
    for (int j = 0; j < J; ++j)
        for (int k = 0; k < nk; ++k) {
            const double v = Ub[k + (size_t)nk * ((size_t)I * j)];
            for (int i = 1; i < I; ++i)
                if (Ub[k + (size_t)nk * (i + (size_t)I * j)] != v)
                    return fail(NS_ERR_PARAM,
                                "the FP32 perturbation form needs a base flow constant along each row "
                                "(Ub(%d,%d,%d) differs from Ub(%d,1,%d))", k + 1, i + 1, j + 1, k + 1, j + 1);
            c->ubh[(size_t)k * J + j] = v;
        }
    std::vector<double> rows(prow_count(P), 0.0);
    const PertRows off = pert_rows(c);
    auto at = [&](const double *dp) { return rows.data() + (dp - c->prow); };
    std::copy(c->ubh.begin(), c->ubh.end(), at(off.ub[0]));
    double *tb = at(off.tb);
    for (int j = 0; j < J; ++j) tb[j] = P.gamma * c->ubh[3 * (size_t)J + j] / c->ubh[j];
    double *outb = c->work + 8 * c->plane;
    int rc;
    // d Ub / dy
    if ((rc = set_ptr_table(c, 0, c->base, nk)) || (rc = set_ptr_table(c, 1, outb, nk))) return rc;
    HIPCHK(launch_deriv((const double *const *)c->dptrs, (double *const *)(c->dptrs + 8), nk, I, J, c->pitch, true,
                        c->mety, 1.0 / P.dn, false, c->stream));
    if ((rc = extract_column(c, outb, nk, at(off.dub[0])))) return rc;
    // d Tb / dy
    {
        std::vector<double> tp(c->plane, 0.0);
        for (int j = 0; j < J; ++j)
            for (int i = 0; i < I; ++i) tp[(size_t)j * c->pitch + i] = tb[j];
        HIPCHK(hipMemcpyAsync(c->work, tp.data(), tp.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        if ((rc = set_ptr_table(c, 0, c->work, 1))) return rc;
        HIPCHK(launch_deriv((const double *const *)c->dptrs, (double *const *)(c->dptrs + 8), 1, I, J, c->pitch,
                            true, c->mety, 1.0 / P.dn, false, c->stream));
        if ((rc = extract_column(c, outb, 1, at(off.dtb)))) return rc;
    }
    // Ly(Ub) - Ub
    HIPCHK(launch_filter(filter_args<double>(c, false, c->base, outb, false, false), false, c->stream));
    double *yc = at(off.ycorr[0]);
    if ((rc = extract_column(c, outb, nk, yc))) return rc;
    for (size_t q = 0; q < (size_t)nk * J; ++q) yc[q] -= c->ubh[q];
    HIPCHK(hipMemcpyAsync(c->prow, rows.data(), rows.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // ccpml rows in perturbation form (cc.f90:38-59 minus the base rows)
    RowEdge re = fp64_row_edge(P);
    auto ub = [&](int k, int j) { return c->ubh[(size_t)k * J + j]; };
    for (int k = 0; k < 3; ++k) {
        re.bot_res[k] = 2.0 * ub(k, 1) - ub(k, 2) - ub(k, 0);
        re.top_res[k] = 2.0 * ub(k, J - 2) - ub(k, J - 3) - ub(k, J - 1);
    }
    re.bot_val[3] = 1.0 / P.gamma - ub(3, 0);
    re.top_val[3] = 1.0 / P.gamma - ub(3, J - 1);
    if (nk >= 5) {
        re.bot_val[4] = 0.0 - ub(4, 0);
        re.top_val[4] = 1.0 - ub(4, J - 1);
    }
    c->re = re;
    return NS_OK;
}

int set_ptr_table(ns_ctx *c, int slot, void *buf, int nk, size_t esz)
{
    void *h[8];
    for (int k = 0; k < 8; ++k) h[k] = k < nk ? static_cast<char *>(buf) + (size_t)k * c->plane * esz : nullptr;
    HIPCHK(hipMemcpyAsync(c->dptrs + 8 * slot, h, sizeof h, hipMemcpyHostToDevice, c->stream));
    // h lives on this stack frame: the copy must have read it before return
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

void release(ns_ctx *c)
{
    if (c->graph) hipGraphExecDestroy(c->graph);
    for (void *p : {c->S[0], c->S[1], c->S[2], c->Wr, (void *)c->base, (void *)c->work, (void *)c->prof,
                    (void *)c->prow})
        if (p) hipFree(p);
    if (c->dptrs) hipFree(c->dptrs);
    if (c->dflag) hipFree(c->dflag);
    for (auto &e : c->ev)
        if (e) hipEventDestroy(e);
    if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
    if (c->fork) hipEventDestroy(c->fork);
    if (c->join) hipEventDestroy(c->join);
    if (c->side) hipStreamDestroy(c->side);
    for (void *p : {(void *)c->adj.arc, (void *)c->adj.jvals, c->adj.ckpt, c->adj.seg, c->adj.stg,
                    c->adj.stgall, c->adj.lam, c->adj.lamW, c->adj.gx, c->adj.gy})
        if (p) hipFree(p);
}

}  // namespace

extern "C" {

int32_t ns_api_version(void) { return NS_API_VERSION; }

const char *ns_last_error(void) { return g_err.c_str(); }

int32_t ns_default_params(ns_params *p)
{
    if (!p) return fail(NS_ERR_PARAM, "null params");
    std::memset(p, 0, sizeof *p);
    // dados.in:1-85 defaults
    p->imax = 521;
    p->jmax = 481;
    p->nk = 5;
    p->D = 90;
    p->dm = (19.5 - (-6.5)) / (double)(float)(521 - 1);
    p->dn = (2.4 - (-2.4)) / (double)(float)(481 - 1);
    p->dt = 0.01;
    p->sigmaf = 0.1;
    p->gamma = 1.4;  // init.f90:361
    p->Re = 10000.0;
    p->Pr = 0.71;
    p->Pe = 1000.0;        // global.f90:38
    p->alpha_cond = 0.5;   // global.f90:39
    p->Tref = 273.15;
    // coefficients.f90:113-123 (N2, Svehla 1995) and :78-79
    p->A_mi = 0.62526577;
    p->B_mi = -31.779652;
    p->C_mi = -1640.7913;
    p->D_mi = 1.7454992;
    p->A_k = 0.85372829;
    p->B_k = 105.18665;
    p->C_k = -12229.753;
    p->D_k = 0.48299104;
    const double T = p->Tref;
    p->K0 = std::exp(p->A_k * std::log(1.0 * T) + p->B_k / T + p->C_k / T / T + p->D_k);
    p->MI0 = std::exp(p->A_mi * std::log(1.0 * T) + p->B_mi / T + p->C_mi / T / T + p->D_mi);
    p->amplitude = 0.8;
    p->wf = 1.0;
    p->r0_acous = 0.03;
    p->m_pt_acous = 0.0;
    p->n_pt_acous = 0.0;
    p->precision = NS_PREC_FP64;
    p->flags = 0;
    p->device = 0;
    return NS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Staging of ns_probe_noalloc without device allocation: the point list and
// the gathered values live in one file-scope device array.  Each device holds
// its own copy of it at its own address; the address is looked up once per
// device (ns_create, with the device current) and kept in scratch_dev.  Host
// code passes that address, never the array's name, to copies and kernels.
// Every call on a device shares that device's copy: calls that overlap on one
// device race.  Nothing in it needs clearing: a call writes the point list
// and every gathered value before it reads them.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t SCRATCH_IJ_BYTES = sizeof(int32_t) * 2 * NS_PROBE_NOALLOC_MAX;
constexpr size_t SCRATCH_OUT_OFFSET = (SCRATCH_IJ_BYTES + 255) / 256 * 256;
constexpr size_t SCRATCH_BYTES = SCRATCH_OUT_OFFSET + sizeof(double) * NKMAX * NS_PROBE_NOALLOC_MAX;
constexpr int SCRATCH_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex scratch_mu;
void *scratch_dev[SCRATCH_MAX_DEVICES] = {};

// the device's address of `scratch` (dev current); once per device
int scratch_setup(int dev)
{
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
        return fail(NS_ERR_PARAM, "device %d beyond the %d the library supports", dev, SCRATCH_MAX_DEVICES);
    std::lock_guard<std::mutex> g(scratch_mu);
    if (scratch_dev[dev]) return NS_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(NS_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    scratch_dev[dev] = ptr;
    return NS_OK;
}

void *scratch_ptr(int dev)
{
    std::lock_guard<std::mutex> g(scratch_mu);
    return (dev >= 0 && dev < SCRATCH_MAX_DEVICES) ? scratch_dev[dev] : nullptr;
}
}  // namespace

extern "C" {

int32_t ns_create(const ns_params *p, ns_ctx **out)
{
    if (!p || !out) return fail(NS_ERR_PARAM, "null argument");
    *out = nullptr;
    const ns_params &P = *p;
    if (P.imax < 16 || P.jmax < 16) return fail(NS_ERR_PARAM, "imax and jmax must be >= 16 (got %d x %d)", P.imax, P.jmax);
    if (P.nk != 4 && P.nk != 5) return fail(NS_ERR_PARAM, "nk must be 4 or 5 (got %d)", P.nk);
    if (P.D < 0 || 2 * P.D > P.imax || 2 * P.D > P.jmax)
        return fail(NS_ERR_PARAM, "buffer width D=%d must satisfy 0 <= 2D <= min(imax, jmax)", P.D);
    if (!(P.dm > 0) || !(P.dn > 0) || !(P.dt > 0)) return fail(NS_ERR_PARAM, "dm, dn, dt must be positive");
    if (!(P.sigmaf >= 0.0 && P.sigmaf <= 1.0)) return fail(NS_ERR_PARAM, "sigmaf must lie in [0,1]");
    if (P.precision != NS_PREC_FP64 && P.precision != NS_PREC_FP32_PERT)
        return fail(NS_ERR_PARAM, "unknown precision mode %d", P.precision);
    if (!(P.gamma > 1.0) || !(P.Re > 0) || !(P.Pr > 0) || !(P.Pe > 0) || !(P.Tref > 0) || !(P.MI0 > 0) || !(P.K0 > 0))
        return fail(NS_ERR_PARAM, "gamma, Re, Pr, Pe, Tref, MI0, K0 must be positive (gamma > 1)");
    if ((P.flags & NS_FLAG_PERIODIC_X) && P.imax < 64)
        return fail(NS_ERR_PARAM, "periodic x needs imax >= 64 (got %d)", P.imax);
    const int W = filter_warmup(P.sigmaf);
    if (W < 0) return fail(NS_ERR_PARAM, "filter recurrence does not decay for sigmaf=%g", P.sigmaf);

    ns_ctx *c = new ns_ctx();
    c->P = P;
    c->pert = P.precision == NS_PREC_FP32_PERT;
    c->esz = c->pert ? sizeof(float) : sizeof(double);
    c->filt_W = W;
    c->filt_L = 256;
    // row pitch: a multiple of 64 elements; FP64 rows whose size is a
    // multiple of 4 KB get 64 more elements (512 B), so that a column's
    // consecutive rows start in different HBM channels (4096^2: FP64 x
    // filter 0.280 -> 0.250 ms, step 6.59 -> 6.43 ms; 2048-wide mixing
    // layer 1.159 -> 1.112 ms).  Every pad tried on FP32 rows (128 B to 4 KB)
    // slowed its stage kernel by 2-10 %, so they keep the multiple of 64.
    c->pitch = (P.imax + 63) / 64 * 64;
    if (!c->pert && ((size_t)c->pitch * sizeof(double)) % 4096 == 0) c->pitch += 64;
    // plane stride: 512 B more than the plane, so that the nk planes' same
    // point starts in different HBM channels (FP32 4096^2 step 3.957 ->
    // 3.928 ms; FP64 unchanged; 2 KB and 8 KB measured no better)
    c->plane = (size_t)c->pitch * P.jmax + 512 / c->esz;
    auto bail = [&](int code) {
        release(c);
        delete c;
        return code;
    };
    hipError_t e = hipSetDevice(P.device);
    if (e != hipSuccess) return bail(fail(NS_ERR_HIP, "hipSetDevice(%d): %s", P.device, hipGetErrorString(e)));
    int src;
    if ((src = scratch_setup(P.device))) return bail(src);
    int lds = 0;
    hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, P.device);
    if (lds < 150 * 1024)
        return bail(fail(NS_ERR_HIP, "device %d offers %d bytes of LDS per workgroup; the stage kernel needs 150 KiB (gfx950)",
                         P.device, lds));
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(fail(NS_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)));
    c->own_stream = true;
    e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join, hipEventDisableTiming);
    if (e == hipSuccess) e = touch_streams(c->stream, c->side, c->fork, c->join);
    if (e != hipSuccess) return bail(fail(NS_ERR_HIP, "side stream/events: %s", hipGetErrorString(e)));
    const size_t sbytes = (size_t)P.nk * c->plane * c->esz;
    for (void **pp : {&c->S[0], &c->S[1], &c->S[2], &c->Wr}) {
        e = hipMalloc(pp, sbytes);
        if (e != hipSuccess) return bail(fail(NS_ERR_NOMEM, "hipMalloc(%zu): %s", sbytes, hipGetErrorString(e)));
        e = hipMemset(*pp, 0, sbytes);
        if (e != hipSuccess) return bail(fail(NS_ERR_HIP, "hipMemset: %s", hipGetErrorString(e)));
    }
    const size_t bbytes = (size_t)P.nk * c->plane * sizeof(double);
    e = hipMalloc(&c->base, bbytes);
    if (e != hipSuccess) return bail(fail(NS_ERR_NOMEM, "hipMalloc(%zu): %s", bbytes, hipGetErrorString(e)));
    e = hipMemset(c->base, 0, bbytes);
    if (e != hipSuccess) return bail(fail(NS_ERR_HIP, "hipMemset: %s", hipGetErrorString(e)));
    e = hipMalloc(&c->prow, prow_count(P) * sizeof(double));
    if (e != hipSuccess) return bail(fail(NS_ERR_NOMEM, "hipMalloc(rows): %s", hipGetErrorString(e)));
    c->re = fp64_row_edge(P);
    e = hipMalloc(&c->work, 16 * c->plane * sizeof(double));
    if (e != hipSuccess) return bail(fail(NS_ERR_NOMEM, "hipMalloc(work): %s", hipGetErrorString(e)));
    e = hipMalloc(&c->prof, profile_count(P) * sizeof(double));
    if (e != hipSuccess) return bail(fail(NS_ERR_NOMEM, "hipMalloc(profiles): %s", hipGetErrorString(e)));
    ProfileLayout L = profile_layout(P);
    c->metx = c->prof + L.metx;
    c->sigx = c->prof + L.sigx;
    c->gx = c->prof + L.gx;
    c->mety = c->prof + L.mety;
    c->sigy = c->prof + L.sigy;
    c->rhob = c->prof + L.rhob;
    c->psib = c->prof + L.psib;
    c->gy = c->prof + L.gy;
    e = hipMalloc(&c->dptrs, 4 * 8 * sizeof(double *));
    if (e != hipSuccess) return bail(fail(NS_ERR_NOMEM, "hipMalloc(ptrs): %s", hipGetErrorString(e)));
    e = hipMalloc(&c->dflag, sizeof(int) * 64);
    if (e != hipSuccess) return bail(fail(NS_ERR_NOMEM, "hipMalloc(flag): %s", hipGetErrorString(e)));
    for (auto &ev : c->ev) {
        e = hipEventCreate(&ev);
        if (e != hipSuccess) return bail(fail(NS_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e)));
    }
    *out = c;
    return NS_OK;
}

int32_t ns_destroy(ns_ctx *ctx)
{
    if (!ctx) return NS_OK;
    hipSetDevice(ctx->P.device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    release(ctx);
    delete ctx;
    return NS_OK;
}

int32_t ns_set_stream(ns_ctx *c, void *s)
{
    if (!c) return fail(NS_ERR_PARAM, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->graph) {
        hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    if (c->own_stream) hipStreamDestroy(c->stream);
    if (s) {
        c->stream = (hipStream_t)s;
        c->own_stream = false;
    } else {
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
    }
    HIPCHK(touch_streams(c->stream, c->side, c->fork, c->join));
    return NS_OK;
}

int32_t ns_synchronize(ns_ctx *c)
{
    if (!c) return fail(NS_ERR_PARAM, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

int32_t ns_set_profiles(ns_ctx *c, const double *meshx, const double *meshy, const double *sigmax,
                        const double *sigmay, const double *m, const double *n, const double *rhob,
                        const double *psib)
{
    if (!c) return fail(NS_ERR_PARAM, "null context");
    const ns_params &P = c->P;
    if (!meshx || !meshy || !sigmax || !sigmay || !m || !n || !rhob)
        return fail(NS_ERR_PARAM, "profile pointer is NULL");
    if (P.nk >= 5 && !psib) return fail(NS_ERR_PARAM, "psib is required when nk == 5");
    const int imax = P.imax, jmax = P.jmax, D = P.D;
    const bool everywhere = P.flags & NS_FLAG_METRIC_EVERYWHERE;
    std::vector<double> h(profile_count(P), 0.0);
    const ProfileLayout L = profile_layout(P);
    double *metx = h.data() + L.metx, *sigx = h.data() + L.sigx, *gx = h.data() + L.gx;
    double *mety = h.data() + L.mety, *sigy = h.data() + L.sigy, *rb = h.data() + L.rhob;
    double *pb = h.data() + L.psib, *gy = h.data() + L.gy;
    const double r2 = P.r0_acous * P.r0_acous;
    const bool xper = P.flags & NS_FLAG_PERIODIC_X;
    for (int i = 0; i < imax; ++i) {
        // metricsX (diff.f90:315-350): meshx applies in the x layers only
        // (periodic x has none)
        const bool strip = !xper && ((i + 1 <= D) || (i + 1 >= imax - D + 1));
        metx[i] = (everywhere || strip) ? meshx[i] : 1.0;
        sigx[i] = sigmax[i];
        const double dx = m[i] - P.m_pt_acous;
        gx[i] = r2 > 0 ? std::exp(-(std::log(2.0) * (dx * dx) / r2)) : 0.0;
    }
    for (int j = 0; j < jmax; ++j) {
        const bool strip = (j + 1 <= D) || (j + 1 >= jmax - D + 1);
        mety[j] = (everywhere || strip) ? meshy[j] : 1.0;
        sigy[j] = sigmay[j];
        rb[j] = rhob[j];
        pb[j] = psib ? psib[j] : 0.0;
        const double dy = n[j] - P.n_pt_acous;
        gy[j] = r2 > 0 ? std::exp(-(std::log(2.0) * (dy * dy) / r2)) : 0.0;
    }
    HIPCHK(hipMemcpyAsync(c->prof, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_profiles = true;
    return NS_OK;
}

int32_t ns_set_base(ns_ctx *c, const double *Ub)
{
    if (!c || !Ub) return fail(NS_ERR_PARAM, "null argument");
    int rc = upload_aos(c, c->base, Ub, c->P.nk);
    if (rc) return rc;
    if (c->pert) {
        if (!c->have_profiles) return fail(NS_ERR_STATE, "the FP32 perturbation form needs ns_set_profiles first");
        if ((rc = setup_perturbation(c, Ub))) return rc;
    }
    c->have_base = true;
    return NS_OK;
}

int32_t ns_put_state(ns_ctx *c, const double *U)
{
    if (!c || !U) return fail(NS_ERR_PARAM, "null argument");
    if (c->pert && !c->have_base) return fail(NS_ERR_STATE, "the FP32 perturbation form needs ns_set_base first");
    int rc = upload_state(c, U);
    if (rc == NS_OK) c->have_state = true;
    return rc;
}

int32_t ns_get_state(ns_ctx *c, double *U)
{
    if (!c || !U) return fail(NS_ERR_PARAM, "null argument");
    return download_state(c, U);
}

int32_t ns_probe(ns_ctx *c, int32_t npts, const int32_t *ij, double *out)
{
    if (!c || !ij || !out || npts < 0) return fail(NS_ERR_PARAM, "bad argument");
    if (npts == 0) return NS_OK;
    for (int t = 0; t < npts; ++t)
        if (ij[2 * t] < 1 || ij[2 * t] > c->P.imax || ij[2 * t + 1] < 1 || ij[2 * t + 1] > c->P.jmax)
            return fail(NS_ERR_PARAM, "probe %d at (%d,%d) lies outside the grid", t, ij[2 * t], ij[2 * t + 1]);
    const int nk = c->P.nk;
    int *dij = nullptr;
    double *dout = nullptr;
    HIPCHK(hipMalloc(&dij, sizeof(int) * 2 * npts));
    HIPCHK(hipMalloc(&dout, sizeof(double) * npts * nk));
    int rc = set_ptr_table(c, 0, c->S[0], nk, c->esz);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(dij, ij, sizeof(int) * 2 * npts, hipMemcpyHostToDevice, c->stream));
    if (c->pert)
        HIPCHK(launch_probe<float>((const float *const *)c->dptrs, nk, c->pitch, npts, dij, dout, c->stream));
    else
        HIPCHK(launch_probe<double>((const double *const *)c->dptrs, nk, c->pitch, npts, dij, dout, c->stream));
    HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * npts * nk, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    hipFree(dij);
    hipFree(dout);
    if (c->pert)
        for (int t = 0; t < npts; ++t)
            for (int k = 0; k < nk; ++k) out[t * nk + k] += c->ubh[(size_t)k * c->P.jmax + (ij[2 * t + 1] - 1)];
    return NS_OK;
}

int32_t ns_probe_noalloc(ns_ctx *c, int32_t npts, const int32_t *ij, double *out)
{
    if (!c || npts < 0 || (npts > 0 && (!ij || !out))) return fail(NS_ERR_PARAM, "bad argument");
    if (npts > NS_PROBE_NOALLOC_MAX)
        return fail(NS_ERR_PARAM, "%d probe points exceed NS_PROBE_NOALLOC_MAX (%d)", npts, NS_PROBE_NOALLOC_MAX);
    if (npts == 0) return NS_OK;
    for (int t = 0; t < npts; ++t)
        if (ij[2 * t] < 1 || ij[2 * t] > c->P.imax || ij[2 * t + 1] < 1 || ij[2 * t + 1] > c->P.jmax)
            return fail(NS_ERR_PARAM, "probe %d at (%d,%d) lies outside the grid", t, ij[2 * t], ij[2 * t + 1]);
    HIPCHK(hipSetDevice(c->P.device));
    unsigned char *base = static_cast<unsigned char *>(scratch_ptr(c->P.device));
    if (!base) return fail(NS_ERR_STATE, "device %d has no probe staging (no context was created on it)", c->P.device);
    int *dij = reinterpret_cast<int *>(base);
    double *dout = reinterpret_cast<double *>(base + SCRATCH_OUT_OFFSET);
    const int nk = c->P.nk;
    int rc = set_ptr_table(c, 0, c->S[0], nk, c->esz);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(dij, ij, sizeof(int) * 2 * npts, hipMemcpyHostToDevice, c->stream));
    if (c->pert)
        HIPCHK(launch_probe<float>((const float *const *)c->dptrs, nk, c->pitch, npts, dij, dout, c->stream));
    else
        HIPCHK(launch_probe<double>((const double *const *)c->dptrs, nk, c->pitch, npts, dij, dout, c->stream));
    HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * npts * nk, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->pert)
        for (int t = 0; t < npts; ++t)
            for (int k = 0; k < nk; ++k) out[t * nk + k] += c->ubh[(size_t)k * c->P.jmax + (ij[2 * t + 1] - 1)];
    return NS_OK;
}

int32_t ns_vorticity(ns_ctx *c, double *omega)
{
    if (!c || !omega) return fail(NS_ERR_PARAM, "null argument");
    const ns_params &P = c->P;
    if (P.nk < 3) return fail(NS_ERR_PARAM, "vorticity needs u and v (nk >= 3)");
    double *d = c->work;  // imax*jmax values
    const size_t n = (size_t)P.imax * P.jmax;
    if (c->pert) {
        const float *q = static_cast<const float *>(c->S[0]);
        HIPCHK(launch_vorticity<float>(q + c->plane, q + 2 * c->plane, pert_rows(c).ub[1], P.imax, P.jmax, c->pitch,
                                       1.0 / P.dm, 1.0 / P.dn, (P.flags & NS_FLAG_PERIODIC_X) != 0, d, c->stream));
    } else {
        const double *U = static_cast<const double *>(c->S[0]);
        HIPCHK(launch_vorticity<double>(U + c->plane, U + 2 * c->plane, nullptr, P.imax, P.jmax, c->pitch,
                                        1.0 / P.dm, 1.0 / P.dn, (P.flags & NS_FLAG_PERIODIC_X) != 0, d, c->stream));
    }
    HIPCHK(hipMemcpyAsync(omega, d, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

int32_t ns_station_energy(ns_ctx *c, int32_t nst, const int32_t *cols, double *ec)
{
    if (!c || nst < 0 || (nst > 0 && (!cols || !ec))) return fail(NS_ERR_PARAM, "bad argument");
    const ns_params &P = c->P;
    if (P.D < 1 || P.jmax - 2 * P.D < 2) return fail(NS_ERR_PARAM, "the energy integral needs 1 <= D < jmax/2");
    if (nst == 0) return NS_OK;
    std::vector<int> c0((size_t)nst);
    for (int t = 0; t < nst; ++t) {
        if (cols[t] < 1 || cols[t] > P.imax) return fail(NS_ERR_PARAM, "station %d at column %d outside 1..%d", t, cols[t], P.imax);
        c0[t] = cols[t] - 1;
    }
    if ((size_t)nst * (sizeof(int) + sizeof(double)) > 16 * c->plane * sizeof(double))
        return fail(NS_ERR_PARAM, "too many stations");
    double *dout = c->work;
    int *dcols = reinterpret_cast<int *>(c->work + nst);
    HIPCHK(hipMemcpyAsync(dcols, c0.data(), sizeof(int) * nst, hipMemcpyHostToDevice, c->stream));
    if (c->pert) {
        const float *q = static_cast<const float *>(c->S[0]);
        const PertRows pr = pert_rows(c);
        HIPCHK(launch_station_energy<float>(q + c->plane, q + 2 * c->plane, pr.ub[1], pr.ub[2], c->pitch, P.jmax, P.D,
                                            P.dn, nst, dcols, dout, c->stream));
    } else {
        const double *U = static_cast<const double *>(c->S[0]);
        HIPCHK(launch_station_energy<double>(U + c->plane, U + 2 * c->plane, nullptr, nullptr, c->pitch, P.jmax, P.D,
                                             P.dn, nst, dcols, dout, c->stream));
    }
    HIPCHK(hipMemcpyAsync(ec, dout, sizeof(double) * nst, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

int32_t ns_row_average(ns_ctx *c, int32_t k, int32_t i0, int32_t i1, double *out)
{
    if (!c || !out) return fail(NS_ERR_PARAM, "null argument");
    const ns_params &P = c->P;
    if (k < 0 || k >= P.nk) return fail(NS_ERR_PARAM, "component %d outside 0..%d", k, P.nk - 1);
    if (i0 < 1 || i1 > P.imax || i1 < i0) return fail(NS_ERR_PARAM, "column range %d..%d outside 1..%d", i0, i1, P.imax);
    double *d = c->work;  // jmax values
    const double *add = c->pert ? pert_rows(c).ub[k] : nullptr;
    if (c->pert)
        HIPCHK(launch_row_mean<float>(static_cast<const float *>(c->S[0]) + (size_t)k * c->plane, c->pitch, P.jmax,
                                      i0 - 1, i1, add, d, c->stream));
    else
        HIPCHK(launch_row_mean<double>(static_cast<const double *>(c->S[0]) + (size_t)k * c->plane, c->pitch,
                                       P.jmax, i0 - 1, i1, add, d, c->stream));
    HIPCHK(hipMemcpyAsync(out, d, sizeof(double) * P.jmax, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

int32_t ns_check_finite(ns_ctx *c)
{
    if (!c) return fail(NS_ERR_PARAM, "null context");
    int rc = set_ptr_table(c, 0, c->S[0], c->P.nk, c->esz);
    if (rc) return rc;
    int zero = 0, flag = 0;
    HIPCHK(hipMemcpyAsync(c->dflag, &zero, sizeof zero, hipMemcpyHostToDevice, c->stream));
    if (c->pert)
        HIPCHK(launch_finite<float>((const float *const *)c->dptrs, c->P.nk, c->P.imax, c->P.jmax, c->pitch, c->dflag,
                                    c->stream));
    else
        HIPCHK(launch_finite<double>((const double *const *)c->dptrs, c->P.nk, c->P.imax, c->P.jmax, c->pitch,
                                     c->dflag, c->stream));
    HIPCHK(hipMemcpyAsync(&flag, c->dflag, sizeof flag, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (flag) return fail(NS_ERR_NONFINITE, "state holds NaN or Inf");
    return NS_OK;
}

int32_t ns_rk_euler(ns_ctx *c, int32_t iter)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (c->stage_variant < 0 && (rc = autotune_stage(c))) return rc;
    rc = enqueue_step(c, iter, false);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

int32_t ns_rk_steps(ns_ctx *c, int32_t iter0, int32_t nsteps)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (nsteps < 0) return fail(NS_ERR_PARAM, "nsteps < 0");
    if (c->stage_variant < 0 && (rc = autotune_stage(c))) return rc;
    if (c->timing || c->P.amplitude != 0.0) {
        // per-step launches: the source amplitude changes with iter
        for (int s = 0; s < nsteps; ++s) {
            rc = enqueue_step(c, iter0 + s, c->timing);
            if (rc) return rc;
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        return NS_OK;
    }
    if (!c->graph) {
        if ((rc = c->pert ? prepare_capture_t<float>(c) : prepare_capture_t<double>(c))) return rc;
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_step(c, iter0, false);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(NS_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) return fail(NS_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(c->graph, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

// ---------------------------------------------------------------------------
// operator-level mirrors
// ---------------------------------------------------------------------------
static int deriv_common(ns_ctx *c, int nk, double *dU, const double *U, bool ydir)
{
    if (!c || !dU || !U) return fail(NS_ERR_PARAM, "null argument");
    if (!c->have_profiles) return fail(NS_ERR_STATE, "ns_set_profiles has not been called");
    if (nk < 1 || nk > 8) return fail(NS_ERR_PARAM, "nk must lie in 1..8");
    double *in = c->work, *outb = c->work + 8 * c->plane;
    int rc = upload_aos(c, in, U, nk);
    if (rc) return rc;
    if ((rc = set_ptr_table(c, 0, in, nk)) || (rc = set_ptr_table(c, 1, outb, nk))) return rc;
    HIPCHK(launch_deriv((const double *const *)c->dptrs, (double *const *)(c->dptrs + 8), nk, c->P.imax, c->P.jmax,
                        c->pitch, ydir, ydir ? c->mety : c->metx, ydir ? 1.0 / c->P.dn : 1.0 / c->P.dm,
                        (c->P.flags & NS_FLAG_PERIODIC_X) != 0, c->stream));
    return download_aos(c, outb, dU, nk);
}

int32_t ns_deropm(ns_ctx *c, int32_t nk, double *dU, const double *U) { return deriv_common(c, nk, dU, U, false); }
int32_t ns_deropn(ns_ctx *c, int32_t nk, double *dU, const double *U) { return deriv_common(c, nk, dU, U, true); }

int32_t ns_transport(ns_ctx *c, const double *T, double *MI, double *K1)
{
    if (!c || !T || !MI || !K1) return fail(NS_ERR_PARAM, "null argument");
    const size_t n = (size_t)c->P.imax * c->P.jmax;
    double *dT = c->work, *dM = c->work + n, *dK = c->work + 2 * n;
    HIPCHK(hipMemcpyAsync(dT, T, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_transport(dT, dM, dK, n, make_transport(c->P), c->stream));
    HIPCHK(hipMemcpyAsync(MI, dM, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(K1, dK, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

int32_t ns_rhs(ns_ctx *c, double *dUdt, const double *U, double ct, int32_t iter)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!dUdt || !U) return fail(NS_ERR_PARAM, "null argument");
    const ns_params &P = c->P;
    double *in = c->work, *outb = c->work + 8 * c->plane;
    if ((rc = upload_aos(c, in, U, P.nk))) return rc;
    double amp = P.amplitude == 0.0 ? 0.0 : P.amplitude * std::sin(P.wf * (((double)iter + ct) * P.dt));
    StageArgs<double> A = stage_args<double>(c, 0, in, outb, amp, false);
    for (int k = 0; k < P.nk; ++k) A.dudt[k] = plane_ptr(outb, c, k);
    HIPCHK(launch_stage(A, P.nk, !(P.flags & NS_FLAG_INVISCID), true, c->stage_variant < 0 ? 0 : c->stage_variant,
                        c->stream, nullptr));
    return download_aos(c, outb, dUdt, P.nk);
}

int32_t ns_rk_steep(ns_ctx *c, double alphat, double betat, double *U, const double *dudt, double *w)
{
    if (!c || !U || !dudt || !w) return fail(NS_ERR_PARAM, "null argument");
    const int nk = c->P.nk;
    double *dU = c->work, *dD = c->work + 5 * c->plane, *dW = c->work + 10 * c->plane;
    int rc;
    if ((rc = upload_aos(c, dU, U, nk)) || (rc = upload_aos(c, dD, dudt, nk)) || (rc = upload_aos(c, dW, w, nk)))
        return rc;
    if ((rc = set_ptr_table(c, 0, dU, nk)) || (rc = set_ptr_table(c, 1, dD, nk)) || (rc = set_ptr_table(c, 2, dW, nk)))
        return rc;
    HIPCHK(launch_rk_steep((double *const *)c->dptrs, (const double *const *)(c->dptrs + 8),
                           (double *const *)(c->dptrs + 16), nk, c->P.imax, c->P.jmax, c->pitch, alphat, betat,
                           c->P.dt, (c->P.flags & NS_FLAG_PERIODIC_X) != 0, c->stream));
    if ((rc = download_aos(c, dU, U, nk))) return rc;
    return download_aos(c, dW, w, nk);
}

int32_t ns_bufferxy(ns_ctx *c, double *U)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!U) return fail(NS_ERR_PARAM, "null argument");
    const int nk = c->P.nk;
    double *dU = c->work;
    if ((rc = upload_aos(c, dU, U, nk))) return rc;
    if ((rc = set_ptr_table(c, 0, dU, nk)) || (rc = set_ptr_table(c, 1, c->base, nk))) return rc;
    HIPCHK(launch_buffer((double *const *)c->dptrs, (const double *const *)(c->dptrs + 8), nk, c->P.imax, c->P.jmax,
                         c->pitch, c->P.D, c->sigx, c->sigy, (c->P.flags & NS_FLAG_PERIODIC_X) != 0, c->stream));
    return download_aos(c, dU, U, nk);
}

static int filter_common(ns_ctx *c, double *U, bool xdir)
{
    if (!c || !U) return fail(NS_ERR_PARAM, "null argument");
    const int nk = c->P.nk;
    double *in = c->work, *outb = c->work + 8 * c->plane;
    int rc;
    if ((rc = upload_aos(c, in, U, nk))) return rc;
    HIPCHK(launch_filter(filter_args<double>(c, xdir, in, outb, false, false), xdir, c->stream));
    return download_aos(c, outb, U, nk);
}

int32_t ns_filter_x(ns_ctx *c, double *U) { return filter_common(c, U, true); }
int32_t ns_filter_y(ns_ctx *c, double *U) { return filter_common(c, U, false); }

int32_t ns_ccpml(ns_ctx *c, double *U)
{
    if (!c || !U) return fail(NS_ERR_PARAM, "null argument");
    if (!c->have_profiles) return fail(NS_ERR_STATE, "ns_set_profiles has not been called");
    const ns_params &P = c->P;
    const int nk = P.nk;
    double *dU = c->work;
    int rc;
    if ((rc = upload_aos(c, dU, U, nk))) return rc;
    // rows j=1, jmax of interior columns come from the y-filter kernel's edge
    // epilogue (run here at sigma = 0, which leaves every line unchanged),
    // then columns i=1, imax with the corners from edge_i_kernel.
    FilterArgs<double> F = filter_args<double>(c, false, dU, c->work + 8 * c->plane, true, false);
    F.sigma = 0.0;
    HIPCHK(launch_filter(F, false, c->stream));
    if (!(P.flags & NS_FLAG_PERIODIC_X))
        HIPCHK(launch_edge_i(edge_args<double>(c, c->work + 8 * c->plane, false), c->stream));
    return download_aos(c, c->work + 8 * c->plane, U, nk);
}

int32_t ns_acousticfont(ns_ctx *c, double ct, int32_t iter, double *s_acou)
{
    if (!c || !s_acou) return fail(NS_ERR_PARAM, "null argument");
    if (!c->have_profiles) return fail(NS_ERR_STATE, "ns_set_profiles has not been called");
    const ns_params &P = c->P;
    const double amp = P.amplitude * std::sin(P.wf * (((double)iter + ct) * P.dt));
    const size_t n = (size_t)P.imax * P.jmax;
    HIPCHK(launch_source(c->work, P.imax, P.jmax, amp, c->gx, c->gy, c->stream));
    HIPCHK(hipMemcpyAsync(s_acou, c->work, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// rhs_euler split the reference's way (derivs.f90, eqeuler.f90,
// coefficients.f90, diff.f90 metricsX/Y) for the drop-in modules of those
// names: host arrays in the reference layout, device planes for the work.
// ---------------------------------------------------------------------------
namespace {

// nplanes pitched FP64 planes for one call, freed with it
struct Planes {
    double *p = nullptr;
    size_t plane = 0;
    ~Planes()
    {
        if (p) (void)hipFree(p);
    }
    int alloc(const ns_ctx *c, int n)
    {
        plane = c->plane;
        if (hipMalloc(&p, (size_t)n * plane * sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return fail(NS_ERR_NOMEM, "hipMalloc of %d planes failed", n);
        }
        return NS_OK;
    }
    double *operator[](int k) const { return p + (size_t)k * plane; }
};

PwArgs pw_args(const ns_ctx *c) { return PwArgs{c->P.imax, c->P.jmax, c->pitch}; }

// derivative of n planes starting at `in` into `out` (x: metricsX factor, y: metricsY)
int deriv_planes(ns_ctx *c, double *in, double *out, int n, bool ydir)
{
    int rc;
    if ((rc = set_ptr_table(c, 0, in, n)) || (rc = set_ptr_table(c, 1, out, n))) return rc;
    HIPCHK(launch_deriv((const double *const *)c->dptrs, (double *const *)(c->dptrs + 8), n, c->P.imax, c->P.jmax,
                        c->pitch, ydir, ydir ? c->mety : c->metx, ydir ? 1.0 / c->P.dn : 1.0 / c->P.dm,
                        (c->P.flags & NS_FLAG_PERIODIC_X) != 0, c->stream));
    return NS_OK;
}

int check_nk(ns_ctx *c, int nk)
{
    if (!c) return fail(NS_ERR_PARAM, "null context");
    if (!c->have_profiles) return fail(NS_ERR_STATE, "ns_set_profiles has not been called");
    if (nk < 4 || nk > NKMAX) return fail(NS_ERR_PARAM, "nk must be 4 or 5 (got %d)", nk);
    return NS_OK;
}

}  // namespace

extern "C" {

int32_t ns_deriv(ns_ctx *c, int32_t nk, const double *U, const double *T, const double *K1, double *dUdm,
                 double *dUdn, double *dTdm, double *dTdn, double *d2Tdm, double *d2Tdn, double *psiDiff,
                 double *d2psidm, double *d2psidn)
{
    int rc = check_nk(c, nk);
    if (rc) return rc;
    if (!U || !T || !K1 || !dUdm || !dUdn || !dTdm || !dTdn || !d2Tdm || !d2Tdn)
        return fail(NS_ERR_PARAM, "null argument");
    if (nk >= 5 && (!psiDiff || !d2psidm || !d2psidn)) return fail(NS_ERR_PARAM, "null species argument (nk = 5)");
    // planes: U, dU/dm, dU/dn (nk each), T, K1, dT/dm, dT/dn, d2T/dm, d2T/dn,
    // psiDiff, d2psi/dm, d2psi/dn
    Planes w;
    if ((rc = w.alloc(c, 3 * nk + 9))) return rc;
    double *u = w[0], *um = w[nk], *un = w[2 * nk];
    double *t = w[3 * nk], *k = w[3 * nk + 1], *tm = w[3 * nk + 2], *tn = w[3 * nk + 3];
    double *t2m = w[3 * nk + 4], *t2n = w[3 * nk + 5], *pd = w[3 * nk + 6], *p2m = w[3 * nk + 7], *p2n = w[3 * nk + 8];
    if ((rc = upload_aos(c, u, U, nk)) || (rc = upload_aos(c, t, T, 1)) || (rc = upload_aos(c, k, K1, 1))) return rc;
    // derivs.f90:17-20, :25-26, kdT :28/:48-56, :32-33
    if ((rc = deriv_planes(c, u, um, nk, false)) || (rc = deriv_planes(c, u, un, nk, true)) ||
        (rc = deriv_planes(c, t, tm, 1, false)) || (rc = deriv_planes(c, t, tn, 1, true)))
        return rc;
    HIPCHK(launch_kdt(pw_args(c), k, tm, tn, c->stream));
    if ((rc = deriv_planes(c, tm, t2m, 1, false)) || (rc = deriv_planes(c, tn, t2n, 1, true))) return rc;
    if (nk >= 5) {  // derivs.f90:36-43
        HIPCHK(launch_psidiff(pw_args(c), u, um + 4 * c->plane, un + 4 * c->plane, c->P.alpha_cond, pd, c->stream));
        if ((rc = deriv_planes(c, pd, p2m, 1, false)) || (rc = deriv_planes(c, pd, p2n, 1, true))) return rc;
    }
    if ((rc = download_aos(c, um, dUdm, nk)) || (rc = download_aos(c, un, dUdn, nk)) ||
        (rc = download_aos(c, tm, dTdm, 1)) || (rc = download_aos(c, tn, dTdn, 1)) ||
        (rc = download_aos(c, t2m, d2Tdm, 1)) || (rc = download_aos(c, t2n, d2Tdn, 1)))
        return rc;
    if (nk >= 5 && ((rc = download_aos(c, pd, psiDiff, 1)) || (rc = download_aos(c, p2m, d2psidm, 1)) ||
                    (rc = download_aos(c, p2n, d2psidn, 1))))
        return rc;
    return NS_OK;
}

int32_t ns_kdt(ns_ctx *c, const double *K1, double *dTdm, double *dTdn)
{
    if (!c || !K1 || !dTdm || !dTdn) return fail(NS_ERR_PARAM, "null argument");
    Planes w;
    int rc;
    if ((rc = w.alloc(c, 3))) return rc;
    if ((rc = upload_aos(c, w[0], K1, 1)) || (rc = upload_aos(c, w[1], dTdm, 1)) || (rc = upload_aos(c, w[2], dTdn, 1)))
        return rc;
    HIPCHK(launch_kdt(pw_args(c), w[0], w[1], w[2], c->stream));
    if ((rc = download_aos(c, w[1], dTdm, 1))) return rc;
    return download_aos(c, w[2], dTdn, 1);
}

int32_t ns_tau_ij(ns_ctx *c, int32_t nk, const double *MI, const double *dUdm, const double *dUdn, double *tau,
                  double *dtaudm, double *dtaudn)
{
    int rc = check_nk(c, nk);
    if (rc) return rc;
    if (!MI || !dUdm || !dUdn || !tau || !dtaudm || !dtaudn) return fail(NS_ERR_PARAM, "null argument");
    Planes w;  // MI, dU/dm (nk), dU/dn (nk), tau (3), dtau/dm (3), dtau/dn (3)
    if ((rc = w.alloc(c, 2 * nk + 10))) return rc;
    double *mi = w[0], *um = w[1], *un = w[1 + nk], *tu = w[1 + 2 * nk], *tm = w[4 + 2 * nk], *tn = w[7 + 2 * nk];
    if ((rc = upload_aos(c, mi, MI, 1)) || (rc = upload_aos(c, um, dUdm, nk)) || (rc = upload_aos(c, un, dUdn, nk)))
        return rc;
    const size_t P = c->plane;
    HIPCHK(launch_tau(pw_args(c), mi, um + P, un + P, um + 2 * P, un + 2 * P, tu, tu + P, tu + 2 * P, c->stream));
    // eqeuler.f90:43-45
    if ((rc = deriv_planes(c, tu, tm, 3, false)) || (rc = deriv_planes(c, tu, tn, 3, true))) return rc;
    if ((rc = download_aos(c, tu, tau, 3)) || (rc = download_aos(c, tm, dtaudm, 3))) return rc;
    return download_aos(c, tn, dtaudn, 3);
}

int32_t ns_eulernonlinear(ns_ctx *c, int32_t nk, double *dUdt, const double *U, const double *dUdm,
                          const double *dUdn, const double *tau, const double *dtaudm, const double *dtaudn,
                          const double *d2Tdm, const double *d2Tdn, const double *d2psidm, const double *d2psidn,
                          const double *s_acou)
{
    int rc = check_nk(c, nk);
    if (rc) return rc;
    if (!dUdt || !U || !dUdm || !dUdn || !tau || !dtaudm || !dtaudn || !d2Tdm || !d2Tdn)
        return fail(NS_ERR_PARAM, "null argument");
    if (nk >= 5 && (!d2psidm || !d2psidn)) return fail(NS_ERR_PARAM, "null species argument (nk = 5)");
    // U, dU/dm, dU/dn, out (nk each), tau, dtau/dm, dtau/dn (3 each), d2T/dm,
    // d2T/dn, d2psi/dm, d2psi/dn, s_acou
    Planes w;
    if ((rc = w.alloc(c, 4 * nk + 14))) return rc;
    const size_t P = c->plane;
    double *u = w[0], *um = w[nk], *un = w[2 * nk], *out = w[3 * nk], *tu = w[4 * nk], *tm = w[4 * nk + 3];
    double *tn = w[4 * nk + 6], *t2m = w[4 * nk + 9], *t2n = w[4 * nk + 10], *p2m = w[4 * nk + 11];
    double *p2n = w[4 * nk + 12], *src = w[4 * nk + 13];
    if ((rc = upload_aos(c, u, U, nk)) || (rc = upload_aos(c, um, dUdm, nk)) || (rc = upload_aos(c, un, dUdn, nk)) ||
        (rc = upload_aos(c, tu, tau, 3)) || (rc = upload_aos(c, tm, dtaudm, 3)) || (rc = upload_aos(c, tn, dtaudn, 3)) ||
        (rc = upload_aos(c, t2m, d2Tdm, 1)) || (rc = upload_aos(c, t2n, d2Tdn, 1)))
        return rc;
    if (nk >= 5 && ((rc = upload_aos(c, p2m, d2psidm, 1)) || (rc = upload_aos(c, p2n, d2psidn, 1)))) return rc;
    if (s_acou)  // s_acou(imax, jmax): unpitched
        HIPCHK(hipMemcpyAsync(src, s_acou, sizeof(double) * c->P.imax * c->P.jmax, hipMemcpyHostToDevice, c->stream));
    EulerPwArgs e{};
    e.a = pw_args(c);
    e.nk = nk;
    for (int k = 0; k < nk; ++k) {
        e.U[k] = u + k * P;
        e.Um[k] = um + k * P;
        e.Un[k] = un + k * P;
        e.out[k] = out + k * P;
    }
    for (int k = 0; k < 3; ++k) {
        e.tau[k] = tu + k * P;
        e.taum[k] = tm + k * P;
        e.taun[k] = tn + k * P;
    }
    e.d2Tm = t2m;
    e.d2Tn = t2n;
    e.d2pm = nk >= 5 ? p2m : nullptr;
    e.d2pn = nk >= 5 ? p2n : nullptr;
    e.src = s_acou ? src : nullptr;
    e.gamma = c->P.gamma;
    e.Re = c->P.Re;
    e.Pr = c->P.Pr;
    e.Pe = c->P.Pe;
    HIPCHK(launch_euler_pw(e, c->stream));
    return download_aos(c, out, dUdt, nk);
}

static int metrics_common(ns_ctx *c, int32_t nk, double *dU, bool ydir)
{
    if (!c || !dU) return fail(NS_ERR_PARAM, "null argument");
    if (!c->have_profiles) return fail(NS_ERR_STATE, "ns_set_profiles has not been called");
    if (nk < 1 || nk > 8) return fail(NS_ERR_PARAM, "nk must lie in 1..8");
    Planes w;
    int rc;
    if ((rc = w.alloc(c, nk)) || (rc = upload_aos(c, w[0], dU, nk)) || (rc = set_ptr_table(c, 0, w[0], nk))) return rc;
    HIPCHK(launch_metric(pw_args(c), (double *const *)c->dptrs, nk, ydir ? c->mety : c->metx, ydir, c->stream));
    return download_aos(c, w[0], dU, nk);
}

int32_t ns_metrics_x(ns_ctx *c, int32_t nk, double *dU) { return metrics_common(c, nk, dU, false); }
int32_t ns_metrics_y(ns_ctx *c, int32_t nk, double *dU) { return metrics_common(c, nk, dU, true); }

int32_t ns_autotune(ns_ctx *c, int32_t *variant, double *ms)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (c->graph) {
        HIPCHK(hipGraphExecDestroy(c->graph));
        c->graph = nullptr;
    }
    if ((rc = autotune_stage(c))) return rc;
    if (variant) *variant = c->stage_variant;
    if (ms)
        for (int v = 0; v < 3; ++v) ms[v] = c->tune_ms[v];
    return NS_OK;
}

int32_t ns_set_variant(ns_ctx *c, int32_t variant)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (variant < 0 || variant > 2) return fail(NS_ERR_PARAM, "stage variant must be 0, 1 or 2");
    if (c->graph) {
        HIPCHK(hipGraphExecDestroy(c->graph));
        c->graph = nullptr;
    }
    c->stage_variant = variant;
    // set up its kernels now, outside any capture the caller may start
    return c->pert ? prepare_capture_t<float>(c) : prepare_capture_t<double>(c);
}

int32_t ns_set_filter_mode(ns_ctx *c, int32_t mode)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (mode < 0 || mode > 1) return fail(NS_ERR_PARAM, "filter mode must be 0 or 1");
    if (c->graph) {
        HIPCHK(hipGraphExecDestroy(c->graph));
        c->graph = nullptr;
    }
    c->filt_mode = mode;
    c->filt_pinned = true;
    return c->pert ? prepare_capture_t<float>(c) : prepare_capture_t<double>(c);
}

int32_t ns_get_filter_mode(ns_ctx *c, int32_t *mode, double *ms)
{
    if (!c || !mode) return fail(NS_ERR_PARAM, "null argument");
    *mode = c->filt_mode;
    if (ms) {
        ms[0] = c->filt_ms;
        ms[1] = c->filt_xy_ms;
    }
    return NS_OK;
}

int32_t ns_enable_timing(ns_ctx *c, int32_t on)
{
    if (!c) return fail(NS_ERR_PARAM, "null context");
    c->timing = on != 0;
    for (int t = 0; t < TC_N; ++t) {
        c->tsum[t] = 0.0;
        c->tcount[t] = 0;
    }
    return NS_OK;
}

int32_t ns_stage_interior(ns_ctx *c, int64_t *points)
{
    if (!c || !points) return fail(NS_ERR_PARAM, "null argument");
    *points = c->interior_points;
    return NS_OK;
}

int32_t ns_kernel_time(ns_ctx *c, const char *name, double *avg_ms, int32_t *count)
{
    if (!c || !name || !avg_ms || !count) return fail(NS_ERR_PARAM, "null argument");
    for (int t = 0; t < TC_N; ++t)
        if (std::strcmp(name, TC_NAMES[t]) == 0) {
            *count = (int32_t)c->tcount[t];
            *avg_ms = c->tcount[t] ? c->tsum[t] / c->tcount[t] : 0.0;
            return NS_OK;
        }
    return fail(NS_ERR_PARAM, "unknown kernel class '%s'", name);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched cases: one context per case holds the storage and the set-up; the
// batch owns per-(stage, case) argument blocks and tile tables in device
// memory so that each kernel class is a single launch over all cases.
// ---------------------------------------------------------------------------
struct ns_batch {
    std::vector<ns_ctx *> cs;
    int n = 0, nk = 0;
    bool pert = false, visc = true;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // general (edge) stage tiles, concurrent with the fast ones
    hipEvent_t fork = nullptr, join = nullptr;
    void *sargs = nullptr, *fxargs = nullptr, *fyargs = nullptr, *eargs = nullptr;
    int4 *stiles = nullptr, *gtiles = nullptr, *fxtiles = nullptr, *fytiles = nullptr;
    int nst = 0, ngt = 0, nfx = 0, nfy = 0;
    bool prepared = false;
    int variant = -1;
    double tune_ms[3] = {0, 0, 0};
    hipGraphExec_t graph = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // marching variants (1: each case's fast zone marked, 2: plain general
    // regions): per-case plans and the work-item table of the layout and
    // chunk height built last, and the chunk height chosen per variant
    BatchMarch *mplans = nullptr;
    int4 *mitems = nullptr;
    int nmi = 0, mlayout = 0, mH = -1;
    bool munit = true;  // every case on unit spacing (else the general code with 1/h factors)
    int march_H[3] = {0, 0, 0};
};

namespace {

void batch_release(ns_batch *b)
{
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (void *p : {b->sargs, b->fxargs, b->fyargs, b->eargs, (void *)b->stiles, (void *)b->gtiles, (void *)b->fxtiles,
                    (void *)b->fytiles, (void *)b->mplans, (void *)b->mitems})
        if (p) (void)hipFree(p);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->side) (void)hipStreamDestroy(b->side);
    for (ns_ctx *c : b->cs)
        if (c) {
            release(c);
            delete c;
        }
}

template <typename T>
int upload_vec(hipStream_t s, void **dst, const std::vector<T> &v)
{
    HIPCHK(hipMalloc(dst, v.size() * sizeof(T)));
    HIPCHK(hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return NS_OK;
}

template <typename S>
int batch_prepare_t(ns_batch *b)
{
    const int n = b->n;
    std::vector<StageArgs<S>> sa((size_t)6 * n);
    std::vector<FilterArgs<S>> fx(n), fy(n);
    std::vector<EdgeArgs<S>> ea(n);
    std::vector<int4> st, gt, tx, ty;
    const int TY = (sizeof(S) == 8 && b->nk >= 5 && b->visc) ? 8 : 16;
    for (int c = 0; c < n; ++c) {
        ns_ctx *x = b->cs[c];
        S *S0 = static_cast<S *>(x->S[0]), *S1 = static_cast<S *>(x->S[1]), *S2 = static_cast<S *>(x->S[2]);
        for (int irk = 0; irk < 6; ++irk) sa[(size_t)irk * n + c] = stage_args<S>(x, irk, S0, S1, 0.0, b->pert);
        fx[c] = filter_args<S>(x, true, S1, S2, false, b->pert);
        fy[c] = filter_args<S>(x, false, S2, S0, true, b->pert);
        ea[c] = edge_args<S>(x, S0, b->pert);
        const int imax = x->P.imax, jmax = x->P.jmax, L = x->filt_L, Lx = x->filt_Lx;
        int fx0, fx1, fy0, fy1;
        const StageArgs<S> &A0 = sa[c];
        fast_range(A0.fast_ok, A0.D, imax, 64, fx0, fx1);
        fast_range(A0.fast_ok, A0.D, jmax, TY, fy0, fy1);
        for (int by = 0; by < (jmax + TY - 1) / TY; ++by)
            for (int bx = 0; bx < (imax + 63) / 64; ++bx) {
                const bool fast = bx >= fx0 && bx < fx1 && by >= fy0 && by < fy1;
                (fast ? st : gt).push_back(make_int4(c, bx, by, 0));
            }
        for (int rb = 0; rb < (jmax + 63) / 64; ++rb)
            for (int ch = 0; ch < (imax + Lx - 1) / Lx; ++ch) tx.push_back(make_int4(c, ch, rb, 0));
        for (int ch = 0; ch < (jmax + L - 1) / L; ++ch)
            for (int cb = 0; cb < (imax + 63) / 64; ++cb) ty.push_back(make_int4(c, cb, ch, 0));
    }
    int rc;
    if ((rc = upload_vec(b->stream, &b->sargs, sa)) || (rc = upload_vec(b->stream, &b->fxargs, fx)) ||
        (rc = upload_vec(b->stream, &b->fyargs, fy)) || (rc = upload_vec(b->stream, &b->eargs, ea)) ||
        (rc = upload_vec(b->stream, (void **)&b->stiles, st)) || (rc = upload_vec(b->stream, (void **)&b->gtiles, gt)) ||
        (rc = upload_vec(b->stream, (void **)&b->fxtiles, tx)) ||
        (rc = upload_vec(b->stream, (void **)&b->fytiles, ty)))
        return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    b->nst = (int)st.size();
    b->ngt = (int)gt.size();
    b->nfx = (int)tx.size();
    b->nfy = (int)ty.size();
    b->prepared = true;
    return NS_OK;
}

// The marching variants' table: each case's plan (batch_march_case), then
// work items -- whole cases to the XCD with the least work so far, largest
// first; per XCD each case's marching blocks (chunks, strips minor), then
// its edge tiles -- interleaved so that block L lands on XCD L % 8, padded
// to equal lengths.  NS_ERR_STATE when a case cannot march (FP32 plane too
// large for 32-bit offsets).
constexpr int BATCH_TILE_NT = 1;  // tile variant: 512-thread tiles (launch_stage_batch block size 1)

template <typename S>
int batch_march_build_t(ns_batch *b, int layout, int H)
{
    if (b->mitems && b->mlayout == layout && b->mH == H) return NS_OK;
    HIPCHK(hipStreamSynchronize(b->stream));  // the old table may be in use
    const int n = b->n;
    std::vector<StageArgs<S>> args(n);
    bool unit = true;  // the fast zones need the PM_HYBRID kernel, which needs unit spacing in every case
    for (int c = 0; c < n; ++c) {
        ns_ctx *x = b->cs[c];
        args[c] = stage_args<S>(x, 0, static_cast<S *>(x->S[0]), static_cast<S *>(x->S[1]), 0.0, b->pert);
        unit = unit && args[c].inv_dm == 1.0 && args[c].inv_dn == 1.0;
    }
    std::vector<BatchMarch> plans(n);
    for (int c = 0; c < n; ++c)
        if (!batch_march_case(args[c], b->nk, b->visc, layout == 1 && unit, H, plans[c]))
            return fail(NS_ERR_STATE, "case %d: the grid is too large for the batched marching kernel", c);
    std::vector<int> order(n);
    for (int c = 0; c < n; ++c) order[c] = c;
    auto pts = [&](int c) { return (long long)b->cs[c]->P.imax * b->cs[c]->P.jmax; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return pts(a) > pts(c); });
    std::vector<std::vector<int4>> mx(NXCD), ex(NXCD);
    long long load[NXCD] = {0};
    for (int c : order) {
        int x = 0;
        for (int q = 1; q < NXCD; ++q)
            if (load[q] < load[x]) x = q;
        load[x] += pts(c);
        const BatchMarch &p = plans[c];
        for (int ch = 0; ch < p.nchunks; ++ch)
            for (int st = 0; st < p.m.nstrips; ++st) mx[x].push_back(make_int4(c, st, ch, 0));
        for (int t = 0; t < p.sides.count() + p.bands.count(); ++t) ex[x].push_back(make_int4(c, t, 0, 1));
    }
    std::vector<int4> items;
    for (auto *lists : {&mx, &ex}) {
        size_t len = 0;
        for (auto &l : *lists) len = std::max(len, l.size());
        for (size_t k = 0; k < len; ++k)
            for (int x = 0; x < NXCD; ++x) items.push_back(k < (*lists)[x].size() ? (*lists)[x][k] : make_int4(0, 0, 0, 2));
    }
    if (b->mplans) (void)hipFree(b->mplans);
    if (b->mitems) (void)hipFree(b->mitems);
    b->mplans = nullptr;
    b->mitems = nullptr;
    b->mH = -1;
    int rc;
    if ((rc = upload_vec(b->stream, (void **)&b->mplans, plans)) ||
        (rc = upload_vec(b->stream, (void **)&b->mitems, items)))
        return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    b->nmi = (int)items.size();
    b->mlayout = layout;
    b->mH = H;
    b->munit = unit;
    return NS_OK;
}

template <typename S>
int batch_enqueue_step_t(ns_batch *b)
{
    for (int irk = 0; irk < 6; ++irk) {
        if (b->variant >= 1) {
            HIPCHK(launch_stage_march_batch(static_cast<const StageArgs<S> *>(b->sargs), b->mplans, b->mitems, b->nmi,
                                            irk, b->n, b->nk, b->visc, b->munit, b->stream));
        } else {
            HIPCHK(hipEventRecord(b->fork, b->stream));
            HIPCHK(hipStreamWaitEvent(b->side, b->fork, 0));
            HIPCHK(launch_stage_batch(static_cast<const StageArgs<S> *>(b->sargs), b->stiles, b->nst, b->gtiles,
                                      b->ngt, irk, b->n, b->nk, b->visc, BATCH_TILE_NT, b->stream, b->side));
            HIPCHK(hipEventRecord(b->join, b->side));
            HIPCHK(hipStreamWaitEvent(b->stream, b->join, 0));
        }
        HIPCHK(launch_filter_batch(static_cast<const FilterArgs<S> *>(b->fxargs), b->fxtiles, b->nfx, b->nk, true,
                                   b->stream));
        HIPCHK(launch_filter_batch(static_cast<const FilterArgs<S> *>(b->fyargs), b->fytiles, b->nfy, b->nk, false,
                                   b->stream));
        HIPCHK(launch_edge_batch(static_cast<const EdgeArgs<S> *>(b->eargs), b->n, b->stream));
    }
    return NS_OK;
}

int batch_enqueue_step(ns_batch *b)
{
    return b->pert ? batch_enqueue_step_t<float>(b) : batch_enqueue_step_t<double>(b);
}

// Time the stage launch's variants on stage 2 of the batch (W dead between
// steps, S1 scratch): 0 = tiles (fast tiles and general tiles on two
// streams), 1 and 2 = the marching kernel with each case's fast zone marked
// or on plain general regions, each at the chunk height (one chunk, 128 or
// 64 rows) that times fastest; keeps the fastest variant.
template <typename S>
int batch_autotune_t(ns_batch *b)
{
    const StageArgs<S> *sa = static_cast<const StageArgs<S> *>(b->sargs);
    // ms per launch over 3 launches after one untimed launch; -1 when the launch fails
    auto time3 = [&](auto launch) -> double {
        if (launch() != hipSuccess) {
            (void)hipGetLastError();
            return -1.0;
        }
        if (hipEventRecord(b->ev0, b->stream) != hipSuccess) return -1.0;
        for (int r = 0; r < 3; ++r)
            if (launch() != hipSuccess) return -1.0;
        float ms = 0.f;
        if (hipEventRecord(b->ev1, b->stream) != hipSuccess || hipEventSynchronize(b->ev1) != hipSuccess ||
            hipEventElapsedTime(&ms, b->ev0, b->ev1) != hipSuccess)
            return -1.0;
        return ms / 3;
    };
    b->tune_ms[0] = time3([&] {
        return launch_stage_batch(sa, b->stiles, b->nst, b->gtiles, b->ngt, 1, b->n, b->nk, b->visc, BATCH_TILE_NT,
                                  b->stream, nullptr);
    });
    for (int v = 1; v <= 2; ++v) {
        b->tune_ms[v] = -1.0;
        for (int H : {0, 128, 64}) {
            if (batch_march_build_t<S>(b, v, H) != NS_OK) break;
            const double t = time3([&] {
                return launch_stage_march_batch(sa, b->mplans, b->mitems, b->nmi, 1, b->n, b->nk, b->visc, b->munit,
                                                b->stream);
            });
            if (t > 0 && (b->tune_ms[v] < 0 || t < b->tune_ms[v])) {
                b->tune_ms[v] = t;
                b->march_H[v] = H;
            }
        }
    }
    int best = -1;
    for (int v = 0; v < 3; ++v)
        if (b->tune_ms[v] > 0 && (best < 0 || b->tune_ms[v] < b->tune_ms[best])) best = v;
    if (best < 0) return fail(NS_ERR_HIP, "batched stage: no variant could be launched");
    b->variant = best;
    if (best >= 1) return batch_march_build_t<S>(b, best, b->march_H[best]);
    return NS_OK;
}

int batch_ready(ns_batch *b)
{
    if (!b) return fail(NS_ERR_PARAM, "null batch");
    for (int c = 0; c < b->n; ++c)
        if (!b->cs[c]->have_profiles || !b->cs[c]->have_base || !b->cs[c]->have_state)
            return fail(NS_ERR_STATE, "case %d has not been set (ns_batch_set_case)", c);
    int rc;
    if (!b->prepared && (rc = b->pert ? batch_prepare_t<float>(b) : batch_prepare_t<double>(b))) return rc;
    if (b->variant < 0 && (rc = b->pert ? batch_autotune_t<float>(b) : batch_autotune_t<double>(b))) return rc;
    return NS_OK;
}

}  // namespace

extern "C" {

int32_t ns_batch_create(const ns_params *params, int32_t ncases, ns_batch **out)
{
    if (!params || !out || ncases < 1) return fail(NS_ERR_PARAM, "bad argument");
    *out = nullptr;
    for (int c = 0; c < ncases; ++c) {
        const ns_params &P = params[c];
        if (P.nk != params[0].nk || P.precision != params[0].precision || P.flags != params[0].flags ||
            P.device != params[0].device)
            return fail(NS_ERR_PARAM, "case %d: all cases of a batch share nk, precision, flags and device", c);
        if (P.amplitude != 0.0) return fail(NS_ERR_PARAM, "case %d: the batched path runs without acoustic source", c);
    }
    ns_batch *b = new ns_batch();
    b->n = ncases;
    b->nk = params[0].nk;
    b->pert = params[0].precision == NS_PREC_FP32_PERT;
    b->visc = !(params[0].flags & NS_FLAG_INVISCID);
    b->cs.assign(ncases, nullptr);
    for (int c = 0; c < ncases; ++c) {
        int rc = ns_create(&params[c], &b->cs[c]);
        if (rc) {
            std::string msg = g_err;
            batch_release(b);
            delete b;
            return fail(rc, "case %d: %s", c, msg.c_str());
        }
    }
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = touch_streams(b->stream, b->side, b->fork, b->join);
    if (e != hipSuccess) {
        batch_release(b);
        delete b;
        return fail(NS_ERR_HIP, "batch stream/events: %s", hipGetErrorString(e));
    }
    *out = b;
    return NS_OK;
}

int32_t ns_batch_set_case(ns_batch *b, int32_t c, const double *meshx, const double *meshy, const double *sigmax,
                          const double *sigmay, const double *m, const double *n, const double *rhob,
                          const double *psib, const double *Ub, const double *U0)
{
    if (!b || c < 0 || c >= b->n) return fail(NS_ERR_PARAM, "bad batch or case index %d", c);
    ns_ctx *x = b->cs[c];
    int rc;
    if ((rc = ns_set_profiles(x, meshx, meshy, sigmax, sigmay, m, n, rhob, psib))) return rc;
    if ((rc = ns_set_base(x, Ub))) return rc;
    if ((rc = ns_put_state(x, U0))) return rc;
    return NS_OK;
}

int32_t ns_batch_steps(ns_batch *b, int32_t iter0, int32_t nsteps)
{
    int rc = batch_ready(b);
    if (rc) return rc;
    (void)iter0;  // no source term in the batched path: steps do not depend on iter
    if (nsteps < 0) return fail(NS_ERR_PARAM, "nsteps < 0");
    if (!b->graph) {
        // the batched filter and edge kernels load before the capture (the
        // stage kernels have run in the autotuner)
        HIPCHK(b->pert ? prepare_step(StageArgs<float>{}, b->nk, b->visc, 0, true)
                       : prepare_step(StageArgs<double>{}, b->nk, b->visc, 0, true));
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        rc = batch_enqueue_step(b);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(NS_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(NS_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    for (int s = 0; s < nsteps; ++s) HIPCHK(hipGraphLaunch(b->graph, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return NS_OK;
}

int32_t ns_batch_get_state(ns_batch *b, int32_t c, double *U)
{
    if (!b || c < 0 || c >= b->n || !U) return fail(NS_ERR_PARAM, "bad batch, case index or pointer");
    HIPCHK(hipStreamSynchronize(b->stream));
    return ns_get_state(b->cs[c], U);
}

int32_t ns_batch_autotune(ns_batch *b, int32_t *variant, double *ms)
{
    if (!b) return fail(NS_ERR_PARAM, "null batch");
    int rc = batch_ready(b);
    if (rc) return rc;
    if (b->graph) {
        HIPCHK(hipGraphExecDestroy(b->graph));
        b->graph = nullptr;
    }
    if ((rc = b->pert ? batch_autotune_t<float>(b) : batch_autotune_t<double>(b))) return rc;
    if (variant) *variant = b->variant;
    if (ms)
        for (int v = 0; v < 3; ++v) ms[v] = b->tune_ms[v];
    return NS_OK;
}

int32_t ns_batch_set_variant(ns_batch *b, int32_t variant)
{
    if (!b) return fail(NS_ERR_PARAM, "null batch");
    if (variant < 0 || variant > 2) return fail(NS_ERR_PARAM, "batch stage variant %d (0..2)", variant);
    int rc = batch_ready(b);  // the first call times every variant, which sets up their kernels
    if (rc) return rc;
    if (b->tune_ms[variant] < 0)
        return fail(NS_ERR_STATE, "batch stage variant %d cannot run these cases", variant);
    if (b->graph) {
        HIPCHK(hipGraphExecDestroy(b->graph));
        b->graph = nullptr;
    }
    if (variant >= 1 && (rc = b->pert ? batch_march_build_t<float>(b, variant, b->march_H[variant])
                                      : batch_march_build_t<double>(b, variant, b->march_H[variant])))
        return rc;
    b->variant = variant;
    return NS_OK;
}

int32_t ns_batch_destroy(ns_batch *b)
{
    if (!b) return NS_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_release(b);
    delete b;
    return NS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Reverse mode: discrete adjoint of the LDDRK6 step (2-D Euler, nk = 4).
// No reference counterpart: the reference has no adjoint.  The forward keeps
// checkpoints every K steps; the backward recomputes each segment's step
// states from its checkpoint, each step's six stage inputs from its step
// state, and runs the backward kernels stage by stage in reverse.
// ---------------------------------------------------------------------------
namespace {

size_t state_bytes(const ns_ctx *c) { return (size_t)c->P.nk * c->plane * c->esz; }

template <typename S>
S *state_slot(void *base, const ns_ctx *c, int n)
{
    return static_cast<S *>(base) + (size_t)n * c->P.nk * c->plane;
}

int grow(void **p, size_t &have, size_t need)
{
    if (need <= have && *p) return NS_OK;
    if (*p) HIPCHK(hipFree(*p));
    *p = nullptr;
    have = 0;
    hipError_t e = hipMalloc(p, need);
    if (e != hipSuccess) return fail(NS_ERR_NOMEM, "hipMalloc(%zu): %s", need, hipGetErrorString(e));
    have = need;
    return NS_OK;
}

int adj_check(ns_ctx *c)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!c->have_state) return fail(NS_ERR_STATE, "no state has been set");
    if (c->P.nk != 4 || !(c->P.flags & NS_FLAG_INVISCID))
        return fail(NS_ERR_PARAM, "the adjoint covers the 2-D Euler equations (nk = 4 with NS_FLAG_INVISCID)");
    if (c->P.flags & NS_FLAG_PERIODIC_X) return fail(NS_ERR_PARAM, "the adjoint does not cover periodic x");
    return NS_OK;
}

int adj_work(ns_ctx *c)
{
    const size_t b = state_bytes(c);
    size_t have;
    int rc;
    for (void **pp : {&c->adj.lam, &c->adj.lamW, &c->adj.gx, &c->adj.gy}) {
        have = *pp ? b : 0;
        if ((rc = grow(pp, have, b))) return rc;
    }
    have = c->adj.stg ? 6 * b : 0;
    return grow(&c->adj.stg, have, 6 * b);
}

hipError_t copy_state(ns_ctx *c, void *dst, const void *src)
{
    return hipMemcpyAsync(dst, src, state_bytes(c), hipMemcpyDeviceToDevice, c->stream);
}

template <typename S>
ArcArgs<S> arc_args(ns_ctx *c, const S *state)
{
    ArcArgs<S> a{};
    a.pts = c->adj.arc;
    a.npts = c->adj.narc;
    a.U = state + 3 * c->plane;
    a.base = c->pert ? nullptr : plane_ptr(c->base, c, 3);
    a.pitch = c->pitch;
    return a;
}

template <typename S>
void set_planes(S **dst, S *buf, const ns_ctx *c)
{
    for (int k = 0; k < NKMAX; ++k) dst[k] = k < c->P.nk ? buf + (size_t)k * c->plane : nullptr;
}

// ccpml^T, Fy^T, Fx^T on lam (result in lam; tmp: a state-sized scratch)
template <typename S>
int adj_linear_back(ns_ctx *c, S *lam, S *tmp, bool edge, bool fy, bool fx)
{
    const ns_params &P = c->P;
    if (edge) {
        AdjEdgeArgs<S> E{};
        set_planes(E.lam, lam, c);
        E.imax = P.imax;
        E.jmax = P.jmax;
        E.pitch = c->pitch;
        E.nk = P.nk;
        HIPCHK(launch_adj_edge(E, c->stream));
    }
    AdjFilterArgs<S> F{};
    F.pitch = c->pitch;
    F.nk = P.nk;
    F.sigma = P.sigmaf;
    F.L = 64;  // short chunks: the transposed sweeps are latency-bound
    F.W = c->filt_W;
    auto run = [&](S *in, S *out, bool ydir) -> int {
        for (int k = 0; k < NKMAX; ++k) {
            F.in[k] = k < P.nk ? in + (size_t)k * c->plane : nullptr;
            F.out[k] = k < P.nk ? out + (size_t)k * c->plane : nullptr;
        }
        F.n = ydir ? P.jmax : P.imax;
        F.m = ydir ? P.imax : P.jmax;
        HIPCHK(launch_adj_filter(F, ydir, c->stream));
        return NS_OK;
    };
    int rc;
    if (fy && fx) {
        if ((rc = run(lam, tmp, true)) || (rc = run(tmp, lam, false))) return rc;
    } else if (fy || fx) {
        if ((rc = run(lam, tmp, fy))) return rc;
        HIPCHK(hipMemcpyAsync(lam, tmp, (size_t)P.nk * c->plane * sizeof(S), hipMemcpyDeviceToDevice, c->stream));
    }
    return NS_OK;
}

// (lam, lamW) <- backward of stage irk with input state Us; gx, gy: work planes
template <typename S>
int adj_stage_back(ns_ctx *c, int irk, S *Us, S *lam, S *lamW, S *gx, S *gy, bool pert, bool rhs_only)
{
    const ns_params &P = c->P;
    if (!rhs_only) {
        int rc = adj_linear_back<S>(c, lam, gx, true, true, true);  // gx is free until the point kernel
        if (rc) return rc;
    }
    AdjPointArgs<S> A{};
    A.st = stage_args<S>(c, irk, Us, Us, 0.0, pert);
    for (int k = 0; k < NKMAX; ++k) {
        const bool on = k < P.nk;
        A.lam[k] = on ? lam + (size_t)k * c->plane : nullptr;
        A.lamW[k] = on && lamW ? lamW + (size_t)k * c->plane : nullptr;
        A.lamWo[k] = on && lamW ? lamW + (size_t)k * c->plane : nullptr;
        A.acc[k] = on ? lam + (size_t)k * c->plane : nullptr;
        A.gx[k] = on ? gx + (size_t)k * c->plane : nullptr;
        A.gy[k] = on ? gy + (size_t)k * c->plane : nullptr;
    }
    A.pert = pert;
    A.rhs_only = rhs_only;
    AdjGatherArgs<S> G{};
    for (int k = 0; k < NKMAX; ++k) {
        G.acc[k] = A.acc[k];
        G.gx[k] = A.gx[k];
        G.gy[k] = A.gy[k];
        G.out[k] = A.acc[k];
    }
    G.imax = P.imax;
    G.jmax = P.jmax;
    G.pitch = c->pitch;
    G.nk = P.nk;
    G.inv_dm = 1.0 / P.dm;
    G.inv_dn = 1.0 / P.dn;
    HIPCHK(launch_adj_stage(A, G, c->stream));
    return NS_OK;
}

// lam <- (d U_{n+1} / d U_n)^T lam for the step from the state at U0
template <typename S>
int adj_step_back(ns_ctx *c, int iter, const S *U0, S *lam)
{
    S *stg = static_cast<S *>(c->adj.stg);
    HIPCHK(copy_state(c, c->S[0], U0));
    for (int irk = 0; irk < 6; ++irk) {
        HIPCHK(copy_state(c, state_slot<S>(stg, c, irk), c->S[0]));
        if (irk < 5) {
            int rc = enqueue_stage(c, irk, iter, false);
            if (rc) return rc;
        }
    }
    // the register W is dead after stage 6 (stage 1 of the next step does not read it)
    HIPCHK(hipMemsetAsync(c->adj.lamW, 0, state_bytes(c), c->stream));
    for (int irk = 5; irk >= 0; --irk) {
        int rc = adj_stage_back<S>(c, irk, state_slot<S>(stg, c, irk), lam, static_cast<S *>(c->adj.lamW),
                                   static_cast<S *>(c->adj.gx), static_cast<S *>(c->adj.gy), c->pert, false);
        if (rc) return rc;
    }
    return NS_OK;
}

// One step back from kept stage inputs: U0 is the step's state (the input
// of stage 1), stg5 the inputs of stages 2-6 (adj_backward_t).
template <typename S>
int adj_step_back_kept(ns_ctx *c, S *U0, void *stg5, S *lam)
{
    // the register W is dead after stage 6 (stage 1 of the next step does not read it)
    HIPCHK(hipMemsetAsync(c->adj.lamW, 0, state_bytes(c), c->stream));
    for (int irk = 5; irk >= 0; --irk) {
        S *in = irk == 0 ? U0 : state_slot<S>(stg5, c, irk - 1);
        int rc = adj_stage_back<S>(c, irk, in, lam, static_cast<S *>(c->adj.lamW), static_cast<S *>(c->adj.gx),
                                   static_cast<S *>(c->adj.gy), c->pert, false);
        if (rc) return rc;
    }
    return NS_OK;
}

template <typename S>
int adj_forward_t(ns_ctx *c, int iter0, int nsteps, int every, double *J)
{
    auto &a = c->adj;
    const int nseg = (nsteps + every - 1) / every;
    const size_t b = state_bytes(c);
    int rc;
    if ((rc = adj_work(c)) || (rc = grow(&a.ckpt, a.ckbytes, (size_t)(nseg + 1) * b)) ||
        (rc = grow((void **)&a.jvals, a.jbytes, sizeof(double) * (size_t)std::max(nsteps, 1))))
        return rc;
    S *S0 = static_cast<S *>(c->S[0]);
    for (int n = 0; n < nsteps; ++n) {
        if (n % every == 0) HIPCHK(copy_state(c, state_slot<S>(a.ckpt, c, n / every), S0));
        if ((rc = enqueue_step(c, iter0 + n, false))) return rc;
        HIPCHK(launch_arc_reduce(arc_args<S>(c, S0), a.jvals + n, c->stream));
    }
    HIPCHK(copy_state(c, state_slot<S>(a.ckpt, c, nseg), S0));
    std::vector<double> jv((size_t)nsteps);
    if (nsteps > 0)
        HIPCHK(hipMemcpyAsync(jv.data(), a.jvals, sizeof(double) * nsteps, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double sum = 0.0;
    for (double v : jv) sum += v;
    if (J) *J = c->P.dt * sum;
    a.iter0 = iter0;
    a.nsteps = nsteps;
    a.every = every;
    a.have_fwd = true;
    return NS_OK;
}

// dJ/dU0 into adj.lam; the state is restored to the final state afterwards
template <typename S>
int adj_backward_t(ns_ctx *c)
{
    auto &a = c->adj;
    const int N = a.nsteps, K = a.every;
    const int nseg = (N + K - 1) / K;
    int rc = grow(&a.seg, a.segbytes, (size_t)(K + 1) * state_bytes(c));
    if (rc) return rc;
    // The segment's forward recompute keeps the inputs of stages 2-6 of every
    // step when they fit in a quarter of the free device memory (and the
    // caller has not switched it off): a step back then runs no forward
    // stage again, and the backward costs one forward step plus the adjoint
    // kernels per step.  Otherwise each step back recomputes its stage
    // inputs from the step state (adj_step_back).  The kept inputs are
    // released when the backward ends.
    const size_t need = (size_t)K * 5 * state_bytes(c);
    bool keep = false;
    if (a.keep) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess && need <= fr / 4) {
            void *p = nullptr;
            if (hipMalloc(&p, need) == hipSuccess) {
                a.stgall = p;
                a.stgallbytes = need;
                keep = true;
            } else {
                (void)hipGetLastError();  // the optional buffer did not fit: not an error
            }
        }
    }
    struct Release {  // also on an error return: wait for the stream, then free
        ns_ctx::Adj &a;
        hipStream_t s;
        ~Release()
        {
            if (!a.stgall) return;
            (void)hipStreamSynchronize(s);
            (void)hipFree(a.stgall);
            a.stgall = nullptr;
            a.stgallbytes = 0;
        }
    } release_kept{a, c->stream};
    S *lam = static_cast<S *>(a.lam);
    S *S0 = static_cast<S *>(c->S[0]);
    HIPCHK(hipMemsetAsync(lam, 0, state_bytes(c), c->stream));
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * K, s1 = std::min(N, s0 + K);
        HIPCHK(copy_state(c, S0, state_slot<S>(a.ckpt, c, sg)));
        HIPCHK(copy_state(c, state_slot<S>(a.seg, c, 0), S0));
        for (int n = s0; n < s1; ++n) {
            if (keep) {
                // stage irk reads its kept input and writes the next one in place
                // (the last stage writes the next step state): no copies
                for (int irk = 0; irk < 6; ++irk) {
                    S *in = irk == 0 ? state_slot<S>(a.seg, c, n - s0) : state_slot<S>(a.stgall, c, (n - s0) * 5 + irk - 1);
                    S *out = irk == 5 ? state_slot<S>(a.seg, c, n - s0 + 1)
                                      : state_slot<S>(a.stgall, c, (n - s0) * 5 + irk);
                    if ((rc = enqueue_stage_t<S>(c, irk, a.iter0 + n, false, in, out))) return rc;
                }
            } else {
                if ((rc = enqueue_step(c, a.iter0 + n, false))) return rc;
                HIPCHK(copy_state(c, state_slot<S>(a.seg, c, n - s0 + 1), S0));
            }
        }
        for (int n = s1 - 1; n >= s0; --n) {
            // J = dt sum_m phi(U_m), m = 1..N: seed dphi/dU at U_{n+1}, then step n back
            HIPCHK(launch_arc_seed(arc_args<S>(c, state_slot<S>(a.seg, c, n - s0 + 1)), lam + 3 * c->plane, c->P.dt,
                                   c->stream));
            rc = keep ? adj_step_back_kept<S>(c, state_slot<S>(a.seg, c, n - s0),
                                              state_slot<S>(a.stgall, c, (n - s0) * 5), lam)
                      : adj_step_back<S>(c, a.iter0 + n, state_slot<S>(a.seg, c, n - s0), lam);
            if (rc) return rc;
        }
    }
    HIPCHK(copy_state(c, S0, state_slot<S>(a.ckpt, c, nseg)));
    // the kept inputs are freed on return: the stream must be done with them
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

// device planes (nk x imax*jmax, FP64) <-> a state-precision pitched buffer
template <typename S>
int dev_to_state(ns_ctx *c, const double *src, S *dst, bool sub_base)
{
    ConvArgs<double, S> a{};
    a.in = src;
    a.out = dst;
    a.ipitch = c->P.imax;
    a.opitch = c->pitch;
    a.imax = c->P.imax;
    a.jmax = c->P.jmax;
    a.nk = c->P.nk;
    a.iplane = (size_t)c->P.imax * c->P.jmax;
    a.oplane = c->plane;
    a.sgn = 1.0;
    if (sub_base && c->pert) {
        const PertRows pr = pert_rows(c);
        for (int k = 0; k < c->P.nk; ++k) a.sub[k] = pr.ub[k];
    }
    HIPCHK(launch_convert(a, c->stream));
    return NS_OK;
}

template <typename S>
int state_to_dev(ns_ctx *c, const S *src, double *dst, bool add_base)
{
    ConvArgs<S, double> a{};
    a.in = src;
    a.out = dst;
    a.ipitch = c->pitch;
    a.opitch = c->P.imax;
    a.imax = c->P.imax;
    a.jmax = c->P.jmax;
    a.nk = c->P.nk;
    a.iplane = c->plane;
    a.oplane = (size_t)c->P.imax * c->P.jmax;
    a.sgn = -1.0;
    if (add_base && c->pert) {
        const PertRows pr = pert_rows(c);
        for (int k = 0; k < c->P.nk; ++k) a.sub[k] = pr.ub[k];
    }
    HIPCHK(launch_convert(a, c->stream));
    return NS_OK;
}

// host AoS (FP64) <-> state-precision pitched buffer (via the FP64 work planes)
template <typename S>
int host_to_buf(ns_ctx *c, const double *host, S *dst)
{
    if (sizeof(S) == sizeof(double)) return upload_aos(c, reinterpret_cast<double *>(dst), host, c->P.nk);
    double *tmp = c->work;
    int rc = upload_aos(c, tmp, host, c->P.nk);
    if (rc) return rc;
    ConvArgs<double, S> a{};
    a.in = tmp;
    a.out = dst;
    a.ipitch = a.opitch = c->pitch;
    a.imax = c->P.imax;
    a.jmax = c->P.jmax;
    a.nk = c->P.nk;
    a.iplane = a.oplane = c->plane;
    HIPCHK(launch_convert(a, c->stream));
    return NS_OK;
}

template <typename S>
int buf_to_host(ns_ctx *c, const S *src, double *host)
{
    if (sizeof(S) == sizeof(double)) return download_aos(c, reinterpret_cast<const double *>(src), host, c->P.nk);
    double *tmp = c->work;
    ConvArgs<S, double> a{};
    a.in = src;
    a.out = tmp;
    a.ipitch = a.opitch = c->pitch;
    a.imax = c->P.imax;
    a.jmax = c->P.jmax;
    a.nk = c->P.nk;
    a.iplane = a.oplane = c->plane;
    HIPCHK(launch_convert(a, c->stream));
    return download_aos(c, tmp, host, c->P.nk);
}

template <typename S>
int vjp_step_t(ns_ctx *c, int iter, const double *U, double *lam)
{
    int rc;
    if ((rc = adj_work(c)) || (rc = grow(&c->adj.seg, c->adj.segbytes, state_bytes(c)))) return rc;
    if ((rc = upload_state(c, U))) return rc;
    HIPCHK(copy_state(c, c->adj.seg, c->S[0]));
    S *l = static_cast<S *>(c->adj.lam);
    if ((rc = host_to_buf<S>(c, lam, l))) return rc;
    if ((rc = adj_step_back<S>(c, iter, static_cast<const S *>(c->adj.seg), l))) return rc;
    HIPCHK(copy_state(c, c->S[0], c->adj.seg));
    return buf_to_host<S>(c, l, lam);
}

}  // namespace

extern "C" {

int32_t ns_adj_set_arc(ns_ctx *c, int32_t npts, const int32_t *ij)
{
    if (!c || npts < 0 || (npts > 0 && !ij)) return fail(NS_ERR_PARAM, "bad argument");
    std::vector<int2> pts((size_t)npts);
    std::vector<char> seen((size_t)c->P.imax * c->P.jmax, 0);
    for (int t = 0; t < npts; ++t) {
        const int i = ij[2 * t], j = ij[2 * t + 1];
        if (i < 1 || i > c->P.imax || j < 1 || j > c->P.jmax)
            return fail(NS_ERR_PARAM, "observer point %d at (%d,%d) lies outside the grid", t, i, j);
        char &s = seen[(size_t)(j - 1) * c->P.imax + (i - 1)];
        if (s) return fail(NS_ERR_PARAM, "observer point %d at (%d,%d) is listed twice", t, i, j);
        s = 1;
        pts[t] = make_int2(i - 1, j - 1);
    }
    if (c->adj.arc) HIPCHK(hipFree(c->adj.arc));
    c->adj.arc = nullptr;
    c->adj.narc = 0;
    if (npts > 0) {
        HIPCHK(hipMalloc(&c->adj.arc, sizeof(int2) * npts));
        HIPCHK(hipMemcpy(c->adj.arc, pts.data(), sizeof(int2) * npts, hipMemcpyHostToDevice));
    }
    c->adj.narc = npts;
    c->adj.have_fwd = false;
    return NS_OK;
}

int32_t ns_adj_set_keep(ns_ctx *c, int32_t keep)
{
    if (!c) return fail(NS_ERR_PARAM, "null context");
    c->adj.keep = keep != 0;
    return NS_OK;
}

int32_t ns_adj_forward(ns_ctx *c, int32_t iter0, int32_t nsteps, int32_t every, double *J)
{
    int rc = adj_check(c);
    if (rc) return rc;
    if (nsteps < 1 || every < 1) return fail(NS_ERR_PARAM, "nsteps and the checkpoint interval must be >= 1");
    if (c->stage_variant < 0 && (rc = autotune_stage(c))) return rc;
    return c->pert ? adj_forward_t<float>(c, iter0, nsteps, every, J) : adj_forward_t<double>(c, iter0, nsteps, every, J);
}

static int adj_backward_common(ns_ctx *c)
{
    int rc = adj_check(c);
    if (rc) return rc;
    if (!c->adj.have_fwd) return fail(NS_ERR_STATE, "ns_adj_forward has not been run");
    return c->pert ? adj_backward_t<float>(c) : adj_backward_t<double>(c);
}

int32_t ns_adj_backward(ns_ctx *c, double *dJdU0)
{
    if (!c || !dJdU0) return fail(NS_ERR_PARAM, "null argument");
    int rc = adj_backward_common(c);
    if (rc) return rc;
    return c->pert ? buf_to_host<float>(c, static_cast<const float *>(c->adj.lam), dJdU0)
                   : buf_to_host<double>(c, static_cast<const double *>(c->adj.lam), dJdU0);
}

int32_t ns_adj_backward_dev(ns_ctx *c, double *dJdU0)
{
    if (!c || !dJdU0) return fail(NS_ERR_PARAM, "null argument");
    int rc = adj_backward_common(c);
    if (rc) return rc;
    rc = c->pert ? state_to_dev<float>(c, static_cast<const float *>(c->adj.lam), dJdU0, false)
                 : state_to_dev<double>(c, static_cast<const double *>(c->adj.lam), dJdU0, false);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

int32_t ns_put_state_dev(ns_ctx *c, const double *U)
{
    if (!c || !U) return fail(NS_ERR_PARAM, "null argument");
    if (c->pert && !c->have_base) return fail(NS_ERR_STATE, "the FP32 perturbation form needs ns_set_base first");
    int rc = c->pert ? dev_to_state<float>(c, U, static_cast<float *>(c->S[0]), true)
                     : dev_to_state<double>(c, U, static_cast<double *>(c->S[0]), true);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_state = true;
    return NS_OK;
}

int32_t ns_get_state_dev(ns_ctx *c, double *U)
{
    if (!c || !U) return fail(NS_ERR_PARAM, "null argument");
    int rc = c->pert ? state_to_dev<float>(c, static_cast<const float *>(c->S[0]), U, true)
                     : state_to_dev<double>(c, static_cast<const double *>(c->S[0]), U, true);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

int32_t ns_vjp_rhs(ns_ctx *c, const double *U, const double *lam_f, double *lam_U)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (!U || !lam_f || !lam_U) return fail(NS_ERR_PARAM, "null argument");
    if (c->P.nk != 4 || !(c->P.flags & NS_FLAG_INVISCID))
        return fail(NS_ERR_PARAM, "the adjoint covers the 2-D Euler equations (nk = 4 with NS_FLAG_INVISCID)");
    if (c->P.flags & NS_FLAG_PERIODIC_X) return fail(NS_ERR_PARAM, "the adjoint does not cover periodic x");
    // FP64 operator mirror: U, lam and the flux adjoints in the FP64 work planes
    double *dU = c->work, *dl = c->work + 4 * c->plane, *gx = c->work + 8 * c->plane, *gy = c->work + 12 * c->plane;
    if ((rc = upload_aos(c, dU, U, 4)) || (rc = upload_aos(c, dl, lam_f, 4))) return rc;
    if ((rc = adj_stage_back<double>(c, 0, dU, dl, nullptr, gx, gy, false, true))) return rc;
    return download_aos(c, dl, lam_U, 4);
}

static int vjp_linear(ns_ctx *c, double *lam, bool edge, bool fy, bool fx)
{
    if (!c || !lam) return fail(NS_ERR_PARAM, "null argument");
    if (c->P.flags & NS_FLAG_PERIODIC_X) return fail(NS_ERR_PARAM, "the adjoint does not cover periodic x");
    if (!c->have_profiles) return fail(NS_ERR_STATE, "ns_set_profiles has not been called");
    double *d = c->work, *tmp = c->work + 8 * c->plane;
    int rc;
    if ((rc = upload_aos(c, d, lam, c->P.nk))) return rc;
    if ((rc = adj_linear_back<double>(c, d, tmp, edge, fy, fx))) return rc;
    return download_aos(c, d, lam, c->P.nk);
}

int32_t ns_vjp_filter_x(ns_ctx *c, double *lam) { return vjp_linear(c, lam, false, false, true); }
int32_t ns_vjp_filter_y(ns_ctx *c, double *lam) { return vjp_linear(c, lam, false, true, false); }
int32_t ns_vjp_ccpml(ns_ctx *c, double *lam) { return vjp_linear(c, lam, true, false, false); }

int32_t ns_vjp_step(ns_ctx *c, int32_t iter, const double *U, double *lam)
{
    int rc = adj_check(c);
    if (rc) return rc;
    if (!U || !lam) return fail(NS_ERR_PARAM, "null argument");
    if (c->stage_variant < 0 && (rc = autotune_stage(c))) return rc;
    rc = c->pert ? vjp_step_t<float>(c, iter, U, lam) : vjp_step_t<double>(c, iter, U, lam);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return NS_OK;
}

}  // extern "C"
