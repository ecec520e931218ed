// ns_kernels.h -- kernel argument blocks for the time-advance path (gfx950).
//
// Two precision modes share the kernels:
//   FP64       S = R = double, the state is the reference's U itself;
//   FP32-pert  S = R = float, the state is q = U - Ub(j): the perturbation
//              about a base flow that is constant along each grid row (the
//              reference's base flows depend on y only, init.f90:368-404).
//              The base rows, their y-derivatives, the base temperature and
//              the y filter's action on the base stay FP64 (per-row arrays).
#pragma once

#include <hip/hip_runtime.h>

#include "ns_device.h"

namespace nsamd {

constexpr int NKMAX = 5;

// Per-row base-flow data of the perturbation form (all FP64, length jmax).
struct PertRows {
    const double *ub[NKMAX];   // Ub(k, j)
    const double *dub[NKMAX];  // d Ub / dy (k, j) including the metric factor
    const double *tb;          // base temperature gamma*pb/rhob
    const double *dtb;         // d Tb / dy including the metric factor
    const double *ycorr[NKMAX];  // Ly(Ub) - Ub: the y filter acting on the base
};

// Arguments of the fused stage kernel (RHS + LDDRK update + buffer zone).
template <typename S>
struct StageArgs {
    const S *U[NKMAX];     // state at the start of the stage (filtered)
    S *Uout[NKMAX];        // updated + buffered state (input of the filters)
    S *W[NKMAX];           // 2N-storage register, updated in place
    const S *base[NKMAX];  // base flow Ub planes (FP64 mode)
    S *dudt[NKMAX];        // RHS output (rhs-only mode)
    const double *metx, *mety;  // effective metric factors per column / row
    const double *sigx, *sigy;  // buffer-zone profiles
    const double *gx, *gy;      // separable source profile
    PertRows pr;                // perturbation form only
    int imax, jmax, pitch, D;
    double inv_dm, inv_dn, dt, alpha, beta;
    double gamma, inv_Re, gm1_Re, inv_RePr, inv_Pe, alpha_cond;
    double amp_src;  // amplitude*sin(wf*(iter+c)*dt) of this stage (0: no source)
    Transport<double> tc;
    int first, last;  // stage 1 (alpha=0: W not read), stage 6 (W not written)
    int fast_ok;      // unit grid, region metrics: interior tiles may take the fast path
};

// ccpml constants of the rows j=1 / j=jmax (cc.f90:38-59): components 0-2
// extrapolate (plus a base residual in the perturbation form), 3-4 are fixed.
struct RowEdge {
    double bot_res[NKMAX], top_res[NKMAX];  // added to 2 q(2) - q(3)
    double bot_val[NKMAX], top_val[NKMAX];  // fixed values (components 3, 4)
};

// Arguments of the in-place selective filter (look-back chunked sweep).
template <typename S>
struct FilterArgs {
    const S *in[NKMAX];
    S *out[NKMAX];
    const double *corr[NKMAX];  // y filter, perturbation form: Ly(Ub) - Ub per row
    int n;          // points along the filtered direction
    int m;          // points across (rows for x, columns for y)
    int pitch;
    int L;  // chunk length
    int W;  // warm-up (look-back) length
    double sigma;
    int nk;
    int edge_bc;  // y filter only: write the j=1 / j=jmax rows of ccpml
    RowEdge re;
};

template <typename S>
struct EdgeArgs {
    S *U[NKMAX];
    const double *rhob, *psib;
    const double *ub[NKMAX];  // perturbation form: base rows (else null)
    RowEdge re;
    int imax, jmax, pitch, nk;
    double gamma;
};

// Range of tiles that may take the fast path (unit grid, no metric or buffer
// work on the tile or its halo): [b0, b1) along
// one direction of n points with tile size T.  Empty when fast_ok is 0.
__host__ __device__ inline void fast_range(int fast_ok, int D, int n, int T, int &b0, int &b1)
{
    const int Dm = D > 4 ? D : 4;
    b0 = (Dm + 3 + T - 1) / T;         // first tile with i0 - 3 >= Dm
    b1 = (n - Dm - 3 - T) / T + 1;     // past the last tile with i0 + T + 3 <= n - Dm
    if (!fast_ok || n - Dm - 3 - T < 0 || b1 < b0) b0 = b1 = 0;
}

// The tiles of an nbx x nby tile grid outside the fast rectangle
// [fx0,fx1) x [fy0,fy1), enumerated as bottom band, top band, then the left
// and right strips row by row.
struct RingMap {
    int nbx, nby, fx0, fx1, fy0, fy1;
    __host__ __device__ int count() const
    {
        return (fy0 + nby - fy1) * nbx + (fy1 - fy0) * (fx0 + nbx - fx1);
    }
    __host__ __device__ void tile(int t, int &bx, int &by) const
    {
        const int nb = fy0 * nbx;
        if (t < nb) {
            by = t / nbx;
            bx = t - by * nbx;
            return;
        }
        t -= nb;
        const int nt = (nby - fy1) * nbx;
        if (t < nt) {
            const int r = t / nbx;
            by = fy1 + r;
            bx = t - r * nbx;
            return;
        }
        t -= nt;
        const int wl = fx0, w = fx0 + nbx - fx1;
        const int r = t / w, c = t - r * w;
        by = fy0 + r;
        bx = c < wl ? c : fx1 + (c - wl);
    }
};

}  // namespace nsamd
