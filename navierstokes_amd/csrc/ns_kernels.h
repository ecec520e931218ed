// ns_kernels.h -- kernel argument blocks and launcher declarations for the
// FP64 time-advance path (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include "ns_device.h"

namespace nsamd {

constexpr int NKMAX = 5;

// Arguments of the fused stage kernel (RHS + LDDRK update + buffer zone).
template <typename R>
struct StageArgs {
    const R *U[NKMAX];     // filtered state at the start of the stage
    R *Uout[NKMAX];        // updated + buffered state (input of the filters)
    R *W[NKMAX];           // 2N-storage register, updated in place
    const R *base[NKMAX];  // base flow Ub
    R *dudt[NKMAX];        // RHS output (rhs-only mode)
    const R *metx, *mety;  // effective metric factors per column / row
    const R *sigx, *sigy;  // buffer-zone profiles
    const R *gx, *gy;      // separable source profile
    int imax, jmax, pitch, D;
    R inv_dm, inv_dn, dt, alpha, beta;
    R gamma, inv_Re, gm1_Re, inv_RePr, inv_Pe, alpha_cond;
    R amp_src;  // amplitude*sin(wf*(iter+c)*dt) of this stage (0: no source)
    Transport<R> tc;
    int first, last;  // stage 1 (alpha=0: W not read), stage 6 (W not written)
};

// Arguments of the in-place selective filter (look-back chunked sweep).
template <typename R>
struct FilterArgs {
    const R *in[NKMAX];
    R *out[NKMAX];
    int n;          // points along the filtered direction
    int m;          // points across (rows for x, columns for y)
    int pitch;
    int L;  // chunk length
    int W;  // warm-up (look-back) length
    R sigma;
    R gamma;
    int nk;
    int edge_bc;  // y filter only: write the j=1 / j=jmax rows of ccpml
};

struct EdgeArgs {
    double *U[NKMAX];
    const double *rhob, *psib;
    int imax, jmax, pitch, nk;
    double gamma;
};

}  // namespace nsamd
