// ns_adjoint.h -- argument blocks of the reverse-mode (adjoint) kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "ns_kernels.h"

namespace nsamd {

// Pointwise backward of one stage: buffer^T, update^T, pointwise RHS^T.
template <typename S>
struct AdjPointArgs {
    StageArgs<S> st;        // forward stage arguments; st.U = the stage's input state
    const S *lam[NKMAX];    // dJ/dU after the buffer (filters and ccpml already transposed)
    const S *lamW[NKMAX];   // dJ/dW' (output register of the stage)
    S *lamWo[NKMAX];        // dJ/dW  (input register); may alias lamW
    S *acc[NKMAX];          // identity + pointwise parts of dJ/dU_s; may alias lam
    S *gx[NKMAX], *gy[NKMAX];  // flux adjoints (metric applied)
    int pert;               // perturbation form: U = Ub(j) + q
    int rhs_only;           // lam holds dJ/dR at every point; acc gets R'^T lam only
};

// dJ/dU_s = acc + D_x^T gx + D_y^T gy
template <typename S>
struct AdjGatherArgs {
    const S *acc[NKMAX];
    const S *gx[NKMAX], *gy[NKMAX];
    S *out[NKMAX];  // may alias acc
    int imax, jmax, pitch, nk;
    double inv_dm, inv_dn;
};

template <typename S>
struct AdjFilterArgs {
    const S *in[NKMAX];
    S *out[NKMAX];  // a different buffer: chunks read each other's inputs
    int n;          // points along the filtered direction
    int m;          // lines
    int pitch, nk;
    int L, W;       // chunk length, warm-up length (as the forward filter)
    double sigma;
};

template <typename S>
struct AdjEdgeArgs {
    S *lam[NKMAX];
    int imax, jmax, pitch, nk;
};

// observer points (0-based (i, j)) and the pressure plane they read
template <typename S>
struct ArcArgs {
    const int2 *pts;
    int npts;
    const S *U;            // component 3 (p, or p' in the perturbation form)
    const double *base;    // pb plane (FP64 form), null in the perturbation form
    int pitch;
};

// plane copy with conversion and pitch change; sub[k] (per row j, may be
// null) enters as out[k](i,j) = Sout(in[k](i,j) - sgn * sub[k][j])
template <typename Sin, typename Sout>
struct ConvArgs {
    const Sin *in;
    Sout *out;
    int ipitch, opitch, imax, jmax, nk;
    size_t iplane, oplane;
    const double *sub[NKMAX];
    double sgn;
};

}  // namespace nsamd
