// ns_kernels.hip -- gfx950 kernels of the mathe1s/NavierStokes time-advance
// hot path, FP64 (reference semantics).
//
// Per LDDRK stage (rk.f90:52-73) the device runs
//   stage_kernel   : RHS (rhs.f90:10-46) + rk_steep (rk.f90:153-169)
//                    + bufferxy (rk.f90:77-151), one LDS-tiled sweep
//   filter_x_kernel: lddfilterx (filtering.f90:9-96)
//   filter_y_kernel: lddfiltery (filtering.f90:98-190) + ccpml rows j=1,jmax
//   edge_i_kernel  : ccpml columns i=1,imax incl. corners (cc.f90:6-68)
// State planes are structure-of-arrays, row-major with i fastest (pitch).
//
// This file is compiled twice: with NS_KERNELS_F64 and -ffp-contract=off
// (the FP64 path keeps the reference's operation order, no fused
// multiply-adds, so it rounds like the reference), and with NS_KERNELS_F32
// (the FP32 perturbation form, contraction allowed).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

#include "ns_kernels.h"

namespace nsamd {

// ---------------------------------------------------------------------------
// Kernel set-up before launch.  The first use of a kernel loads its code
// object, and a launch path may set its dynamic-LDS limit and query its
// occupancy and the CU count.  Those runtime calls must not run inside a
// stream capture (a kernel first loaded inside a capture aborted the HIP
// runtime), so each (device, kernel) is set up once, outside any capture,
// and remembered here.  A launch path that reaches a kernel not yet set up
// while its stream is capturing returns hipErrorStreamCaptureUnsupported
// (the library reports NS_ERR_STATE) instead of making those calls.
// ---------------------------------------------------------------------------
struct KernelSetup {
    int ncu = 0;     // compute units of the device
    int perblk = 0;  // resident blocks per CU at the set-up's block size and LDS
};

static hipError_t kernel_ready(const void *kern, size_t dyn_smem, int threads, hipStream_t s,
                               KernelSetup *out = nullptr)
{
    static std::mutex mu;
    static std::map<std::pair<int, const void *>, KernelSetup> done;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = done.find({dev, kern});
        if (it != done.end()) {
            if (out) *out = it->second;
            return hipSuccess;
        }
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (s && hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
        return hipErrorStreamCaptureUnsupported;
    if (dyn_smem > 0 &&
        (e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_smem)) != hipSuccess)
        return e;
    hipFuncAttributes fa;
    if ((e = hipFuncGetAttributes(&fa, kern)) != hipSuccess) return e;
    KernelSetup k;
    if (hipDeviceGetAttribute(&k.ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || k.ncu <= 0)
        k.ncu = 256;
    if (threads > 0 && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k.perblk, kern, threads, dyn_smem) !=
                            hipSuccess ||
                        k.perblk < 1))
        k.perblk = 1;
    (void)hipGetLastError();
    std::lock_guard<std::mutex> g(mu);
    done[{dev, kern}] = k;
    if (out) *out = k;
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Stage kernel: LDS-tiled RHS with nested stencils, RK update and buffer.
//
// Tile TX x TY outputs.  The U window holds the tile plus a 6-point halo
// (3 for the first derivatives, 3 more for the derivatives of tau and k dT);
// at the domain edges the one-sided stencils reach 6 points inward, handled by
// shifting the window/level-1 boxes (see tile_ranges).
// ---------------------------------------------------------------------------
constexpr int STX = 64;

struct Range {
    int l0, l1, w0, w1;  // level-1 box [l0,l1), U window [w0,w1) (global, 0-based)
};

__device__ __forceinline__ Range tile_ranges(int t0, int t1, int n)
{
    Range r;
    r.l0 = max(0, t0 - 3);
    if (t1 - 1 >= n - 3) r.l0 = min(r.l0, n - 7);
    r.l1 = min(n, t1 + 3);
    if (t0 <= 2) r.l1 = max(r.l1, 7);
    r.w0 = max(0, r.l0 - 3);
    if (r.l1 - 1 >= n - 3) r.w0 = min(r.w0, n - 7);
    r.w1 = min(n, r.l1 + 3);
    if (r.l0 <= 2) r.w1 = max(r.w1, 7);
    return r;
}

// Interior DRP derivative on a unit grid (dm = dn = 1): the reference's
// ((a1 d1)/1 + (a2 d2)/1) + (a3 d3)/1 without the exact divisions by 1.
template <typename R, typename F>
__device__ __forceinline__ R drp_interior_unit(const F &f, int q)
{
    const R t1 = R(DRP_A1) * (f(q + 1) - f(q - 1));
    const R t2 = R(DRP_A2) * (f(q + 2) - f(q - 2));
    const R t3 = R(DRP_A3) * (f(q + 3) - f(q - 3));
    return (t1 + t2) + t3;
}

// Element g of plane p.  With a 32-bit offset the byte offset is formed in
// 32 bits, so the compiler addresses through the plane's base in SGPRs plus a
// 32-bit VGPR offset instead of a 64-bit address per plane and point (the
// caller guarantees planes below 4 GiB).
template <typename T>
__device__ __forceinline__ T *elt(T *p, size_t g)
{
    return p + g;
}
template <typename T>
__device__ __forceinline__ T *elt(T *p, uint32_t g)
{
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return reinterpret_cast<T *>(reinterpret_cast<B *>(p) + g * (uint32_t)sizeof(T));
}

// ---------------------------------------------------------------------------
// Per-point pieces of the stage, shared by the tile and the marching kernels,
// which therefore compute every value with the same operations.  A point
// mode PM says what the caller guarantees about the points it computes:
//   PM_GEN : nothing (one-sided stencils near the edges, metric factors,
//            buffer layers and the edge test point by point);
//   PM_MID : the point and its 3-point level-1 halo lie at least 6 points
//            from every edge, so every stencil is the interior one;
//   PM_MIDU: PM_MID on a unit grid (dm = dn = 1);
//   PM_FAST: PM_MIDU outside the metric/buffer layers.
// The stronger modes drop branches and multiplications by exactly 1.0, so
// every result bit is the one PM_GEN computes.
// ---------------------------------------------------------------------------
constexpr int PM_GEN = 0, PM_MID = 1, PM_MIDU = 2, PM_FAST = 3;

// per: the line is periodic (x with xper): the interior stencil everywhere,
// the accessor wraps
template <typename R, int PM, typename F>
__device__ __forceinline__ R dv(const F &f, int q, int n, R inv_h, bool per = false)
{
    if constexpr (PM >= PM_MIDU) return drp_interior_unit<R>(f, q);
    else if constexpr (PM == PM_MID) return drp_interior<R>(f, q, inv_h);
    else return per ? drp_interior<R>(f, q, inv_h) : drp_deriv<R>(f, q, n, inv_h);
}

template <typename R, int PM>
__device__ __forceinline__ R met(R d, R m)
{
    if constexpr (PM == PM_FAST) return d;
    else return d * m;
}

// Per-row base data of the perturbation form at one row, loaded ahead by a
// caller whose row is wave-uniform (null: read from the per-row arrays).
template <typename R, int NK>
struct RowCache {
    R ub[NK], dub[NK], tb, dtb;
};

template <typename S, typename R, int NK>
__device__ __forceinline__ void load_rows(const StageArgs<S> &A, int gy, RowCache<R, NK> &rc)
{
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        rc.ub[k] = R(A.pr.ub[k][gy]);
        rc.dub[k] = R(A.pr.dub[k][gy]);
    }
    rc.tb = A.pr.tb ? R(A.pr.tb[gy]) : R(0);
    rc.dtb = A.pr.dtb ? R(A.pr.dtb[gy]) : R(0);
}

// level-1 slots: 0..4 = tau_xx, tau_xy, tau_yy, k dT/dx, k dT/dy (viscous),
// then the species flux rho alpha (psi_x + psi_y) (nk = 5)
template <int NK, bool VISC>
struct L1 {
    static constexpr int PSI = VISC ? 5 : 0;
    static constexpr int N = (VISC ? 5 : 0) + (NK >= 5 ? 1 : 0);
};

// Level-1 fields at (gx, gy) (eqeuler.f90:34-41, derivs.f90:25-56, :40);
// mx, my: the metric factors there (metx[gx], mety[gy]; unused in PM_FAST).
// ux(k) / uy(k): accessor of stored component k along x at row gy / along y
// at column gx; tx / ty: the same for T (T - Tb in the perturbation form);
// tval: that value at the point.
template <typename S, typename R, int NK, bool VISC, bool PERT, int PM, typename UXF, typename UYF, typename TXF,
          typename TYF>
__device__ __forceinline__ void level1_point(const StageArgs<S> &A, const Transport<R> &tc, int gx, int gy, R mx,
                                             R my, const UXF &ux, const UYF &uy, const TXF &tx, const TYF &ty,
                                             R tval, R *out, const RowCache<R, NK> *rc = nullptr)
{
    const int imax = A.imax, jmax = A.jmax;
    const R inv_dm = R(A.inv_dm), inv_dn = R(A.inv_dn);
    if constexpr (NK >= 5) {
        const R psx = met<R, PM>(dv<R, PM>(ux(4), gx, imax, inv_dm, A.xper), mx);
        R psy = met<R, PM>(dv<R, PM>(uy(4), gy, jmax, inv_dn), my);
        R rho = ux(0)(gx);
        if constexpr (PERT) {
            psy = psy + (rc ? rc->dub[4] : R(A.pr.dub[4][gy]));
            rho = rho + (rc ? rc->ub[0] : R(A.pr.ub[0][gy]));
        }
        out[L1<NK, VISC>::PSI] = rho * R(A.alpha_cond) * (psx + psy);  // derivs.f90:40
    }
    if constexpr (VISC) {
        const R uxv = met<R, PM>(dv<R, PM>(ux(1), gx, imax, inv_dm, A.xper), mx);
        const R vxv = met<R, PM>(dv<R, PM>(ux(2), gx, imax, inv_dm, A.xper), mx);
        R uyv = met<R, PM>(dv<R, PM>(uy(1), gy, jmax, inv_dn), my);
        R vyv = met<R, PM>(dv<R, PM>(uy(2), gy, jmax, inv_dn), my);
        const R Tx = met<R, PM>(dv<R, PM>(tx, gx, imax, inv_dm, A.xper), mx);
        R Ty = met<R, PM>(dv<R, PM>(ty, gy, jmax, inv_dn), my);
        if constexpr (PERT) {
            uyv = uyv + (rc ? rc->dub[1] : R(A.pr.dub[1][gy]));
            vyv = vyv + (rc ? rc->dub[2] : R(A.pr.dub[2][gy]));
            Ty += rc ? rc->dtb : R(A.pr.dtb[gy]);
            tval += rc ? rc->tb : R(A.pr.tb[gy]);
        }
        R mu, kk;
        transport_eval<R>(tc, tval, mu, kk);
        const R c23mu = R(2.0 / 3.0) * mu;
        const R div = uxv + vyv;
        out[0] = -(c23mu * div) + R(2) * mu * uxv;  // eqeuler.f90:37
        out[1] = mu * (uyv + vxv);                  // :38
        out[2] = -(c23mu * div) + R(2) * mu * vyv;  // :39
        out[3] = Tx * kk;                           // derivs.f90:55
        out[4] = Ty * kk;                           // :54
    }
}

// RHS at (gx, gy) (eqeuler.f90:61-97): f[], the stored state qc[] and the
// full state uc[] there; mx, my as for level1_point.  lx(l) / ly(l): accessor of level-1 field l along x
// at row gy / along y at column gx; lv(l): its value at the point.
template <typename S, typename R, int NK, bool VISC, bool PERT, int PM, typename UXF, typename UYF, typename LXF,
          typename LYF, typename LVF>
__device__ __forceinline__ void rhs_point(const StageArgs<S> &A, int gx, int gy, R mx, R my, const UXF &ux,
                                          const UYF &uy, const LXF &lx, const LYF &ly, const LVF &lv, R *qc, R *uc,
                                          R *f, const RowCache<R, NK> *rc = nullptr)
{
    const int imax = A.imax, jmax = A.jmax;
    const R inv_dm = R(A.inv_dm), inv_dn = R(A.inv_dn), gam = R(A.gamma);
    R fx[NK], fy[NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        qc[k] = ux(k)(gx);
        uc[k] = qc[k];
        fx[k] = met<R, PM>(dv<R, PM>(ux(k), gx, imax, inv_dm, A.xper), mx);
        fy[k] = met<R, PM>(dv<R, PM>(uy(k), gy, jmax, inv_dn), my);
        if constexpr (PERT) {
            uc[k] = qc[k] + (rc ? rc->ub[k] : R(A.pr.ub[k][gy]));
            fy[k] = fy[k] + (rc ? rc->dub[k] : R(A.pr.dub[k][gy]));
        }
    }
    const R rho = uc[0], u = uc[1], v = uc[2], p = uc[3];
    const R irho = R(1) / rho;
    const R src = (A.amp_src != 0.0) ? R(A.amp_src * (A.gx[gx] * A.gy[gy])) : R(0);
    // eqeuler.f90:63-66
    f[0] = -(u * fx[0] + rho * fx[1] + v * fy[0] + rho * fy[2]);
    if constexpr (VISC) {
        const R txx_x = met<R, PM>(dv<R, PM>(lx(0), gx, imax, inv_dm, A.xper), mx);
        const R txy_x = met<R, PM>(dv<R, PM>(lx(1), gx, imax, inv_dm, A.xper), mx);
        const R ktx_x = met<R, PM>(dv<R, PM>(lx(3), gx, imax, inv_dm, A.xper), mx);
        const R txy_y = met<R, PM>(dv<R, PM>(ly(1), gy, jmax, inv_dn), my);
        const R tyy_y = met<R, PM>(dv<R, PM>(ly(2), gy, jmax, inv_dn), my);
        const R kty_y = met<R, PM>(dv<R, PM>(ly(4), gy, jmax, inv_dn), my);
        const R t0 = lv(0), t1 = lv(1), t2 = lv(2);
        const R vre = R(A.inv_Re) * irho;
        // eqeuler.f90:68-92
        f[1] = -(u * fx[1] + irho * fx[3] + v * fy[1] - vre * (txx_x + txy_y));
        f[2] = -(u * fx[2] + v * fy[2] + irho * fy[3] - vre * (txy_x + tyy_y));
        f[3] = -(gam * p * fx[1] + u * fx[3] + gam * p * fy[2] + v * fy[3] - src -
                 R(A.gm1_Re) * (fx[1] * t0 + fx[2] * t1 + fy[1] * t1 + fy[2] * t2) -
                 R(A.inv_RePr) * (ktx_x + kty_y));
    } else {
        f[1] = -(u * fx[1] + irho * fx[3] + v * fy[1]);
        f[2] = -(u * fx[2] + v * fy[2] + irho * fy[3]);
        f[3] = -(gam * p * fx[1] + u * fx[3] + gam * p * fy[2] + v * fy[3] - src);
    }
    if constexpr (NK >= 5) {
        constexpr int P = L1<NK, VISC>::PSI;
        const R psd_x = met<R, PM>(dv<R, PM>(lx(P), gx, imax, inv_dm, A.xper), mx);
        const R psd_y = met<R, PM>(dv<R, PM>(ly(P), gy, jmax, inv_dn), my);
        f[4] = -(u * fx[4] + v * fy[4] - R(A.inv_Pe) * irho * (psd_x + psd_y));  // eqeuler.f90:94-95
    }
}

// Data of one output point loaded ahead by a caller: W, and at buffer
// points the buffer sigmas and (FP64 form) the base state.
template <typename S, int NK>
struct StorePre {
    S w[NK];
    S b[NK];
    double sx, sy;
};

// W is loaded unconditionally (gx, gy must lie in the grid) so that the
// number of loads in flight is known to the compiler.
template <typename S, int NK, bool PERT, int PM, typename IX = size_t>
__device__ __forceinline__ void store_prefetch(const StageArgs<S> &A, int gx, int gy, StorePre<S, NK> &p)
{
    const IX g = (IX)gy * (IX)A.pitch + (IX)gx;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        const S w = *elt(A.W[k], g);
        p.w[k] = A.first ? S(0) : w;
    }
    if constexpr (PM != PM_FAST) {
        const bool ybuf = gy < A.D || gy >= A.jmax - A.D;
        const bool xbuf = !A.xper && (gx < A.D || gx >= A.imax - A.D);
        p.sy = ybuf ? A.sigy[gy] : 1.0;
        p.sx = xbuf ? A.sigx[gx] : 1.0;
#pragma unroll
        for (int k = 0; k < NK; ++k) p.b[k] = (!PERT && (ybuf || xbuf)) ? *elt(A.base[k], g) : S(0);
    }
}

// LDDRK update at interior points (rk.f90:160-167) and the buffer layers
// (rk.f90:81-108), or the RHS alone (RHS_ONLY).  pre: the point's data
// loaded ahead by store_prefetch (null: load it here).
template <typename S, typename R, int NK, bool RHS_ONLY, bool PERT, int PM, typename IX = size_t>
__device__ __forceinline__ void stage_store(const StageArgs<S> &A, int gx, int gy, const R *qc, const R *uc,
                                            const R *f, const StorePre<S, NK> *pre = nullptr)
{
    const int imax = A.imax, jmax = A.jmax;
    const IX g = (IX)gy * (IX)A.pitch + (IX)gx;
    if constexpr (RHS_ONLY) {
#pragma unroll
        for (int k = 0; k < NK; ++k) *elt(A.dudt[k], g) = S(f[k]);
        return;
    }
    // interior points (rk.f90:161-162; periodic x: every column)
    const bool interior =
        PM != PM_GEN || ((A.xper || (gx >= 1 && gx <= imax - 2)) && gy >= 1 && gy <= jmax - 2);
    const bool ybuf = PM != PM_FAST && (gy < A.D || gy >= jmax - A.D);
    const bool xbuf = PM != PM_FAST && !A.xper && (gx < A.D || gx >= imax - A.D);
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        if constexpr (PERT) {
            // stage accumulation of the state in FP64 registers
            double qn = double(qc[k]);
            if (interior) {
                const double w0 = double(pre ? pre->w[k] : *elt(A.W[k], g));
                const double wn = A.first ? A.dt * double(f[k]) : A.alpha * w0 + A.dt * double(f[k]);
                if (!A.last) *elt(A.W[k], g) = S(wn);
                qn = qn + A.beta * wn;
            }
            // Ub + s (U - Ub) - Ub = s q
            if (ybuf) qn = (pre ? pre->sy : A.sigy[gy]) * qn;
            if (xbuf) qn = (pre ? pre->sx : A.sigx[gx]) * qn;
            *elt(A.Uout[k], g) = S(qn);
        } else {
            R un = uc[k];
            if (interior) {
                const R w0 = R(pre ? pre->w[k] : *elt(A.W[k], g));
                const R wn = A.first ? R(A.dt) * f[k] : R(A.alpha) * w0 + R(A.dt) * f[k];
                if (!A.last) *elt(A.W[k], g) = S(wn);
                un = uc[k] + R(A.beta) * wn;
            }
            if (ybuf || xbuf) {  // y layer first, then x
                const R ub = R(pre ? pre->b[k] : *elt(A.base[k], g));
                if (ybuf) un = ub + R(pre ? pre->sy : A.sigy[gy]) * (un - ub);
                if (xbuf) un = ub + R(pre ? pre->sx : A.sigx[gx]) * (un - ub);
            }
            *elt(A.Uout[k], g) = S(un);
        }
    }
}

// ---------------------------------------------------------------------------
// Tile version: one block computes the outputs [i0,i1) x [j0,j1) (at most
// TX x TY) from an LDS box holding U (and T) on the tile plus a 6-point halo
// and the level-1 fields on the tile plus a 3-point halo.
// ---------------------------------------------------------------------------
template <typename S, typename R, int NK, bool VISC, int TY, bool RHS_ONLY, bool PERT, int NT, int PM, int TX = STX>
__device__ __forceinline__ void stage_tile(const StageArgs<S> &A, int i0, int i1, int j0, int j1)
{
    static_assert(NT % TX == 0, "the block covers whole tile rows");
    constexpr int WX = TX + 12, WY = TY + 12;
    constexpr int LX = TX + 6, LY = TY + 6;
    constexpr int NL = L1<NK, VISC>::N;
    extern __shared__ char smem_raw[];
    R *sU = reinterpret_cast<R *>(smem_raw);  // NK x WY x WX   (U, or q in the perturbation form)
    R *sT = sU + NK * WY * WX;                // WY x WX (VISC)  (T, or T - Tb)
    R *sL = sT + (VISC ? WY * WX : 0);        // level-1 fields: LY x LX each

    const int imax = A.imax, jmax = A.jmax, pitch = A.pitch;
    // periodic x: the windows are not shifted near the x ends; their columns
    // wrap when loaded
    const Range rx = A.xper ? Range{i0 - 3, i1 + 3, i0 - 6, i1 + 6} : tile_ranges(i0, i1, imax);
    const Range ry = tile_ranges(j0, j1, jmax);
    const int tid = threadIdx.x;
    const R gam = R(A.gamma);

    // phase 1: U window (and T = gamma p / rho, rhs.f90:20)
    const int wxn = rx.w1 - rx.w0, wyn = ry.w1 - ry.w0;
    for (int idx = tid; idx < wyn * WX; idx += NT) {
        const int wy = idx / WX, wx = idx - wy * WX;
        if (wx < wxn) {
            const size_t g = (size_t)(ry.w0 + wy) * pitch + wrap_col(rx.w0 + wx, imax);
#pragma unroll
            for (int k = 0; k < NK; ++k) sU[(k * WY + wy) * WX + wx] = R(A.U[k][g]);
            if constexpr (VISC) {
                const R q0 = sU[(0 * WY + wy) * WX + wx], q3 = sU[(3 * WY + wy) * WX + wx];
                if constexpr (PERT) {
                    // T - Tb = gamma (p' rhob - pb rho') / (rho rhob): no cancellation
                    const int gy = ry.w0 + wy;
                    const R rb = R(A.pr.ub[0][gy]), pb = R(A.pr.ub[3][gy]);
                    sT[wy * WX + wx] = gam * (q3 * rb - pb * q0) / ((rb + q0) * rb);
                } else {
                    sT[wy * WX + wx] = gam * q3 / q0;
                }
            }
        }
    }
    __syncthreads();

    const int ox = rx.w0, oy = ry.w0;
    auto ux_at = [=](int gy) {  // component accessors along x at global row gy
        return [=](int k) {
            const R *row = sU + (k * WY + (gy - oy)) * WX;
            return [row, ox](int q) { return row[q - ox]; };
        };
    };
    auto uy_at = [=](int gx) {  // component accessors along y at global column gx
        return [=](int k) {
            const R *col = sU + k * WY * WX + (gx - ox);
            return [col, oy](int q) { return col[(q - oy) * WX]; };
        };
    };

    // phase 2: level-1 quantities (viscous stresses, k grad T, species flux)
    if constexpr (NL > 0) {
        const Transport<R> tc{R(A.tc.A_mi), R(A.tc.Bt_mi), R(A.tc.Ct_mi), R(A.tc.D_mi), R(A.tc.inv_mi0),
                              R(A.tc.A_k),  R(A.tc.Bt_k),  R(A.tc.Ct_k),  R(A.tc.D_k),  R(A.tc.inv_k0),
                              R(A.tc.Tref)};
        const int lxn = rx.l1 - rx.l0, lyn = ry.l1 - ry.l0;
        for (int idx = tid; idx < lyn * LX; idx += NT) {
            const int ly = idx / LX, lx = idx - ly * LX;
            if (lx >= lxn) continue;
            const int gx = rx.l0 + lx, gy = ry.l0 + ly;
            const R *trow = sT + (gy - oy) * WX;
            const R *tcol = sT + (gx - ox);
            R out[NL];
            level1_point<S, R, NK, VISC, PERT, PM>(
                A, tc, gx, gy, PM == PM_FAST ? R(1) : R(A.metx[wrap_col(gx, imax)]),
                PM == PM_FAST ? R(1) : R(A.mety[gy]),
                ux_at(gy), uy_at(gx), [trow, ox](int q) { return trow[q - ox]; },
                [tcol, oy](int q) { return tcol[(q - oy) * WX]; }, VISC ? trow[gx - ox] : R(0), out);
            const int o = ly * LX + lx;
#pragma unroll
            for (int l = 0; l < NL; ++l) sL[l * LY * LX + o] = out[l];
        }
        __syncthreads();
    }

    // phase 3: RHS at the tile's points, then the stage update
    const int gx = i0 + tid % TX;
    if (gx >= i1) return;
    const int lx0 = rx.l0, ly0 = ry.l0;
    for (int gy = j0 + tid / TX; gy < j1; gy += NT / TX) {
        auto lxa = [=](int l) {
            const R *row = sL + (l * LY + (gy - ly0)) * LX;
            return [row, lx0](int q) { return row[q - lx0]; };
        };
        auto lya = [=](int l) {
            const R *col = sL + l * LY * LX + (gx - lx0);
            return [col, ly0](int q) { return col[(q - ly0) * LX]; };
        };
        auto lv = [=](int l) { return sL[l * LY * LX + (gy - ly0) * LX + (gx - lx0)]; };
        R qc[NK], uc[NK], f[NK];
        // the update's loads (W, buffer sigmas, base state) issued ahead of
        // the RHS: issued in the update, each would wait behind the stores
        // of the planes before it, which may alias it
        StorePre<S, NK> spre;
        if constexpr (!RHS_ONLY) store_prefetch<S, NK, PERT, PM>(A, gx, gy, spre);
        rhs_point<S, R, NK, VISC, PERT, PM>(A, gx, gy, PM == PM_FAST ? R(1) : R(A.metx[gx]),
                                            PM == PM_FAST ? R(1) : R(A.mety[gy]), ux_at(gy), uy_at(gx), lxa, lya,
                                            lv, qc, uc, f);
        stage_store<S, R, NK, RHS_ONLY, PERT, PM>(A, gx, gy, qc, uc, f, RHS_ONLY ? nullptr : &spre);
    }
}

// interior tiles of the fast rectangle (tile indices from bx0, by0)
template <typename S, typename R, int NK, bool VISC, int TY, bool PERT, int NT>
__global__ __launch_bounds__(NT) void stage_kernel_fast(StageArgs<S> A, int bx0, int by0)
{
    const int i0 = (bx0 + (int)blockIdx.x) * STX, j0 = (by0 + (int)blockIdx.y) * TY;
    stage_tile<S, R, NK, VISC, TY, false, PERT, NT, PM_FAST>(A, i0, i0 + STX, j0, j0 + TY);
}

// Edge tiles around the interior region in one launch: the first
// sides.count() blocks take SX x SY tiles over the side strips beside it (a
// few columns wide: wide tiles there would be mostly empty), the rest BX x BY
// tiles over the bands below and above it (the whole grid when there is no
// interior region).  One launch, so that band tiles need not wait for the
// last side tile as a second launch on the same stream would.
//   large (512 threads): 16 x 4 TY and 64 x TY, the same LDS box transposed,
//     most of a CU's LDS;
//   small (256 threads): 16 x 16 and 32 x 8, about a third of it, so one
//     fits on a CU beside a marching block.
template <bool SMALL, int TY>
struct EdgeShape {
    static constexpr int SX = SMALL ? SMALL_SX : 16, SY = SMALL ? SMALL_SY : 4 * TY;
    static constexpr int BX = SMALL ? SMALL_BX : STX, BY = SMALL ? SMALL_BY : TY;
    static constexpr int NT = SMALL ? 256 : 512;
};
// edge tile t of the side strips, then of the bands
template <typename S, typename R, int NK, bool VISC, int TY, bool RHS_ONLY, bool PERT, bool SMALL>
__device__ __forceinline__ void edge_block(const StageArgs<S> &A, const RectSet &sides, const RectSet &bands, int t)
{
    using E = EdgeShape<SMALL, TY>;
    int i0, i1, j0, j1;
    if (t < sides.count()) {
        sides.tile(t, i0, i1, j0, j1);
        stage_tile<S, R, NK, VISC, E::SY, RHS_ONLY, PERT, E::NT, PM_GEN, E::SX>(A, i0, i1, j0, j1);
    } else {
        bands.tile(t - sides.count(), i0, i1, j0, j1);
        stage_tile<S, R, NK, VISC, E::BY, RHS_ONLY, PERT, E::NT, PM_GEN, E::BX>(A, i0, i1, j0, j1);
    }
}

template <typename S, typename R, int NK, bool VISC, int TY, bool RHS_ONLY, bool PERT, bool SMALL>
__global__ __launch_bounds__((EdgeShape<SMALL, TY>::NT)) void stage_kernel_edges(StageArgs<S> A, RectSet sides,
                                                                                RectSet bands)
{
    edge_block<S, R, NK, VISC, TY, RHS_ONLY, PERT, SMALL>(A, sides, bands, (int)blockIdx.x);
}

// Batched launch: one block per entry of a tile table spanning every case;
// args holds one StageArgs per (stage, case), so a captured graph replays the
// whole batch without argument updates.  Fast and general tiles are listed in
// separate tables and run as separate launches.
template <typename S, typename R, int NK, bool VISC, int TY, bool PERT, int NT, bool FAST>
__global__ __launch_bounds__(NT) void stage_kernel_batch(const StageArgs<S> *args, const int4 *tiles, int stage,
                                                         int ncases)
{
    const int4 t = tiles[blockIdx.x];
    // a copy: through a reference into the table, every store of the tile
    // could alias the arguments, and their fields would be reloaded
    const StageArgs<S> A = args[(size_t)stage * ncases + t.x];
    const int i0 = t.y * STX, j0 = t.z * TY;
    stage_tile<S, R, NK, VISC, TY, false, PERT, NT, FAST ? PM_FAST : PM_GEN>(A, i0, min(i0 + STX, A.imax), j0,
                                                                        min(j0 + TY, A.jmax));
}

// ---------------------------------------------------------------------------
// Marching version (interior region; PM_FAST: the fast zone, else the general
// region of march_region, PM_MID(U), with buffers and metrics point by
// point).  A block
// owns a strip of MTX = 58
// output columns, so its level-1 rows are exactly 64 wide (one wave per row)
// and its U window rows 70 wide, and marches up a chunk of rows RB at a time
// (RB waves).  Window rows and level-1 rows live in LDS rings: every row is
// loaded, and every level-1 row computed, once per chunk.  Per step:
//   (a) store the RB window rows prefetched during the previous step (+ T),
//       barrier, issue the loads of the next step's rows into registers;
//   (b) wave w computes level-1 row base+3+w, barrier;
//   (c) wave w computes the outputs of row base+w.
// The window ring holds 2 RB + 9 rows: (c) reads rows base-3 .. base+RB+2,
// (b) rows up to base+RB+5, and (a) of the next step may already store rows
// base+RB+6 .. base+2RB+5 while other waves are still in (c).  The level-1
// ring holds RB + 6 rows: (b) of the next step runs only after the barrier
// that follows (a), when every wave has left (c).  Row indices are wave-uniform, so the ring slot
// arithmetic runs on the scalar unit.
// ---------------------------------------------------------------------------
constexpr int MTX = 58, MLX = 64, MWX = 70;

template <int RB>
struct MarchRing {
    static constexpr int WR = 2 * RB + 9;
    static constexpr int LR = RB + 6;
};

template <typename R>
static size_t march_smem(int NK, bool visc, int RB)
{
    const int WR = 2 * RB + 9, LR = RB + 6;
    const int NL = (visc ? 5 : 0) + (NK >= 5 ? 1 : 0);
    return sizeof(R) * ((size_t)(NK + (visc ? 1 : 0)) * WR * MWX + (size_t)NL * LR * MLX);
}

// one block's march: output columns [i0, i0 + nout) (nout <= 58), rows [y0, y1)
template <typename S, typename R, int NK, bool VISC, bool PERT, int RB, int PM>
__device__ __forceinline__ void march_body(const StageArgs<S> A, const int i0, const int nout, const int y0,
                                           const int y1)
{
    constexpr int NT = RB * 64, WR = MarchRing<RB>::WR, LR = MarchRing<RB>::LR;
    constexpr int NL = L1<NK, VISC>::N;
    constexpr int PF = (RB * MWX + NT - 1) / NT;  // prefetched window points per thread
    // element offsets: 32-bit in the FP32 form (128 instead of 151 VGPRs in
    // the fast-zone kernel, 4 waves per SIMD instead of 3; 6 % off the FP32
    // step), 64-bit in FP64 (32-bit offsets measured 2.7 % slower there)
    using IX = typename std::conditional<sizeof(S) == 4, uint32_t, size_t>::type;
    extern __shared__ char smem_raw[];
    R *sU = reinterpret_cast<R *>(smem_raw);  // NK x WR x MWX
    R *sT = sU + NK * WR * MWX;               // WR x MWX (VISC)
    R *sL = sT + (VISC ? WR * MWX : 0);       // NL x LR x MLX

    const int pitch = A.pitch;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ox = i0 - 6, lx0 = i0 - 3;  // window / level-1 column origins
    const int wy0 = y0 - 6, ly0 = y0 - 3;  // first window / level-1 row of the chunk
    const R gam = R(A.gamma);

    // window columns past the grid (partial last strip) load the last column;
    // periodic x: every window column wraps around the period
    const int cmax = A.imax - 1 - ox;
    const bool xper = A.xper;
    auto load_pt = [&](int r, int c, R *v) {
        const IX g = (IX)r * (IX)pitch + (IX)(xper ? wrap_col(ox + c, A.imax) : ox + (c < cmax ? c : cmax));
#pragma unroll
        for (int k = 0; k < NK; ++k) v[k] = R(*elt(A.U[k], g));
    };
    auto put_pt = [&](int r, int c, const R *v) {
        const int s = (r - wy0) % WR;
#pragma unroll
        for (int k = 0; k < NK; ++k) sU[(k * WR + s) * MWX + c] = v[k];
        if constexpr (VISC) {
            if constexpr (PERT) {
                const R rb = R(A.pr.ub[0][r]), pb = R(A.pr.ub[3][r]);
                sT[s * MWX + c] = gam * (v[3] * rb - pb * v[0]) / ((rb + v[0]) * rb);
            } else {
                sT[s * MWX + c] = gam * v[3] / v[0];
            }
        }
    };
    // accessors.  r, r0: wave-uniform rows; a y-accessor centred on row r0
    // takes the slot of r0 once and reaches rows r0 +- 6 by a wrapped offset
    // (the offsets are compile-time constants after inlining)
    auto wrap = [](int s, int n) { return s < 0 ? s + n : (s >= n ? s - n : s); };
    auto ux_at = [=](int r) {
        const int s = (r - wy0) % WR;
        return [=](int k) {
            const R *row = sU + (k * WR + s) * MWX;
            return [row, ox](int q) { return row[q - ox]; };
        };
    };
    auto uy_at = [=](int gx, int r0) {
        const int s0 = (r0 - wy0) % WR;
        return [=](int k) {
            const R *col = sU + k * WR * MWX + (gx - ox);
            return [=](int q) { return col[wrap(s0 + (q - r0), WR) * MWX]; };
        };
    };
    auto tx_at = [=](int r) {
        const R *row = sT + ((r - wy0) % WR) * MWX;
        return [row, ox](int q) { return row[q - ox]; };
    };
    auto ty_at = [=](int gx, int r0) {
        const R *col = sT + (gx - ox);
        const int s0 = (r0 - wy0) % WR;
        return [=](int q) { return col[wrap(s0 + (q - r0), WR) * MWX]; };
    };
    const Transport<R> tc{R(A.tc.A_mi), R(A.tc.Bt_mi), R(A.tc.Ct_mi), R(A.tc.D_mi), R(A.tc.inv_mi0),
                          R(A.tc.A_k),  R(A.tc.Bt_k),  R(A.tc.Ct_k),  R(A.tc.D_k),  R(A.tc.inv_k0),
                          R(A.tc.Tref)};
    // metric factors of this lane's level-1 and output columns (fixed for
    // the whole march)
    constexpr bool MET = PM != PM_FAST;
    const int cl = xper ? wrap_col(lx0 + lane, A.imax) : min(lx0 + lane, A.imax - 1);
    const int co = xper ? wrap_col(i0 + lane, A.imax) : min(i0 + lane, A.imax - 1);
    const R mxl = MET ? R(A.metx[cl]) : R(1);
    const R mxo = MET ? R(A.metx[co]) : R(1);
    // wave-uniform row r, column lx0 + lane; my: mety[r]
    auto level1_row = [&](int r, const RowCache<R, NK> *rc, R my) {
        const int gx = lx0 + lane;
        R out[NL > 0 ? NL : 1];
        level1_point<S, R, NK, VISC, PERT, PM>(A, tc, gx, r, mxl, my, ux_at(r), uy_at(gx, r), tx_at(r),
                                               ty_at(gx, r), VISC ? tx_at(r)(gx) : R(0), out, rc);
        const int s = (r - ly0) % LR;
#pragma unroll
        for (int l = 0; l < NL; ++l) sL[(l * LR + s) * MLX + lane] = out[l];
    };

    // prologue: window rows y0-6 .. y0+5, level-1 rows y0-3 .. y0+2
    for (int t = tid; t < 12 * MWX; t += NT) {
        const int r = wy0 + t / MWX, c = t % MWX;
        R v[NK];
        load_pt(r, c, v);
        put_pt(r, c, v);
    }
    __syncthreads();
    if constexpr (NL > 0) {
        for (int rr = wave; rr < 6; rr += RB) level1_row(ly0 + rr, nullptr, MET ? R(A.mety[ly0 + rr]) : R(1));
    }
    // window rows r0 .. r0+RB-1 into registers: every thread issues its PF
    // loads (clamped to the last point / row of the grid; the extra values
    // are never stored), so the count of loads in flight is static
    R pre[PF][NK];
    auto prefetch = [&](int r0) {
#pragma unroll
        for (int p = 0; p < PF; ++p) {
            const int t = min(tid + p * NT, RB * MWX - 1);
            load_pt(min(r0 + t / MWX, A.jmax - 1), t % MWX, pre[p]);
        }
    };
    prefetch(y0 + 6);

    for (int base = y0; base < y1; base += RB) {
        // (a) window rows base+6 .. base+RB+5, loaded during the previous step
#pragma unroll
        for (int p = 0; p < PF; ++p) {
            const int t = tid + p * NT;
            if (t < RB * MWX) put_pt(base + 6 + t / MWX, t % MWX, pre[p]);
        }
        __syncthreads();
        // this step's loads, issued after the wait of (a) so that (a) never
        // waits for them: per-row data of the level-1 and output rows (base
        // state of the perturbation form, metric factors), the store data of
        // the outputs, then the next step's window rows
        RowCache<R, NK> rcl, rco;
        if constexpr (PERT) {
            load_rows<S, R, NK>(A, base + 3 + wave, rcl);
            load_rows<S, R, NK>(A, base + wave, rco);
        }
        const R myl = MET ? R(A.mety[base + 3 + wave]) : R(1), myo = MET ? R(A.mety[base + wave]) : R(1);
        StorePre<S, NK> spre;
        store_prefetch<S, NK, PERT, PM, IX>(A, min(i0 + lane, A.imax - 1), base + wave, spre);
        prefetch(base + RB + 6);
        // (b) level-1 row base+3+wave
        if constexpr (NL > 0) {
            level1_row(base + 3 + wave, PERT ? &rcl : nullptr, myl);
            __syncthreads();
        }
        // (c) outputs of row base+wave
        const int gy = base + wave;
        if (lane < nout) {
            const int gx = i0 + lane;
            const int ls = (gy - ly0) % LR;
            auto lxa = [=](int l) {
                const R *row = sL + (l * LR + ls) * MLX;
                return [row, lx0](int q) { return row[q - lx0]; };
            };
            auto lya = [=](int l) {
                const R *col = sL + l * LR * MLX + (gx - lx0);
                return [=](int q) { return col[wrap(ls + (q - gy), LR) * MLX]; };
            };
            auto lv = [=](int l) { return sL[(l * LR + ls) * MLX + (gx - lx0)]; };
            R qc[NK], uc[NK], f[NK];
            rhs_point<S, R, NK, VISC, PERT, PM>(A, gx, gy, mxo, myo, ux_at(gy), uy_at(gx, gy), lxa, lya, lv, qc,
                                                uc, f, PERT ? &rco : nullptr);
            stage_store<S, R, NK, false, PERT, PM, IX>(A, gx, gy, qc, uc, f, &spre);
        }
    }
}

// PM_HYBRID: blocks whose outputs lie in the fast zone take PM_FAST, the
// others PM_MIDU (the same bits, as for every point mode)
constexpr int PM_HYBRID = 4;

// FP32 blocks of 4 rows per step take 36 KB of LDS (nk = 4), so 4 fit a
// CU: the registers must allow 4 waves per SIMD as well (the general-region
// kernel then keeps 128 VGPRs instead of 137; nk = 5 would spill)
template <typename S, int RB, int NK>
constexpr int march_waves_per_simd()
{
    return (sizeof(S) == 4 && RB == 4 && NK == 4) ? 4 : 1;
}

// the marching block of (strip, chunk)
template <typename S, typename R, int NK, bool VISC, bool PERT, int RB, int PM>
__device__ __forceinline__ void march_block_at(const StageArgs<S> &A, const MarchMap &m, int strip, int chunk)
{
    const int i0 = m.xa + strip * MTX;
    const int nout = min(MTX, m.xe - i0);
    if constexpr (PM == PM_HYBRID) {
        int y0, y1;
        march_chunk_rows(m, chunk, y0, y1);
        if (march_block_fast(m, i0, y0, y1))
            march_body<S, R, NK, VISC, PERT, RB, PM_FAST>(A, i0, nout, y0, y1);
        else
            march_body<S, R, NK, VISC, PERT, RB, PM_MIDU>(A, i0, nout, y0, y1);
    } else {
        // plain layout (y_a = ya, y_b = yb): chunks of H rows
        const int y0 = m.ya + chunk * m.H;
        march_body<S, R, NK, VISC, PERT, RB, PM>(A, i0, nout, y0, min(y0 + m.H, m.yb));
    }
}

// marching block L of T (march_tile_of)
template <typename S, typename R, int NK, bool VISC, bool PERT, int RB, int PM>
__device__ __forceinline__ void march_block(const StageArgs<S> &A, const MarchMap &m, int L, int T)
{
    int strip, chunk;
    march_tile_of(L, T, m.nstrips, strip, chunk);
    march_block_at<S, R, NK, VISC, PERT, RB, PM>(A, m, strip, chunk);
}

template <typename S, typename R, int NK, bool VISC, bool PERT, int RB, int PM>
__global__ __launch_bounds__(RB * 64, (march_waves_per_simd<S, RB, NK>())) void stage_march_kernel(StageArgs<S> A,
                                                                                           MarchMap m)
{
    march_block<S, R, NK, VISC, PERT, RB, PM>(A, m, (int)(blockIdx.x + blockIdx.y * gridDim.x),
                                              (int)(gridDim.x * gridDim.y));
}

// The marching blocks and, after them in the same launch, the small edge
// tiles around the marching region (same 256-thread blocks; the dynamic LDS
// is the larger of the two layouts, the march's).  Blocks start in order, so
// the edge tiles take the slots the marching blocks leave and then those
// they free: one launch per stage, without a second stream's fork and join.
template <typename S, typename R, int NK, bool VISC, bool PERT, int RB, int PM, int TY>
__global__ __launch_bounds__(RB * 64, (march_waves_per_simd<S, RB, NK>())) void stage_march_edges_kernel(
    StageArgs<S> A, MarchMap m, RectSet sides, RectSet bands, int nmarch)
{
    static_assert(RB * 64 == EdgeShape<true, TY>::NT, "marching blocks and small edge tiles share the block size");
    const int L = (int)blockIdx.x;
    if (L < nmarch)
        march_block<S, R, NK, VISC, PERT, RB, PM>(A, m, L, nmarch);
    else
        edge_block<S, R, NK, VISC, TY, false, PERT, true>(A, sides, bands, L - nmarch);
}

template <typename R>
static size_t stage_smem(int NK, bool visc, int TY, int TX = STX)
{
    const int WX = TX + 12, WY = TY + 12, LX = TX + 6, LY = TY + 6;
    const int NL = (visc ? 5 : 0) + (NK >= 5 ? 1 : 0);
    return sizeof(R) * ((size_t)NK * WY * WX + (visc ? (size_t)WY * WX : 0) + (size_t)NL * LY * LX);
}

// The marching region: output columns [xa, xa + 58 n) and rows [ya, yb),
// rows a multiple of RB (march_region, march_hybrid); chunk height H from a
// simple occupancy model (blocks per wave of the GPU x rows per block).
// False when there is no such region (e is then hipSuccess) or when the
// marching kernel could not be set up (e says why).
template <typename S, int NK, bool VISC, bool PERT, int RB, int PM>
static bool march_plan(const StageArgs<S> &A, MarchMap &m, int &nchunks, bool &one_wave, hipStream_t s,
                       hipError_t &e)
{
    constexpr bool FAST = PM == PM_FAST;
    e = hipSuccess;
    if ((FAST || PM == PM_HYBRID) && !A.fast_ok) return false;
    // the FP32 marching kernel forms element offsets in 32 bits
    if (sizeof(S) == 4 && (size_t)A.pitch * (size_t)A.jmax * sizeof(S) >= ((size_t)1 << 32)) return false;
    if constexpr (PM == PM_HYBRID) {
        if (!march_hybrid(A.D, A.imax, A.jmax, RB, m, A.xper)) return false;
    } else {
        if (!march_region(FAST, A.D, A.imax, A.jmax, RB, m, A.xper)) return false;
    }
    // chunks of H rows over [y_a, y_b); the edge parts are one chunk each
    // (march_chunks, march_chunk_rows)
    const int rows = m.y_b - m.y_a;
    const int edges = (m.y_a > m.ya ? 1 : 0) + (m.yb > m.y_b ? 1 : 0);
    KernelSetup ks;
    if ((e = kernel_ready((const void *)stage_march_kernel<S, S, NK, VISC, PERT, RB, PM>, march_smem<S>(NK, VISC, RB),
                          RB * 64, s, &ks)) != hipSuccess)
        return false;
    const long slots = (long)ks.ncu * ks.perblk;
    long best = -1;
    for (int nc = 1; nc <= rows / (4 * RB); ++nc) {
        const int H = ((rows + nc - 1) / nc + RB - 1) / RB * RB;
        const int n = (rows + H - 1) / H + edges;
        const long blocks = (long)m.nstrips * n;
        const long cost = ((blocks + slots - 1) / slots) * (H + 8);
        if (best < 0 || cost < best || (cost == best && n > nchunks)) {
            best = cost;
            m.H = H;
            nchunks = n;
        }
    }
    one_wave = (long)m.nstrips * nchunks <= slots;
    return true;
}

// The stage: the interior region on stream s -- the marching kernel (MARCH)
// or fast tiles -- and the general
// tile kernel on the rest of the grid in one launch (stream s2 when given, so
// the two run concurrently; the caller orders s2 after s and joins it back).
// The RHS-only mode runs the general tiles everywhere.
template <typename S, int NK, bool VISC, bool RHS_ONLY, bool PERT, int NT, int RB, bool MARCH, bool GENERAL = false>
static hipError_t launch_stage_t(const StageArgs<S> &A, hipStream_t s, hipStream_t s2, StageHooks *tm)
{
    using R = S;
    constexpr int TY = (sizeof(S) == 8 && NK >= 5 && VISC) ? 8 : 16;
    MarchMap m{};
    int nchunks = 0;
    int fx0 = 0, fx1 = 0, fy0 = 0, fy1 = 0;
    bool inner = false;
    int pm = PM_GEN;
    bool one_wave = false;  // every marching block resident at once
    int xa = 0, xb = 0, ya = 0, yb = 0;  // interior region; the edge tiles cover the rest
    if constexpr (MARCH && !RHS_ONLY) {
        // the fast zone (GENERAL: the general region with its fast blocks)
        // when there is one, else the general region (interior stencils;
        // unit spacing drops the 1/h factors)
        const int gpm = (A.inv_dm == 1.0 && A.inv_dn == 1.0) ? PM_MIDU : PM_MID;
        pm = gpm;
        hipError_t pe = hipSuccess;
        if (A.fast_ok) {
            if constexpr (GENERAL) {
                pm = PM_HYBRID;
                inner = march_plan<S, NK, VISC, PERT, RB, PM_HYBRID>(A, m, nchunks, one_wave, s, pe);
            } else {
                pm = PM_FAST;
                inner = march_plan<S, NK, VISC, PERT, RB, PM_FAST>(A, m, nchunks, one_wave, s, pe);
            }
        }
        if (pe != hipSuccess) return pe;
        if (!inner) {
            pm = gpm;
            inner = pm == PM_MIDU ? march_plan<S, NK, VISC, PERT, RB, PM_MIDU>(A, m, nchunks, one_wave, s, pe)
                                  : march_plan<S, NK, VISC, PERT, RB, PM_MID>(A, m, nchunks, one_wave, s, pe);
            if (pe != hipSuccess) return pe;
        }
        if (inner) xa = m.xa, xb = m.xe, ya = m.ya, yb = m.yb;
    } else {
        fast_range(RHS_ONLY ? 0 : A.fast_ok, A.D, A.imax, STX, fx0, fx1);
        fast_range(RHS_ONLY ? 0 : A.fast_ok, A.D, A.jmax, TY, fy0, fy1);
        inner = fx1 > fx0 && fy1 > fy0;
        if (inner) xa = fx0 * STX, xb = fx1 * STX, ya = fy0 * TY, yb = fy1 * TY;
    }
    hipError_t e;
    using EL = EdgeShape<false, TY>;
    using ES = EdgeShape<true, TY>;
    if constexpr (MARCH && !RHS_ONLY && RB * 64 == ES::NT) {
        // one launch: the marching blocks, then the small edge tiles
        // (stage_march_edges_kernel); the whole grid is its work.  The edge
        // tiles fill the slots the march leaves and frees, without the second
        // stream's fork and join; which form is faster depends on the edge
        // work and on the waves of marching blocks (4096^2: FP32 with its
        // 44-point ring 4.08 -> 4.00 ms per step, FP64 with its 6-point band
        // 6.63 -> 6.69 ms), so the autotuner picks (same_launch).
        const bool plan = tm && tm->plan_only;
        if (inner && tm && (tm->same_launch || plan)) {
            const RectSet sides = RectSet::around(A.imax, A.jmax, ES::SX, ES::SY, xa, xb, ya, yb, 12);
            const RectSet bands = RectSet::around(A.imax, A.jmax, ES::BX, ES::BY, xa, xb, ya, yb, 3);
            size_t sm = march_smem<R>(NK, VISC, RB);
            sm = std::max(sm, stage_smem<R>(NK, VISC, ES::SY, ES::SX));
            sm = std::max(sm, stage_smem<R>(NK, VISC, ES::BY, ES::BX));
            void (*comb)(StageArgs<S>, MarchMap, RectSet, RectSet, int) =
                pm == PM_MIDU ? stage_march_edges_kernel<S, R, NK, VISC, PERT, RB, PM_MIDU, TY>
                              : stage_march_edges_kernel<S, R, NK, VISC, PERT, RB, PM_MID, TY>;
            if constexpr (GENERAL) {
                if (pm == PM_HYBRID) comb = stage_march_edges_kernel<S, R, NK, VISC, PERT, RB, PM_HYBRID, TY>;
            } else {
                if (pm == PM_FAST) comb = stage_march_edges_kernel<S, R, NK, VISC, PERT, RB, PM_FAST, TY>;
            }
            if ((e = kernel_ready((const void *)comb, sm, RB * 64, s)) != hipSuccess) return e;
            if (!plan) {
                tm->interior_points = (long long)A.imax * A.jmax;
                const int nmarch = m.nstrips * nchunks;
                if (tm->before && (e = hipEventRecord(tm->before, s)) != hipSuccess) return e;
                hipLaunchKernelGGL(comb, dim3(nmarch + sides.count() + bands.count()), dim3(RB * 64), sm, s, A, m,
                                   sides, bands, nmarch);
                if ((e = hipGetLastError()) != hipSuccess) return e;
                if (tm->after && (e = hipEventRecord(tm->after, s)) != hipSuccess) return e;
                return hipSuccess;
            }
        }
    }
    // edge tiles (stage_kernel_edges): side strips, then the bands.  Small
    // tiles when the marching blocks all run in one wave: LDS freed by a
    // finished marching block is then not needed again, and a small edge
    // tile can start there beside the other one.  With more waves that LDS
    // goes to the next marching blocks, and large tiles wait for whole CUs.
    const bool small = inner && one_wave;
    const RectSet sides = small ? RectSet::around(A.imax, A.jmax, ES::SX, ES::SY, xa, xb, ya, yb, 12)
                                : RectSet::around(A.imax, A.jmax, EL::SX, EL::SY, xa, xb, ya, yb, inner ? 12 : 0);
    const RectSet bands = small ? RectSet::around(A.imax, A.jmax, ES::BX, ES::BY, xa, xb, ya, yb, 3)
                                : RectSet::around(A.imax, A.jmax, EL::BX, EL::BY, xa, xb, ya, yb, inner ? 3 : 15);
    const size_t smem = stage_smem<R>(NK, VISC, TY);
    size_t esm, sb;
    if (small) {
        esm = stage_smem<R>(NK, VISC, ES::SY, ES::SX);
        sb = stage_smem<R>(NK, VISC, ES::BY, ES::BX);
    } else {
        esm = stage_smem<R>(NK, VISC, EL::SY, EL::SX);
        sb = stage_smem<R>(NK, VISC, EL::BY, EL::BX);
    }
    esm = sb > esm ? sb : esm;
    void (*gen)(StageArgs<S>, RectSet, RectSet) = stage_kernel_edges<S, R, NK, VISC, TY, RHS_ONLY, PERT, false>;
    if constexpr (MARCH && !RHS_ONLY) {
        if (small) gen = stage_kernel_edges<S, R, NK, VISC, TY, RHS_ONLY, PERT, true>;
    }
    const int ent = small ? ES::NT : EL::NT;
    if ((e = kernel_ready((const void *)gen, esm, ent, s)) != hipSuccess) return e;
    // An edge tile needs most of a CU's LDS, so it cannot share a CU with
    // interior blocks.  The interior kernel is launched first; the edge
    // tiles, launched after it on s2, take CUs as the interior blocks drain
    // them instead of holding CUs the long-running interior blocks wait for.
    auto edges = [&](hipStream_t st) -> hipError_t {
        const int n = sides.count() + bands.count();
        if (n > 0) {
            hipLaunchKernelGGL(gen, dim3(n), dim3(ent), esm, st, A, sides, bands);
            return hipGetLastError();
        }
        return hipSuccess;
    };
    if (tm) {
        if constexpr (MARCH && !RHS_ONLY)
            tm->interior_points = inner ? (long long)(m.xe - m.xa) * (m.yb - m.ya) : 0;
        else
            tm->interior_points = inner ? (long long)(fx1 - fx0) * STX * (fy1 - fy0) * TY : 0;
    }
    // the interior kernel and its LDS attribute; every runtime call that may
    // load a kernel happens before the first launch (plan_only stops there)
    void (*march)(StageArgs<S>, MarchMap) = nullptr;
    size_t msm = 0;
    if constexpr (!RHS_ONLY && MARCH) {
        march = pm == PM_MIDU ? stage_march_kernel<S, R, NK, VISC, PERT, RB, PM_MIDU>
                              : stage_march_kernel<S, R, NK, VISC, PERT, RB, PM_MID>;
        if constexpr (GENERAL) {
            if (pm == PM_HYBRID) march = stage_march_kernel<S, R, NK, VISC, PERT, RB, PM_HYBRID>;
        } else {
            if (pm == PM_FAST) march = stage_march_kernel<S, R, NK, VISC, PERT, RB, PM_FAST>;
        }
        msm = march_smem<R>(NK, VISC, RB);
        if (inner && (e = kernel_ready((const void *)march, msm, RB * 64, s)) != hipSuccess) return e;
    } else if constexpr (!RHS_ONLY) {
        if (inner && (e = kernel_ready((const void *)stage_kernel_fast<S, R, NK, VISC, TY, PERT, NT>, smem, NT, s)) !=
                         hipSuccess)
            return e;
    }
    if (tm && tm->plan_only) return hipSuccess;
    if (!inner) return edges(s);
    // the edge tiles on s2 start after the work queued on s so far (fork)
    // and s continues after them (join)
    const bool side = s2 && tm && tm->fork && tm->join;
    if (side) {
        if ((e = hipEventRecord(tm->fork, s)) != hipSuccess) return e;
        if ((e = hipStreamWaitEvent(s2, tm->fork, 0)) != hipSuccess) return e;
    }
    if (tm && tm->before && (e = hipEventRecord(tm->before, s)) != hipSuccess) return e;
    if constexpr (!RHS_ONLY && MARCH) {
        hipLaunchKernelGGL(march, dim3(m.nstrips, nchunks), dim3(RB * 64), msm, s, A, m);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (tm && tm->after && (e = hipEventRecord(tm->after, s)) != hipSuccess) return e;
    } else if constexpr (!RHS_ONLY) {
        hipLaunchKernelGGL((stage_kernel_fast<S, R, NK, VISC, TY, PERT, NT>), dim3(fx1 - fx0, fy1 - fy0), dim3(NT),
                           smem, s, A, fx0, fy0);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (tm && tm->after && (e = hipEventRecord(tm->after, s)) != hipSuccess) return e;
    }
    if (!side) return edges(s);
    if ((e = edges(s2)) != hipSuccess) return e;
    if ((e = hipEventRecord(tm->join, s2)) != hipSuccess) return e;
    return hipStreamWaitEvent(s, tm->join, 0);
}

// variants (same arithmetic):
//   FP64: 0 -> marching 4 rows/step over the general region [6, n-6)^2
//         (buffers and metrics point by point; blocks inside the fast zone
//         take the fast code), the thin edge band concurrent on s2;
//         1 -> marching 4 rows/step over the fast zone (the general region
//         when there is none), the edge ring concurrent on s2;
//         2 -> as 0 with 8 rows/step
//   FP32: 0 -> interior tiles (512 threads) over the fast zone, general tiles
//         elsewhere; 1 -> marching 4 rows/step over the general region;
//         2 -> as 0 of FP64
// FP64 results are bitwise identical across variants (the general per-point
// code at unit metric and spacing performs the same operations); in FP32
// (contraction allowed) the tile and marching kernels may round differently.
template <typename S, int NK, bool VISC, bool RHS_ONLY, bool PERT>
static hipError_t launch_stage_v(const StageArgs<S> &A, int variant, hipStream_t s, hipStream_t s2, StageHooks *tm)
{
    if constexpr (sizeof(S) == 8) {
        switch (variant) {
        case 1: return launch_stage_t<S, NK, VISC, RHS_ONLY, PERT, 512, 4, true>(A, s, s2, tm);
        case 2: return launch_stage_t<S, NK, VISC, RHS_ONLY, PERT, 512, 8, true, true>(A, s, s2, tm);
        default: return launch_stage_t<S, NK, VISC, RHS_ONLY, PERT, 512, 4, true, true>(A, s, s2, tm);
        }
    } else {
        switch (variant) {
        case 1: return launch_stage_t<S, NK, VISC, RHS_ONLY, PERT, 512, 4, true, true>(A, s, s2, tm);
        case 2: return launch_stage_t<S, NK, VISC, RHS_ONLY, PERT, 512, 4, true>(A, s, s2, tm);
        default: return launch_stage_t<S, NK, VISC, RHS_ONLY, PERT, 512, 4, false>(A, s, s2, tm);
        }
    }
}

template <typename S, bool PERT>
static hipError_t launch_stage_p(const StageArgs<S> &A, int nk, bool visc, bool rhs_only, int variant, hipStream_t s,
                                 hipStream_t s2, StageHooks *tm)
{
    if (nk == 4) {
        if (visc)
            return rhs_only ? launch_stage_v<S, 4, true, true, PERT>(A, variant, s, s2, tm)
                            : launch_stage_v<S, 4, true, false, PERT>(A, variant, s, s2, tm);
        return rhs_only ? launch_stage_v<S, 4, false, true, PERT>(A, variant, s, s2, tm)
                        : launch_stage_v<S, 4, false, false, PERT>(A, variant, s, s2, tm);
    }
    if (visc)
        return rhs_only ? launch_stage_v<S, 5, true, true, PERT>(A, variant, s, s2, tm)
                        : launch_stage_v<S, 5, true, false, PERT>(A, variant, s, s2, tm);
    return rhs_only ? launch_stage_v<S, 5, false, true, PERT>(A, variant, s, s2, tm)
                    : launch_stage_v<S, 5, false, false, PERT>(A, variant, s, s2, tm);
}

template <typename S, int NK, bool VISC, bool PERT, int NT, bool FAST>
static hipError_t launch_stage_batch_t(const StageArgs<S> *args, const int4 *tiles, int ntiles, int stage, int ncases,
                                       hipStream_t s)
{
    using R = S;
    constexpr int TY = (sizeof(S) == 8 && NK >= 5 && VISC) ? 8 : 16;
    if (ntiles <= 0) return hipSuccess;
    size_t smem = stage_smem<R>(NK, VISC, TY);
    auto kern = stage_kernel_batch<S, R, NK, VISC, TY, PERT, NT, FAST>;
    hipError_t e = kernel_ready((const void *)kern, smem, NT, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(ntiles), dim3(NT), smem, s, args, tiles, stage, ncases);
    return hipGetLastError();
}

template <typename S, bool PERT, int NK, bool VISC, int NT>
static hipError_t launch_stage_batch_n(const StageArgs<S> *args, const int4 *ftiles, int nf, const int4 *gtiles,
                                       int ng, int stage, int ncases, hipStream_t s, hipStream_t s2)
{
    hipError_t e = launch_stage_batch_t<S, NK, VISC, PERT, NT, false>(args, gtiles, ng, stage, ncases, s2 ? s2 : s);
    if (e != hipSuccess) return e;
    return launch_stage_batch_t<S, NK, VISC, PERT, NT, true>(args, ftiles, nf, stage, ncases, s);
}

template <typename S, bool PERT, int NK, bool VISC>
static hipError_t launch_stage_batch_v(const StageArgs<S> *args, const int4 *ftiles, int nf, const int4 *gtiles,
                                       int ng, int stage, int ncases, int variant, hipStream_t s,
                                       hipStream_t s2)
{
    switch (variant) {
    case 1: return launch_stage_batch_n<S, PERT, NK, VISC, 512>(args, ftiles, nf, gtiles, ng, stage, ncases, s, s2);
    case 2: return launch_stage_batch_n<S, PERT, NK, VISC, 1024>(args, ftiles, nf, gtiles, ng, stage, ncases, s, s2);
    default: return launch_stage_batch_n<S, PERT, NK, VISC, 256>(args, ftiles, nf, gtiles, ng, stage, ncases, s, s2);
    }
}

template <typename S, bool PERT>
static hipError_t launch_stage_batch_p(const StageArgs<S> *args, const int4 *ftiles, int nf, const int4 *gtiles,
                                       int ng, int stage, int ncases, int nk, bool visc, int variant, hipStream_t s,
                                       hipStream_t s2)
{
    if (nk == 4)
        return visc ? launch_stage_batch_v<S, PERT, 4, true>(args, ftiles, nf, gtiles, ng, stage, ncases, variant, s, s2)
                    : launch_stage_batch_v<S, PERT, 4, false>(args, ftiles, nf, gtiles, ng, stage, ncases, variant, s, s2);
    return visc ? launch_stage_batch_v<S, PERT, 5, true>(args, ftiles, nf, gtiles, ng, stage, ncases, variant, s, s2)
                : launch_stage_batch_v<S, PERT, 5, false>(args, ftiles, nf, gtiles, ng, stage, ncases, variant, s, s2);
}

#if defined(NS_KERNELS_F64)
hipError_t launch_stage_batch(const StageArgs<double> *args, const int4 *ftiles, int nf, const int4 *gtiles, int ng,
                              int stage, int ncases, int nk, bool visc, int variant, hipStream_t s, hipStream_t s2)
{
    return launch_stage_batch_p<double, false>(args, ftiles, nf, gtiles, ng, stage, ncases, nk, visc, variant, s,
                                                 s2);
}
#endif
#if defined(NS_KERNELS_F32)
hipError_t launch_stage_batch(const StageArgs<float> *args, const int4 *ftiles, int nf, const int4 *gtiles, int ng,
                              int stage, int ncases, int nk, bool visc, int variant, hipStream_t s, hipStream_t s2)
{
    return launch_stage_batch_p<float, true>(args, ftiles, nf, gtiles, ng, stage, ncases, nk, visc, variant, s,
                                                 s2);
}
#endif

// ---------------------------------------------------------------------------
// Batched marching stage: one launch per stage over every case of a batch.
// Each block takes one work item (case, a, b, kind) of a table the host
// builds (batch_march_case gives each case's plan): kind 0 marches strip a,
// chunk b of the case's region (as stage_march_kernel, 4 rows per step),
// kind 1 computes edge tile a of the case's frame (BatchEdgeShape), kind 2
// is padding.
// Blocks go to the XCDs in turn (block L to XCD L % 8), and the table gives
// each XCD whole cases, so a case's neighbouring strips and consecutive
// chunks share one L2; the marching items come first, and the edge tiles
// after them take the slots the march leaves.  PM_HYBRID: blocks inside a
// case's fast zone take the fast code, the others PM_MIDU (a plain plan
// marks no fast zone); PM_MID: the general code with 1/h factors everywhere
// (batches with a case of non-unit spacing).
// ---------------------------------------------------------------------------
constexpr int BRB = 4;  // rows per marching step in batches

template <int NK, bool VISC>
constexpr bool batch_frame_tiles()
{
    return L1<NK, VISC>::N == 0;
}

template <typename S, typename R, int NK, bool VISC, bool PERT, int PM>
__global__ __launch_bounds__(BRB * 64, (march_waves_per_simd<S, BRB, NK>())) void stage_march_batch_kernel(
    const StageArgs<S> *args, const BatchMarch *plans, const int4 *items, int stage, int ncases)
{
    using E = BatchEdgeShape<batch_frame_tiles<NK, VISC>()>;
    const int4 t = items[blockIdx.x];
    if (t.w == 2) return;
    // copies, as in stage_kernel_batch
    const StageArgs<S> A = args[(size_t)stage * ncases + t.x];
    if (t.w == 0) {
        const MarchMap m = plans[t.x].m;
        march_block_at<S, R, NK, VISC, PERT, BRB, PM>(A, m, t.y, t.z);
    } else {
        const RectSet sides = plans[t.x].sides, bands = plans[t.x].bands;
        int i0, i1, j0, j1;
        if (t.y < sides.count()) {
            sides.tile(t.y, i0, i1, j0, j1);
            stage_tile<S, R, NK, VISC, E::SY, false, PERT, BRB * 64, PM_GEN, E::SX>(A, i0, i1, j0, j1);
        } else {
            bands.tile(t.y - sides.count(), i0, i1, j0, j1);
            stage_tile<S, R, NK, VISC, E::BY, false, PERT, BRB * 64, PM_GEN, E::BX>(A, i0, i1, j0, j1);
        }
    }
}

// The plan of one case (batch_march_plan, 4 rows per step): hybrid allows
// its fast zone to be marked (unit spacing and fast-zone conditions are
// checked here).  False when an FP32 plane is too large for 32-bit offsets.
template <typename S>
static bool batch_march_case_t(const StageArgs<S> &A, int nk, bool visc, bool hybrid, int H, BatchMarch &p)
{
    if (sizeof(S) == 4 && (size_t)A.pitch * (size_t)A.jmax * sizeof(S) >= ((size_t)1 << 32)) return false;
    const bool unit = A.inv_dm == 1.0 && A.inv_dn == 1.0;
    hybrid = hybrid && unit && A.fast_ok;
    if (nk == 4 && !visc)
        batch_march_plan<batch_frame_tiles<4, false>()>(A.imax, A.jmax, A.D, A.xper, hybrid, BRB, H, p);
    else
        batch_march_plan<false>(A.imax, A.jmax, A.D, A.xper, hybrid, BRB, H, p);
    return true;
}

template <typename S, bool PERT, int NK, bool VISC, int PM>
static hipError_t launch_stage_march_batch_t(const StageArgs<S> *args, const BatchMarch *plans, const int4 *items,
                                             int nitems, int stage, int ncases, hipStream_t s)
{
    using R = S;
    using E = BatchEdgeShape<batch_frame_tiles<NK, VISC>()>;
    size_t sm = march_smem<R>(NK, VISC, BRB);
    sm = std::max(sm, stage_smem<R>(NK, VISC, E::SY, E::SX));
    sm = std::max(sm, stage_smem<R>(NK, VISC, E::BY, E::BX));
    auto kern = stage_march_batch_kernel<S, R, NK, VISC, PERT, PM>;
    hipError_t e = kernel_ready((const void *)kern, sm, BRB * 64, s);
    if (e != hipSuccess || nitems <= 0) return e;
    hipLaunchKernelGGL(kern, dim3(nitems), dim3(BRB * 64), sm, s, args, plans, items, stage, ncases);
    return hipGetLastError();
}

template <typename S, bool PERT, int NK, bool VISC>
static hipError_t launch_stage_march_batch_v(const StageArgs<S> *args, const BatchMarch *plans, const int4 *items,
                                             int nitems, int stage, int ncases, bool unit, hipStream_t s)
{
    return unit ? launch_stage_march_batch_t<S, PERT, NK, VISC, PM_HYBRID>(args, plans, items, nitems, stage, ncases, s)
                : launch_stage_march_batch_t<S, PERT, NK, VISC, PM_MID>(args, plans, items, nitems, stage, ncases, s);
}

template <typename S, bool PERT>
static hipError_t launch_stage_march_batch_p(const StageArgs<S> *args, const BatchMarch *plans, const int4 *items,
                                             int nitems, int stage, int ncases, int nk, bool visc, bool unit,
                                             hipStream_t s)
{
    if (nk == 4)
        return visc ? launch_stage_march_batch_v<S, PERT, 4, true>(args, plans, items, nitems, stage, ncases, unit, s)
                    : launch_stage_march_batch_v<S, PERT, 4, false>(args, plans, items, nitems, stage, ncases, unit, s);
    return visc ? launch_stage_march_batch_v<S, PERT, 5, true>(args, plans, items, nitems, stage, ncases, unit, s)
                : launch_stage_march_batch_v<S, PERT, 5, false>(args, plans, items, nitems, stage, ncases, unit, s);
}

#if defined(NS_KERNELS_F64)
bool batch_march_case(const StageArgs<double> &A, int nk, bool visc, bool hybrid, int H, BatchMarch &p)
{
    return batch_march_case_t(A, nk, visc, hybrid, H, p);
}
hipError_t launch_stage_march_batch(const StageArgs<double> *args, const BatchMarch *plans, const int4 *items,
                                    int nitems, int stage, int ncases, int nk, bool visc, bool unit, hipStream_t s)
{
    return launch_stage_march_batch_p<double, false>(args, plans, items, nitems, stage, ncases, nk, visc, unit, s);
}
#endif
#if defined(NS_KERNELS_F32)
bool batch_march_case(const StageArgs<float> &A, int nk, bool visc, bool hybrid, int H, BatchMarch &p)
{
    return batch_march_case_t(A, nk, visc, hybrid, H, p);
}
hipError_t launch_stage_march_batch(const StageArgs<float> *args, const BatchMarch *plans, const int4 *items,
                                    int nitems, int stage, int ncases, int nk, bool visc, bool unit, hipStream_t s)
{
    return launch_stage_march_batch_p<float, true>(args, plans, items, nitems, stage, ncases, nk, visc, unit, s);
}
#endif

#if defined(NS_KERNELS_F64)
hipError_t launch_stage(const StageArgs<double> &A, int nk, bool visc, bool rhs_only, int variant, hipStream_t s,
                        hipStream_t s2, StageHooks *tm)
{
    return launch_stage_p<double, false>(A, nk, visc, rhs_only, variant, s, s2, tm);
}
#endif

#if defined(NS_KERNELS_F32)
hipError_t launch_stage(const StageArgs<float> &A, int nk, bool visc, bool rhs_only, int variant, hipStream_t s,
                        hipStream_t s2, StageHooks *tm)
{
    return launch_stage_p<float, true>(A, nk, visc, rhs_only, variant, s, s2, tm);
}
#endif

// ---------------------------------------------------------------------------
// In-place selective filter as a chunked sweep with look-back.
//
// The reference sweeps each line in place (filtering.f90): a point's update
// reads the already-updated points behind it, so the interior is a linear
// recurrence with 6 feedback taps sigma*d(-1..-6).  Its homogeneous response
// decays like 0.33^n at sigma=0.2, so a chunk starting at position a can start
// the recurrence W points earlier from unfiltered values: the error left at a
// is below 1e-16 of the filter correction for the W chosen by the host.
// Chunks that start within W+12 of the left end run the exact sweep from the
// start (left block + interior); the chunk that holds the right end also runs
// the descending right block (filtering.f90:81-94).
//
// One lane filters one line.  Window w[13] holds positions q-6..q+6 with
// slot (u+s)%13 <-> position q-6+s at unroll step u, so no register moves.
// ---------------------------------------------------------------------------

template <int ROT, typename R>
__device__ __forceinline__ void unrot13(const R (&w)[13], R (&v)[13])
{
#pragma unroll
    for (int s = 0; s < 13; ++s) v[s] = w[(ROT + s) % 13];
}

template <typename R>
__device__ __forceinline__ void unrotate(int rot, const R (&w)[13], R (&v)[13])
{
    switch (rot) {
    case 0: unrot13<0>(w, v); break;
    case 1: unrot13<1>(w, v); break;
    case 2: unrot13<2>(w, v); break;
    case 3: unrot13<3>(w, v); break;
    case 4: unrot13<4>(w, v); break;
    case 5: unrot13<5>(w, v); break;
    case 6: unrot13<6>(w, v); break;
    case 7: unrot13<7>(w, v); break;
    case 8: unrot13<8>(w, v); break;
    case 9: unrot13<9>(w, v); break;
    case 10: unrot13<10>(w, v); break;
    case 11: unrot13<11>(w, v); break;
    default: unrot13<12>(w, v); break;
    }
}

struct LineEnds {
    double v1, v2;      // final values at positions 1, 2 (exact-start chunk)
    double vn2, vn3;    // final values at positions n-2, n-3 (last chunk)
};

// IO policy contract: load(p) for strictly increasing p (p < n; RING: p may
// start at -6, which reads the unfiltered point n + p), emit(e, v) for
// strictly increasing e (positions outside [a, b) are dropped), finish() at
// the end; RING also needs load_head(p), p in [0, 12), a direct read of the
// unfiltered point p.  PD: groups of 13 loads issued ahead of use (2 for
// direct global loads, so that a load has about 26 steps of work to arrive;
// 1 where the IO prefetches itself).
//
// RING (periodic x, no reference counterpart): the interior update sweeps
// around the ring from position 0: the exact-start chunk begins at q = 0
// with the unfiltered points n-6..n-1 on its left; the chunk holding the
// end updates positions n-6..n-1 last, reading on its right the points 0..5
// as the start of the sweep updated them, which it recomputes from the
// unfiltered points n-6..n-1 and 0..11 (the same operations, so the same
// bits).
template <typename R, bool YDIR, int PD, typename IO, bool RING = false>
__device__ __forceinline__ void filter_line(IO &io, R sigma, int n, int a, int b, int W, LineEnds &ends)
{
    const bool last = (b == n);
    const bool exact = (a <= W + 12);
    R w[13];
    int q;
    if (RING && exact) {
#pragma unroll
        for (int s = 0; s < 13; ++s) w[s] = io.load(s - 6);
        q = 0;
    } else if (exact) {
#pragma unroll
        for (int s = 0; s < 13; ++s) w[s] = io.load(s);
        // left block, ascending (filtering.f90:36-49)
#pragma unroll
        for (int qq = 1; qq <= 6; ++qq) {
            R Du = R(filt_dfb<YDIR>(0)) * w[qq - 1];
#pragma unroll
            for (int s = 1; s < 7; ++s) Du = Du + R(filt_dfb<YDIR>(s)) * w[qq - 1 + s];
            w[qq] = w[qq] - sigma * Du;
        }
        ends.v1 = (double)w[1];
        ends.v2 = (double)w[2];
        q = 6;
    } else {
        const int q0 = a - W;
#pragma unroll
        for (int s = 0; s < 13; ++s) w[s] = io.load(q0 - 6 + s);
        q = q0;
    }
    // RING: the chunk holding the end stops its loop at n - 7 as well, and
    // updates n-6..n-1 after it against the head
    const int qe = last ? n - 7 : b - 1;
    // loads for the steps of the next PD groups of 13 are in flight.  Every
    // load is issued, at a clamped position: a value fetched past the line
    // end or past the last step is never used, and branch-free loads let
    // the compiler keep later loads outstanding while waiting for earlier ones.
    const int lim = qe + 6 < n - 1 ? qe + 6 : n - 1;  // last position a refill uses
    auto ld = [&](int p) { return io.load(p < lim ? p : lim); };
    R nb[13], nn[13];
#pragma unroll
    for (int u = 0; u < 13; ++u) nb[u] = ld(q + 7 + u);
    if constexpr (PD == 2) {
#pragma unroll
        for (int u = 0; u < 13; ++u) nn[u] = ld(q + 20 + u);
    }
    // the next burst of loads (positions q + 13 PD + 7 ..) into nb / nn; cur:
    // the burst this block of 13 steps consumes
    auto refill = [&](R (&cur)[13]) {
#pragma unroll
        for (int u = 0; u < 13; ++u) cur[u] = nb[u];
        const int qn = q + 13 * PD;
        if constexpr (PD == 2) {
#pragma unroll
            for (int u = 0; u < 13; ++u) nb[u] = nn[u];
#pragma unroll
            for (int u = 0; u < 13; ++u) nn[u] = ld(qn + 7 + u);
        } else {
            io.burst(qn + 7);
#pragma unroll
            for (int u = 0; u < 13; ++u) nb[u] = ld(qn + 7 + u);
        }
    };
    // Whole blocks of 13 steps that end before qe need no exit test per
    // step: the warm-up blocks emit nothing (every position they finish lies
    // before a), the block that reaches a emits through the checked path,
    // and the blocks after it emit unchecked (positions in [a, b) there).
    // The same update, in the same order, as the general loop below.
    auto bulk = [&](auto mode) {
        constexpr int M = decltype(mode)::value;  // 0: no emits, 1: checked, 2: unchecked
        R cur[13];
        refill(cur);
#pragma unroll
        for (int u = 0; u < 13; ++u) {
            R Du = R(filt_d(0)) * w[u % 13];
#pragma unroll
            for (int s = 1; s < 13; ++s) Du = Du + R(filt_d(s)) * w[(u + s) % 13];
            w[(u + 6) % 13] = w[(u + 6) % 13] - sigma * Du;
            if constexpr (M == 1) io.emit(q - 6 + u, w[u % 13]);
            if constexpr (M == 2) io.emit_fast(q - 6 + u, w[u % 13]);
            w[u % 13] = cur[u];
        }
        q += 13;
    };
    while (q + 13 <= qe && q + 6 < a) bulk(std::integral_constant<int, 0>{});
    if (q + 13 <= qe && q - 6 < a) bulk(std::integral_constant<int, 1>{});
    while (q + 13 <= qe) bulk(std::integral_constant<int, 2>{});
    int rot = 0;
    for (;;) {
        R cur[13];
#pragma unroll
        for (int u = 0; u < 13; ++u) cur[u] = nb[u];
        const int qn = q + 13 * PD;
        if constexpr (PD == 2) {
#pragma unroll
            for (int u = 0; u < 13; ++u) nb[u] = nn[u];
#pragma unroll
            for (int u = 0; u < 13; ++u) nn[u] = ld(qn + 7 + u);
        } else {
            io.burst(qn + 7);  // the IO may refill its source before a burst of 13 loads
#pragma unroll
            for (int u = 0; u < 13; ++u) nb[u] = ld(qn + 7 + u);
        }
        bool done = false;
#pragma unroll
        for (int u = 0; u < 13; ++u) {
            if (!done) {
                // interior 13-point update (filtering.f90:69-79)
                R Du = R(filt_d(0)) * w[u % 13];
#pragma unroll
                for (int s = 1; s < 13; ++s) Du = Du + R(filt_d(s)) * w[(u + s) % 13];
                w[(u + 6) % 13] = w[(u + 6) % 13] - sigma * Du;
                if (q == qe) {
                    rot = u;
                    done = true;
                } else {
                    io.emit(q - 6, w[u % 13]);
                    w[u % 13] = cur[u];
                    ++q;
                }
            }
        }
        if (done) break;
    }
    R v[13];  // positions qe-6 .. qe+6
    unrotate<R>(rot, w, v);
    if (RING && last) {
        // the head: positions 0..5 as the sweep's start updated them, from
        // the unfiltered positions n-6..n-1 (v[7..12], not yet updated here)
        // and 0..11
        R hv[18];  // positions -6 .. 11
#pragma unroll
        for (int s = 0; s < 6; ++s) hv[s] = v[7 + s];
#pragma unroll
        for (int s = 0; s < 12; ++s) hv[6 + s] = io.load_head(s);
#pragma unroll
        for (int qq = 0; qq < 6; ++qq) {
            R Du = R(filt_d(0)) * hv[qq];
#pragma unroll
            for (int s = 1; s < 13; ++s) Du = Du + R(filt_d(s)) * hv[qq + s];
            hv[qq + 6] = hv[qq + 6] - sigma * Du;
        }
        // positions n-6 .. n-1 in order; ext[s]: position n-13+s, the
        // points past the end are the updated head
        R ext[19];
#pragma unroll
        for (int s = 0; s < 13; ++s) ext[s] = v[s];
#pragma unroll
        for (int s = 0; s < 6; ++s) ext[13 + s] = hv[6 + s];
#pragma unroll
        for (int qq = 7; qq <= 12; ++qq) {
            R Du = R(filt_d(0)) * ext[qq - 6];
#pragma unroll
            for (int s = 1; s < 13; ++s) Du = Du + R(filt_d(s)) * ext[qq - 6 + s];
            ext[qq] = ext[qq] - sigma * Du;
        }
#pragma unroll
        for (int s = 0; s < 13; ++s) io.emit(qe - 6 + s, ext[s]);
    } else if (last) {
        // right block, descending, mirrored stencil (filtering.f90:81-94)
#pragma unroll
        for (int qq = 11; qq >= 6; --qq) {
            R Du = R(filt_dfb<YDIR>(0)) * v[qq + 1];
#pragma unroll
            for (int s = 0; s <= 5; ++s) Du = Du + R(filt_dfb<YDIR>(s + 1)) * v[qq - s];
            v[qq] = v[qq] - sigma * Du;
        }
        ends.vn2 = (double)v[11];
        ends.vn3 = (double)v[10];
#pragma unroll
        for (int s = 0; s < 13; ++s) io.emit(qe - 6 + s, v[s]);
    } else {
#pragma unroll
        for (int s = 0; s < 7; ++s) io.emit(qe - 6 + s, v[s]);
    }
    io.finish();
}

// x direction: lanes are rows; LDS-staged 64 x TW tiles keep global access
// coalesced (lanes read/write TW-column row segments).  The x sweep takes
// one 128-byte line per row segment (16 columns in FP64, 32 in FP32: the
// 64-byte FP32 segments fetched each line twice, FP32 x sweep 0.172 ->
// 0.152 ms at 4096^2); the fused sweep's x phase keeps 16 columns (its
// registers do not allow 32).
template <typename R, int TW = 16>
struct XIO {
    static constexpr int T = TW;
    static constexpr int RP = 64 / T;  // rows per pass
    const R *in;
    R *out;
    int pitch, n, m, row0, a, b, lane;
    bool ring;  // periodic line: positions below 0 wrap to n + p
    R (*tin)[T + 1];
    R (*tout)[T + 1];
    int tc;  // start of the tile in LDS
    int oc;  // start of the output tile being filled (-1 none)
    R pre[T];  // the next tile, in flight (a second one costs more occupancy than it saves)

    __device__ __forceinline__ void fetch(R (&dst)[T], int t)  // issue loads of tile t into registers
    {
        // clamped indices, every load issued: values outside the grid land in
        // tile slots that are never read
        const int c = lane % T, rs = lane / T;
        int gi = t + c;
        gi = gi < 0 ? (ring ? gi + n : 0) : (gi < n ? gi : n - 1);
#pragma unroll
        for (int rr = 0; rr < T; ++rr) {
            int gj = row0 + rr * RP + rs;
            gj = gj < m ? gj : m - 1;
            dst[rr] = in[(size_t)gj * pitch + gi];
        }
    }
    __device__ __forceinline__ void commit()
    {
        const int c = lane % T, rs = lane / T;
        __syncthreads();
#pragma unroll
        for (int rr = 0; rr < T; ++rr) tin[rr * RP + rs][c] = pre[rr];
        __syncthreads();
    }
    __device__ __forceinline__ void start(int p0)
    {
        tc = p0 & ~(T - 1);
        fetch(pre, tc);
        commit();
        fetch(pre, tc + T);
        oc = -1;
    }
    __device__ __forceinline__ R load(int p)
    {
        if (p >= tc + T) {  // uniform: all lanes stream the same positions
            commit();
            tc += T;
            fetch(pre, tc + T);
        }
        return tin[lane][p - tc];
    }
    __device__ __forceinline__ void flush()
    {
        if (oc < 0) return;
        const int c = lane % T, rs = lane / T;
        __syncthreads();
#pragma unroll
        for (int rr = 0; rr < T; ++rr) {
            const int gj = row0 + rr * RP + rs, gi = oc + c;
            if (gj < m && gi >= a && gi < b) out[(size_t)gj * pitch + gi] = tout[rr * RP + rs][c];
        }
        __syncthreads();
    }
    __device__ __forceinline__ void emit(int e, R v)
    {
        const int t = e & ~(T - 1);
        if (t != oc) {
            flush();
            oc = t;
        }
        tout[lane][e - oc] = v;
    }
    __device__ __forceinline__ void emit_fast(int e, R v) { emit(e, v); }
    __device__ __forceinline__ void finish() { flush(); }
    __device__ __forceinline__ void burst(int) {}
    __device__ __forceinline__ R load_head(int p)  // this lane's row, position p (unstaged)
    {
        const int gj = row0 + lane < m ? row0 + lane : m - 1;
        return in[(size_t)gj * pitch + p];
    }
};

template <typename R>
using XSweepIO = XIO<R, 128 / sizeof(R)>;

template <typename R>
__device__ __forceinline__ void filter_x_tile(const FilterArgs<R> &F, int chunk, int rowblock, int k)
{
    constexpr int T = XSweepIO<R>::T;
    __shared__ R tin[64][T + 1];
    __shared__ R tout[64][T + 1];
    int a, b;
    if (!chunk_bounds(chunk, F.L, F.n, a, b)) return;
    XSweepIO<R> io;
    io.in = F.in[k];
    io.out = F.out[k];
    io.pitch = F.pitch;
    io.n = F.n;
    io.m = F.m;
    io.row0 = rowblock * 64;
    io.a = a;
    io.b = b;
    io.lane = threadIdx.x;
    io.tin = tin;
    io.tout = tout;
    io.ring = F.xper;
    const bool exact = (a <= F.W + 12);
    LineEnds ends;
    if (F.xper) {
        io.start(exact ? -6 : a - F.W - 6);
        filter_line<R, false, 1, XSweepIO<R>, true>(io, R(F.sigma), F.n, a, b, F.W, ends);
    } else {
        io.start(exact ? 0 : a - F.W - 6);
        filter_line<R, false, 1>(io, R(F.sigma), F.n, a, b, F.W, ends);
    }
}

template <typename R>
__global__ __launch_bounds__(64) void filter_x_kernel(FilterArgs<R> F)
{
    filter_x_tile<R>(F, blockIdx.x, blockIdx.y, blockIdx.z);
}

// batched: tiles[blockIdx.x] = (case, chunk, row block), blockIdx.y = variable
template <typename R>
__global__ __launch_bounds__(64) void filter_x_kernel_batch(const FilterArgs<R> *args, const int4 *tiles)
{
    const int4 t = tiles[blockIdx.x];
    const FilterArgs<R> &F = args[t.x];
    if ((int)blockIdx.y >= F.nk) return;
    filter_x_tile<R>(F, t.y, t.z, blockIdx.y);
}

// y direction: lanes are columns; rows are contiguous across lanes.
// CORR: the perturbation form adds Ly(Ub) - Ub of the row (corr, FP64);
// FULL: every lane's column lies in the grid
template <typename R, bool CORR, bool FULL>
struct YIO {
    const R *in;
    R *out;
    const double *corr;  // perturbation form: Ly(Ub) - Ub of this row (null in FP64)
    int pitch, col, a, b;
    bool active;
    int colc;  // col clamped into the grid: inactive lanes load a valid value and ignore it
    __device__ __forceinline__ R load(int p) { return in[(size_t)p * pitch + colc]; }
    __device__ __forceinline__ R corrected(int e, R v) const
    {
        if constexpr (CORR) return R(double(v) + corr[e]);
        else return v;
    }
    __device__ __forceinline__ void emit(int e, R v)
    {
        if ((FULL || active) && e >= a && e < b) out[(size_t)e * pitch + col] = corrected(e, v);
    }
    // e in [a, b)
    __device__ __forceinline__ void emit_fast(int e, R v)
    {
        if (FULL || active) out[(size_t)e * pitch + col] = corrected(e, v);
    }
    __device__ __forceinline__ void finish() {}
    __device__ __forceinline__ void burst(int) {}
    __device__ __forceinline__ R load_head(int p) { return in[(size_t)p * pitch + colc]; }
};

template <typename R, bool CORR, bool FULL>
__device__ __forceinline__ void filter_y_lines(const FilterArgs<R> &F, int col, int k, int a, int b, LineEnds &ends)
{
    YIO<R, CORR, FULL> io;
    io.in = F.in[k];
    io.out = F.out[k];
    io.corr = F.corr[k];
    io.pitch = F.pitch;
    io.col = col;
    io.colc = col < F.m ? col : F.m - 1;
    io.a = a;
    io.b = b;
    io.active = col < F.m;
    filter_line<R, true, 2>(io, R(F.sigma), F.n, a, b, F.W, ends);
}

template <typename R>
__device__ __forceinline__ void filter_y_tile(const FilterArgs<R> &F, int colblock, int chunk, int k)
{
    int a, b;
    if (!chunk_bounds(chunk, F.L, F.n, a, b)) return;
    const int col = colblock * 64 + (int)(threadIdx.x & 63);
    LineEnds ends;
    const bool full = colblock * 64 + 64 <= F.m;
    if (F.corr[k]) {
        if (full) filter_y_lines<R, true, true>(F, col, k, a, b, ends);
        else filter_y_lines<R, true, false>(F, col, k, a, b, ends);
    } else {
        if (full) filter_y_lines<R, false, true>(F, col, k, a, b, ends);
        else filter_y_lines<R, false, false>(F, col, k, a, b, ends);
    }
    const bool active = col < F.m;
    // ccpml rows j=1 and j=jmax for columns 2..imax-1 (cc.f90:46-59; periodic
    // x: every column, there are no inflow / outflow columns)
    if (F.edge_bc && active && (F.xper || (col >= 1 && col <= F.m - 2))) {
        R *o = F.out[k];
        const int jm = F.n;
        const double *cr = F.corr[k];
        if (a == 0) {
            double val;
            if (k < 3) {
                const double v1 = cr ? ends.v1 + cr[1] : ends.v1, v2 = cr ? ends.v2 + cr[2] : ends.v2;
                val = 2.0 * v1 - v2 + F.re.bot_res[k];
            } else {
                val = F.re.bot_val[k];
            }
            o[col] = R(val);
        }
        if (b == jm) {
            double val;
            if (k < 3) {
                const double v2 = cr ? ends.vn2 + cr[jm - 2] : ends.vn2;
                const double v3 = cr ? ends.vn3 + cr[jm - 3] : ends.vn3;
                val = 2.0 * v2 - v3 + F.re.top_res[k];
            } else {
                val = F.re.top_val[k];
            }
            o[(size_t)(jm - 1) * F.pitch + col] = R(val);
        }
    }
}

template <typename R>
__global__ __launch_bounds__(64 * filter_y_waves<R>()) void filter_y_kernel(FilterArgs<R> F)
{
    constexpr int NW = filter_y_waves<R>();
    const int cb = (int)blockIdx.x * NW + (int)(threadIdx.x >> 6);
    if (cb * 64 < F.m) filter_y_tile<R>(F, cb, blockIdx.y, blockIdx.z);
}

// batched: tiles[blockIdx.x] = (case, column block, chunk), blockIdx.y =
// variable; one wave per block (4-wave blocks over adjacent column blocks,
// as the single-grid FP32 sweep takes, measured no faster here: 204.9
// against 203.5 us per FP32 launch on the 256-case sweep)
template <typename R>
__global__ __launch_bounds__(64) void filter_y_kernel_batch(const FilterArgs<R> *args, const int4 *tiles)
{
    const int4 t = tiles[blockIdx.x];
    const FilterArgs<R> &F = args[t.x];
    if ((int)blockIdx.y >= F.nk) return;
    filter_y_tile<R>(F, t.y, t.z, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Both filters in one sweep (lddfilterx then lddfiltery, filtering.f90:9-190)
// for the plane of one variable: one wave owns the 64 columns [c0, c0+64) of
// a chunk of rows.  It walks its chunk's rows in groups of XY_G = 52; for
// each group the x phase filters the group's rows along x (lanes = rows,
// loads staged through LDS for coalescing, the x recurrence started W points
// left of c0 from unfiltered values, or exactly from the left end) and
// leaves the strip's x-filtered values in an LDS tile; the y phase (lanes =
// columns) continues each column's y recurrence through the tile and writes
// the filtered rows, plus the ccpml rows at the ends, once.  The y phase
// loads in bursts of 13 positions, and the groups are aligned so that a
// burst that starts past the tile is exactly the first of the next group:
// the tile is refilled there and nowhere else.  The state is read and
// written once per stage instead of twice; the price is the x recurrence's
// W + 12 extra columns per strip (from L2 for neighbouring strips) and, per
// row chunk, the y look-back rows.
// ---------------------------------------------------------------------------
constexpr int XY_G = 52;  // rows per group: 4 bursts of 13

// x phase IO: lanes = rows row0 + lane; emits the strip's columns into the tile
template <typename R>
struct XTIO : XIO<R> {
    R (*tt)[65];  // XY_G rows x 64 columns (+1 pad)
    int c0;
    __device__ __forceinline__ void emit(int e, R v)
    {
        if (e >= c0 && e < c0 + 64 && this->lane < XY_G) tt[this->lane][e - c0] = v;
    }
    __device__ __forceinline__ void emit_fast(int e, R v) { emit(e, v); }
    __device__ __forceinline__ void finish() {}
};

// y phase IO: lane = column c0 + lane; positions are rows, read from the tile
// (groups of XY_G rows start at the sweep's first load; the sweep's bursts of
// 13 start there too, so a group boundary is always a burst's start)
template <typename R, bool RING>
struct YTIO {
    const FilterXYArgs<R> *F;
    int k, c0, lane, col, a, b;
    bool active;
    R (*tt)[65];
    R (*tin)[XIO<R>::T + 1];
    int g0, gend;  // the tile holds rows [g0, gend)
    R *out;
    const double *corr;
    __device__ __forceinline__ void fill(int g)  // x-filter rows [g, g + XY_G) of the strip into the tile
    {
        XTIO<R> x;
        x.in = F->in[k];
        x.out = nullptr;
        x.pitch = F->pitch;
        x.n = F->imax;
        x.m = F->jmax;
        x.row0 = g;
        x.lane = lane;
        x.ring = RING;
        x.tin = tin;
        x.tout = nullptr;
        x.tt = tt;
        x.c0 = c0;
        const int xb = c0 + 64 < F->imax ? c0 + 64 : F->imax;
        x.a = c0;
        x.b = xb;
        const bool exact = F->x_exact || c0 <= F->W + 12;
        LineEnds ends;
        __syncthreads();  // every lane has read the previous group from the tile
        x.start(exact ? (RING ? -6 : 0) : c0 - F->W - 6);
        filter_line<R, false, 1, XTIO<R>, RING>(x, R(F->sigma), F->imax, exact ? 0 : c0, xb, F->W, ends);
        __syncthreads();  // the tile is complete
        g0 = g;
        gend = g + XY_G;
    }
    __device__ __forceinline__ R load(int p)
    {
        int r = p - g0;
        r = r < 0 ? 0 : (r < XY_G ? r : XY_G - 1);  // clamped positions are never used
        return tt[r][lane];
    }
    __device__ __forceinline__ void burst(int p)
    {
        if (p >= gend && p < F->jmax) fill(gend);
    }
    __device__ __forceinline__ void emit(int e, R v)
    {
        if (active && e >= a && e < b) out[(size_t)e * F->pitch + col] = corr ? R(double(v) + corr[e]) : v;
    }
    __device__ __forceinline__ void emit_fast(int e, R v) { emit(e, v); }
    __device__ __forceinline__ void finish() {}
    __device__ __forceinline__ R load_head(int) { return R(0); }
};

template <typename R, bool RING>
__device__ __forceinline__ void filter_xy_tile(const FilterXYArgs<R> &F, int strip, int chunk, int k,
                                               R (*tin)[XIO<R>::T + 1], R (*tt)[65])
{
    int a, b;
    if (!chunk_bounds(chunk, F.Ly, F.jmax, a, b)) return;
    const int c0 = strip * 64;
    YTIO<R, RING> io;
    io.F = &F;
    io.k = k;
    io.c0 = c0;
    io.lane = threadIdx.x;
    io.col = c0 + (int)threadIdx.x;
    io.active = io.col < F.imax;
    io.a = a;
    io.b = b;
    io.tt = tt;
    io.tin = tin;
    io.out = F.out[k];
    io.corr = F.corr[k];
    io.gend = 0;
    // the first group starts at the y sweep's first load (filter_line:
    // row 0 for the exact start, else a - W - 6)
    const bool exact = a <= F.W + 12;
    io.fill(exact ? 0 : a - F.W - 6);
    LineEnds ends;
    filter_line<R, true, 1, YTIO<R, RING>>(io, R(F.sigma), F.jmax, a, b, F.W, ends);
    // ccpml rows j=1 and j=jmax (cc.f90:46-59), as filter_y_tile
    const int col = io.col;
    if (io.active && (F.xper || (col >= 1 && col <= F.imax - 2))) {
        R *o = F.out[k];
        const int jm = F.jmax;
        const double *cr = F.corr[k];
        if (a == 0) {
            double val;
            if (k < 3) {
                const double v1 = cr ? ends.v1 + cr[1] : ends.v1, v2 = cr ? ends.v2 + cr[2] : ends.v2;
                val = 2.0 * v1 - v2 + F.re.bot_res[k];
            } else {
                val = F.re.bot_val[k];
            }
            o[col] = R(val);
        }
        if (b == jm) {
            double val;
            if (k < 3) {
                const double v2 = cr ? ends.vn2 + cr[jm - 2] : ends.vn2;
                const double v3 = cr ? ends.vn3 + cr[jm - 3] : ends.vn3;
                val = 2.0 * v2 - v3 + F.re.top_res[k];
            } else {
                val = F.re.top_val[k];
            }
            o[(size_t)(jm - 1) * F.pitch + col] = R(val);
        }
    }
}

// One block per (64-column strip, row chunk, component), numbered like the
// marching blocks (march_tile_of): each XCD runs a contiguous run of tiles,
// strips minor, so neighbouring strips, whose x recurrences re-read each
// other's W + 6 look-back columns, share that XCD's L2.
template <typename R>
__global__ __launch_bounds__(64) void filter_xy_kernel(FilterXYArgs<R> F, int nstrips, int nchunks)
{
    __shared__ R tin[64][XIO<R>::T + 1];
    __shared__ R tt[XY_G][65];
    int strip, ck;
    march_tile_of((int)blockIdx.x, (int)gridDim.x, nstrips, strip, ck);
    const int chunk = ck % nchunks, k = ck / nchunks;
    if (F.xper)
        filter_xy_tile<R, true>(F, strip, chunk, k, tin, tt);
    else
        filter_xy_tile<R, false>(F, strip, chunk, k, tin, tt);
}

template <typename S>
static hipError_t launch_filter_xy_t(const FilterXYArgs<S> &F, hipStream_t s)
{
    hipError_t e = kernel_ready((const void *)filter_xy_kernel<S>, 0, 64, s);
    if (e != hipSuccess) return e;
    const int nstrips = (F.imax + 63) / 64, nchunks = (F.jmax + F.Ly - 1) / F.Ly;
    hipLaunchKernelGGL(filter_xy_kernel<S>, dim3(nstrips * nchunks * F.nk), dim3(64), 0, s, F, nstrips, nchunks);
    return hipGetLastError();
}

#if defined(NS_KERNELS_F64)
hipError_t launch_filter_xy(const FilterXYArgs<double> &F, hipStream_t s) { return launch_filter_xy_t(F, s); }
#endif
#if defined(NS_KERNELS_F32)
hipError_t launch_filter_xy(const FilterXYArgs<float> &F, hipStream_t s) { return launch_filter_xy_t(F, s); }
#endif

template <typename S>
static hipError_t launch_filter_t(const FilterArgs<S> &F, bool xdir, hipStream_t s)
{
    const int nchunks = (F.n + F.L - 1) / F.L;
    hipError_t e = kernel_ready(xdir ? (const void *)filter_x_kernel<S> : (const void *)filter_y_kernel<S>, 0,
                                xdir ? 64 : 64 * filter_y_waves<S>(), s);
    if (e != hipSuccess) return e;
    if (xdir) {
        dim3 grid(nchunks, (F.m + 63) / 64, F.nk);
        hipLaunchKernelGGL(filter_x_kernel<S>, grid, dim3(64), 0, s, F);
    } else {
        constexpr int NW = filter_y_waves<S>();
        dim3 grid((F.m + 64 * NW - 1) / (64 * NW), nchunks, F.nk);
        hipLaunchKernelGGL(filter_y_kernel<S>, grid, dim3(64 * NW), 0, s, F);
    }
    return hipGetLastError();
}

template <typename S>
static hipError_t launch_filter_batch_t(const FilterArgs<S> *args, const int4 *tiles, int ntiles, int nk, bool xdir,
                                        hipStream_t s)
{
    const int nt = 64;
    hipError_t e = kernel_ready(
        xdir ? (const void *)filter_x_kernel_batch<S> : (const void *)filter_y_kernel_batch<S>, 0, nt, s);
    if (e != hipSuccess) return e;
    if (xdir)
        hipLaunchKernelGGL(filter_x_kernel_batch<S>, dim3(ntiles, nk), dim3(nt), 0, s, args, tiles);
    else
        hipLaunchKernelGGL(filter_y_kernel_batch<S>, dim3(ntiles, nk), dim3(nt), 0, s, args, tiles);
    return hipGetLastError();
}

#if defined(NS_KERNELS_F64)
hipError_t launch_filter(const FilterArgs<double> &F, bool xdir, hipStream_t s) { return launch_filter_t(F, xdir, s); }
hipError_t launch_filter_batch(const FilterArgs<double> *args, const int4 *tiles, int ntiles, int nk, bool xdir,
                               hipStream_t s)
{
    return launch_filter_batch_t(args, tiles, ntiles, nk, xdir, s);
}
#endif
#if defined(NS_KERNELS_F32)
hipError_t launch_filter(const FilterArgs<float> &F, bool xdir, hipStream_t s) { return launch_filter_t(F, xdir, s); }
hipError_t launch_filter_batch(const FilterArgs<float> *args, const int4 *tiles, int ntiles, int nk, bool xdir,
                               hipStream_t s)
{
    return launch_filter_batch_t(args, tiles, ntiles, nk, xdir, s);
}
#endif

// ---------------------------------------------------------------------------
// ccpml columns i=1 (inflow) and i=imax (outflow) for every row, with the
// corner rows j=1, jmax computed from those columns' values (cc.f90:23-59).
// ---------------------------------------------------------------------------
template <typename S>
__device__ __forceinline__ void edge_i_point(const EdgeArgs<S> &E, int j, int side)
{
    if (j >= E.jmax || E.xper) return;
    const int imax = E.imax, jmax = E.jmax, pitch = E.pitch;
    const double ig = 1.0 / E.gamma;
    auto at = [&](int k, int i, int jj) { return double(E.U[k][(size_t)jj * pitch + i]); };
    auto ub = [&](int k, int jj) { return E.ub[k] ? E.ub[k][jj] : 0.0; };
    // value of the column after the inflow/outflow assignment at row jj
    // (total value minus the base row in the perturbation form)
    auto colval = [&](int k, int jj) -> double {
        if (side == 0) {
            switch (k) {
            case 0: return E.rhob[jj] - ub(0, jj);
            case 1: return 2.0 * at(1, 1, jj) - at(1, 2, jj);
            case 2: return 0.0 - ub(2, jj);
            case 3: return ig - ub(3, jj);
            default: return (E.psib ? E.psib[jj] : 0.0) - ub(4, jj);
            }
        }
        if (k == 3) return ig - ub(3, jj);
        return 2.0 * at(k, imax - 2, jj) - at(k, imax - 3, jj);
    };
    const int i = side == 0 ? 0 : imax - 1;
    double vals[NKMAX];
    for (int k = 0; k < E.nk; ++k) {
        double v;
        if (j == 0) {  // lower boundary overwrites the corner (cc.f90:54-59)
            v = (k < 3) ? 2.0 * colval(k, 1) - colval(k, 2) + E.re.bot_res[k] : E.re.bot_val[k];
        } else if (j == jmax - 1) {  // upper boundary (cc.f90:38-43)
            v = (k < 3) ? 2.0 * colval(k, jmax - 2) - colval(k, jmax - 3) + E.re.top_res[k] : E.re.top_val[k];
        } else {
            v = colval(k, j);
        }
        vals[k] = v;
    }
    for (int k = 0; k < E.nk; ++k) E.U[k][(size_t)j * pitch + i] = S(vals[k]);
}

template <typename S>
__global__ void edge_i_kernel(EdgeArgs<S> E)
{
    // blockIdx.y: 0 inflow column, 1 outflow column
    edge_i_point<S>(E, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

// batched: blockIdx.x = case, blockIdx.y = column, blockIdx.z = block of
// 256 rows (rows past the case's last are skipped; larger cases loop)
template <typename S>
__global__ void edge_i_kernel_batch(const EdgeArgs<S> *args)
{
    const EdgeArgs<S> &E = args[blockIdx.x];
    for (int j = (int)(blockIdx.z * blockDim.x + threadIdx.x); j < E.jmax; j += (int)(gridDim.z * blockDim.x))
        edge_i_point<S>(E, j, blockIdx.y);
}

template <typename S>
static hipError_t launch_edge_t(const EdgeArgs<S> &E, hipStream_t s)
{
    hipError_t e = kernel_ready((const void *)edge_i_kernel<S>, 0, 256, s);
    if (e != hipSuccess) return e;
    dim3 grid((E.jmax + 255) / 256, 2);
    hipLaunchKernelGGL(edge_i_kernel<S>, grid, dim3(256), 0, s, E);
    return hipGetLastError();
}

template <typename S>
static hipError_t launch_edge_batch_t(const EdgeArgs<S> *args, int ncases, hipStream_t s)
{
    hipError_t e = kernel_ready((const void *)edge_i_kernel_batch<S>, 0, 256, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(edge_i_kernel_batch<S>, dim3(ncases, 2, 2), dim3(256), 0, s, args);
    return hipGetLastError();
}

#if defined(NS_KERNELS_F64)
hipError_t launch_edge_i(const EdgeArgs<double> &E, hipStream_t s) { return launch_edge_t(E, s); }
hipError_t launch_edge_batch(const EdgeArgs<double> *args, int ncases, hipStream_t s)
{
    return launch_edge_batch_t(args, ncases, s);
}
#endif
#if defined(NS_KERNELS_F32)
hipError_t launch_edge_i(const EdgeArgs<float> &E, hipStream_t s) { return launch_edge_t(E, s); }
hipError_t launch_edge_batch(const EdgeArgs<float> *args, int ncases, hipStream_t s)
{
    return launch_edge_batch_t(args, ncases, s);
}
#endif

// Host-side set-up of a step before it is captured into a graph.  A kernel's
// code object loads at its first use, and a marching plan queries the device
// the first time; such runtime calls must not run inside stream capture.
// This makes the stage launch's set-up without launching (plan_only), and
// fetches the attributes of the filter and edge kernels of the single-case
// (batch = false) or batched step, which loads them.
template <typename S>
static hipError_t prepare_step_t(const StageArgs<S> &A, int nk, bool visc, int variant, bool batch)
{
    hipError_t e;
    if (!batch) {
        StageHooks h;
        h.plan_only = true;
        if ((e = launch_stage(A, nk, visc, false, variant, nullptr, nullptr, &h)) != hipSuccess) return e;
    }
    const void *kerns[4];
    if (batch) {
        kerns[0] = (const void *)filter_x_kernel_batch<S>;
        kerns[1] = (const void *)filter_y_kernel_batch<S>;
        kerns[2] = (const void *)edge_i_kernel_batch<S>;
        kerns[3] = nullptr;
    } else {
        kerns[0] = (const void *)filter_x_kernel<S>;
        kerns[1] = (const void *)filter_y_kernel<S>;
        kerns[2] = (const void *)edge_i_kernel<S>;
        kerns[3] = (const void *)filter_xy_kernel<S>;
    }
    const int threads[4] = {64, batch ? 64 : 64 * filter_y_waves<S>(), 256, 64};
    for (int q = 0; q < 4; ++q)
        if (kerns[q] && (e = kernel_ready(kerns[q], 0, threads[q], nullptr)) != hipSuccess) return e;
    return hipSuccess;
}

#if defined(NS_KERNELS_F64)
hipError_t prepare_step(const StageArgs<double> &A, int nk, bool visc, int variant, bool batch)
{
    return prepare_step_t(A, nk, visc, variant, batch);
}
#endif
#if defined(NS_KERNELS_F32)
hipError_t prepare_step(const StageArgs<float> &A, int nk, bool visc, int variant, bool batch)
{
    return prepare_step_t(A, nk, visc, variant, batch);
}
#endif

#if defined(NS_KERNELS_F64)

// ---------------------------------------------------------------------------
// Operator-level kernels for the parity API (one thread per point).
// ---------------------------------------------------------------------------
__global__ void deriv_kernel(const double *const *in, double *const *out, int nk, int imax, int jmax, int pitch,
                             int ydir, const double *met, double inv_h, int xper)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= imax) return;
    for (int k = 0; k < nk; ++k) {
        const double *u = in[k];
        double d;
        if (!ydir) {
            const double *row = u + (size_t)j * pitch;
            if (xper)  // periodic x: the interior stencil, wrapped
                d = drp_interior<double>([row, imax](int q) { return row[wrap_col(q, imax)]; }, i, inv_h) * met[i];
            else
                d = drp_deriv<double>([row](int q) { return row[q]; }, i, imax, inv_h) * met[i];
        } else {
            const double *col = u + i;
            d = drp_deriv<double>([col, pitch](int q) { return col[(size_t)q * pitch]; }, j, jmax, inv_h) * met[j];
        }
        out[k][(size_t)j * pitch + i] = d;
    }
}

hipError_t launch_deriv(const double *const *in, double *const *out, int nk, int imax, int jmax, int pitch, bool ydir,
                        const double *met, double inv_h, bool xper, hipStream_t s)
{
    dim3 grid((imax + 127) / 128, jmax);
    hipLaunchKernelGGL(deriv_kernel, grid, dim3(128), 0, s, in, out, nk, imax, jmax, pitch, (int)ydir, met, inv_h,
                       (int)xper);
    return hipGetLastError();
}

__global__ void transport_kernel(const double *T, double *MI, double *K1, size_t n, Transport<double> tc)
{
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    double mu, kk;
    transport_eval<double>(tc, T[q], mu, kk);
    MI[q] = mu;
    K1[q] = kk;
}

hipError_t launch_transport(const double *T, double *MI, double *K1, size_t n, const Transport<double> &tc,
                            hipStream_t s)
{
    hipLaunchKernelGGL(transport_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, T, MI, K1, n, tc);
    return hipGetLastError();
}

// rk_steep on nk planes (rk.f90:160-167)
__global__ void rk_steep_kernel(double *const *U, const double *const *dudt, double *const *W, int nk, int imax,
                                int jmax, int pitch, double alpha, double beta, double dt, int xper)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= imax || (!xper && (i < 1 || i > imax - 2)) || j < 1 || j > jmax - 2) return;
    const size_t g = (size_t)j * pitch + i;
    for (int k = 0; k < nk; ++k) {
        const double wn = alpha * W[k][g] + dt * dudt[k][g];
        W[k][g] = wn;
        U[k][g] = U[k][g] + beta * wn;
    }
}

hipError_t launch_rk_steep(double *const *U, const double *const *dudt, double *const *W, int nk, int imax, int jmax,
                           int pitch, double alpha, double beta, double dt, bool xper, hipStream_t s)
{
    dim3 grid((imax + 127) / 128, jmax);
    hipLaunchKernelGGL(rk_steep_kernel, grid, dim3(128), 0, s, U, dudt, W, nk, imax, jmax, pitch, alpha, beta, dt,
                       (int)xper);
    return hipGetLastError();
}

// bufferxy (rk.f90:77-151)
__global__ void buffer_kernel(double *const *U, const double *const *base, int nk, int imax, int jmax, int pitch,
                              int D, const double *sigx, const double *sigy, int xper)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= imax) return;
    const bool ybuf = j < D || j >= jmax - D, xbuf = !xper && (i < D || i >= imax - D);
    if (!ybuf && !xbuf) return;
    const size_t g = (size_t)j * pitch + i;
    for (int k = 0; k < nk; ++k) {
        double un = U[k][g];
        const double ub = base[k][g];
        if (ybuf) un = ub + sigy[j] * (un - ub);
        if (xbuf) un = ub + sigx[i] * (un - ub);
        U[k][g] = un;
    }
}

hipError_t launch_buffer(double *const *U, const double *const *base, int nk, int imax, int jmax, int pitch, int D,
                         const double *sigx, const double *sigy, bool xper, hipStream_t s)
{
    dim3 grid((imax + 127) / 128, jmax);
    hipLaunchKernelGGL(buffer_kernel, grid, dim3(128), 0, s, U, base, nk, imax, jmax, pitch, D, sigx, sigy,
                       (int)xper);
    return hipGetLastError();
}

// acousticfont (rk.f90:186-191), separable form amp * gx(i) * gy(j)
__global__ void source_kernel(double *s_out, int imax, int jmax, double amp, const double *gx, const double *gy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= imax) return;
    s_out[(size_t)j * imax + i] = amp * (gx[i] * gy[j]);
}

hipError_t launch_source(double *s_out, int imax, int jmax, double amp, const double *gx, const double *gy,
                         hipStream_t s)
{
    dim3 grid((imax + 127) / 128, jmax);
    hipLaunchKernelGGL(source_kernel, grid, dim3(128), 0, s, s_out, imax, jmax, amp, gx, gy);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pieces of rhs_euler as the reference splits them (derivs.f90, eqeuler.f90,
// diff.f90 metricsX/Y), one thread per point on pitched planes; they serve
// the drop-in modules of the same names.  Each keeps the reference's
// expression order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool pw_point(const PwArgs &a, size_t &g, int &i, int &j)
{
    i = blockIdx.x * blockDim.x + threadIdx.x;
    j = blockIdx.y;
    g = (size_t)j * a.pitch + i;
    return i < a.imax;
}

// kdT (derivs.f90:48-56): dTdm *= k, dTdn *= k
__global__ void kdt_kernel(PwArgs a, const double *k, double *dTdm, double *dTdn)
{
    size_t g;
    int i, j;
    if (!pw_point(a, g, i, j)) return;
    dTdn[g] = dTdn[g] * k[g];
    dTdm[g] = dTdm[g] * k[g];
}

// psiDiff = rho alpha (dpsi/dm + dpsi/dn) (derivs.f90:36-40)
__global__ void psidiff_kernel(PwArgs a, const double *rho, const double *pm, const double *pn, double alpha,
                               double *out)
{
    size_t g;
    int i, j;
    if (!pw_point(a, g, i, j)) return;
    out[g] = rho[g] * alpha * (pm[g] + pn[g]);
}

// tau_ij (eqeuler.f90:34-41) from MI and the u, v derivatives
__global__ void tau_kernel(PwArgs a, const double *mi, const double *ux, const double *uy, const double *vx,
                           const double *vy, double *txx, double *txy, double *tyy)
{
    size_t g;
    int i, j;
    if (!pw_point(a, g, i, j)) return;
    const double m = mi[g];
    txx[g] = -(2.0 / 3.0 * m * (ux[g] + vy[g])) + 2.0 * m * ux[g];
    txy[g] = m * (uy[g] + vx[g]);
    tyy[g] = -(2.0 / 3.0 * m * (ux[g] + vy[g])) + 2.0 * m * vy[g];
}

// EULERNONLINEAR (eqeuler.f90:61-97)
__global__ void euler_pw_kernel(EulerPwArgs e)
{
    size_t g;
    int i, j;
    if (!pw_point(e.a, g, i, j)) return;
    const double u0 = e.U[0][g], u1 = e.U[1][g], u2 = e.U[2][g], u3 = e.U[3][g];
    const double gamma = e.gamma, Re = e.Re, Pr = e.Pr;
    const double s = e.src ? e.src[(size_t)j * e.a.imax + i] : 0.0;
    auto dm = [&](int k) { return e.Um[k][g]; };
    auto dn = [&](int k) { return e.Un[k][g]; };
    e.out[0][g] = -(u1 * dm(0) + u0 * dm(1) + u2 * dn(0) + u0 * dn(2));
    e.out[1][g] = -(u1 * dm(1) + 1.0 / u0 * dm(3) + u2 * dn(1) - 1.0 / Re / u0 * (e.taum[0][g] + e.taun[1][g]));
    e.out[2][g] = -(u1 * dm(2) + u2 * dn(2) + 1.0 / u0 * dn(3) - 1.0 / Re / u0 * (e.taum[1][g] + e.taun[2][g]));
    e.out[3][g] = -(gamma * u3 * dm(1) + u1 * dm(3) + gamma * u3 * dn(2) + u2 * dn(3) - s -
                    (gamma - 1.0) / Re *
                        (+dm(1) * e.tau[0][g] + dm(2) * e.tau[1][g] + dn(1) * e.tau[1][g] + dn(2) * e.tau[2][g]) -
                    1.0 / Re / Pr * (e.d2Tm[g] + e.d2Tn[g]));
    if (e.nk >= 5)
        e.out[4][g] = -(u1 * dm(4) + u2 * dn(4) - 1.0 / e.Pe / u0 * (e.d2pm[g] + e.d2pn[g]));
}

// metricsX / metricsY (diff.f90:276-350): du *= the point's effective factor
__global__ void metric_kernel(PwArgs a, double *const *dU, int nk, const double *met, int ydir)
{
    size_t g;
    int i, j;
    if (!pw_point(a, g, i, j)) return;
    const double f = ydir ? met[j] : met[i];
    for (int k = 0; k < nk; ++k) dU[k][g] = dU[k][g] * f;
}

static dim3 pw_grid(const PwArgs &a) { return dim3((a.imax + 127) / 128, a.jmax); }

hipError_t launch_kdt(const PwArgs &a, const double *k, double *dTdm, double *dTdn, hipStream_t s)
{
    hipLaunchKernelGGL(kdt_kernel, pw_grid(a), dim3(128), 0, s, a, k, dTdm, dTdn);
    return hipGetLastError();
}

hipError_t launch_psidiff(const PwArgs &a, const double *rho, const double *pm, const double *pn, double alpha,
                          double *out, hipStream_t s)
{
    hipLaunchKernelGGL(psidiff_kernel, pw_grid(a), dim3(128), 0, s, a, rho, pm, pn, alpha, out);
    return hipGetLastError();
}

hipError_t launch_tau(const PwArgs &a, const double *mi, const double *ux, const double *uy, const double *vx,
                      const double *vy, double *txx, double *txy, double *tyy, hipStream_t s)
{
    hipLaunchKernelGGL(tau_kernel, pw_grid(a), dim3(128), 0, s, a, mi, ux, uy, vx, vy, txx, txy, tyy);
    return hipGetLastError();
}

hipError_t launch_euler_pw(const EulerPwArgs &e, hipStream_t s)
{
    hipLaunchKernelGGL(euler_pw_kernel, pw_grid(e.a), dim3(128), 0, s, e);
    return hipGetLastError();
}

hipError_t launch_metric(const PwArgs &a, double *const *dU, int nk, const double *met, bool ydir, hipStream_t s)
{
    hipLaunchKernelGGL(metric_kernel, pw_grid(a), dim3(128), 0, s, a, dU, nk, met, (int)ydir);
    return hipGetLastError();
}

#endif  // NS_KERNELS_F64 (operator-level kernels)

// probes (euler.f90:156-161): gather nk values at npts points
template <typename S>
__global__ void probe_kernel(const S *const *U, int nk, int pitch, int npts, const int *ij, double *out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts * nk) return;
    const int pnt = t / nk, k = t - pnt * nk;
    const int i = ij[2 * pnt] - 1, j = ij[2 * pnt + 1] - 1;
    out[t] = double(U[k][(size_t)j * pitch + i]);
}

template <typename S>
hipError_t launch_probe(const S *const *U, int nk, int pitch, int npts, const int *ij, double *out, hipStream_t s)
{
    const int n = npts * nk;
    hipLaunchKernelGGL(probe_kernel<S>, dim3((n + 255) / 256), dim3(256), 0, s, U, nk, pitch, npts, ij, out);
    return hipGetLastError();
}
#if defined(NS_KERNELS_F64)
template hipError_t launch_probe<double>(const double *const *, int, int, int, const int *, double *, hipStream_t);
#endif
#if defined(NS_KERNELS_F32)
template hipError_t launch_probe<float>(const float *const *, int, int, int, const int *, double *, hipStream_t);
#endif

// Vorticity of the output path (outt.f90:49-55): omega = dv/dm - du/dn with
// the DRP operators of diffall.f90 (no metric factors), from the state; in
// the perturbation form U = Ub(j) + q, and d Ub/dn is taken on the FP64 base
// row with the same operator.  out: jmax x imax, row-major.
template <typename S>
__global__ __launch_bounds__(256) void vorticity_kernel(const S *u, const S *v, const double *ub1, int imax, int jmax,
                                                        int pitch, double inv_dm, double inv_dn, int xper, double *out)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= imax || j >= jmax) return;
    const S *vrow = v + (size_t)j * pitch;
    const S *ucol = u + i;
    const double dvdm =
        xper ? drp_interior<double>([vrow, imax](int q) { return double(vrow[wrap_col(q, imax)]); }, i, inv_dm)
             : drp_deriv<double>([vrow](int q) { return double(vrow[q]); }, i, imax, inv_dm);
    double dudn = drp_deriv<double>([ucol, pitch](int q) { return double(ucol[(size_t)q * pitch]); }, j, jmax, inv_dn);
    if (ub1) dudn = dudn + drp_deriv<double>([ub1](int q) { return ub1[q]; }, j, jmax, inv_dn);
    out[(size_t)j * imax + i] = dvdm - dudn;
}

template <typename S>
hipError_t launch_vorticity(const S *u, const S *v, const double *ub1, int imax, int jmax, int pitch, double inv_dm,
                            double inv_dn, bool xper, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(vorticity_kernel<S>, dim3((imax + 63) / 64, (jmax + 3) / 4), dim3(256), 0, s, u, v, ub1, imax,
                       jmax, pitch, inv_dm, inv_dn, (int)xper, out);
    return hipGetLastError();
}
#if defined(NS_KERNELS_F64)
template hipError_t launch_vorticity<double>(const double *, const double *, const double *, int, int, int, double,
                                             double, bool, double *, hipStream_t);
#endif
#if defined(NS_KERNELS_F32)
template hipError_t launch_vorticity<float>(const float *, const float *, const double *, int, int, int, double,
                                            double, bool, double *, hipStream_t);
#endif

// Kinetic energy at x-stations (outt.f90:92-110): for column i,
//   ec = (2 sum_{j=D+1}^{jmax-D-1} k_j + k_D + k_{jmax-D}) dn / 2,
//   k_j = (u_j^2 + v_j^2) / 2  (1-based j; the full velocity, base included)
// One block per station, FP64 sums in a fixed order.
template <typename S>
__global__ __launch_bounds__(256) void station_energy_kernel(const S *u, const S *v, const double *ub1,
                                                             const double *ub2, int pitch, int jmax, int D,
                                                             double dn, const int *cols, double *out)
{
    __shared__ double part[256];
    const int i = cols[blockIdx.x];
    auto ke = [&](int j) {  // j: 0-based row
        double uu = double(u[(size_t)j * pitch + i]), vv = double(v[(size_t)j * pitch + i]);
        if (ub1) {
            uu += ub1[j];
            vv += ub2[j];
        }
        return 0.5 * (uu * uu + vv * vv);
    };
    double s = 0.0;
    for (int j = D + (int)threadIdx.x; j <= jmax - D - 2; j += 256) s += ke(j);  // 1-based D+1 .. jmax-D-1
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = (2.0 * part[0] + (ke(D - 1) + ke(jmax - D - 1))) * dn / 2.0;
}

template <typename S>
hipError_t launch_station_energy(const S *u, const S *v, const double *ub1, const double *ub2, int pitch, int jmax,
                                 int D, double dn, int nst, const int *cols, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(station_energy_kernel<S>, dim3(nst), dim3(256), 0, s, u, v, ub1, ub2, pitch, jmax, D, dn,
                       cols, out);
    return hipGetLastError();
}
#if defined(NS_KERNELS_F64)
template hipError_t launch_station_energy<double>(const double *, const double *, const double *, const double *, int,
                                                  int, int, double, int, const int *, double *, hipStream_t);
#endif
#if defined(NS_KERNELS_F32)
template hipError_t launch_station_energy<float>(const float *, const float *, const double *, const double *, int,
                                                 int, int, double, int, const int *, double *, hipStream_t);
#endif

// Row averages of one plane: out[j] = mean over i in [i0, i1) of U(i, j)
// (+ add[j] when given: the base row of the perturbation form).  One block
// per row, FP64 sums in a fixed order (deterministic).
template <typename S>
__global__ __launch_bounds__(256) void row_mean_kernel(const S *U, int pitch, int i0, int i1, const double *add,
                                                       double *out)
{
    __shared__ double part[256];
    const int j = blockIdx.x;
    const S *row = U + (size_t)j * pitch;
    double s = 0.0;
    for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) s += double(row[i]);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = part[0] / double(i1 - i0) + (add ? add[j] : 0.0);
}

template <typename S>
hipError_t launch_row_mean(const S *U, int pitch, int jmax, int i0, int i1, const double *add, double *out,
                           hipStream_t s)
{
    hipLaunchKernelGGL(row_mean_kernel<S>, dim3(jmax), dim3(256), 0, s, U, pitch, i0, i1, add, out);
    return hipGetLastError();
}
#if defined(NS_KERNELS_F64)
template hipError_t launch_row_mean<double>(const double *, int, int, int, int, const double *, double *,
                                            hipStream_t);
#endif
#if defined(NS_KERNELS_F32)
template hipError_t launch_row_mean<float>(const float *, int, int, int, int, const double *, double *, hipStream_t);
#endif

// non-finite scan: flag = 1 if any value is NaN/Inf
template <typename S>
__global__ void finite_kernel(const S *const *U, int nk, int imax, int jmax, int pitch, int *flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= imax) return;
    for (int k = 0; k < nk; ++k) {
        const S v = U[k][(size_t)j * pitch + i];
        if (!isfinite(v)) {
            atomicOr(flag, 1);
            return;
        }
    }
}

template <typename S>
hipError_t launch_finite(const S *const *U, int nk, int imax, int jmax, int pitch, int *flag, hipStream_t s)
{
    dim3 grid((imax + 127) / 128, jmax);
    hipLaunchKernelGGL(finite_kernel<S>, grid, dim3(128), 0, s, U, nk, imax, jmax, pitch, flag);
    return hipGetLastError();
}
#if defined(NS_KERNELS_F64)
template hipError_t launch_finite<double>(const double *const *, int, int, int, int, int *, hipStream_t);
#endif
#if defined(NS_KERNELS_F32)
template hipError_t launch_finite<float>(const float *const *, int, int, int, int, int *, hipStream_t);
#endif

#if defined(NS_KERNELS_F64)
// fill kernel (graph-safe zeroing)
__global__ void fill_kernel(double *p, size_t n, double v)
{
    const size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (q < n) p[q] = v;
}

hipError_t launch_fill(double *p, size_t n, double v, hipStream_t s)
{
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

#endif  // NS_KERNELS_F64

}  // namespace nsamd
