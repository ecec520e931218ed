// ns_adjoint.hip -- discrete adjoint (reverse mode) of the LDDRK6 step for the
// 2-D Euler equations (nk = 4, inviscid), gfx950.
//
// The forward stage (rk.f90:52-73) is
//   f = R(U_s);  W' = alpha W + dt f;  U' = U_s + beta W'   (interior points)
//   U'' = buffer(U');  U''' = Fy(Fx(U''));  U_{s+1} = ccpml(U''')
// Reverse mode takes (lam = dJ/dU_{s+1}, lamW = dJ/dW') back to
// (dJ/dU_s, dJ/dW) with one backward kernel per forward operator:
//   adj_edge_kernel     ccpml^T   (cc.f90:6-68 assignments in reverse order)
//   adj_filter_kernel   F^T       (filtering.f90:9-190 row updates in reverse
//                                  order, each transposed)
//   adj_point_kernel    buffer^T, update^T and the pointwise part of R'(U_s)^T;
//                       also the flux adjoints G = metric * dR/d(dU/dxi)^T lam
//   adj_gather_kernel   D^T G     (transposed DRP operator incl. the one-sided
//                                  rows, diff.f90:6-125)
// Compiled twice like ns_kernels.hip: FP64 (S = double) and the FP32
// perturbation form (S = float, U = Ub(j) + q).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ns_adjoint.h"

namespace nsamd {

// ---------------------------------------------------------------------------
// ccpml^T.  The forward sets, in order, the inflow column i=1, the outflow
// column i=imax, the top row j=jmax and the bottom row j=1 (cc.f90:23-59);
// components 0-2 (and 4 at the outflow) extrapolate as 2 u(2) - u(3), the
// rest are constants.  The transpose runs the four sets in reverse order
// (phase 0 = bottom ... 3 = inflow): for an extrapolated
// value, lam(2) += 2 lam(1), lam(3) -= lam(1); every overwritten value's
// adjoint is then zero.
// ---------------------------------------------------------------------------
template <typename S>
__device__ __forceinline__ void adj_edge_item(const AdjEdgeArgs<S> &E, int phase, int t)
{
    const int imax = E.imax, jmax = E.jmax, pitch = E.pitch;
    auto at = [&](int k, int i, int j) -> S & { return E.lam[k][(size_t)j * pitch + i]; };
    auto extrap = [&](int k, int i0, int j0, int i1, int j1, int i2, int j2) {
        const S v = at(k, i0, j0);
        at(k, i1, j1) += S(2) * v;
        at(k, i2, j2) -= v;
        at(k, i0, j0) = S(0);
    };
    if (phase == 0 || phase == 1) {  // bottom (j=1) / top (j=jmax) rows, over i
        if (t >= imax) return;
        const int j0 = phase == 0 ? 0 : jmax - 1, d = phase == 0 ? 1 : -1;
        for (int k = 0; k < 3; ++k) extrap(k, t, j0, t, j0 + d, t, j0 + 2 * d);
        for (int k = 3; k < E.nk; ++k) at(k, t, j0) = S(0);
    } else if (phase == 2) {  // outflow column i=imax, over j
        if (t >= jmax) return;
        const int i0 = imax - 1;
        for (int k = 0; k < E.nk; ++k) {
            if (k == 3) at(k, i0, t) = S(0);
            else extrap(k, i0, t, i0 - 1, t, i0 - 2, t);
        }
    } else {  // inflow column i=1: only u extrapolates
        if (t >= jmax) return;
        for (int k = 0; k < E.nk; ++k) {
            if (k == 1) extrap(k, 0, t, 1, t, 2, t);
            else at(k, 0, t) = S(0);
        }
    }
}

// Sets first..last in order, thread t taking position t of each.  A row
// set at position t reads and writes only column t, a column set only row
// t, so one thread can run both row sets at its column (one launch), then
// both column sets at its row (a second launch: they read corner values
// the row sets changed).
template <typename S>
__global__ void adj_edge_kernel(AdjEdgeArgs<S> E, int first, int last)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    for (int phase = first; phase <= last; ++phase) adj_edge_item<S>(E, phase, t);
}

// ---------------------------------------------------------------------------
// F^T on one line of n points (0-based positions).  The forward
// (filtering.f90:36-94) applies, in order,
//   L: q = 1..6        u(q) -= sigma * sum_{s=-1..5} dfb(s) u(q+s)
//   I: q = 6..n-7      u(q) -= sigma * sum_{s=-6..6} d(s)   u(q+s)
//   R: q = n-2..n-7    u(q) -= sigma * sum_{s=-1..5} dfb(s) u(q-s)
// each a rank-one update x <- (I - sigma e_q a^T) x.  Its transpose
// lam <- (I - sigma a e_q^T) lam is "t = sigma lam(q); lam(c) -= a_c t", and
// the transposed sweep runs the updates in reverse order: R ascending,
// I descending, L descending.
//
// Chunked like the forward filter: one lane computes the outputs [a, b) of
// a line.  After op q (descending) position q+6 is final (later ops touch
// q+5 at most), so the lane sweeps from q0 down to a-6 and emits q+6 when it
// lies in [a, b).  Near the right end q0 = n-7 after R^T (exact); elsewhere
// q0 = b+6+W starts from the unmodified input: the ops above q0 are missing,
// and their effect on t_q decays like the forward recurrence's (|h_n| below
// 1e-19 in total past W).  The chunk holding position 0 ends with L^T.
// A 13-value window in registers (rotating, unrolled by 13) holds
// positions q-6..q+6.
// ---------------------------------------------------------------------------
template <typename S>
struct LineRef {
    S *p;
    size_t st;
    __host__ __device__ __forceinline__ S &operator()(int q) const { return p[(size_t)q * st]; }
};

template <typename S>
__host__ __device__ __forceinline__ S line_rd(const LineRef<const S> &r, int q)
{
    return r(q);
}
template <typename S>
__host__ __device__ __forceinline__ void line_wr(const LineRef<S> &r, int q, S v)
{
    r(q) = v;
}

// x direction: lanes are rows, so reading one position of 64 rows touches
// 64 cache lines.  XDescIO stages T-position tiles of the wave's 64 rows in
// LDS (each row segment one 128-byte line, loaded coalesced), one tile
// ahead in registers, in the order the transposed sweep reads: every read
// lies below all earlier ones (the first window is read from its top).
// Outputs collect in a second tile, written back coalesced when the sweep
// moves below it; only positions in [a, b) are written, and the sweep emits
// every one of them while its tile is current (the left end's last 12 in
// the tile of position 0).
template <typename S>
struct XDescIO {
    static constexpr int T = 128 / sizeof(S);
    static constexpr int RP = 64 / T;  // rows per pass
    const S *in;
    S *out;
    int pitch, n, m, row0, a, b, lane;
    S (*tin)[T + 1];
    S (*tout)[T + 1];
    int tc;  // tile in LDS: [tc, tc + T)
    int oc;  // output tile being filled: [oc, oc + T); -1 none
    S pre[T];  // the tile below, in flight
    __device__ __forceinline__ void fetch(S (&dst)[T], int t)
    {
        // clamped indices, every load issued: values outside the grid land
        // in slots that are never read
        const int c = lane % T, rs = lane / T;
        int gi = t + c;
        gi = gi < 0 ? 0 : (gi < n ? gi : n - 1);
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
    __device__ __forceinline__ void start(int p_top)  // the first read is position p_top
    {
        tc = p_top & ~(T - 1);
        fetch(pre, tc);
        commit();
        fetch(pre, tc - T);
        oc = -1;
    }
    __device__ __forceinline__ S load(int p)
    {
        if (p < tc) {  // uniform: all lanes read the same positions
            commit();
            tc -= T;
            fetch(pre, tc - T);
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
    __device__ __forceinline__ void emit(int e, S v)
    {
        const int t = e & ~(T - 1);
        if (t != oc) {
            flush();
            oc = t;
        }
        tout[lane][e - oc] = v;
    }
};

template <typename S>
__device__ __forceinline__ S line_rd(XDescIO<S> &io, int q)
{
    return io.load(q);
}
template <typename S>
__device__ __forceinline__ void line_wr(XDescIO<S> &io, int q, S v)
{
    io.emit(q, v);
}

template <typename S, bool YDIR, typename IN, typename OUT>
__host__ __device__ __forceinline__ void adj_filter_line(IN &In, OUT &Out, int n, S sigma, int a, int b, int W)
{
    const bool exact_right = b + 6 + W >= n - 7;
    int q = exact_right ? n - 7 : b + 6 + W;
    const int qs = a - 6 > 6 ? a - 6 : 6;
    S w[13];
    // window = positions q-6 .. q+6 (slot k = position q-6+k), read from the top
#pragma unroll
    for (int k = 12; k >= 0; --k) w[k] = line_rd(In, q - 6 + k);
    if (exact_right) {
        // R^T: ops n-7 .. n-2 ascending; slot of position p is p-(n-13)
#pragma unroll
        for (int qq = 6; qq <= 11; ++qq) {
            const S t = sigma * w[qq];
#pragma unroll
            for (int s = -1; s <= 5; ++s) w[qq - s] = w[qq - s] - S(filt_dfb<YDIR>(s + 1)) * t;
        }
    }
    // I^T from q down to qs.  At step u position q+s sits in slot (6+s-u) mod 13.
    // Whole blocks of 13 steps that stop above qs need no exit test per
    // step, and their 13 loads can all be issued up front: the blocks whose
    // emitted positions (q+6 .. q-6) all lie at or above b emit nothing, the
    // block that reaches b emits through the bounds test, the blocks after it
    // emit unchecked.  The same operations, in the same order, as the
    // general loop below.
    auto bulk = [&](auto mode) {
        constexpr int M = decltype(mode)::value;  // 0: no emits, 1: checked, 2: unchecked
#pragma unroll
        for (int u = 0; u < 13; ++u) {
            const S t = sigma * w[(6 - u + 13) % 13];
#pragma unroll
            for (int s = -6; s <= 6; ++s) {
                const int sl = (6 + s - u + 26) % 13;
                w[sl] = w[sl] - S(filt_d(s + 6)) * t;
            }
            const int sl = (12 - u + 13) % 13;  // position q-u+6, final now
            const int e = q - u + 6;
            if constexpr (M == 1) {
                if (e >= a && e < b) line_wr(Out, e, w[sl]);
            }
            if constexpr (M == 2) line_wr(Out, e, w[sl]);
            w[sl] = line_rd(In, q - u - 7);
        }
        q -= 13;
    };
    while (q - 13 >= qs && q - 6 >= b) bulk(std::integral_constant<int, 0>{});
    if (q - 13 >= qs && q + 6 >= b) bulk(std::integral_constant<int, 1>{});
    while (q - 13 >= qs && q - 6 >= a) bulk(std::integral_constant<int, 2>{});
    int rot = 0;
    for (;;) {
        bool done = false;
#pragma unroll
        for (int u = 0; u < 13; ++u) {
            if (!done) {
                const S t = sigma * w[(6 - u + 13) % 13];
#pragma unroll
                for (int s = -6; s <= 6; ++s) {
                    const int sl = (6 + s - u + 26) % 13;
                    w[sl] = w[sl] - S(filt_d(s + 6)) * t;
                }
                const int sl = (12 - u + 13) % 13;  // position q+6, final now
                if (q + 6 >= a && q + 6 < b) line_wr(Out, q + 6, w[sl]);
                if (q == qs) {
                    rot = u;
                    done = true;
                } else {
                    w[sl] = line_rd(In, q - 7);
                    --q;
                }
            }
        }
        if (done) break;
    }
    if (qs == 6) {
        // window = positions 0..12, position p in slot (p - rot) mod 13; L^T on 0..11
        S v[13];
#pragma unroll
        for (int k = 0; k < 13; ++k) {
            S x = w[0];
#pragma unroll
            for (int r = 0; r < 13; ++r)
                if (r == rot) x = w[(k - r + 13) % 13];
            v[k] = x;
        }
#pragma unroll
        for (int qq = 6; qq >= 1; --qq) {
            const S t = sigma * v[qq];
#pragma unroll
            for (int s = -1; s <= 5; ++s) v[qq + s] = v[qq + s] - S(filt_dfb<YDIR>(s + 1)) * t;
        }
#pragma unroll
        for (int k = 0; k < 12; ++k)
            if (k >= a && k < b) line_wr(Out, k, v[k]);
    }
}

template <typename S, bool YDIR>
__global__ __launch_bounds__(64) void adj_filter_kernel(AdjFilterArgs<S> F)
{
    const int k = blockIdx.z;
    if (k >= F.nk) return;
    int a, b;
    if (!chunk_bounds(blockIdx.y, F.L, F.n, a, b)) return;
    if constexpr (YDIR) {  // line = column i, positions along j (coalesced across lanes)
        const int line = blockIdx.x * 64 + threadIdx.x;
        if (line >= F.m) return;
        LineRef<const S> In;
        LineRef<S> Out;
        In.p = F.in[k] + line;
        Out.p = F.out[k] + line;
        In.st = Out.st = (size_t)F.pitch;
        adj_filter_line<S, true>(In, Out, F.n, S(F.sigma), a, b, F.W);
    } else {  // line = row j: every lane takes part in the tile loads (rows past the grid clamp)
        constexpr int T = XDescIO<S>::T;
        __shared__ S tin[64][T + 1];
        __shared__ S tout[64][T + 1];
        XDescIO<S> io;
        io.in = F.in[k];
        io.out = F.out[k];
        io.pitch = F.pitch;
        io.n = F.n;
        io.m = F.m;
        io.row0 = blockIdx.x * 64;
        io.a = a;
        io.b = b;
        io.lane = threadIdx.x;
        io.tin = tin;
        io.tout = tout;
        const bool exact_right = b + 6 + F.W >= F.n - 7;
        io.start((exact_right ? F.n - 7 : b + 6 + F.W) + 6);
        adj_filter_line<S, false>(io, io, F.n, S(F.sigma), a, b, F.W);
        io.flush();
    }
}

// ---------------------------------------------------------------------------
// Pointwise backward of one stage at (i, j), given lam = dJ/dU'' (after the
// buffer, i.e. F^T ccpml^T already applied) and lamW = dJ/dW':
//   buffer^T : lam' = sy * sx * lam in the layers (rk.f90:81-108)
//   update^T : Wt = lamW + beta lam' ; lam_f = dt Wt ; lamW_out = alpha Wt
//              (interior points, rk.f90:160-167; elsewhere lam_f = 0)
//   RHS^T    : with L = lam_f and the Euler RHS (eqeuler.f90:63-84)
//     f0 = -(u rx + r ux + v ry + r vy)      f1 = -(u ux + px/r + v uy)
//     f2 = -(u vx + v vy + py/r)             f3 = -(g p ux + u px + g p vy + v py)
//   the pointwise part  lam_U += (df/dU)^T L  and the flux adjoints
//   G_x = metx * (df/d(U_x))^T L, G_y = mety * (df/d(U_y))^T L, which the
//   gather kernel turns into D_x^T G_x + D_y^T G_y.
// ---------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void adj_point_kernel(AdjPointArgs<S> P)
{
    using R = S;
    const StageArgs<S> &A = P.st;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= A.imax || j >= A.jmax) return;
    const int imax = A.imax, jmax = A.jmax, pitch = A.pitch;
    const size_t g = (size_t)j * pitch + i;
    const bool ybuf = j < A.D || j >= jmax - A.D;
    const bool xbuf = i < A.D || i >= imax - A.D;
    R bs = R(1);
    if (ybuf) bs = bs * R(A.sigy[j]);
    if (xbuf) bs = bs * R(A.sigx[i]);
    const bool interior = P.rhs_only || (i >= 1 && i <= imax - 2 && j >= 1 && j <= jmax - 2);
    // every load of the point is issued before its first store: a load
    // after a store to another plane would wait for it (the planes may alias
    // as far as the compiler knows), and the planes' round trips would run
    // one after another
    R lamin[4], lamw[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        lamin[k] = P.lam[k][g];
        lamw[k] = (!P.rhs_only && interior) ? P.lamW[k][g] : R(0);
    }
    R lamp[4], Lf[4], lwo[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        lwo[k] = R(0);
        if (P.rhs_only) {  // R'^T alone, at every point (one-sided rows at the edges)
            lamp[k] = R(0);
            Lf[k] = lamin[k];
            continue;
        }
        lamp[k] = bs * lamin[k];
        if (interior) {
            const R wt = lamw[k] + R(A.beta) * lamp[k];
            Lf[k] = R(A.dt) * wt;
            lwo[k] = A.first ? R(0) : R(A.alpha) * wt;
        } else {
            Lf[k] = R(0);
        }
    }
    if (!interior) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            P.lamWo[k][g] = lwo[k];
            P.acc[k][g] = lamp[k];
            P.gx[k][g] = R(0);
            P.gy[k][g] = R(0);
        }
        return;
    }
    // forward quantities at the point (same stencils and metrics as the stage)
    const R mx = R(A.metx[i]), my = R(A.mety[j]);
    const R inv_dm = R(A.inv_dm), inv_dn = R(A.inv_dn);
    R uc[4], fx[4], fy[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const S *pl = A.U[k];
        const R *row = pl + (size_t)j * pitch;
        const R *col = pl + i;
        fx[k] = drp_deriv<R>([row](int q) { return row[q]; }, i, imax, inv_dm) * mx;
        fy[k] = drp_deriv<R>([col, pitch](int q) { return col[(size_t)q * pitch]; }, j, jmax, inv_dn) * my;
        R b = R(0);
        if (P.pert) {
            b = R(A.pr.ub[k][j]);
            fy[k] = fy[k] + R(A.pr.dub[k][j]);
        }
        uc[k] = pl[g] + b;
    }
    const R r = uc[0], u = uc[1], v = uc[2], p = uc[3], gam = R(A.gamma);
    const R ir = R(1) / r;
    const R div = fx[1] + fy[2];
    // pointwise part (df/dU)^T L
    R acc[4];
    acc[0] = -Lf[0] * div + (Lf[1] * fx[3] + Lf[2] * fy[3]) * ir * ir;
    acc[1] = -Lf[0] * fx[0] - Lf[1] * fx[1] - Lf[2] * fx[2] - Lf[3] * fx[3];
    acc[2] = -Lf[0] * fy[0] - Lf[1] * fy[1] - Lf[2] * fy[2] - Lf[3] * fy[3];
    acc[3] = -Lf[3] * gam * div;
    // flux adjoints
    R gx[4], gy[4];
    gx[0] = -Lf[0] * u;
    gx[1] = -Lf[0] * r - Lf[1] * u - Lf[3] * gam * p;
    gx[2] = -Lf[2] * u;
    gx[3] = -Lf[1] * ir - Lf[3] * u;
    gy[0] = -Lf[0] * v;
    gy[1] = -Lf[1] * v;
    gy[2] = -Lf[0] * r - Lf[2] * v - Lf[3] * gam * p;
    gy[3] = -Lf[2] * ir - Lf[3] * v;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!P.rhs_only) P.lamWo[k][g] = lwo[k];
        P.acc[k][g] = lamp[k] + acc[k];
        P.gx[k][g] = mx * gx[k];
        P.gy[k][g] = my * gy[k];
    }
}

// (D^T g)(c) on a line of n points: the interior rows r in [3, n-4] carry
// a_m (u(r+m) - u(r-m)) / h, rows 0..2 and n-3..n-1 the one-sided stencils.
template <typename R, typename F>
__host__ __device__ __forceinline__ R drp_transpose(const F &g, int c, int n, R inv_h)
{
    R s = R(0);
    const R a[3] = {R(DRP_A1), R(DRP_A2), R(DRP_A3)};
#pragma unroll
    for (int m = 1; m <= 3; ++m) {
        const int rl = c - m, rr = c + m;  // row rl has +a_m at column c, row rr has -a_m
        if (rl >= 3 && rl <= n - 4) s = s + a[m - 1] * g(rl);
        if (rr >= 3 && rr <= n - 4) s = s - a[m - 1] * g(rr);
    }
    if (c <= 6) {
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            R coef = R(0);
#pragma unroll
            for (int m = 0; m < 7; ++m)
                if (m == c) coef = R(drp_left(r, m));
            s = s + coef * g(r);
        }
    }
    if (c >= n - 7) {
        const int m0 = c - (n - 7);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            R coef = R(0);
#pragma unroll
            for (int m = 0; m < 7; ++m)
                if (m == m0) coef = R(drp_right(r, m));
            s = s + coef * g(n - 3 + r);
        }
    }
    return s * inv_h;
}

template <typename S>
__global__ __launch_bounds__(256) void adj_gather_kernel(AdjGatherArgs<S> G)
{
    using R = S;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= G.imax || j >= G.jmax) return;
    const int pitch = G.pitch;
    const size_t g = (size_t)j * pitch + i;
    // all planes' loads before the first store (see adj_point_kernel)
    R o[NKMAX];
#pragma unroll
    for (int k = 0; k < NKMAX; ++k) {
        if (k >= G.nk) break;
        const R *rowx = G.gx[k] + (size_t)j * pitch;
        const R *coly = G.gy[k] + i;
        const R dx = drp_transpose<R>([rowx](int q) { return rowx[q]; }, i, G.imax, R(G.inv_dm));
        const R dy = drp_transpose<R>([coly, pitch](int q) { return coly[(size_t)q * pitch]; }, j, G.jmax,
                                      R(G.inv_dn));
        o[k] = G.acc[k][g] + dx + dy;
    }
#pragma unroll
    for (int k = 0; k < NKMAX; ++k) {
        if (k >= G.nk) break;
        G.out[k][g] = o[k];
    }
}

// ---------------------------------------------------------------------------
// Observer objective phi(U) = sum over the arc points of (p - pb)^2: one
// block reduces it in a fixed order (deterministic), and its gradient
// 2 (p - pb) (times a weight) is added to lam at the same points.
// ---------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void arc_reduce_kernel(ArcArgs<S> a, double *out)
{
    __shared__ double part[256];
    double s = 0.0;
    for (int t = threadIdx.x; t < a.npts; t += 256) {
        const int2 ij = a.pts[t];
        const size_t g = (size_t)ij.y * a.pitch + ij.x;
        double d = (double)a.U[g];
        if (a.base) d -= a.base[g];
        s += d * d;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0];
}

template <typename S>
__global__ __launch_bounds__(256) void arc_seed_kernel(ArcArgs<S> a, S *lam, double weight)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= a.npts) return;
    const int2 ij = a.pts[t];
    const size_t g = (size_t)ij.y * a.pitch + ij.x;
    double d = (double)a.U[g];
    if (a.base) d -= a.base[g];
    lam[g] = lam[g] + S(2.0 * weight * d);
}

template <typename Sin, typename Sout>
__global__ __launch_bounds__(256) void convert_planes_kernel(ConvArgs<Sin, Sout> a)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= a.imax || j >= a.jmax) return;
    for (int k = 0; k < a.nk; ++k) {
        double v = (double)a.in[(size_t)k * a.iplane + (size_t)j * a.ipitch + i];
        if (a.sub[k]) v -= a.sgn * a.sub[k][j];
        a.out[(size_t)k * a.oplane + (size_t)j * a.opitch + i] = Sout(v);
    }
}

template <typename Sin, typename Sout>
static hipError_t launch_convert_t(const ConvArgs<Sin, Sout> &a, hipStream_t s)
{
    hipLaunchKernelGGL((convert_planes_kernel<Sin, Sout>), dim3((a.imax + 63) / 64, (a.jmax + 3) / 4), dim3(256), 0,
                       s, a);
    return hipGetLastError();
}

template <typename S>
static hipError_t launch_adj_stage_t(const AdjPointArgs<S> &P, const AdjGatherArgs<S> &G, hipStream_t s)
{
    const dim3 grid((P.st.imax + 63) / 64, (P.st.jmax + 3) / 4);
    hipLaunchKernelGGL(adj_point_kernel<S>, grid, dim3(256), 0, s, P);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(adj_gather_kernel<S>, grid, dim3(256), 0, s, G);
    return hipGetLastError();
}

template <typename S>
static hipError_t launch_adj_filter_t(const AdjFilterArgs<S> &F, bool ydir, hipStream_t s)
{
    const dim3 grid((F.m + 63) / 64, (F.n + F.L - 1) / F.L, F.nk);
    if (ydir) hipLaunchKernelGGL((adj_filter_kernel<S, true>), grid, dim3(64), 0, s, F);
    else hipLaunchKernelGGL((adj_filter_kernel<S, false>), grid, dim3(64), 0, s, F);
    return hipGetLastError();
}

template <typename S>
static hipError_t launch_adj_edge_t(const AdjEdgeArgs<S> &E, hipStream_t s)
{
    for (int first = 0; first < 4; first += 2) {
        const int n = first < 2 ? E.imax : E.jmax;
        hipLaunchKernelGGL(adj_edge_kernel<S>, dim3((n + 255) / 256), dim3(256), 0, s, E, first, first + 1);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename S>
static hipError_t launch_arc_reduce_t(const ArcArgs<S> &a, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(arc_reduce_kernel<S>, dim3(1), dim3(256), 0, s, a, out);
    return hipGetLastError();
}

template <typename S>
static hipError_t launch_arc_seed_t(const ArcArgs<S> &a, S *lam, double weight, hipStream_t s)
{
    if (a.npts <= 0) return hipSuccess;
    hipLaunchKernelGGL(arc_seed_kernel<S>, dim3((a.npts + 255) / 256), dim3(256), 0, s, a, lam, weight);
    return hipGetLastError();
}

#if defined(NS_KERNELS_F64)
#define NS_ADJ_T double
#else
#define NS_ADJ_T float
#endif
hipError_t launch_adj_stage(const AdjPointArgs<NS_ADJ_T> &P, const AdjGatherArgs<NS_ADJ_T> &G, hipStream_t s)
{
    return launch_adj_stage_t(P, G, s);
}
hipError_t launch_adj_filter(const AdjFilterArgs<NS_ADJ_T> &F, bool ydir, hipStream_t s)
{
    return launch_adj_filter_t(F, ydir, s);
}
hipError_t launch_adj_edge(const AdjEdgeArgs<NS_ADJ_T> &E, hipStream_t s) { return launch_adj_edge_t(E, s); }
hipError_t launch_arc_reduce(const ArcArgs<NS_ADJ_T> &a, double *out, hipStream_t s)
{
    return launch_arc_reduce_t(a, out, s);
}
hipError_t launch_arc_seed(const ArcArgs<NS_ADJ_T> &a, NS_ADJ_T *lam, double weight, hipStream_t s)
{
    return launch_arc_seed_t(a, lam, weight, s);
}
hipError_t launch_convert(const ConvArgs<double, NS_ADJ_T> &a, hipStream_t s) { return launch_convert_t(a, s); }
#if defined(NS_KERNELS_F32)
hipError_t launch_convert(const ConvArgs<float, double> &a, hipStream_t s) { return launch_convert_t(a, s); }
#endif
#undef NS_ADJ_T

}  // namespace nsamd
