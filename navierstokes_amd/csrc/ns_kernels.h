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
    int xper;         // periodic in x: wrapped x stencils, no x buffer layers, every column updated
};

// Column x of a grid of n columns, wrapped into [0, n) (periodic x; x may
// lie up to one period outside).
__host__ __device__ inline int wrap_col(int x, int n) { return x < 0 ? x + n : (x >= n ? x - n : x); }

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
    int xper;     // periodic x: the x filter sweeps around the ring; the y
                  // filter's ccpml rows cover every column
    RowEdge re;
};

// Pointwise pieces of rhs_euler on pitched planes (operator-level API).
struct PwArgs {
    int imax, jmax, pitch;
};

// EULERNONLINEAR (eqeuler.f90:61-97): U, dU/dm, dU/dn nk planes each; tau,
// dtau/dm, dtau/dn 3 planes each; d2p* (nk = 5 only) and src (s_acou,
// imax x jmax unpitched) may be null
struct EulerPwArgs {
    PwArgs a;
    int nk;
    const double *U[NKMAX], *Um[NKMAX], *Un[NKMAX];
    const double *tau[3], *taum[3], *taun[3];
    const double *d2Tm, *d2Tn, *d2pm, *d2pn, *src;
    double gamma, Re, Pr, Pe;
    double *out[NKMAX];
};

// Both filters in one sweep (filter_xy_kernel): x filter from `in`, y filter
// (with the ccpml rows) into `out`.
template <typename S>
struct FilterXYArgs {
    const S *in[NKMAX];
    S *out[NKMAX];
    const double *corr[NKMAX];  // perturbation form: Ly(Ub) - Ub per row (null in FP64)
    int imax, jmax, pitch;
    int Ly;       // row chunk length
    int W;        // look-back of both recurrences
    int x_exact;  // every strip starts its x recurrence at the left end (small grids: the reference order)
    int xper;     // periodic x: the x recurrence runs around the ring
    double sigma;
    int nk;
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
    int xper;  // periodic x: no inflow / outflow columns (the kernel does nothing)
};

// Filter chunk bounds: chunks of length L; a chunk that would end within 16
// points of the line end absorbs the rest; a chunk starting there is empty.
__host__ __device__ inline bool chunk_bounds(int c, int L, int n, int &a, int &b)
{
    a = c * L;
    b = a + L < n ? a + L : n;
    if (a > 0 && a > n - 16) return false;
    if (b > n - 16) b = n;
    return true;
}

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

// Up to four rectangles [x0,x1) x [y0,y1), each cut into tw x ty tiles from
// its own origin: the grid around an inner rectangle (bottom band, top band,
// left and right strips), or the whole grid when the inner one is empty.
// `which` selects rectangles (bit 0 bottom, 1 top, 2 left, 3 right), so the
// bands and the side strips can use differently shaped tiles.
struct RectSet {
    int x0[4], x1[4], y0[4], y1[4];
    int tx[4];      // tiles across each rectangle
    int start[5];   // first tile of each rectangle (start[4] = total)
    int tw, ty;     // tile width and height
    __host__ __device__ int count() const { return start[4]; }
    __host__ __device__ void tile(int t, int &i0, int &i1, int &j0, int &j1) const
    {
        int r = 0;
        while (r < 3 && t >= start[r + 1]) ++r;
        t -= start[r];
        const int by = t / tx[r], bx = t - by * tx[r];
        i0 = x0[r] + bx * tw;
        i1 = i0 + tw < x1[r] ? i0 + tw : x1[r];
        j0 = y0[r] + by * ty;
        j1 = j0 + ty < y1[r] ? j0 + ty : y1[r];
    }
    static RectSet around(int imax, int jmax, int tw, int ty, int xa, int xb, int ya, int yb, int which = 15)
    {
        RectSet s{};
        s.tw = tw;
        s.ty = ty;
        if (xb <= xa || yb <= ya) xa = xb = ya = yb = 0;
        const int r[4][4] = {{0, imax, 0, ya}, {0, imax, yb, jmax}, {0, xa, ya, yb}, {xb, imax, ya, yb}};
        int n = 0;
        for (int q = 0; q < 4; ++q) {
            int X0 = r[q][0], X1 = r[q][1], Y0 = r[q][2], Y1 = r[q][3];
            if (q == 0 && xa == xb) Y1 = jmax;  // empty inner rectangle: the whole grid
            s.x0[q] = X0;
            s.x1[q] = X1;
            s.y0[q] = Y0;
            s.y1[q] = Y1;
            const bool empty = X1 <= X0 || Y1 <= Y0 || (q > 0 && xa == xb) || !((which >> q) & 1);
            s.tx[q] = empty ? 1 : (X1 - X0 + tw - 1) / tw;
            s.start[q] = n;
            if (!empty) n += s.tx[q] * ((Y1 - Y0 + ty - 1) / ty);
        }
        s.start[4] = n;
        return s;
    }
};

// Optional hooks of a stage launch: events recorded on the interior
// kernel's stream right before and after its launch, and the number of
// output points that kernel covers (the whole grid when the edge tiles run
// in the same launch); plan_only: do the host-side set-up (launch plan, LDS
// attributes, which load the kernels) and launch nothing.
struct StageHooks {
    hipEvent_t before = nullptr, after = nullptr;
    long long interior_points = 0;
    bool plan_only = false;
    // events for running the edge tiles on the second stream: without them
    // (or without a second stream) everything runs on the first
    hipEvent_t fork = nullptr, join = nullptr;
    // marching stages: the small edge tiles in the marching launch instead
    // (stage_march_edges_kernel); plan_only sets up both forms
    bool same_launch = false;
};

// y filter sweep blocks: 4 waves on adjacent column blocks, which walk down
// the same rows together (1 KB of each FP32 row, 2 KB of each FP64 row at a
// time: FP32 y sweep 0.128 -> 0.116 ms at 4096^2; FP64 step 6.444 -> 6.401
// ms once its rows were padded, where before the pad 4-wave blocks had
// measured slower in FP64)
template <typename R>
constexpr int filter_y_waves()
{
    return 4;
}

struct MarchMap {
    int xa, xe, nstrips, ya, yb, H;  // output columns [xa, xe): nstrips strips of 58, the last one partial
    int y_a, y_b;
    int fx0, fx1;
};

// One case of a batched marching stage (stage_march_batch_kernel): its
// marching region and chunks, and the small edge tiles around the region
// (the whole grid as band tiles when there is no region: nstrips = 0).
struct BatchMarch {
    MarchMap m;
    RectSet sides, bands;
    int nchunks;
};

__host__ __device__ inline int march_chunks(const MarchMap &m)
{
    return (m.y_a > m.ya ? 1 : 0) + (m.y_b - m.y_a + m.H - 1) / m.H + (m.yb > m.y_b ? 1 : 0);
}

// rows [y0, y1) of chunk c (blockIdx.y): [ya, y_a) when not empty, then
// chunks of H rows over [y_a, y_b), then [y_b, yb) when not empty
__host__ __device__ inline void march_chunk_rows(const MarchMap &m, int c, int &y0, int &y1)
{
    if (m.y_a > m.ya) {
        if (c == 0) {
            y0 = m.ya;
            y1 = m.y_a;
            return;
        }
        --c;
    }
    const int nmid = (m.y_b - m.y_a + m.H - 1) / m.H;
    if (c < nmid) {
        y0 = m.y_a + c * m.H;
        y1 = y0 + m.H < m.y_b ? y0 + m.H : m.y_b;
    } else {
        y0 = m.y_b;
        y1 = m.yb;
    }
}

// Workgroups of a launch go to the 8 XCDs in turn (block L to XCD L % 8),
// and each XCD has its own L2.  Block L of T marches tile k = (chunk,
// strip), strips minor, with the tiles split into 8 contiguous runs, run x
// for XCD x: the strips an XCD runs side by side share their 6-column
// halos, and a chunk's 12 prologue rows are the previous chunk's last rows,
// in the same L2.  A bijection of [0, T) for any T.
constexpr int NXCD = 8;
__host__ __device__ inline void march_tile_of(int L, int T, int nstrips, int &strip, int &chunk)
{
    const int x = L % NXCD, q = L / NXCD;
    const int a = T / NXCD, b = T % NXCD;  // run x holds a + (x < b) tiles
    const int k = x * a + (x < b ? x : b) + q;
    chunk = k / nstrips;
    strip = k - chunk * nstrips;
}

// hybrid layout: the block with outputs [i0, i0 + 58) x [y0, y1) lies in the
// fast zone and takes the fast per-point code
__host__ __device__ inline bool march_block_fast(const MarchMap &m, int i0, int y0, int y1)
{
    return i0 >= m.fx0 && i0 + 58 <= m.fx1 && y0 >= m.y_a && y1 <= m.y_b;
}

// Output columns / rows of the marching region for strips of 58 (the last
// one partial: its lanes past xe compute from loads clamped to the grid and
// store nothing) and steps of rb rows, rows a multiple of rb.  fast: every output's 3-point level-1 halo
// inside [max(D,4), n - max(D,4)) in both directions (no buffer, metric or
// one-sided stencil there).  General: outputs in [6, n - 6), so the U window
// (6-point halo) stays inside the grid and every derivative taken in it (U at
// the level-1 rows and columns, 3 points beyond the outputs) is an interior
// stencil (3 <= q <= n - 4); buffers and metrics are then handled point by
// point.
// False when the region is too small.
// Periodic x (xper): every column is an output column (the window columns
// wrap around the period) and there are no x layers.
constexpr int GEN_LO = 6;  // first output row / column of the general region
__host__ __device__ inline bool march_region(bool fast, int D, int imax, int jmax, int rb, MarchMap &m,
                                             bool xper = false)
{
    const int lo = fast ? (D > 4 ? D : 4) + 3 : GEN_LO;
    m.xa = xper ? 0 : lo;
    m.xe = xper ? imax : imax - lo;
    m.nstrips = m.xe > m.xa ? (m.xe - m.xa + 57) / 58 : 0;
    m.ya = lo;
    const int rows = jmax - lo - m.ya;
    m.yb = m.ya + (rows > 0 ? rows / rb * rb : 0);
    m.H = 0;
    m.y_a = m.ya;
    m.y_b = m.yb;
    m.fx0 = m.fx1 = 0;
    return m.xe - m.xa >= 58 && m.yb - m.ya >= 4 * rb;
}

// General region with the fast zone marked (unit grids outside the layers):
// columns as the general region; rows split at the fast zone's bounds, each
// part a multiple of rb.  False when there is no fast zone of at least 4 rb
// rows; then the plain general region applies.
__host__ __device__ inline bool march_hybrid(int D, int imax, int jmax, int rb, MarchMap &m, bool xper = false)
{
    MarchMap f;
    if (!march_region(true, D, imax, jmax, rb, f, xper) || !march_region(false, D, imax, jmax, rb, m, xper))
        return false;
    const int lo = f.ya;            // fast zone rows/columns [lo, n - lo)
    m.ya = GEN_LO + (lo - GEN_LO) % rb;  // lo - ya: a multiple of rb
    m.y_a = lo;
    m.y_b = lo + (jmax - 2 * lo) / rb * rb;
    const int top = jmax - GEN_LO - m.y_b;
    m.yb = m.y_b + (top > 0 ? top / rb * rb : 0);
    m.fx0 = xper ? 0 : lo;
    m.fx1 = xper ? imax : imax - lo;
    return m.y_b - m.y_a >= 4 * rb;
}

// Small edge tiles (256 threads, beside marching blocks): 16 x 16 over the
// side strips, 32 x 8 over the bands.
constexpr int SMALL_SX = 16, SMALL_SY = 16, SMALL_BX = 32, SMALL_BY = 8;

// Edge tiles of batched marching stages (256 threads).  Without level-1
// rows (Euler, nk = 4) the frame around each case's marching region, 6
// columns or rows wide, takes 8 x 64 tiles on the side strips and 64 x 8 on
// the bands: a quarter of the small tiles' count for the same points, in
// less LDS than the march.  With level-1 rows such tiles would need more
// LDS than the march, and the small shapes stay.
template <bool FRAME>
struct BatchEdgeShape {
    static constexpr int SX = FRAME ? 8 : SMALL_SX, SY = FRAME ? 64 : SMALL_SY;
    static constexpr int BX = FRAME ? 64 : SMALL_BX, BY = FRAME ? 8 : SMALL_BY;
};

// The plan of one case of a batched marching stage: the general region with
// the fast zone marked (march_hybrid) when `hybrid` allows it and the case
// has one, else the plain general region (march_region); chunks of H rows
// between the edge chunks (H <= 0: one chunk); the edge tiles around the
// region (BatchEdgeShape<frame>), or the whole grid as band tiles when the
// case is too small to march (nstrips = 0, nchunks = 0).
template <bool FRAME>
inline void batch_march_plan(int imax, int jmax, int D, bool xper, bool hybrid, int rb, int H, BatchMarch &p)
{
    using E = BatchEdgeShape<FRAME>;
    MarchMap m{};
    bool ok = hybrid && march_hybrid(D, imax, jmax, rb, m, xper);
    if (!ok) ok = march_region(false, D, imax, jmax, rb, m, xper);
    int xa = 0, xb = 0, ya = 0, yb = 0;
    if (ok) {
        const int rows = m.y_b - m.y_a;
        const int h = H > 0 && H < rows ? H : rows;
        m.H = (h + rb - 1) / rb * rb;
        xa = m.xa, xb = m.xe, ya = m.ya, yb = m.yb;
    } else {
        m = MarchMap{};
    }
    p.m = m;
    p.nchunks = ok ? march_chunks(m) : 0;
    p.sides = RectSet::around(imax, jmax, E::SX, E::SY, xa, xb, ya, yb, ok ? 12 : 0);
    p.bands = RectSet::around(imax, jmax, E::BX, E::BY, xa, xb, ya, yb, ok ? 3 : 15);
}

}  // namespace nsamd
