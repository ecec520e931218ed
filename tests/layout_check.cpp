// Host-side check of the stage's launch layouts (navierstokes_amd/csrc/ns_kernels.h):
// for many grids, buffer widths, rows per step and chunk heights, the
// marching region (fast zone, general region, general region with fast
// blocks) plus the edge tiles around it (RectSet::around: wide tiles over the
// bands below and above, narrow tall tiles over the side strips) must cover
// every grid point exactly once, every marching block must keep its U window and
// level-1 rows inside the grid and step a whole number of RB rows, and
// every block that takes the fast per-point code must have its outputs'
// 3-point level-1 halo inside the fast zone; and the XCD numbering of the
// marching blocks (march_tile_of) must be a bijection onto the tiles with
// each XCD's tiles consecutive; and a batched case's marching blocks and
// small edge tiles (batch_march_plan) must cover its grid once.  Prints "ok
// <layouts>" or the first failure, and exits non-zero on failure.
#include <cstdio>
#include <vector>

#include "ns_kernels.h"

using namespace nsamd;

static int fail(const char *what, int imax, int jmax, int D, int rb, int mode, int H)
{
    std::printf("FAIL %s: imax=%d jmax=%d D=%d rb=%d mode=%d H=%d\n", what, imax, jmax, D, rb, mode, H);
    return 1;
}

// mode 0: fast zone; 1: general region; 2: general region with fast blocks;
// xper: periodic x (every column marched, windows wrap within one period)
static int check(int imax, int jmax, int D, int rb, int mode, int H, long &layouts, bool xper = false)
{
    MarchMap m{};
    const bool ok = mode == 0 ? march_region(true, D, imax, jmax, rb, m, xper)
                    : mode == 1 ? march_region(false, D, imax, jmax, rb, m, xper)
                                : march_hybrid(D, imax, jmax, rb, m, xper);
    if (!ok) return 0;
    if ((m.y_b - m.y_a) < rb || H % rb) return 0;
    m.H = H;
    ++layouts;
    const int Dm = D > 4 ? D : 4;
    std::vector<int> cnt((size_t)imax * jmax, 0);
    const int nch = march_chunks(m);
    if (m.nstrips != (m.xe - m.xa + 57) / 58 || m.xe > imax) return fail("strips", imax, jmax, D, rb, mode, H);
    for (int s = 0; s < m.nstrips; ++s)
        for (int c = 0; c < nch; ++c) {
            const int i0 = m.xa + 58 * s;
            const int ie = i0 + 58 < m.xe ? i0 + 58 : m.xe;  // stored outputs [i0, ie); the last strip may be partial
            int y0, y1;
            march_chunk_rows(m, c, y0, y1);
            if (y1 <= y0 || (y1 - y0) % rb) return fail("chunk rows", imax, jmax, D, rb, mode, H);
            // the stored outputs' window: rows y0-6 .. y1+5 (level-1 rows
            // y0-3 .. y1+2), columns i0-6 .. ie+5 (lanes past ie load clamped
            // columns and store nothing)
            if (ie <= i0 || y0 - 6 < 0 || y1 + 5 > jmax - 1)
                return fail("window outside the grid", imax, jmax, D, rb, mode, H);
            if (xper) {
                // the whole 70-column window (lanes past ie too) wraps into
                // the grid with one period: wrap_col's range
                if (i0 - 6 < -imax || i0 + 63 >= 2 * imax || wrap_col(i0 - 6, imax) < 0 ||
                    wrap_col(i0 + 63, imax) >= imax)
                    return fail("periodic window beyond one period", imax, jmax, D, rb, mode, H);
            } else if (i0 - 6 < 0 || ie + 5 > imax - 1) {
                return fail("window outside the grid", imax, jmax, D, rb, mode, H);
            }
            // every derivative the stored outputs depend on is the interior
            // one (periodic x: always along x)
            if (y0 - 3 < 3 || y1 + 2 > jmax - 4 || (!xper && (i0 - 3 < 3 || ie + 2 > imax - 4)))
                return fail("one-sided stencil in a marching block", imax, jmax, D, rb, mode, H);
            const bool fast = mode == 0 || (mode == 2 && march_block_fast(m, i0, y0, y1));
            if (fast && (y0 - 3 < Dm || y1 + 3 > jmax - Dm ||
                         (!xper && (i0 - 3 < Dm || ie + 3 > imax - Dm))))
                return fail("fast block outside the fast zone", imax, jmax, D, rb, mode, H);
            if (xper && (m.xa != 0 || m.xe != imax)) return fail("periodic columns", imax, jmax, D, rb, mode, H);
            // a hybrid block on the fast code covers whole fast-zone columns
            if (mode == 2 && march_block_fast(m, i0, y0, y1) && ie != i0 + 58)
                return fail("partial strip on the fast code", imax, jmax, D, rb, mode, H);
            for (int j = y0; j < y1; ++j)
                for (int i = i0; i < ie; ++i) ++cnt[(size_t)j * imax + i];
        }
    // the stage launch's edge sets: bands below and above the marching
    // region in 64 x ty tiles, side strips beside it in 16 x 4ty tiles
    // (large), or 32 x 8 and 16 x 16 (small)
    const int shapes[3][4] = {{64, 8, 16, 32}, {64, 16, 16, 64}, {32, 8, 16, 16}};
    for (const auto &sh : shapes) {
        std::vector<int> c2 = cnt;
        const int xa = m.xa, xb = m.xe;
        const RectSet sets[2] = {RectSet::around(imax, jmax, sh[0], sh[1], xa, xb, m.ya, m.yb, 3),
                                 RectSet::around(imax, jmax, sh[2], sh[3], xa, xb, m.ya, m.yb, 12)};
        for (const RectSet &rs : sets)
            for (int t = 0; t < rs.count(); ++t) {
                int i0, i1, j0, j1;
                rs.tile(t, i0, i1, j0, j1);
                if (i0 < 0 || j0 < 0 || i1 > imax || j1 > jmax || i1 - i0 > rs.tw || j1 - j0 > rs.ty || i1 <= i0 ||
                    j1 <= j0)
                    return fail("edge tile", imax, jmax, D, rb, mode, H);
                for (int j = j0; j < j1; ++j)
                    for (int i = i0; i < i1; ++i) ++c2[(size_t)j * imax + i];
            }
        for (int v : c2)
            if (v != 1) return fail("coverage", imax, jmax, D, rb, mode, H);
    }
    return 0;
}

// no interior region: the bands' set alone tiles the whole grid, the side set is empty
static int check_whole(int imax, int jmax, int ty)
{
    const RectSet all = RectSet::around(imax, jmax, 64, ty, 0, 0, 0, 0, 15);
    const RectSet none = RectSet::around(imax, jmax, 16, 4 * ty, 0, 0, 0, 0, 0);
    if (none.count() != 0) return fail("side tiles without an interior", imax, jmax, 0, 0, -1, 0);
    std::vector<int> c((size_t)imax * jmax, 0);
    for (int t = 0; t < all.count(); ++t) {
        int i0, i1, j0, j1;
        all.tile(t, i0, i1, j0, j1);
        for (int j = j0; j < j1; ++j)
            for (int i = i0; i < i1; ++i) ++c[(size_t)j * imax + i];
    }
    for (int v : c)
        if (v != 1) return fail("whole-grid coverage", imax, jmax, 0, 0, -1, 0);
    return 0;
}

// the XCD-aware renumbering of marching blocks is a bijection onto the
// (chunk, strip) tiles, and XCD x's tiles are consecutive
static int check_tiles(int nstrips, int nchunks)
{
    const int T = nstrips * nchunks;
    std::vector<int> seen(T, 0);
    std::vector<int> last(NXCD, -1);
    for (int L = 0; L < T; ++L) {
        int s, c;
        march_tile_of(L, T, nstrips, s, c);
        if (s < 0 || s >= nstrips || c < 0 || c >= nchunks) return fail("tile index", nstrips, nchunks, 0, 0, -2, 0);
        const int k = c * nstrips + s;
        ++seen[k];
        if (last[L % NXCD] >= 0 && k != last[L % NXCD] + 1) return fail("tile run", nstrips, nchunks, 0, 0, -2, 0);
        last[L % NXCD] = k;
    }
    for (int v : seen)
        if (v != 1) return fail("tile bijection", nstrips, nchunks, 0, 0, -2, 0);
    return 0;
}

// batched marching stage (batch_march_plan): one case's marching blocks
// plus its small edge tiles cover the grid once, windows stay inside the
// grid, and on a plain plan (no fast zone marked) the chunk rows of the
// general-code path (ya + c H, clipped at yb) are those of march_chunk_rows
template <bool FRAME>
static int check_batch(int imax, int jmax, int D, bool hybrid, int H, long &layouts, bool xper = false)
{
    const int rb = 4;
    using E = BatchEdgeShape<FRAME>;
    BatchMarch p{};
    batch_march_plan<FRAME>(imax, jmax, D, xper, hybrid, rb, H, p);
    const MarchMap &m = p.m;
    ++layouts;
    std::vector<int> cnt((size_t)imax * jmax, 0);
    const bool plain = m.y_a == m.ya && m.y_b == m.yb;
    if (p.nchunks != (m.nstrips ? march_chunks(m) : 0)) return fail("batch chunks", imax, jmax, D, rb, 10 + hybrid, H);
    for (int c = 0; c < p.nchunks; ++c)
        for (int s = 0; s < m.nstrips; ++s) {
            int y0, y1;
            march_chunk_rows(m, c, y0, y1);
            if (plain) {
                const int z0 = m.ya + c * m.H, z1 = z0 + m.H < m.yb ? z0 + m.H : m.yb;
                if (z0 != y0 || z1 != y1) return fail("batch plain chunk rows", imax, jmax, D, rb, 10 + hybrid, H);
            } else if (!hybrid) {
                return fail("batch plan marks a fast zone", imax, jmax, D, rb, 10 + hybrid, H);
            }
            const int i0 = m.xa + 58 * s, ie = i0 + 58 < m.xe ? i0 + 58 : m.xe;
            if (y1 <= y0 || (y1 - y0) % rb || ie <= i0 || y0 - 6 < 0 || y1 + 5 > jmax - 1 ||
                (!xper && (i0 - 6 < 0 || ie + 5 > imax - 1)) ||
                (xper && (wrap_col(i0 - 6, imax) < 0 || wrap_col(i0 + 63, imax) >= imax)))
                return fail("batch marching window", imax, jmax, D, rb, 10 + hybrid, H);
            for (int j = y0; j < y1; ++j)
                for (int i = i0; i < ie; ++i) ++cnt[(size_t)j * imax + i];
        }
    for (int t = 0; t < p.sides.count() + p.bands.count(); ++t) {
        int i0, i1, j0, j1;
        if (t < p.sides.count())
            p.sides.tile(t, i0, i1, j0, j1);
        else
            p.bands.tile(t - p.sides.count(), i0, i1, j0, j1);
        const bool side = t < p.sides.count();
        if (i1 - i0 > (side ? E::SX : E::BX) || j1 - j0 > (side ? E::SY : E::BY) || i0 < 0 || j0 < 0 || i1 > imax ||
            j1 > jmax)
            return fail("batch edge tile", imax, jmax, D, rb, 10 + hybrid, H);
        for (int j = j0; j < j1; ++j)
            for (int i = i0; i < i1; ++i) ++cnt[(size_t)j * imax + i];
    }
    for (int v : cnt)
        if (v != 1) return fail("batch coverage", imax, jmax, D, rb, 10 + hybrid, H);
    return 0;
}

int main()
{
    long layouts = 0;
    for (int ns = 1; ns <= 80; ++ns)
        for (int nc = 1; nc <= 70; ++nc)
            if (check_tiles(ns, nc)) return 1;
    const int sizes[] = {40, 77, 128, 150, 201, 256, 301, 330, 512, 700, 1031};
    const int Ds[] = {4, 5, 6, 8, 12, 20, 30, 40, 60};
    for (int imax : sizes)
        for (int jmax : sizes)
            for (int D : Ds)
                for (int rb : {4, 6, 8})
                    for (int mode = 0; mode < 3; ++mode)
                        for (int H : {rb, 4 * rb, 7 * rb, 16 * rb, 50 * rb, 4096})
                            for (bool xper : {false, true})
                                if ((!xper || imax >= 64) &&
                                    check(imax, jmax, D, rb, mode, H - H % rb, layouts, xper))
                                    return 1;
    for (int imax : sizes)
        for (int jmax : sizes)
            for (int ty : {8, 16})
                if (check_whole(imax, jmax, ty)) return 1;
    const int bsizes[] = {40, 64, 69, 70, 96, 128, 160, 192, 256, 384, 512, 201};
    for (int imax : bsizes)
        for (int jmax : bsizes)
            for (int D : {4, 8, 12, 20, 30})
                for (bool hybrid : {false, true})
                    for (int H : {0, 4, 64, 128, 130})
                        for (bool xper : {false, true})
                            if ((!xper || imax >= 64) && (check_batch<false>(imax, jmax, D, hybrid, H, layouts, xper) ||
                                                          check_batch<true>(imax, jmax, D, hybrid, H, layouts, xper)))
                                return 1;
    std::printf("ok %ld\n", layouts);
    return 0;
}
