// Cell grid of a ring: a 32 x 32 grid over the ring's bounding box whose
// cells are marked "boundary" (some edge's bounding box meets the closed
// cell) or, when no edge comes near, "inside" / "outside" the ring.  Used by
// the batched intersects join to decide a candidate pair from the smaller
// polygon's bounding box without streaming the larger one's edges.
//
// Cell geometry.  Along x, edges E(i) = x0 + i * wx (i = 0..32, fp64, no
// fused multiply-add: the library is built with -ffp-contract=off, so build
// and query compute the same doubles) with wx = (x1 - x0) / 32.  Cell i is
// the closed interval [E(i), E(i+1)], except that cell 0 extends to -inf and
// cell 31 to +inf, so every finite value lies in some cell.  Same along y.
//
// Index searches.  A grid is only built when wx >= 2^-40 max(|x0|, |x1|)
// (and likewise in y) and 1e-250 < wx, max(|x0|, |x1|) < 1e250.  Then each
// E(i) lies within (32 + 2^41) 2^-53 wx < 2^-11 wx of x0 + i wx, and the
// first guess t = (v - x0) (1 / wx) (three roundings) within 2^-40 of
// (v - x0) / wx for v inside the box (outside it the guess is clamped to the
// end cell, which is then the answer), so floor(t) is within one of the
// exact index and one comparison with E on each side settles it.
//
// Marks.  An edge with bounding box [mnx, mxx] x [mny, mxy] marks every cell
// whose closed range meets that box: columns cell_lo(mnx)..cell_hi(mxx),
// rows cell_lo(mny)..cell_hi(mxy).  A cell left unmarked is met by no edge,
// so the whole closed cell lies on one side of the ring (location is constant
// on connected sets that miss the boundary), and that side is the side of its
// corner p = (E(i), F(j)).  Side of p: under the half-open rule
// (ay > y) != (by > y) a closed ring crosses the line y = F(j) an even number
// of times, and none of the crossings is at p (p is off the boundary), so the
// crossings left of p have the parity of those right of p -- the even-odd
// location.  A crossing edge of an unmarked cell lies wholly left
// (mxx < E(i)) or wholly right (mnx > E(i+1)) of it, since otherwise its box
// would meet the cell; so the left crossings are the crossing edges with
// mxx < E(i).  Edge e toggles, in row j, the cells i >= k_e with
// k_e = #{i : E(i) <= mxx}, and a prefix XOR along the row gives each
// cell's parity.
//
// Decisions for a ring A (finite coordinates) against such a ring B: if
// every cell met by A's bounding box is unmarked and inside, A lies in B's
// interior and its first vertex is inside B; if every one is unmarked and
// outside, no edge of B comes near A's box, A's first vertex is outside B and
// B's first vertex (on B's boundary) is outside A's box.  These are exactly
// the facts on which the exact intersects test (ring_intersects_polys: first
// vertex of A located in B, B's edges near A tested against A's edges, B's
// first vertex located in A) returns true, respectively false, so the
// decision is the same boolean.  Otherwise A's vertices are looked up: A's
// boundary is one connected polyline, so either it meets B's boundary (then
// the test's second step finds a touching edge pair and answers true) or all
// of it lies on one side of B, the side of any of its vertices.  Hence a
// vertex in an unmarked inside cell decides true, and a vertex in an unmarked
// outside cell means that the first step (A's first vertex located in B) can
// be skipped: it would find the first vertex outside, off the boundary,
// unless a touching pair exists, which the second step then finds.
#pragma once

#include "vs_exact.h"

namespace vs {

constexpr int kGrid = 32;

struct GridAxes {
    double x0, y0;  // E(0), F(0)
    double wx, wy;  // cell widths (> 0)
    double ix, iy;  // 1 / wx, 1 / wy: first guesses only
};

VS_HD double cell_edge(double base, double w, int i) { return base + (double)i * w; }

VS_HD int cell_guess(double v, double base, double inv) {
    double t = (v - base) * inv;
    t = t > 0.0 ? (t < (double)(kGrid - 1) ? t : (double)(kGrid - 1)) : 0.0;
    return (int)t;
}

// Lowest cell holding v: the smallest i with v <= E(i+1), else 31.
VS_HD int cell_lo(double v, double base, double w, double inv) {
    const int g = cell_guess(v, base, inv);
    if (g > 0 && v <= cell_edge(base, w, g)) return g - 1;
    if (g < kGrid - 1 && v > cell_edge(base, w, g + 1)) return g + 1;
    return g;
}

// Highest cell holding v: the largest i with v >= E(i), else 0.
VS_HD int cell_hi(double v, double base, double w, double inv) {
    const int g = cell_guess(v, base, inv);
    if (g < kGrid - 1 && v >= cell_edge(base, w, g + 1)) return g + 1;
    if (g > 0 && v < cell_edge(base, w, g)) return g - 1;
    return g;
}

// Bits i0..i1 of a row (0 <= i0 <= i1 <= 31).
VS_HD uint32_t span_mask(int i0, int i1) {
    const uint32_t hi = i1 >= kGrid - 1 ? 0xFFFFFFFFu : ((1u << (i1 + 1)) - 1u);
    return hi & ~((1u << i0) - 1u);
}

VS_HD uint32_t prefix_xor32(uint32_t t) {
    t ^= t << 1;
    t ^= t << 2;
    t ^= t << 4;
    t ^= t << 8;
    t ^= t << 16;
    return t;
}

// Grid over the box (x0, y0, x1, y1); false when a side is empty, the
// coordinates are not finite or too large, or a cell is too narrow for its
// offset (the conditions under which the index searches above are exact).
VS_HD bool grid_axis_ok(double lo, double hi, double w) {
    const double m = fmax(fabs(lo), fabs(hi));
    return m < 1e250 && w > 1e-250 && w < 1e250 && w * 1099511627776.0 >= m;  // 2^40
}

VS_HD bool grid_setup(double x0, double y0, double x1, double y1, GridAxes& g) {
    if (!(fabs(x0) < INFINITY && fabs(y0) < INFINITY && fabs(x1) < INFINITY && fabs(y1) < INFINITY)) return false;
    g.x0 = x0;
    g.y0 = y0;
    g.wx = (x1 - x0) / kGrid;
    g.wy = (y1 - y0) / kGrid;
    if (!(grid_axis_ok(x0, x1, g.wx) && grid_axis_ok(y0, y1, g.wy))) return false;
    g.ix = 1.0 / g.wx;
    g.iy = 1.0 / g.wy;
    return true;
}

// Edge a->b: the cell rectangle its box meets (columns i0..i1, rows j0..j1)
// and k = #{i : E(i) <= max x}, the first column whose corner lies strictly
// right of the edge.  Rows j0..j1 hold every j with F(j) in [min y, max y].
struct GridEdge {
    int i0, i1, j0, j1, k;
};

VS_HD GridEdge grid_edge(const GridAxes& g, double ax, double ay, double bx, double by) {
    const double mnx = fmin(ax, bx), mxx = fmax(ax, bx), mny = fmin(ay, by), mxy = fmax(ay, by);
    GridEdge e;
    e.i0 = cell_lo(mnx, g.x0, g.wx, g.ix);
    e.i1 = cell_hi(mxx, g.x0, g.wx, g.ix);
    e.j0 = cell_lo(mny, g.y0, g.wy, g.iy);
    e.j1 = cell_hi(mxy, g.y0, g.wy, g.iy);
    e.k = (mxx >= cell_edge(g.x0, g.wx, e.i1)) ? e.i1 + 1 : e.i1;  // E(0) = x0 <= mxx: always i1 + 1 in a build
    return e;
}

// Does edge a->b cross the line y = F(j) under the half-open rule?
VS_HD bool grid_crosses(const GridAxes& g, int j, double ay, double by) {
    const double f = cell_edge(g.y0, g.wy, j);
    return (ay > f) != (by > f);
}

// Candidate decision for ring A against the grid of ring B: (ax0, ay0, ax1,
// ay1) = A's bounding box, (fx, fy) = A's first vertex, x[0..n), y[0..n) =
// A's coordinates (1 <= n), read only when the box does not decide.
// Returns 1 intersects, 0 does not, 2 undecided but the first step of the
// exact test may be skipped, 3 undecided (also when the first vertex, or for
// the vertex test any coordinate, is not finite).  cells[j] = (boundary bits, inside bits) of grid row j.  Loads
// are issued in groups: the first eight grid rows of the box at once (rows
// past the box repeat its last row, which changes none of the folds), then
// the vertices eight at a time (indices past n - 1 repeat the last vertex,
// which changes neither test), then each vertex's row.
VS_HD int grid_decide(const GridAxes& g, const uint2* cells, double ax0, double ay0, double ax1, double ay1,
                      double fx, double fy, const double* x, const double* y, int n) {
    if (!(fabs(fx) < INFINITY && fabs(fy) < INFINITY)) return 3;
    const int i0 = cell_lo(ax0, g.x0, g.wx, g.ix), i1 = cell_hi(ax1, g.x0, g.wx, g.ix);
    const int j0 = cell_lo(ay0, g.y0, g.wy, g.iy), j1 = cell_hi(ay1, g.y0, g.wy, g.iy);
    const uint32_t m = span_mask(i0, i1);
    uint2 rows[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) rows[u] = cells[j0 + u <= j1 ? j0 + u : j1];
    uint32_t marks = 0u, all_in = m, any_in = 0u;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        marks |= rows[u].x & m;
        all_in &= rows[u].y;
        any_in |= rows[u].y & m;
    }
    for (int j = j0 + 8; j <= j1; ++j) {
        const uint2 c = cells[j];
        marks |= c.x & m;
        all_in &= c.y;
        any_in |= c.y & m;
    }
    if (!marks) {
        if (all_in == m) return 1;
        if (any_in == 0u) return 0;
    }
    bool finite = true, in = false, out = false;
    for (int k0 = 0; k0 < n; k0 += 8) {
        double vx[8], vy[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int k = k0 + u < n ? k0 + u : n - 1;
            vx[u] = x[k];
            vy[u] = y[k];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) finite = finite && fabs(vx[u]) < INFINITY && fabs(vy[u]) < INFINITY;
        if (!finite) break;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int ci = cell_lo(vx[u], g.x0, g.wx, g.ix), cj = cell_lo(vy[u], g.y0, g.wy, g.iy);
            const uint2 c = cells[cj];
            if (!((c.x >> ci) & 1u)) {
                if ((c.y >> ci) & 1u) in = true;
                else out = true;
            }
        }
    }
    if (!finite) return 3;
    return in ? 1 : (out ? 2 : 3);
}

// The grid of one ring (n coordinates, closed, finite, box (x0, y0, x1, y1)),
// one edge after another; the device build applies the same grid_edge /
// grid_crosses marks with LDS atomics.  Returns false when the ring gets no
// grid (not closed, fewer than 4 coordinates, a non-finite coordinate, an
// empty side).
inline bool grid_build_serial(const double* x, const double* y, int64_t n, double x0, double y0, double x1,
                              double y1, GridAxes& g, uint2* cells) {
    if (n < 4 || !grid_setup(x0, y0, x1, y1, g)) return false;
    if (!(x[0] == x[n - 1] && y[0] == y[n - 1])) return false;
    uint32_t bnd[kGrid] = {0}, tog[kGrid] = {0};
    for (int64_t e = 1; e < n; ++e) {
        const double ax = x[e - 1], ay = y[e - 1], bx = x[e], by = y[e];
        if (!(fabs(ax) < INFINITY && fabs(ay) < INFINITY && fabs(bx) < INFINITY && fabs(by) < INFINITY)) return false;
        const GridEdge ge = grid_edge(g, ax, ay, bx, by);
        const uint32_t m = span_mask(ge.i0, ge.i1);
        for (int j = ge.j0; j <= ge.j1; ++j) {
            bnd[j] |= m;
            if (ge.k < kGrid && grid_crosses(g, j, ay, by)) tog[j] ^= 1u << ge.k;
        }
    }
    for (int j = 0; j < kGrid; ++j) cells[j] = make_uint2(bnd[j], prefix_xor32(tog[j]));
    return true;
}

}  // namespace vs
