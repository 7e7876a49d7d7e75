// Exact polygon x polygon relations beyond intersects: within / contains /
// covers / covered_by / contains_properly / touches / overlaps / crosses for
// two polygonal rows (replaces the DE-9IM mask refine of the reference,
// predicates/polygon.py:19-40, 294-370 and polygon_kernels.py:630-830).
//
// Every boundary edge of one row is cut at the other row's vertices that lie
// on it; between two consecutive cut points the open piece meets the other
// boundary nowhere unless the edges cross properly (reported separately), so
// the location of its midpoint (exact: the midpoint of two fp64 points
// enters every orientation as a sum of two exact determinants) classifies the
// whole piece: interior, exterior, or on the other boundary -- then with the
// other row's interior on the same side or the opposite side (ring
// orientation from the lowest vertex, shells vs holes).  From the pieces of
// both boundaries (A vs B, B vs A):
//   interiors meet     proper crossing, or a piece of one boundary inside the
//                      other, or a shared piece with both interiors on one side
//   A within B (closed) no proper crossing, no piece of dA outside B, no piece
//                      of dB inside A, and every part of A shows a piece inside
//                      B or a same-side shared piece (its interior, which dB
//                      does not enter, is then inside B)
//   A properly in B    A within B and the boundaries never touch
// and within = covered_by, contains = covers (areal rows), touches =
// intersects without interiors meeting, overlaps = interiors meet and neither
// is within the other, crosses = false (two areal rows never cross).
// Valid rings (closed, simple, shells then holes per part) are assumed, as in
// the reference.
#pragma once

#include "vs_distance.h"
#include "vs_exact.h"
#include "vs_polys.h"

namespace vs {

// Exact sign of orient(a, b, p) + orient(a, b, q) = 2 orient(a, b, (p+q)/2),
// orient(a, b, x) = (bx - ax)(xy - ay) - (by - ay)(xx - ax).
__host__ __device__ __noinline__ inline int orient_mid_exact(double ax, double ay, double bx, double by, double px,
                                                             double py, double qx, double qy) {
    double uxh, uxl, uyh, uyl;
    ef_two_diff(bx, ax, uxh, uxl);
    ef_two_diff(by, ay, uyh, uyl);
    ExpAcc acc;
    acc.n = 0;
    const double xs[2] = {px, qx}, ys[2] = {py, qy};
    for (int k = 0; k < 2; ++k) {
        double wxh, wxl, wyh, wyl;
        ef_two_diff(xs[k], ax, wxh, wxl);
        ef_two_diff(ys[k], ay, wyh, wyl);
        acc_add_prod(acc, uxh, wyh);
        acc_add_prod(acc, uxh, wyl);
        acc_add_prod(acc, uxl, wyh);
        acc_add_prod(acc, uxl, wyl);
        acc_add_prod(acc, -uyh, wxh);
        acc_add_prod(acc, -uyh, wxl);
        acc_add_prod(acc, -uyl, wxh);
        acc_add_prod(acc, -uyl, wxl);
    }
    acc.n = ef_compress(acc.e, acc.n);
    return exp_sign(acc.e, acc.n);
}

VS_HD int orient_mid(double ax, double ay, double bx, double by, double px, double py, double qx, double qy) {
    const double ux = bx - ax, uy = by - ay;
    const double t1 = ux * (py - ay), t2 = uy * (px - ax), t3 = ux * (qy - ay), t4 = uy * (qx - ax);
    const double s = (t1 - t2) + (t3 - t4);
    const double bound = 16.0 * kDistEps * (fabs(t1) + fabs(t2) + fabs(t3) + fabs(t4));
    if (s > bound) return 1;
    if (s < -bound) return -1;
    return orient_mid_exact(ax, ay, bx, by, px, py, qx, qy);
}

// Exact sign of (p + q) - 2 c (midpoint coordinate against c).
VS_HD int mid_cmp(double p, double q, double c) {
    double s, t;
    ef_two_sum(p, q, s, t);
    double h, l;
    ef_two_diff(s, 2.0 * c, h, l);
    double e[3];
    int n = 0;
    n = ef_grow(e, n, t);
    n = ef_grow(e, n, l);
    n = ef_grow(e, n, h);
    return exp_sign(e, n);
}

// Location of the midpoint of p and q in polygonal row `row` of P:
// 0 outside, 1 on the boundary, 2 inside (even-odd per part, as the point
// location of the join).
VS_HD int mid_location(double px, double py, double qx, double qy, const Polys& P, int64_t row) {
    // rounded midpoint and a margin beyond its rounding, for cheap skips of
    // edges wholly above, below or left of it (their exact tests agree)
    const double mx = 0.5 * (px + qx), my = 0.5 * (py + qy);
    const double tx = 0x1p-50 * (fabs(px) + fabs(qx)) + 0x1p-1000, ty = 0x1p-50 * (fabs(py) + fabs(qy)) + 0x1p-1000;
    for (int32_t pp = P.geom[row]; pp < P.geom[row + 1]; ++pp) {
        int parity = 0;
        for (int32_t r = P.part[pp]; r < P.part[pp + 1]; ++r)
            for (int32_t c = P.ring[r] + 1; c < P.ring[r + 1]; ++c) {
                const double ax = P.x[c - 1], ay = P.y[c - 1], bx = P.x[c], by = P.y[c];
                if (fmin(ay, by) > my + ty || fmax(ay, by) < my - ty || fmax(ax, bx) < mx - tx) continue;
                const int ya = mid_cmp(py, qy, ay), yb = mid_cmp(py, qy, by);  // sign(2 my - a.y) etc.
                const bool crosses = (ya < 0) != (yb < 0);  // exactly one endpoint above the midpoint
                const bool in_y = (ya >= 0 || yb >= 0) && (ya <= 0 || yb <= 0);
                const bool in_x = (mid_cmp(px, qx, ax) >= 0 || mid_cmp(px, qx, bx) >= 0) &&
                                  (mid_cmp(px, qx, ax) <= 0 || mid_cmp(px, qx, bx) <= 0);
                if (!crosses && !(in_x && in_y)) continue;
                const int o = orient_mid(ax, ay, bx, by, px, py, qx, qy);
                if (o == 0 && in_x && in_y) return 1;
                if (crosses && ((o > 0) == (yb < 0))) parity ^= 1;  // b above: edge goes up
            }
        if (parity) return 2;
    }
    return 0;
}

// Interior of the polygon on the left of ring r's direction: shells counter-
// clockwise, holes clockwise.  The turn at the lowest (then leftmost) vertex
// of a simple ring has the ring's orientation.
VS_HD bool ring_interior_left(const Polys& P, int32_t r, bool shell) {
    const int32_t c0 = P.ring[r], c1 = P.ring[r + 1] - 1;  // last coordinate closes the ring
    const int32_t m = c1 - c0;
    if (m < 3) return shell;
    int32_t k = c0;
    for (int32_t c = c0 + 1; c < c1; ++c)
        if (P.y[c] < P.y[k] || (P.y[c] == P.y[k] && P.x[c] < P.x[k])) k = c;
    int32_t a = k, b = k;  // neighbours distinct from the vertex (repeated points skipped)
    for (int32_t s = 1; s < m; ++s) {
        a = c0 + ((k - c0 - s) % m + m) % m;
        if (P.x[a] != P.x[k] || P.y[a] != P.y[k]) break;
    }
    for (int32_t s = 1; s < m; ++s) {
        b = c0 + (k - c0 + s) % m;
        if (P.x[b] != P.x[k] || P.y[b] != P.y[k]) break;
    }
    const bool ccw = orient2d(P.x[a], P.y[a], P.x[k], P.y[k], P.x[b], P.y[b]) > 0;
    return shell ? ccw : !ccw;
}

struct RelSide {
    bool cross;     // some edges cross properly
    bool touch;     // the boundaries meet (filled by the first pass only)
    bool in_int;    // a boundary piece of X lies in the interior of Y
    bool in_ext;    // a boundary piece of X lies outside Y
    bool same;      // a shared piece with both interiors on one side
    bool all_in;    // every part of X shows a piece inside Y or a same-side shared piece
};

// The next vertex of Y strictly between cur and b on edge (a, b) (nearest to
// cur), or b.  `ax_major`: compare along x (else y).
VS_HD void next_cut(double ax, double ay, double bx, double by, double cx, double cy, const Polys& Y, int64_t ry,
                    bool ax_major, double& nx, double& ny) {
    nx = bx;
    ny = by;
    const bool inc = ax_major ? bx > ax : by > ay;
    const int32_t v0 = Y.ring[Y.part[Y.geom[ry]]], v1 = Y.ring[Y.part[Y.geom[ry + 1]]];
    for (int32_t v = v0; v < v1; ++v) {
        const double x = Y.x[v], y = Y.y[v];
        const double t = ax_major ? x : y, tc = ax_major ? cx : cy, tn = ax_major ? nx : ny;
        const bool after = inc ? (t > tc && t < tn) : (t < tc && t > tn);
        if (!after) continue;
        if (fmin(ax, bx) > x || x > fmax(ax, bx) || fmin(ay, by) > y || y > fmax(ay, by)) continue;
        if (orient2d(ax, ay, bx, by, x, y) != 0) continue;
        nx = x;
        ny = y;
    }
}

// One boundary edge c of X (ring interior on the left when left_x) against
// row ry of Y: crossings, contact (when `first`), and its pieces.
VS_HD void relate_edge(const Polys& X, int32_t c, bool left_x, const Polys& Y, int64_t ry, bool first, RelSide& S,
                       bool& part_in) {
    const double ax = X.x[c - 1], ay = X.y[c - 1], bx = X.x[c], by = X.y[c];
    if (ax == bx && ay == by) return;
    // proper crossings (and, on the first pass, any contact)
    for (int32_t q = Y.part[Y.geom[ry]]; q < Y.part[Y.geom[ry + 1]]; ++q)
        for (int32_t d = Y.ring[q] + 1; d < Y.ring[q + 1]; ++d) {
            const double cx = Y.x[d - 1], cy = Y.y[d - 1], dx = Y.x[d], dy = Y.y[d];
            if (fmax(ax, bx) < fmin(cx, dx) || fmax(cx, dx) < fmin(ax, bx) || fmax(ay, by) < fmin(cy, dy) ||
                fmax(cy, dy) < fmin(ay, by))
                continue;
            if (first && !S.touch && segments_touch(ax, ay, bx, by, cx, cy, dx, dy)) S.touch = true;
            const int o1 = orient2d(ax, ay, bx, by, cx, cy), o2 = orient2d(ax, ay, bx, by, dx, dy);
            if (o1 * o2 >= 0) continue;
            const int o3 = orient2d(cx, cy, dx, dy, ax, ay), o4 = orient2d(cx, cy, dx, dy, bx, by);
            if (o3 * o4 < 0) S.cross = true;
        }
    // pieces between consecutive cut points
    const bool ax_major = ax != bx;
    double px = ax, py = ay;
    for (;;) {
        double qx, qy;
        next_cut(ax, ay, bx, by, px, py, Y, ry, ax_major, qx, qy);
        const int loc = mid_location(px, py, qx, qy, Y, ry);
        if (loc == 2) {
            S.in_int = true;
            part_in = true;
        } else if (loc == 0) {
            S.in_ext = true;
        } else {
            // the Y edge holding the midpoint: interiors on one side?
            const double mx = 0.5 * (px + qx), my = 0.5 * (py + qy);
            const double tx = 0x1p-50 * (fabs(px) + fabs(qx)) + 0x1p-1000;
            const double ty = 0x1p-50 * (fabs(py) + fabs(qy)) + 0x1p-1000;
            bool found = false;
            for (int32_t q = Y.part[Y.geom[ry]]; q < Y.part[Y.geom[ry + 1]] && !found; ++q)
                for (int32_t d = Y.ring[q] + 1; d < Y.ring[q + 1]; ++d) {
                    const double cx = Y.x[d - 1], cy = Y.y[d - 1], dx = Y.x[d], dy = Y.y[d];
                    if (cx == dx && cy == dy) continue;
                    if (fmin(cx, dx) > mx + tx || fmax(cx, dx) < mx - tx || fmin(cy, dy) > my + ty ||
                        fmax(cy, dy) < my - ty)
                        continue;
                    if (orient_mid(cx, cy, dx, dy, px, py, qx, qy) != 0) continue;
                    if (mid_cmp(px, qx, fmin(cx, dx)) < 0 || mid_cmp(px, qx, fmax(cx, dx)) > 0 ||
                        mid_cmp(py, qy, fmin(cy, dy)) < 0 || mid_cmp(py, qy, fmax(cy, dy)) > 0)
                        continue;
                    // X's piece lies on this Y edge; ring q's role: the first
                    // ring of its part is the shell
                    found = true;
                    int32_t qp = Y.geom[ry];
                    while (qp + 1 < Y.geom[ry + 1] && Y.part[qp + 1] <= q) ++qp;
                    const bool ly = ring_interior_left(Y, q, q == Y.part[qp]);
                    const bool same_dir = dot_diff_sign(bx, ax, by, ay, dx, cx, dy, cy) > 0;
                    if (same_dir ? (ly == left_x) : (ly != left_x)) {
                        S.same = true;
                        part_in = true;
                    }
                    break;
                }
        }
        if (qx == bx && qy == by) break;
        px = qx;
        py = qy;
    }
}

// Pieces of X's boundary against row ry of Y.
VS_HD RelSide relate_side(const Polys& X, int64_t rx, const Polys& Y, int64_t ry, bool first) {
    RelSide S = {false, false, false, false, false, true};
    for (int32_t pp = X.geom[rx]; pp < X.geom[rx + 1]; ++pp) {
        bool part_in = false;
        for (int32_t r = X.part[pp]; r < X.part[pp + 1]; ++r) {
            const bool left_x = ring_interior_left(X, r, r == X.part[pp]);
            for (int32_t c = X.ring[r] + 1; c < X.ring[r + 1]; ++c) relate_edge(X, c, left_x, Y, ry, first, S, part_in);
        }
        if (!part_in) S.all_in = false;
    }
    return S;
}

struct Relation {
    bool interiors_meet;
    bool a_in_b;         // A within B (closed)
    bool b_in_a;         // B within A (closed)
    bool touch;          // the boundaries meet
};

VS_HD Relation relate_polygons(const Polys& A, int64_t ra, const Polys& B, int64_t rb) {
    const RelSide sa = relate_side(A, ra, B, rb, true);
    const RelSide sb = relate_side(B, rb, A, ra, false);
    Relation R;
    const bool cross = sa.cross || sb.cross;
    R.interiors_meet = cross || sa.in_int || sb.in_int || sa.same || sb.same;
    R.a_in_b = !cross && !sa.in_ext && !sb.in_int && sa.all_in;
    R.b_in_a = !cross && !sb.in_ext && !sa.in_int && sb.all_in;
    R.touch = sa.touch;
    return R;
}

VS_HD double4 row_box(const Polys& P, int64_t row) {
    double4 b = make_double4(INFINITY, INFINITY, -INFINITY, -INFINITY);
    for (int32_t c = P.ring[P.part[P.geom[row]]]; c < P.ring[P.part[P.geom[row + 1]]]; ++c) {
        b.x = fmin(b.x, P.x[c]);
        b.y = fmin(b.y, P.y[c]);
        b.z = fmax(b.z, P.x[c]);
        b.w = fmax(b.w, P.y[c]);
    }
    return b;
}

// Some edge of A crosses some edge of B properly (interior points of both).
VS_HD bool any_proper_crossing(const Polys& A, int64_t ra, const Polys& B, int64_t rb) {
    for (int32_t r = A.part[A.geom[ra]]; r < A.part[A.geom[ra + 1]]; ++r)
        for (int32_t c = A.ring[r] + 1; c < A.ring[r + 1]; ++c) {
            const double ax = A.x[c - 1], ay = A.y[c - 1], bx = A.x[c], by = A.y[c];
            for (int32_t q = B.part[B.geom[rb]]; q < B.part[B.geom[rb + 1]]; ++q)
                for (int32_t d = B.ring[q] + 1; d < B.ring[q + 1]; ++d) {
                    const double cx = B.x[d - 1], cy = B.y[d - 1], dx = B.x[d], dy = B.y[d];
                    if (fmax(ax, bx) < fmin(cx, dx) || fmax(cx, dx) < fmin(ax, bx) || fmax(ay, by) < fmin(cy, dy) ||
                        fmax(cy, dy) < fmin(ay, by))
                        continue;
                    if (orient2d(ax, ay, bx, by, cx, cy) * orient2d(ax, ay, bx, by, dx, dy) >= 0) continue;
                    if (orient2d(cx, cy, dx, dy, ax, ay) * orient2d(cx, cy, dx, dy, bx, by) < 0) return true;
                }
        }
    return false;
}

// predicate from the full relation (intersects already known for touches).
VS_HD bool relation_decides(int pred, const Relation& R) {
    switch (pred) {
        case VS_PRED_WITHIN:
        case VS_PRED_COVERED_BY: return R.a_in_b;
        case VS_PRED_CONTAINS:
        case VS_PRED_COVERS: return R.b_in_a;
        case VS_PRED_CONTAINS_PROPERLY: return R.b_in_a && !R.touch;
        case VS_PRED_TOUCHES: return !R.interiors_meet;
        case VS_PRED_OVERLAPS: return R.interiors_meet && !R.a_in_b && !R.b_in_a;
        default: return false;
    }
}

// Cheap necessary conditions of predicate(A row, B row) for two polygonal
// rows (codes of vsjoin.h): box containment for the containment predicates;
// intersects, then a proper crossing -- which settles every relation -- for
// touches / overlaps.  1 true, 0 false, -1 undecided (the full relation
// decides, with intersects known to hold for touches).
VS_HD int polygon_predicate_fast(int pred, const Polys& A, int64_t ra, const Polys& B, int64_t rb) {
    if (pred == VS_PRED_INTERSECTS) return polygons_intersect_auto(A, ra, B, rb) ? 1 : 0;
    if (pred == VS_PRED_CROSSES) return 0;
    if (pred == VS_PRED_WITHIN || pred == VS_PRED_COVERED_BY || pred == VS_PRED_CONTAINS ||
        pred == VS_PRED_COVERS || pred == VS_PRED_CONTAINS_PROPERLY) {
        const bool a_inner = pred == VS_PRED_WITHIN || pred == VS_PRED_COVERED_BY;
        const double4 in = a_inner ? row_box(A, ra) : row_box(B, rb);
        const double4 out = a_inner ? row_box(B, rb) : row_box(A, ra);
        if (!(out.x <= in.x && in.z <= out.z && out.y <= in.y && in.w <= out.w)) return 0;
        if (any_proper_crossing(A, ra, B, rb)) return 0;
        return -1;
    }
    // touches, overlaps
    if (!polygons_intersect_auto(A, ra, B, rb)) return 0;
    if (any_proper_crossing(A, ra, B, rb)) return pred == VS_PRED_OVERLAPS ? 1 : 0;
    return -1;
}

// predicate(A row, B row), one thread.
VS_HD bool polygon_predicate(int pred, const Polys& A, int64_t ra, const Polys& B, int64_t rb) {
    const int f = polygon_predicate_fast(pred, A, ra, B, rb);
    if (f >= 0) return f != 0;
    return relation_decides(pred, relate_polygons(A, ra, B, rb));
}

// ---------------------------------------------------------------------------
// One wave per pair, for pairs of large rows: lanes split the first row's
// edges (each lane against every edge of the other row), votes decide.
// ---------------------------------------------------------------------------
__device__ inline int32_t row_coords(const Polys& P, int64_t row) {
    return P.ring[P.part[P.geom[row + 1]]] - P.ring[P.part[P.geom[row]]];
}

// mode 0: some pair of edges touches (segments_touch); 1: some pair crosses
// properly.
__device__ inline bool wave_any_edge_pair(int mode, const Polys& A, int64_t ra, const Polys& B, int64_t rb,
                                          int lane) {
    // the row's rings are contiguous coordinates: lanes take its edges
    // (c - 1, c) across every ring at once, c not the start of a ring; a row
    // of at most 32 (16) positions gives each edge 2 (4) lanes, which split
    // the other row's edges by stride
    const int32_t r0 = A.part[A.geom[ra]], r1 = A.part[A.geom[ra + 1]];
    if (r1 <= r0) return false;
    const int32_t e0 = A.ring[r0] + 1, e1 = A.ring[r1];
    const int split = e1 - e0 <= 16 ? 4 : (e1 - e0 <= 32 ? 2 : 1);
    const int sub = lane % split;
    for (int32_t base = e0; base < e1; base += 64 / split) {
        const int32_t c = base + lane / split;
        bool hit = false;
        bool edge = c < e1;
        for (int32_t r = r0 + 1; edge && r < r1; ++r)
            if (A.ring[r] == c) edge = false;
        if (edge) {
            const double ax = A.x[c - 1], ay = A.y[c - 1], bx = A.x[c], by = A.y[c];
            for (int32_t q = B.part[B.geom[rb]]; q < B.part[B.geom[rb + 1]] && !hit; ++q)
                for (int32_t d = B.ring[q] + 1 + sub; d < B.ring[q + 1]; d += split) {
                    const double cx = B.x[d - 1], cy = B.y[d - 1], dx = B.x[d], dy = B.y[d];
                    if (fmax(ax, bx) < fmin(cx, dx) || fmax(cx, dx) < fmin(ax, bx) ||
                        fmax(ay, by) < fmin(cy, dy) || fmax(cy, dy) < fmin(ay, by))
                        continue;
                    if (mode == 0) {
                        if (segments_touch(ax, ay, bx, by, cx, cy, dx, dy)) {
                            hit = true;
                            break;
                        }
                    } else if (orient2d(ax, ay, bx, by, cx, cy) * orient2d(ax, ay, bx, by, dx, dy) < 0 &&
                               orient2d(cx, cy, dx, dy, ax, ay) * orient2d(cx, cy, dx, dy, bx, by) < 0) {
                        hit = true;
                        break;
                    }
                }
        }
        if (__ballot(hit) != 0) return true;
    }
    return false;
}

// Exact intersects (the rule of polygons_intersect) by one wave.
__device__ inline bool wave_intersects(const Polys& A, int64_t ra, const Polys& B, int64_t rb, int lane) {
    if (row_coords(A, ra) == 0 || row_coords(B, rb) == 0) return false;
    if (wave_any_edge_pair(0, A, ra, B, rb, lane)) return true;
    bool in = false;
    for (int32_t r = A.part[A.geom[ra]] + lane; r < A.part[A.geom[ra + 1]]; r += 64) {
        const int32_t c0 = A.ring[r];
        if (A.ring[r + 1] > c0 && point_location_f64(A.x[c0], A.y[c0], B.x, B.y, B.geom, B.part, B.ring, rb) != 0)
            in = true;
    }
    for (int32_t r = B.part[B.geom[rb]] + lane; r < B.part[B.geom[rb + 1]]; r += 64) {
        const int32_t c0 = B.ring[r];
        if (B.ring[r + 1] > c0 && point_location_f64(B.x[c0], B.y[c0], A.x, A.y, A.geom, A.part, A.ring, ra) != 0)
            in = true;
    }
    return __ballot(in) != 0;
}

__device__ inline double4 wave_row_box(const Polys& P, int64_t row, int lane) {
    double4 b = make_double4(INFINITY, INFINITY, -INFINITY, -INFINITY);
    for (int32_t c = P.ring[P.part[P.geom[row]]] + lane; c < P.ring[P.part[P.geom[row + 1]]]; c += 64) {
        b.x = fmin(b.x, P.x[c]);
        b.y = fmin(b.y, P.y[c]);
        b.z = fmax(b.z, P.x[c]);
        b.w = fmax(b.w, P.y[c]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        b.x = fmin(b.x, __shfl_xor(b.x, o, 64));
        b.y = fmin(b.y, __shfl_xor(b.y, o, 64));
        b.z = fmax(b.z, __shfl_xor(b.z, o, 64));
        b.w = fmax(b.w, __shfl_xor(b.w, o, 64));
    }
    return b;
}

// ring_interior_left by one wave (every lane returns the answer): lanes scan
// strided vertices for the lowest (then leftmost, then first) one and a
// shuffle reduction picks it -- the vertex the serial scan finds.
__device__ inline bool wave_ring_interior_left(const Polys& P, int32_t r, bool shell, int lane) {
    const int32_t c0 = P.ring[r], c1 = P.ring[r + 1] - 1;  // last coordinate closes the ring
    const int32_t m = c1 - c0;
    if (m < 3) return shell;
    int32_t k = c1;  // c1: no vertex yet
    double ky = INFINITY, kx = INFINITY;
    for (int32_t c = c0 + lane; c < c1; c += 64) {
        const double y = P.y[c], x = P.x[c];
        if (k == c1 || y < ky || (y == ky && x < kx)) {
            k = c;
            ky = y;
            kx = x;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int32_t k2 = __shfl_xor(k, o, 64);
        const double y2 = __shfl_xor(ky, o, 64), x2 = __shfl_xor(kx, o, 64);
        if (k2 != c1 && (k == c1 || y2 < ky || (y2 == ky && (x2 < kx || (x2 == kx && k2 < k))))) {
            k = k2;
            ky = y2;
            kx = x2;
        }
    }
    k = __shfl(k, 0, 64);  // one vertex for every lane (NaN coordinates compare unordered)
    int32_t a = k, b = k;  // neighbours distinct from the vertex (repeated points skipped)
    for (int32_t s = 1; s < m; ++s) {
        a = c0 + ((k - c0 - s) % m + m) % m;
        if (P.x[a] != P.x[k] || P.y[a] != P.y[k]) break;
    }
    for (int32_t s = 1; s < m; ++s) {
        b = c0 + (k - c0 + s) % m;
        if (P.x[b] != P.x[k] || P.y[b] != P.y[k]) break;
    }
    const bool ccw = orient2d(P.x[a], P.y[a], P.x[k], P.y[k], P.x[b], P.y[b]) > 0;
    return shell ? ccw : !ccw;
}

// relate_side by one wave: lanes split each ring's edges, votes merge.
__device__ inline RelSide wave_relate_side(const Polys& X, int64_t rx, const Polys& Y, int64_t ry, bool first,
                                           int lane) {
    RelSide S = {false, false, false, false, false, true};
    bool all_in = true;
    for (int32_t pp = X.geom[rx]; pp < X.geom[rx + 1]; ++pp) {
        bool part_in = false;
        for (int32_t r = X.part[pp]; r < X.part[pp + 1]; ++r) {
            const bool left_x = wave_ring_interior_left(X, r, r == X.part[pp], lane);
            for (int32_t c = X.ring[r] + 1 + lane; c < X.ring[r + 1]; c += 64)
                relate_edge(X, c, left_x, Y, ry, first, S, part_in);
        }
        if (__ballot(part_in) == 0) all_in = false;
    }
    RelSide W;
    W.cross = __ballot(S.cross) != 0;
    W.touch = __ballot(S.touch) != 0;
    W.in_int = __ballot(S.in_int) != 0;
    W.in_ext = __ballot(S.in_ext) != 0;
    W.same = __ballot(S.same) != 0;
    W.all_in = all_in;
    return W;
}

__device__ inline Relation wave_relate_polygons(const Polys& A, int64_t ra, const Polys& B, int64_t rb, int lane) {
    const RelSide sa = wave_relate_side(A, ra, B, rb, true, lane);
    const RelSide sb = wave_relate_side(B, rb, A, ra, false, lane);
    Relation R;
    const bool cross = sa.cross || sb.cross;
    R.interiors_meet = cross || sa.in_int || sb.in_int || sa.same || sb.same;
    R.a_in_b = !cross && !sa.in_ext && !sb.in_int && sa.all_in;
    R.b_in_a = !cross && !sb.in_ext && !sa.in_int && sb.all_in;
    R.touch = sa.touch;
    return R;
}

// polygon_predicate_fast by one wave (every lane returns the answer).
__device__ inline int wave_polygon_predicate_fast(int pred, const Polys& A, int64_t ra, const Polys& B, int64_t rb,
                                                  int lane) {
    if (pred == VS_PRED_INTERSECTS) return wave_intersects(A, ra, B, rb, lane) ? 1 : 0;
    if (pred == VS_PRED_CROSSES) return 0;
    if (pred == VS_PRED_WITHIN || pred == VS_PRED_COVERED_BY || pred == VS_PRED_CONTAINS ||
        pred == VS_PRED_COVERS || pred == VS_PRED_CONTAINS_PROPERLY) {
        const bool a_inner = pred == VS_PRED_WITHIN || pred == VS_PRED_COVERED_BY;
        const double4 in = a_inner ? wave_row_box(A, ra, lane) : wave_row_box(B, rb, lane);
        const double4 out = a_inner ? wave_row_box(B, rb, lane) : wave_row_box(A, ra, lane);
        if (!(out.x <= in.x && in.z <= out.z && out.y <= in.y && in.w <= out.w)) return 0;
        if (wave_any_edge_pair(1, A, ra, B, rb, lane)) return 0;
        return -1;
    }
    // touches, overlaps
    if (!wave_intersects(A, ra, B, rb, lane)) return 0;
    if (wave_any_edge_pair(1, A, ra, B, rb, lane)) return pred == VS_PRED_OVERLAPS ? 1 : 0;
    return -1;
}

}  // namespace vs
