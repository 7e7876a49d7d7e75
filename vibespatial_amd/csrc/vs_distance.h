// Exact "within distance d" predicates: |p - q| <= d for points, segments and
// polygonal rows, decided on the exact real values of the fp64 inputs.
//
// The reference decides dwithin as distance <= threshold on fp64 distances
// (spatial/nearest.py:2855-3016, point_distance_kernels.py point-segment
// distance; GEOS on its CPU path, query_utils.py:505); rounding there can flip
// a pair whose distance is within an ulp of d.  Here every decision is exact:
//   point-point    d^2 - |p - q|^2 >= 0
//   point-segment  closest point by the signs of u.w and u.(b - p)
//                  (u = b - a, w = p - a): an end point, or the interior, where
//                  the test is d^2 |u|^2 - (u x w)^2 >= 0
// Each is first evaluated in fp64 with a forward error bound; only values
// inside the bound are recomputed as expansions (Dekker/Knuth error-free
// transforms, Shewchuk's 1997 growth and compression).  Exact for finite
// inputs whose products neither overflow nor underflow.  NaN coordinates or a
// NaN / negative d never match; d = +inf matches every non-empty pair.
#pragma once

#include "vs_exact.h"
#include "vs_polys.h"

namespace vs {

constexpr double kDistEps = 1.1102230246251565e-16;  // 2^-53
constexpr int kAccCap = 64;                            // accumulator capacity

// Shewchuk's Compress: e[0, m) nonoverlapping, increasing magnitude -> the
// same value with the fewest components (in place); returns the new length.
VS_HD int ef_compress(double* e, int m) {
    if (m <= 1) return m;
    double g[kAccCap];
    int bottom = m - 1;
    double q_run = e[m - 1];
    for (int i = m - 2; i >= 0; --i) {
        const double s = q_run + e[i];
        const double t = e[i] - (s - q_run);  // |q_run| >= |e[i]|
        if (t != 0.0) {
            g[bottom--] = s;
            q_run = t;
        } else {
            q_run = s;
        }
    }
    int top = 0;
    for (int i = bottom + 1; i < m; ++i) {
        const double s = g[i] + q_run;
        const double t = q_run - (s - g[i]);  // |g[i]| >= |q_run|
        if (t != 0.0) e[top++] = t;
        q_run = s;
    }
    e[top++] = q_run;
    return top;
}

// Growing accumulator with compression before it can overflow its capacity.
struct ExpAcc {
    double e[kAccCap];
    int n;
};

VS_HD void acc_add(ExpAcc& a, double b) {
    if (b == 0.0) return;
    if (a.n >= kAccCap - 8) a.n = ef_compress(a.e, a.n);
    if (a.n < kAccCap - 1) a.n = ef_grow(a.e, a.n, b);
}

VS_HD void acc_add_prod(ExpAcc& a, double x, double y) {
    double p, t;
    ef_two_prod(x, y, p, t);
    acc_add(a, t);
    acc_add(a, p);
}

// Fixed-size expansion of at most 16 terms (no compression needed).
VS_HD void grow_prod16(double* e, int& n, double x, double y) {
    double p, t;
    ef_two_prod(x, y, p, t);
    n = ef_grow(e, n, t);
    n = ef_grow(e, n, p);
}

VS_HD int exp_sign(const double* e, int n) {
    if (n == 0) return 0;
    const double top = e[n - 1];
    return (top > 0.0) - (top < 0.0);
}

VS_HD bool finite6(double a, double b, double c, double d, double e, double f) {
    const double m = 1.79769313486231570815e308;
    return fabs(a) <= m && fabs(b) <= m && fabs(c) <= m && fabs(d) <= m && fabs(e) <= m && fabs(f) <= m;
}

// Exact sign of u.v with u = (u1 - u0), v = (v1 - v0) per axis.
__host__ __device__ __noinline__ inline int dot_diff_sign_exact(double ux1, double ux0, double uy1, double uy0,
                                                                double vx1, double vx0, double vy1, double vy0) {
    double uxh, uxl, uyh, uyl, vxh, vxl, vyh, vyl;
    ef_two_diff(ux1, ux0, uxh, uxl);
    ef_two_diff(uy1, uy0, uyh, uyl);
    ef_two_diff(vx1, vx0, vxh, vxl);
    ef_two_diff(vy1, vy0, vyh, vyl);
    double e[16];
    int n = 0;
    grow_prod16(e, n, uxh, vxh);
    grow_prod16(e, n, uxh, vxl);
    grow_prod16(e, n, uxl, vxh);
    grow_prod16(e, n, uxl, vxl);
    grow_prod16(e, n, uyh, vyh);
    grow_prod16(e, n, uyh, vyl);
    grow_prod16(e, n, uyl, vyh);
    grow_prod16(e, n, uyl, vyl);
    return exp_sign(e, n);
}

// Filtered sign of (u1 - u0).(v1 - v0).
VS_HD int dot_diff_sign(double ux1, double ux0, double uy1, double uy0, double vx1, double vx0, double vy1,
                        double vy0) {
    const double t1 = (ux1 - ux0) * (vx1 - vx0), t2 = (uy1 - uy0) * (vy1 - vy0);
    const double t = t1 + t2;
    const double bound = 5.0 * kDistEps * (fabs(t1) + fabs(t2));
    if (t > bound) return 1;
    if (t < -bound) return -1;
    return dot_diff_sign_exact(ux1, ux0, uy1, uy0, vx1, vx0, vy1, vy0);
}

__host__ __device__ __noinline__ inline bool point_point_within_exact(double px, double py, double qx, double qy,
                                                                     double d) {
    double xh, xl, yh, yl;
    ef_two_diff(px, qx, xh, xl);
    ef_two_diff(py, qy, yh, yl);
    double e[16];
    int n = 0;
    grow_prod16(e, n, d, d);
    grow_prod16(e, n, -xh, xh);
    grow_prod16(e, n, -2.0 * xh, xl);
    grow_prod16(e, n, -xl, xl);
    grow_prod16(e, n, -yh, yh);
    grow_prod16(e, n, -2.0 * yh, yl);
    grow_prod16(e, n, -yl, yl);
    return exp_sign(e, n) >= 0;
}

// |p - q| <= d; d finite and >= 0, coordinates finite.
VS_HD bool point_point_within_fin(double px, double py, double qx, double qy, double d) {
    const double dx = px - qx, dy = py - qy;
    const double s = dx * dx + dy * dy, dd = d * d;
    const double bound = 6.0 * kDistEps * (s + dd);
    const double diff = dd - s;
    if (diff > bound) return true;
    if (diff < -bound) return false;
    return point_point_within_exact(px, py, qx, qy, d);
}

// Interior case of point-segment: d^2 |u|^2 - (u x w)^2 >= 0 exactly.
__host__ __device__ __noinline__ inline bool perp_within_exact(double px, double py, double ax, double ay,
                                                              double bx, double by, double d) {
    double uxh, uxl, uyh, uyl, wxh, wxl, wyh, wyl;
    ef_two_diff(bx, ax, uxh, uxl);
    ef_two_diff(by, ay, uyh, uyl);
    ef_two_diff(px, ax, wxh, wxl);
    ef_two_diff(py, ay, wyh, wyl);
    double c[16];  // u x w = ux wy - uy wx
    int nc = 0;
    grow_prod16(c, nc, uxh, wyh);
    grow_prod16(c, nc, uxh, wyl);
    grow_prod16(c, nc, uxl, wyh);
    grow_prod16(c, nc, uxl, wyl);
    grow_prod16(c, nc, -uyh, wxh);
    grow_prod16(c, nc, -uyh, wxl);
    grow_prod16(c, nc, -uyl, wxh);
    grow_prod16(c, nc, -uyl, wxl);
    double l[16];  // |u|^2
    int nl = 0;
    grow_prod16(l, nl, uxh, uxh);
    grow_prod16(l, nl, 2.0 * uxh, uxl);
    grow_prod16(l, nl, uxl, uxl);
    grow_prod16(l, nl, uyh, uyh);
    grow_prod16(l, nl, 2.0 * uyh, uyl);
    grow_prod16(l, nl, uyl, uyl);
    double ddh, ddl;
    ef_two_prod(d, d, ddh, ddl);
    ExpAcc a;
    a.n = 0;
    for (int k = 0; k < nl; ++k) {
        acc_add_prod(a, ddh, l[k]);
        acc_add_prod(a, ddl, l[k]);
    }
    for (int i = 0; i < nc; ++i)
        for (int j = 0; j < nc; ++j) acc_add_prod(a, -c[i], c[j]);
    a.n = ef_compress(a.e, a.n);
    return exp_sign(a.e, a.n) >= 0;
}

// Distance from p to segment a-b is <= d; d finite and >= 0, coordinates finite.
VS_HD bool point_segment_within_fin(double px, double py, double ax, double ay, double bx, double by, double d) {
    if (dot_diff_sign(bx, ax, by, ay, px, ax, py, ay) <= 0) return point_point_within_fin(px, py, ax, ay, d);
    if (dot_diff_sign(bx, ax, by, ay, bx, px, by, py) <= 0) return point_point_within_fin(px, py, bx, by, d);
    const double ux = bx - ax, uy = by - ay, wx = px - ax, wy = py - ay;
    const double c1 = ux * wy, c2 = uy * wx;
    const double c = c1 - c2;
    const double ec = 5.0 * kDistEps * (fabs(c1) + fabs(c2));
    const double l = ux * ux + uy * uy, dd = d * d;
    const double D = dd * l - c * c;
    const double bound =
        1.01 * (8.0 * kDistEps * (dd * l + c * c) + 2.0 * fabs(c) * ec + ec * ec + kDistEps * fabs(D));
    if (D > bound) return true;
    if (D < -bound) return false;
    return perp_within_exact(px, py, ax, ay, bx, by, d);
}

// Public forms with the NaN / negative / infinite distance rules.
VS_HD bool point_point_within(double px, double py, double qx, double qy, double d) {
    if (!(d >= 0.0) || px != px || py != py || qx != qx || qy != qy) return false;
    if (d == INFINITY) return true;
    if (!finite6(px, py, qx, qy, 0.0, 0.0)) return false;
    return point_point_within_fin(px, py, qx, qy, d);
}

VS_HD bool point_segment_within(double px, double py, double ax, double ay, double bx, double by, double d) {
    if (!(d >= 0.0) || !finite6(px, py, ax, ay, bx, by)) return false;
    if (d == INFINITY) return true;
    return point_segment_within_fin(px, py, ax, ay, bx, by, d);
}

// Segment a-b cannot be within d of p: p lies more than d outside the
// segment's box on one axis (gap measured in fp64; the limit carries a
// relative margin above the rounding of one subtraction, so a rejected
// segment is truly farther than d).
VS_HD bool segment_far(double px, double py, double ax, double ay, double bx, double by, double lim) {
    const double gx = fmax(fmin(ax, bx) - px, px - fmax(ax, bx));
    const double gy = fmax(fmin(ay, by) - py, py - fmax(ay, by));
    return gx > lim || gy > lim;
}

VS_HD double far_limit(double d) { return d * (1.0 + 8.881784197001252e-16); }  // d (1 + 2^-50)

// Polygonal row `row` of P is within d of point p: a boundary segment within
// d, or p inside a part (even-odd over the part's rings; a point on the
// boundary is at distance 0 and found by either test).  One pass over the
// edges; segments farther than d on an axis are skipped before the exact
// distance test.
__device__ inline bool point_polygon_within(double px, double py, const Polys& P, int64_t row, double d) {
    if (!(d >= 0.0) || px != px || py != py) return false;
    const int32_t p0 = P.geom[row], p1 = P.geom[row + 1];
    if (P.ring[P.part[p1]] <= P.ring[P.part[p0]]) return false;  // empty
    if (d == INFINITY) return true;
    const double lim = far_limit(d);
    for (int32_t pp = p0; pp < p1; ++pp) {
        int parity = 0;
        for (int32_t r = P.part[pp]; r < P.part[pp + 1]; ++r)
            for (int32_t c = P.ring[r] + 1; c < P.ring[r + 1]; ++c) {
                const double ax = P.x[c - 1], ay = P.y[c - 1], bx = P.x[c], by = P.y[c];
                if (!segment_far(px, py, ax, ay, bx, by, lim) && point_segment_within(px, py, ax, ay, bx, by, d))
                    return true;
                if (edge_step_f64(px, py, ax, ay, bx, by, parity)) return true;  // on the boundary
            }
        if (parity) return true;  // inside this part
    }
    return false;
}

// Rows ra of A and rb of B are within d: they intersect, or (disjoint closed
// regions) some vertex of one is within d of a boundary segment of the other,
// where the distance between the boundaries is attained.
__device__ inline bool polygons_within(const Polys& A, int64_t ra, const Polys& B, int64_t rb, double d) {
    if (!(d >= 0.0)) return false;
    const int32_t a0 = A.part[A.geom[ra]], a1 = A.part[A.geom[ra + 1]];
    const int32_t b0 = B.part[B.geom[rb]], b1 = B.part[B.geom[rb + 1]];
    if (A.ring[a1] <= A.ring[a0] || B.ring[b1] <= B.ring[b0]) return false;  // empty
    if (d == INFINITY) return true;
    if (polygons_intersect_auto(A, ra, B, rb)) return true;
    const double lim = far_limit(d);
    for (int32_t va = A.ring[a0]; va < A.ring[a1]; ++va) {
        const double px = A.x[va], py = A.y[va];
        for (int32_t r = b0; r < b1; ++r)
            for (int32_t c = B.ring[r] + 1; c < B.ring[r + 1]; ++c) {
                const double ax = B.x[c - 1], ay = B.y[c - 1], bx = B.x[c], by = B.y[c];
                if (!segment_far(px, py, ax, ay, bx, by, lim) && point_segment_within(px, py, ax, ay, bx, by, d))
                    return true;
            }
    }
    for (int32_t vb = B.ring[b0]; vb < B.ring[b1]; ++vb) {
        const double px = B.x[vb], py = B.y[vb];
        for (int32_t r = a0; r < a1; ++r)
            for (int32_t c = A.ring[r] + 1; c < A.ring[r + 1]; ++c) {
                const double ax = A.x[c - 1], ay = A.y[c - 1], bx = A.x[c], by = A.y[c];
                if (!segment_far(px, py, ax, ay, bx, by, lim) && point_segment_within(px, py, ax, ay, bx, by, d))
                    return true;
            }
    }
    return false;
}

}  // namespace vs
