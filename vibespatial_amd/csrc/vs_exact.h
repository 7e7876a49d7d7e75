// Exact 2D orientation and the point/edge classification used by the refine
// kernels.  __host__ __device__ so the same code is exported to the host for
// unit tests (vs_host_orient2d, vs_host_point_location).
//
// Arithmetic: Dekker/Knuth error-free transforms (two-sum, two-diff, fma
// two-product) and Shewchuk's (1997) expansion growth with zero elimination.
// Stage A is Shewchuk's filter bound (3 + 16 eps) eps with eps = 2^-53; any
// determinant inside it is evaluated exactly as a 16-term expansion, so the
// returned sign is exact for every fp64 input whose products neither overflow
// nor underflow.  (The reference's vs_orient2d, orient2d.py:37-84, stops at a
// rounded tail sum and is not exact in general; this is not a copy of it.)
//
// Sign convention follows src/vibespatial/cuda/device_functions/orient2d.py:37-84:
//   orient(a, b, c) = sign((ax-cx)(by-cy) - (ay-cy)(bx-cx)), +1 when a,b,c turn left.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define VS_HD __host__ __device__ __forceinline__

namespace vs {

VS_HD void ef_two_sum(double a, double b, double& s, double& e) {
    const double x = a + b;
    const double bv = x - a;
    const double av = x - bv;
    e = (a - av) + (b - bv);
    s = x;
}

VS_HD void ef_two_diff(double a, double b, double& d, double& e) {
    const double x = a - b;
    const double bv = a - x;
    const double av = x + bv;
    e = (a - av) + (bv - b);
    d = x;
}

VS_HD void ef_two_prod(double a, double b, double& p, double& e) {
    p = a * b;
    e = fma(a, b, -p);
}

// e[0..n) nonoverlapping, increasing magnitude; add b; returns the new length.
VS_HD int ef_grow(double* e, int n, double b) {
    double q = b;
    int m = 0;
    for (int i = 0; i < n; ++i) {
        double s, h;
        ef_two_sum(q, e[i], s, h);
        q = s;
        if (h != 0.0) e[m++] = h;
    }
    if (q != 0.0 || m == 0) e[m++] = q;
    return m;
}

// Exact sign of the full 16-term expansion of the determinant.
__host__ __device__ __noinline__ inline int orient2d_expansion(double ax, double ay, double bx,
                                                               double by, double cx, double cy) {
    double acx, acxt, bcx, bcxt, acy, acyt, bcy, bcyt;
    ef_two_diff(ax, cx, acx, acxt);
    ef_two_diff(bx, cx, bcx, bcxt);
    ef_two_diff(ay, cy, acy, acyt);
    ef_two_diff(by, cy, bcy, bcyt);
    const double lf[4] = {acx, acx, acxt, acxt};
    const double ls[4] = {bcy, bcyt, bcy, bcyt};
    const double rf[4] = {acy, acy, acyt, acyt};
    const double rs[4] = {bcx, bcxt, bcx, bcxt};
    double e[34];
    int n = 0;
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
        double p, t;
        ef_two_prod(lf[k], ls[k], p, t);
        n = ef_grow(e, n, t);
        n = ef_grow(e, n, p);
        ef_two_prod(rf[k], rs[k], p, t);
        n = ef_grow(e, n, -t);
        n = ef_grow(e, n, -p);
    }
    const double top = e[n - 1];
    return (top > 0.0) - (top < 0.0);
}

// Stage-A filtered fp64 orientation, exact.
VS_HD int orient2d(double ax, double ay, double bx, double by, double cx, double cy) {
    const double dl = (ax - cx) * (by - cy);
    const double dr = (ay - cy) * (bx - cx);
    const double det = dl - dr;
    const double bound = 3.3306690738754716e-16 * (fabs(dl) + fabs(dr));
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return orient2d_expansion(ax, ay, bx, by, cx, cy);
}

// One edge a->b of a ring against point p, following the half-open crossing
// rule and the closed-bbox boundary rule of
// src/vibespatial/cuda/device_functions/point_in_ring.py:148-188.
// Returns 1 if p is on the edge, else 0 and toggles `parity` when the
// rightward ray from p crosses the edge.
VS_HD int edge_step_f64(double px, double py, double ax, double ay, double bx, double by,
                        int& parity) {
    const bool crosses = (ay > py) != (by > py);
    const double minx = fmin(ax, bx), maxx = fmax(ax, bx);
    const bool in_box = py >= fmin(ay, by) && py <= fmax(ay, by) && px >= minx && px <= maxx;
    if (!crosses && !in_box) return 0;
    if (crosses && px < minx) {
        parity ^= 1;
        return 0;
    }
    if (crosses && px > maxx) return 0;
    const int o = orient2d(ax, ay, bx, by, px, py);
    if (in_box && o == 0) return 1;
    if (crosses && ((o > 0) == (by > ay))) parity ^= 1;
    return 0;
}

// Exact segment touch-or-cross (polygon_kernels.py:131-161 semantics).
VS_HD bool segments_touch(double ax, double ay, double bx, double by, double cx, double cy,
                          double dx, double dy) {
    if (fmax(ax, bx) < fmin(cx, dx) || fmax(cx, dx) < fmin(ax, bx) || fmax(ay, by) < fmin(cy, dy) ||
        fmax(cy, dy) < fmin(ay, by))
        return false;
    const int o1 = orient2d(ax, ay, bx, by, cx, cy);
    const int o2 = orient2d(ax, ay, bx, by, dx, dy);
    const int o3 = orient2d(cx, cy, dx, dy, ax, ay);
    const int o4 = orient2d(cx, cy, dx, dy, bx, by);
    if (o1 * o2 < 0 && o3 * o4 < 0) return true;
    if (o1 == 0 && cx >= fmin(ax, bx) && cx <= fmax(ax, bx) && cy >= fmin(ay, by) && cy <= fmax(ay, by))
        return true;
    if (o2 == 0 && dx >= fmin(ax, bx) && dx <= fmax(ax, bx) && dy >= fmin(ay, by) && dy <= fmax(ay, by))
        return true;
    if (o3 == 0 && ax >= fmin(cx, dx) && ax <= fmax(cx, dx) && ay >= fmin(cy, dy) && ay <= fmax(cy, dy))
        return true;
    if (o4 == 0 && bx >= fmin(cx, dx) && bx <= fmax(cx, dx) && by >= fmin(cy, dy) && by <= fmax(cy, dy))
        return true;
    return false;
}

// Location code of p in polygonal row `row` (0 outside, 1 boundary, 2 interior):
// XOR over the rings of a part, first part with boundary or interior decides
// (point_relations_kernels.py:37-95).  Rings are closed, so the edges of a ring
// are (c-1, c) for c in (ring_start, ring_end).
VS_HD int point_location_f64(double px, double py, const double* x, const double* y,
                             const int32_t* geom, const int32_t* part, const int32_t* ring,
                             int64_t row) {
    for (int32_t pp = geom[row]; pp < geom[row + 1]; ++pp) {
        int parity = 0;
        for (int32_t rr = part[pp]; rr < part[pp + 1]; ++rr) {
            const int32_t c0 = ring[rr], c1 = ring[rr + 1];
            for (int32_t c = c0 + 1; c < c1; ++c)
                if (edge_step_f64(px, py, x[c - 1], y[c - 1], x[c], y[c], parity)) return 1;
        }
        if (parity) return 2;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// 64-bit Hilbert key: the public-domain prefix-scan method
// (rawrunprotected/hilbert_curves, the same method as the reference's 16-level
// encoder, api/tools/hilbert_curve.py:189-247) widened to 32-bit coordinates:
// one more scan round (shift 16).  Top 32 bits == the 16-level code of the top
// 16 bits of (x, y).
// ---------------------------------------------------------------------------
VS_HD uint64_t spread32(uint32_t v) {
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

VS_HD uint64_t hilbert64(uint32_t x, uint32_t y) {
    const uint32_t F = 0xFFFFFFFFu;
    uint32_t a = x ^ y, b = F ^ a, c = F ^ (x | y), d = x & (y ^ F);
    uint32_t A = a | (b >> 1), B = (a >> 1) ^ a;
    uint32_t C = ((c >> 1) ^ (b & (d >> 1))) ^ c, D = ((a & (c >> 1)) ^ (d >> 1)) ^ d;
#pragma unroll
    for (int s = 2; s <= 16; s <<= 1) {
        a = A; b = B; c = C; d = D;
        if (s < 16) {
            A = (a & (a >> s)) ^ (b & (b >> s));
            B = (a & (b >> s)) ^ (b & ((a ^ b) >> s));
        }
        C ^= (a & (c >> s)) ^ (b & (d >> s));
        D ^= (b & (c >> s)) ^ ((a ^ b) & (d >> s));
    }
    a = C ^ (C >> 1);
    b = D ^ (D >> 1);
    const uint32_t i0 = x ^ y;
    const uint32_t i1 = b | (F ^ (i0 | a));
    return (spread32(i1) << 1) | spread32(i0);
}

// Quantise a coordinate to `bits` bits over [lo, lo + span] (clamped).
VS_HD uint32_t quantize(double v, double lo, double scale, int bits) {
    double q = (v - lo) * scale;
    const double top = (double)((bits >= 32) ? 4294967295.0 : (double)((1ull << bits) - 1));
    q = q > 0.0 ? (q < top ? q : top) : 0.0;  // NaN (empty extent) -> 0
    return (uint32_t)q;
}

// Grid quantisation of the key space: `bits` per axis over an extent.
struct KeyScale {
    double x0, y0, sx, sy;
};

__device__ __forceinline__ KeyScale key_scale(const double4* ext, int bits) {
    const double4 e = ext[0];
    const double cells = (bits >= 32) ? 4294967295.0 : (double)((1ull << bits) - 1);
    KeyScale k;
    k.x0 = e.x;
    k.y0 = e.y;
    k.sx = (e.z > e.x) ? cells / (e.z - e.x) : 0.0;
    k.sy = (e.w > e.y) ? cells / (e.w - e.y) : 0.0;
    return k;
}

}  // namespace vs
