// Polygon x polygon intersects on the exact predicates (shared by the pair
// refine and the batched join).
#pragma once

#include "vs_exact.h"

namespace vs {

// Polygonal rows in GeoArrow-style offsets: geom -> parts -> rings -> coords.
struct Polys {
    const double* x;
    const double* y;
    const int32_t* geom;
    const int32_t* part;
    const int32_t* ring;
};

// Exact `intersects` of row ra of A and row rb of B
// (src/vibespatial/predicates/polygon_kernels.py:163-228): some boundary
// segment of A touches or crosses one of B, else a point of one lies in the
// other (location != 0).
__device__ inline bool polygons_intersect(const Polys& A, int64_t ra, const Polys& B, int64_t rb) {
    const int32_t ap0 = A.geom[ra], ap1 = A.geom[ra + 1];
    const int32_t bp0 = B.geom[rb], bp1 = B.geom[rb + 1];
    const int32_t ar0 = A.part[ap0], ar1 = A.part[ap1];
    const int32_t br0 = B.part[bp0], br1 = B.part[bp1];
    if (A.ring[ar1] <= A.ring[ar0] || B.ring[br1] <= B.ring[br0]) return false;
    for (int32_t ar = ar0; ar < ar1; ++ar)
        for (int32_t ai = A.ring[ar] + 1; ai < A.ring[ar + 1]; ++ai) {
            const double ax0 = A.x[ai - 1], ay0 = A.y[ai - 1], ax1 = A.x[ai], ay1 = A.y[ai];
            for (int32_t br = br0; br < br1; ++br)
                for (int32_t bi = B.ring[br] + 1; bi < B.ring[br + 1]; ++bi)
                    if (segments_touch(ax0, ay0, ax1, ay1, B.x[bi - 1], B.y[bi - 1], B.x[bi], B.y[bi]))
                        return true;
        }
    // No boundary segments touch: every ring of A is a closed curve that does
    // not meet the boundary of B, so all its points share one location in B
    // (the even-odd location is constant off the boundary).  The first vertex
    // of each ring therefore decides what the reference's loop over every
    // vertex decides (polygon_kernels.py:202-226), and likewise for B in A.
    for (int32_t ar = ar0; ar < ar1; ++ar) {
        const int32_t c0 = A.ring[ar];
        if (A.ring[ar + 1] > c0 &&
            point_location_f64(A.x[c0], A.y[c0], B.x, B.y, B.geom, B.part, B.ring, rb) != 0)
            return true;
    }
    for (int32_t br = br0; br < br1; ++br) {
        const int32_t c0 = B.ring[br];
        if (B.ring[br + 1] > c0 &&
            point_location_f64(B.x[c0], B.y[c0], A.x, A.y, A.geom, A.part, A.ring, ra) != 0)
            return true;
    }
    return false;
}

// Same predicate when one side A is a single ring (one part, one ring), with
// the loops turned around: the other polygon's edges are streamed to locate
// A's first vertex in B (inside or on the boundary: true at once); only if
// some B edge's box meets A's box are they streamed again, each tested
// against A's edges; then the first vertex of each of B's rings is located
// in A unless it lies outside A's bounds.  The bounds checks use exact
// min/max, so they only skip edges that cannot touch A.  Segment touch is
// symmetric and the location checks are those of polygons_intersect, so the
// answer is the same boolean.
__device__ inline bool single_ring_of(const Polys& P, int64_t r, int32_t& c0, int32_t& n) {
    const int32_t p0 = P.geom[r];
    if (P.geom[r + 1] - p0 != 1) return false;
    const int32_t r0 = P.part[p0];
    if (P.part[p0 + 1] - r0 != 1) return false;
    c0 = P.ring[r0];
    n = P.ring[r0 + 1] - c0;
    return n >= 1;
}

// Ring coordinates k = 0..n-1 at x[k * stride], y[k * stride] (global memory
// with stride 1, or a lane's column of an LDS tile).
struct RingRef {
    const double* x;
    const double* y;
    int32_t stride;
    __device__ __forceinline__ double X(int32_t k) const { return x[k * stride]; }
    __device__ __forceinline__ double Y(int32_t k) const { return y[k * stride]; }
};

// skip_first: A's first vertex is known to lie outside every part of B and
// off its boundary (pass 1's outcome apart from `near`), so pass 1 is not
// run and pass 2 tests every edge's box.
__device__ inline bool ring_intersects_polys(const RingRef& A, int32_t na, const Polys& B, int64_t rb,
                                             bool skip_first = false) {
    const int32_t bp0 = B.geom[rb], bp1 = B.geom[rb + 1];
    if (B.ring[B.part[bp1]] <= B.ring[B.part[bp0]]) return false;
    double mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
    for (int32_t k = 0; k < na; ++k) {
        mnx = fmin(mnx, A.X(k));
        mny = fmin(mny, A.Y(k));
        mxx = fmax(mxx, A.X(k));
        mxy = fmax(mxy, A.Y(k));
    }
    // Pass 1: location of A's first vertex in each part of B (inside or on
    // the boundary decides at once), noting whether any B edge comes near A.
    const double fx = A.X(0), fy = A.Y(0);
    bool near = skip_first;
    for (int32_t bp = skip_first ? bp1 : bp0; bp < bp1; ++bp) {
        int parity = 0;
        for (int32_t br = B.part[bp]; br < B.part[bp + 1]; ++br) {
            const int32_t e0 = B.ring[br], e1 = B.ring[br + 1];
            if (e1 <= e0) continue;
            double px = B.x[e0], py = B.y[e0];
            // four coordinates loaded together per step (memory-level parallelism)
            for (int32_t c0 = e0 + 1; c0 < e1; c0 += 4) {
                double qxs[4], qys[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int32_t c = c0 + u < e1 ? c0 + u : e1 - 1;
                    qxs[u] = B.x[c];
                    qys[u] = B.y[c];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (c0 + u >= e1) break;
                    const double qx = qxs[u], qy = qys[u];
                    near |= fmax(px, qx) >= mnx && fmin(px, qx) <= mxx && fmax(py, qy) >= mny && fmin(py, qy) <= mxy;
                    if (edge_step_f64(fx, fy, px, py, qx, qy, parity)) return true;  // on B's boundary
                    px = qx;
                    py = qy;
                }
            }
        }
        if (parity) return true;  // inside this part of B
    }
    // Pass 2 (only when some B edge's box meets A's box): boundary contacts.
    if (near) {
        for (int32_t br = B.part[bp0]; br < B.part[bp1]; ++br) {
            const int32_t e0 = B.ring[br], e1 = B.ring[br + 1];
            if (e1 <= e0) continue;
            double px = B.x[e0], py = B.y[e0];
            for (int32_t c0 = e0 + 1; c0 < e1; c0 += 4) {
                double qxs[4], qys[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int32_t c = c0 + u < e1 ? c0 + u : e1 - 1;
                    qxs[u] = B.x[c];
                    qys[u] = B.y[c];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (c0 + u >= e1) break;
                    const double qx = qxs[u], qy = qys[u];
                    if (fmax(px, qx) >= mnx && fmin(px, qx) <= mxx && fmax(py, qy) >= mny && fmin(py, qy) <= mxy) {
                        for (int32_t k = 1; k < na; ++k)
                            if (segments_touch(A.X(k - 1), A.Y(k - 1), A.X(k), A.Y(k), px, py, qx, qy)) return true;
                    }
                    px = qx;
                    py = qy;
                }
            }
        }
    }
    for (int32_t br = B.part[bp0]; br < B.part[bp1]; ++br) {
        const int32_t e0 = B.ring[br];
        if (B.ring[br + 1] <= e0) continue;
        const double px = B.x[e0], py = B.y[e0];
        if (px < mnx || px > mxx || py < mny || py > mxy) continue;  // outside A's bounds: location 0
        int parity = 0;
        for (int32_t k = 1; k < na; ++k)
            if (edge_step_f64(px, py, A.X(k - 1), A.Y(k - 1), A.X(k), A.Y(k), parity)) return true;
        if (parity) return true;
    }
    return false;
}

__device__ inline bool polygons_intersect_ring(const Polys& A, int32_t a0, int32_t na, const Polys& B, int64_t rb) {
    return ring_intersects_polys(RingRef{A.x + a0, A.y + a0, 1}, na, B, rb);
}

// polygons_intersect, streaming the larger side when the smaller is one ring.
__device__ inline bool polygons_intersect_auto(const Polys& A, int64_t ra, const Polys& B, int64_t rb) {
    int32_t ca, na, cb, nb;
    const bool sa = single_ring_of(A, ra, ca, na), sb = single_ring_of(B, rb, cb, nb);
    if (sa && (!sb || na <= nb)) return polygons_intersect_ring(A, ca, na, B, rb);
    if (sb) return polygons_intersect_ring(B, cb, nb, A, ra);
    return polygons_intersect(A, ra, B, rb);
}

}  // namespace vs
