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

}  // namespace vs
