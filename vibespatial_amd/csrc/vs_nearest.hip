// Nearest-neighbour join support: fp64 distances of candidate pairs and the
// per-query selection of the nearest ones (ties kept), driven by an
// expanding-radius box search on the host side.
//
// Replaces the reference's nearest refine (spatial/nearest.py:3278
// _nearest_refine_gpu, distances in _compute_pair_distances_gpu :2228 and
// _compute_mixed_distances_gpu :2482, reduction in _nearest_from_distances
// :3237).  The distance of two geometries is the fp64 value GEOS computes:
//   point - point     sqrt(dx*dx + dy*dy)
//   point - segment   GEOS Distance::pointToSegment: the projection parameter
//                     r = ((p - a).(b - a)) / |b - a|^2 decides the end point
//                     (r <= 0, r >= 1) or the perpendicular
//                     |((a.y - p.y)(b.x - a.x) - (a.x - p.x)(b.y - a.y)) / |b - a|^2| * sqrt(|b - a|^2)
//   point - polygon   0 when the point is not outside (exact location), else
//                     the least point - segment distance over the boundary
//   polygon - polygon 0 when they intersect (exact test), else the least
//                     vertex - segment distance either way
// Empty geometries and NaN points are at +inf (never nearest).  Ties are
// equal fp64 distances, as in GEOS's nearest query; the reference's GPU
// reduction instead keeps distances within numpy.isclose of the minimum.
//
// A round of the search (host side): every open query's box grown by its
// radius -> box query over the tree -> vs_pair_distances -> vs_nearest_select,
// which keeps, for each query whose least candidate distance is <= its radius,
// the pairs at that least distance and closes it; other queries get the next
// radius.  Any tree geometry at fp64 distance <= r from a query lies in the
// query box grown by r (1 + 2^-40) (the search radius), so a closed query's
// pairs are complete.
#include "vs_exact.h"
#include "vs_internal.h"
#include "vs_polys.h"

namespace vs {

int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, long long* d_total, void* ws,
                       hipStream_t s);
int64_t scan_workspace_bytes(int64_t n);

__device__ __forceinline__ double point_dist_f64(double px, double py, double qx, double qy) {
    const double dx = px - qx, dy = py - qy;
    return sqrt(dx * dx + dy * dy);
}

__device__ __forceinline__ double segment_dist_f64(double px, double py, double ax, double ay, double bx,
                                                   double by) {
    if (ax == bx && ay == by) return point_dist_f64(px, py, ax, ay);
    const double ux = bx - ax, uy = by - ay;
    const double len2 = ux * ux + uy * uy;
    const double r = ((px - ax) * ux + (py - ay) * uy) / len2;
    if (r <= 0.0) return point_dist_f64(px, py, ax, ay);
    if (r >= 1.0) return point_dist_f64(px, py, bx, by);
    const double s = ((ay - py) * ux - (ax - px) * uy) / len2;
    return fabs(s) * sqrt(len2);
}

__device__ __forceinline__ bool poly_empty(const Polys& P, int64_t row) {
    return P.ring[P.part[P.geom[row + 1]]] <= P.ring[P.part[P.geom[row]]];
}

// Least distance from p to the boundary segments of polygonal row `row`
// (min with `best`).  A segment whose box lies more than best (1 + 2^-30)
// away on an axis is skipped: its computed distance (a few ulps from the
// true one) exceeds best, so the minimum is the one over every segment.
__device__ double boundary_dist_f64(double px, double py, const Polys& P, int64_t row, double best) {
    const int32_t r0 = P.part[P.geom[row]], r1 = P.part[P.geom[row + 1]];
    for (int32_t r = r0; r < r1; ++r)
        for (int32_t c = P.ring[r] + 1; c < P.ring[r + 1]; ++c) {
            const double ax = P.x[c - 1], ay = P.y[c - 1], bx = P.x[c], by = P.y[c];
            const double lim = best * (1.0 + 0x1p-30);
            if (px - fmax(ax, bx) > lim || fmin(ax, bx) - px > lim || py - fmax(ay, by) > lim ||
                fmin(ay, by) - py > lim)
                continue;
            best = fmin(best, segment_dist_f64(px, py, ax, ay, bx, by));
        }
    return best;
}

__device__ double point_polygon_dist_f64(double px, double py, const Polys& P, int64_t row) {
    if (px != px || py != py || poly_empty(P, row)) return INFINITY;
    if (point_location_f64(px, py, P.x, P.y, P.geom, P.part, P.ring, row) != 0) return 0.0;
    return boundary_dist_f64(px, py, P, row, INFINITY);
}

__device__ double polygons_dist_f64(const Polys& A, int64_t ra, const Polys& B, int64_t rb) {
    if (poly_empty(A, ra) || poly_empty(B, rb)) return INFINITY;
    if (polygons_intersect_auto(A, ra, B, rb)) return 0.0;
    double best = INFINITY;
    const int32_t a0 = A.ring[A.part[A.geom[ra]]], a1 = A.ring[A.part[A.geom[ra + 1]]];
    for (int32_t v = a0; v < a1; ++v) best = boundary_dist_f64(A.x[v], A.y[v], B, rb, best);
    const int32_t b0 = B.ring[B.part[B.geom[rb]]], b1 = B.ring[B.part[B.geom[rb + 1]]];
    for (int32_t v = b0; v < b1; ++v) best = boundary_dist_f64(B.x[v], B.y[v], A, ra, best);
    return best;
}

// Kinds: 0 point rows (x[row], y[row]), 1 polygonal rows.  exclusive: a
// point pair with equal coordinates is at +inf (equal geometries excluded).
// Pairs are visited in right-row order (order[t]) when given.
__global__ __launch_bounds__(256) void pair_distances_kernel(int left_kind, int right_kind,
                                                            const int64_t* __restrict__ left,
                                                            const int64_t* __restrict__ right, int64_t n,
                                                            const int32_t* __restrict__ order, int exclusive,
                                                            Polys L, Polys R, double* __restrict__ out) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const int64_t i = order ? (int64_t)order[t] : t;
        const int64_t l = left[i], r = right[i];
        double d;
        if (left_kind == 0 && right_kind == 0) {
            const double px = L.x[l], py = L.y[l], qx = R.x[r], qy = R.y[r];
            if (px != px || py != py || qx != qx || qy != qy) d = INFINITY;
            else if (exclusive && px == qx && py == qy) d = INFINITY;
            else d = point_dist_f64(px, py, qx, qy);
        } else if (left_kind == 0) {
            d = point_polygon_dist_f64(L.x[l], L.y[l], R, r);
        } else if (right_kind == 0) {
            d = point_polygon_dist_f64(R.x[r], R.y[r], L, l);
        } else {
            d = polygons_dist_f64(L, l, R, r);
        }
        out[i] = d;
    }
}

__global__ __launch_bounds__(256) void right_order_keys_kernel(const int64_t* __restrict__ right, int64_t n,
                                                              uint32_t* __restrict__ key, int32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        key[i] = (uint32_t)right[i];
        idx[i] = (int32_t)i;
    }
}

int64_t pair_distances_workspace_bytes(int64_t n) {
    Carver c(nullptr);
    c.take<uint32_t>(n);
    c.take<int32_t>(n);
    c.take<uint32_t>(n);
    c.take<int32_t>(n);
    c.take<char>(sort_workspace_bytes_u32(n));
    return c.bytes;
}

// Non-negative doubles order as their bit patterns.
__global__ __launch_bounds__(256) void nearest_init_kernel(int64_t nq, unsigned long long* __restrict__ mind,
                                                          unsigned long long* __restrict__ first,
                                                          unsigned long long* __restrict__ open) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q < nq) {
        mind[q] = (unsigned long long)__double_as_longlong(INFINITY);
        first[q] = ~0ull;
    }
    if (q == 0) *open = 0;
}

__global__ __launch_bounds__(256) void nearest_min_kernel(const int64_t* __restrict__ left,
                                                         const double* __restrict__ dist, int64_t n,
                                                         unsigned long long* __restrict__ mind) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const double d = dist[i];
        if (d < INFINITY) atomicMin(&mind[left[i]], (unsigned long long)__double_as_longlong(d));
    }
}

// keep = the pair is at its query's least distance and that distance is
// within the query's radius; first = the least right row among kept pairs.
__global__ __launch_bounds__(256) void nearest_keep_kernel(const int64_t* __restrict__ left,
                                                          const int64_t* __restrict__ right,
                                                          const double* __restrict__ dist, int64_t n,
                                                          const double* __restrict__ radius,
                                                          const unsigned long long* __restrict__ mind,
                                                          unsigned long long* __restrict__ first, int return_all,
                                                          int32_t* __restrict__ keep) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t q = left[i];
    const double d = dist[i];
    const double m = __longlong_as_double((long long)mind[q]);
    const bool k = d < INFINITY && d == m && m <= radius[q];
    keep[i] = k ? 1 : 0;
    if (k && !return_all) atomicMin(&first[q], (unsigned long long)right[i]);
}

__global__ __launch_bounds__(256) void nearest_first_kernel(const int64_t* __restrict__ left,
                                                           const int64_t* __restrict__ right, int64_t n,
                                                           const unsigned long long* __restrict__ first,
                                                           int32_t* __restrict__ keep) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n && keep[i] && (unsigned long long)right[i] != first[left[i]]) keep[i] = 0;
}

__global__ __launch_bounds__(256) void nearest_compact_kernel(const int64_t* __restrict__ left,
                                                             const int64_t* __restrict__ right,
                                                             const double* __restrict__ dist, int64_t n,
                                                             const int32_t* __restrict__ keep,
                                                             const int32_t* __restrict__ pos,
                                                             int64_t* __restrict__ out_l, int64_t* __restrict__ out_r,
                                                             double* __restrict__ out_d) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n && keep[i]) {
        out_l[pos[i]] = left[i];
        out_r[pos[i]] = right[i];
        if (out_d) out_d[pos[i]] = dist[i];
    }
}

// Next radius per query: NaN once closed (nearest found within the radius,
// or the final radius searched without a match), else radius * growth, or
// final_radius once that reaches switch_radius; search = the radius the
// query box is grown by, radius * (1 + 2^-40).
constexpr double kSearchSlack = 1.0 + 0x1p-40;

__global__ __launch_bounds__(256) void nearest_radius_kernel(int64_t nq, double* __restrict__ radius,
                                                            double* __restrict__ search,
                                                            const unsigned long long* __restrict__ mind,
                                                            double growth, double switch_radius,
                                                            double final_radius,
                                                            unsigned long long* __restrict__ open) {
    __shared__ int cnt[4];
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int o = 0;
    if (q < nq) {
        const double r = radius[q];
        if (r == r) {
            const double m = __longlong_as_double((long long)mind[q]);
            double nr;
            if (m <= r || r >= final_radius) {
                nr = NAN;
            } else {
                nr = r * growth;
                if (!(nr < switch_radius)) nr = final_radius;
                o = 1;
            }
            radius[q] = nr;
            search[q] = nr * kSearchSlack;
        }
    }
    const int w = wave_sum_i32(o);
    if ((threadIdx.x & 63) == 0) cnt[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int t = cnt[0] + cnt[1] + cnt[2] + cnt[3];
        if (t) atomicAdd(open, (unsigned long long)t);
    }
}

__global__ void nearest_result_kernel(const long long* __restrict__ total, const unsigned long long* __restrict__ open,
                                      int64_t* __restrict__ result) {
    if (threadIdx.x == 0) {
        result[0] = *total;
        result[1] = (int64_t)*open;
    }
}

struct NearestWs {
    unsigned long long* mind;
    unsigned long long* first;
    unsigned long long* open;
    int32_t* keep;
    int32_t* pos;
    long long* total;
    void* scan_ws;
};

static NearestWs carve_nearest(Carver& c, int64_t n_pairs, int64_t n_query) {
    NearestWs W;
    W.mind = c.take<unsigned long long>(n_query);
    W.first = c.take<unsigned long long>(n_query);
    W.open = c.take<unsigned long long>(1);
    W.keep = c.take<int32_t>(n_pairs);
    W.pos = c.take<int32_t>(n_pairs);
    W.total = c.take<long long>(1);
    W.scan_ws = c.take<char>(scan_workspace_bytes(n_pairs));
    return W;
}

int64_t nearest_select_workspace_bytes(int64_t n_pairs, int64_t n_query) {
    Carver c(nullptr);
    carve_nearest(c, n_pairs, n_query);
    return c.bytes;
}

}  // namespace vs

extern "C" {

int64_t vs_pair_distances_workspace_bytes(int64_t n_pairs) { return vs::pair_distances_workspace_bytes(n_pairs); }

int vs_pair_distances(int32_t left_kind, int32_t right_kind, const int64_t* left, const int64_t* right,
                      int64_t n_pairs, int64_t n_right, int32_t exclusive, const double* lx, const double* ly,
                      const int32_t* l_geom, const int32_t* l_part, const int32_t* l_ring, const double* rx,
                      const double* ry, const int32_t* r_geom, const int32_t* r_part, const int32_t* r_ring,
                      double* out_distance, void* workspace, int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    VS_REQUIRE((left_kind == 0 || left_kind == 1) && (right_kind == 0 || right_kind == 1),
               "vs_pair_distances: kinds must be 0 (points) or 1 (polygons)");
    VS_REQUIRE(n_pairs >= 0 && n_pairs < (int64_t)INT32_MAX, "vs_pair_distances: n_pairs out of range");
    VS_REQUIRE(n_right >= 0 && n_right <= (int64_t)INT32_MAX, "vs_pair_distances: n_right out of range");
    VS_REQUIRE(workspace_bytes >= vs::pair_distances_workspace_bytes(n_pairs),
               "vs_pair_distances: workspace too small");
    if (n_pairs == 0) return VS_OK;
    VS_REQUIRE(left && right && out_distance, "vs_pair_distances: null pointer");
    hipStream_t s = vs::stream_of(stream);
    vs::Carver c(workspace);
    uint32_t* key_in = c.take<uint32_t>(n_pairs);
    int32_t* idx_in = c.take<int32_t>(n_pairs);
    uint32_t* key_out = c.take<uint32_t>(n_pairs);
    int32_t* order = c.take<int32_t>(n_pairs);
    void* sort_ws = c.take<char>(vs::sort_workspace_bytes_u32(n_pairs));
    const vs::Polys L = {lx, ly, l_geom, l_part, l_ring};
    const vs::Polys R = {rx, ry, r_geom, r_part, r_ring};
    const int32_t* ord = nullptr;
    if (right_kind == 1) {  // lanes of a wave share a polygon
        hipLaunchKernelGGL(vs::right_order_keys_kernel, dim3((unsigned)vs::ceil_div(n_pairs, 256)), dim3(256), 0, s,
                           right, n_pairs, key_in, idx_in);
        VS_LAUNCH_CHECK();
        int bits = 0;
        while (bits < 31 && ((int64_t)1 << bits) < n_right) ++bits;
        VS_TRY(vs::sort_pairs_into_u32(key_in, idx_in, key_out, order, n_pairs, 0, bits, sort_ws,
                                       vs::sort_workspace_bytes_u32(n_pairs), s));
        ord = order;
    }
    const int grid = (int)std::min<int64_t>(vs::ceil_div(n_pairs, 256), 16384);
    hipLaunchKernelGGL(vs::pair_distances_kernel, dim3(grid), dim3(256), 0, s, left_kind, right_kind, left, right,
                       n_pairs, ord, exclusive ? 1 : 0, L, R, out_distance);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int64_t vs_nearest_select_workspace_bytes(int64_t n_pairs, int64_t n_query) {
    return vs::nearest_select_workspace_bytes(n_pairs, n_query);
}

int vs_nearest_select(const int64_t* left, const int64_t* right, const double* distance, int64_t n_pairs,
                      int64_t n_query, double* radius, double* search, double growth, double switch_radius, double final_radius,
                      int32_t return_all, int64_t* out_left, int64_t* out_right, double* out_distance,
                      int64_t* d_result, void* workspace, int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    VS_REQUIRE(n_pairs >= 0 && n_pairs < (int64_t)INT32_MAX, "vs_nearest_select: n_pairs out of range");
    VS_REQUIRE(n_query >= 0 && n_query < (int64_t)INT32_MAX, "vs_nearest_select: n_query out of range");
    VS_REQUIRE(workspace_bytes >= vs::nearest_select_workspace_bytes(n_pairs, n_query),
               "vs_nearest_select: workspace too small");
    VS_REQUIRE(d_result != nullptr && (n_query == 0 || (radius != nullptr && search != nullptr)),
               "vs_nearest_select: null pointer");
    VS_REQUIRE(growth > 1.0, "vs_nearest_select: growth must exceed 1");
    hipStream_t s = vs::stream_of(stream);
    vs::Carver c(workspace);
    const vs::NearestWs W = vs::carve_nearest(c, n_pairs, n_query);
    hipLaunchKernelGGL(vs::nearest_init_kernel, dim3((unsigned)std::max<int64_t>(1, vs::ceil_div(n_query, 256))),
                       dim3(256), 0, s, n_query, W.mind, W.first, W.open);
    VS_LAUNCH_CHECK();
    const unsigned gp = (unsigned)vs::ceil_div(n_pairs, 256);
    if (n_pairs > 0) {
        hipLaunchKernelGGL(vs::nearest_min_kernel, dim3(gp), dim3(256), 0, s, left, distance, n_pairs, W.mind);
        VS_LAUNCH_CHECK();
        hipLaunchKernelGGL(vs::nearest_keep_kernel, dim3(gp), dim3(256), 0, s, left, right, distance, n_pairs, radius,
                           W.mind, W.first, return_all ? 1 : 0, W.keep);
        VS_LAUNCH_CHECK();
        if (!return_all) {
            hipLaunchKernelGGL(vs::nearest_first_kernel, dim3(gp), dim3(256), 0, s, left, right, n_pairs, W.first,
                               W.keep);
            VS_LAUNCH_CHECK();
        }
    }
    VS_TRY(vs::exclusive_scan_i32(W.keep, W.pos, n_pairs, W.total, W.scan_ws, s));
    if (n_pairs > 0) {
        hipLaunchKernelGGL(vs::nearest_compact_kernel, dim3(gp), dim3(256), 0, s, left, right, distance, n_pairs,
                           W.keep, W.pos, out_left, out_right, out_distance);
        VS_LAUNCH_CHECK();
    }
    if (n_query > 0) {
        hipLaunchKernelGGL(vs::nearest_radius_kernel, dim3((unsigned)vs::ceil_div(n_query, 256)), dim3(256), 0, s,
                           n_query, radius, search, W.mind, growth, switch_radius, final_radius, W.open);
        VS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(vs::nearest_result_kernel, dim3(1), dim3(64), 0, s, W.total, W.open, d_result);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

}  // extern "C"
