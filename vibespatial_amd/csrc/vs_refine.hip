// Pair refine over candidate pairs, order-preserving compaction, pair lexsort.
//
// Used by the routes that are not the fused point-in-polygon join:
//   * polygon x polygon intersects after a bbox query, one thread per pair:
//     any edge pair touches or crosses (exact orientation + collinear
//     containment), else a vertex of one located in the other
//     (src/vibespatial/predicates/polygon_kernels.py:131-228); the other
//     polygon x polygon predicates through the exact relation of vs_relate.h;
//   * point x polygon location for swapped predicates (contains / covers of a
//     polygon over tree points), predicate mapping of
//     src/vibespatial/predicates/point_relations.py:100-129;
//   * dwithin after a query with distance-expanded boxes, exact distance
//     decisions (replaces _dwithin_refine_gpu, spatial/nearest.py:2855-3016).
// Compaction keeps the input order, so lexsorted candidates give lexsorted
// output (replaces d_left[keep], spatial/query_utils.py:867-868).
#include "vs_distance.h"
#include "vs_exact.h"
#include "vs_internal.h"
#include "vs_polys.h"
#include "vs_relate.h"

namespace vs {

int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, long long* d_total, void* ws,
                       hipStream_t s);
int64_t scan_workspace_bytes(int64_t n);
int sort_pairs_u64(uint64_t*, int32_t*, int64_t, int, int, void*, int64_t, hipStream_t);
int64_t sort_workspace_bytes_u64(int64_t n);

__device__ __forceinline__ int loc_keep(int pred, int loc) {
    switch (pred) {
        case VS_PRED_WITHIN:
        case VS_PRED_CONTAINS: return loc == 2;
        case VS_PRED_TOUCHES: return loc == 1;
        default: return loc != 0;
    }
}

// mode 0: left = points (px, py), right = polygons B; keep = pred(location).
// mode 1: left = polygons A, right = polygons B; keep = pred(A, B), exact
//         (intersects, or the relations of vs_relate.h).
// mode 2: left = polygons B, right = points (px, py); keep = pred(location of
//         the right point in the left polygon) -- contains / covers / touches /
//         intersects of a query polygon over tree points.
// Pairs are visited in order[t] when given (polygon pairs grouped by right
// row, so the lanes of a wave mostly share one right polygon).  Polygon pairs
// whose coordinate counts multiply to more than the heavy-pair threshold go to `heavy`
// (count in *n_heavy) for refine_heavy_kernel, one wave per pair; pairs the
// cheap tests leave undecided go to `undecided` (count in *n_undecided) for
// refine_relate_kernel, so this kernel carries none of the full relation's
// registers.  At 2^8 small pairs (box x box: 25) stay one per lane; one lane
// walking the edge pairs of two many-vertex polygons leaves the rest of its
// wave idle, so those go one wave per pair (100k x 100k star polygons,
// overlaps: 49 ms per step at 2^15, 24 at 2^12, 15.7 at 2^8, flat below).
// The caller can move the threshold (bits 8-15 of vs_refine_pairs' mode), so
// tests run the same pairs through both routes.
constexpr int kHeavyPairLog2 = 8;

__global__ __launch_bounds__(256) void refine_pairs_kernel(int mode, int pred, const int64_t* __restrict__ left,
                                                          const int64_t* __restrict__ right, int64_t n,
                                                          const int32_t* __restrict__ order,
                                                          const double* __restrict__ px,
                                                          const double* __restrict__ py, Polys A, Polys B,
                                                          int32_t* __restrict__ keep, int32_t* __restrict__ heavy,
                                                          unsigned long long* __restrict__ n_heavy,
                                                          int32_t* __restrict__ undecided,
                                                          unsigned long long* __restrict__ n_undecided,
                                                          int64_t heavy_pair) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const int64_t i = order ? (int64_t)order[t] : t;
        const int64_t l = left[i], r = right[i];
        int k;
        if (mode == 0 || mode == 2) {
            const int64_t pi = mode == 0 ? l : r, gi = mode == 0 ? r : l;
            const double x = px[pi], y = py[pi];
            const int loc = (x == x && y == y) ? point_location_f64(x, y, B.x, B.y, B.geom, B.part, B.ring, gi) : 0;
            k = loc_keep(pred, loc);
        } else if ((int64_t)row_coords(A, l) * row_coords(B, r) > heavy_pair) {
            heavy[atomicAdd(n_heavy, 1ull)] = (int32_t)i;
            k = 0;  // decided by refine_heavy_kernel
        } else {
            const int f = polygon_predicate_fast(pred, A, l, B, r);
            if (f < 0) undecided[atomicAdd(n_undecided, 1ull)] = (int32_t)i;
            k = f > 0 ? 1 : 0;
        }
        keep[i] = k;
    }
}

// One wave per heavy polygon pair (grid-stride over the device count).
__global__ __launch_bounds__(256) void refine_heavy_kernel(int pred, const int64_t* __restrict__ left,
                                                          const int64_t* __restrict__ right,
                                                          const int32_t* __restrict__ heavy,
                                                          const unsigned long long* __restrict__ n_heavy, Polys A,
                                                          Polys B, int32_t* __restrict__ keep,
                                                          int32_t* __restrict__ undecided,
                                                          unsigned long long* __restrict__ n_undecided) {
    const int lane = threadIdx.x & 63;
    const int64_t m = (int64_t)*n_heavy;
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); w < m; w += waves) {
        const int64_t i = heavy[w];
        const int f = wave_polygon_predicate_fast(pred, A, left[i], B, right[i], lane);
        if (lane == 0) {
            keep[i] = f > 0 ? 1 : 0;
            if (f < 0) undecided[atomicAdd(n_undecided, 1ull)] = (int32_t)i;
        }
    }
}

// The full relation for the pairs the cheap tests left undecided, one wave
// per pair (grid-stride over the device count; lanes split the edges).
__global__ __launch_bounds__(256) void refine_relate_kernel(int pred, const int64_t* __restrict__ left,
                                                           const int64_t* __restrict__ right,
                                                           const int32_t* __restrict__ undecided,
                                                           const unsigned long long* __restrict__ n_undecided,
                                                           Polys A, Polys B, int32_t* __restrict__ keep) {
    const int lane = threadIdx.x & 63;
    const int64_t m = (int64_t)*n_undecided;
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); w < m; w += waves) {
        const int64_t i = undecided[w];
        const bool k = relation_decides(pred, wave_relate_polygons(A, left[i], B, right[i], lane));
        if (lane == 0) keep[i] = k ? 1 : 0;
    }
}

__global__ void zero_count_kernel(unsigned long long* c) {
    if (threadIdx.x < 2) c[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void compact_pairs_kernel(const int64_t* __restrict__ left,
                                                           const int64_t* __restrict__ right, int64_t n,
                                                           const int32_t* __restrict__ keep,
                                                           const int32_t* __restrict__ pos,
                                                           int64_t* __restrict__ out_l,
                                                           int64_t* __restrict__ out_r) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n && keep[i]) {
        out_l[pos[i]] = left[i];
        out_r[pos[i]] = right[i];
    }
}

__global__ void total_to_result_kernel(const long long* __restrict__ total, int64_t* __restrict__ result) {
    if (threadIdx.x == 0) result[0] = *total;
}

// Exact dwithin per pair; kinds: 0 point rows (x[row], y[row]), 1 polygonal
// rows.  d = distance[left row].
// Pairs are visited in right-row order (order[t]), so the lanes of a wave
// mostly share one right geometry: same coordinates, same loop lengths.
__global__ __launch_bounds__(256) void dwithin_pairs_kernel(int left_kind, int right_kind,
                                                           const int64_t* __restrict__ left,
                                                           const int64_t* __restrict__ right, int64_t n,
                                                           const int32_t* __restrict__ order,
                                                           const double* __restrict__ distance, Polys L, Polys R,
                                                           int32_t* __restrict__ keep) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const int64_t i = order[t];
        const int64_t l = left[i], r = right[i];
        const double d = distance[l];
        bool k;
        if (left_kind == 0 && right_kind == 0) k = point_point_within(L.x[l], L.y[l], R.x[r], R.y[r], d);
        else if (left_kind == 0) k = point_polygon_within(L.x[l], L.y[l], R, r, d);
        else if (right_kind == 0) k = point_polygon_within(R.x[r], R.y[r], L, l, d);
        else k = polygons_within(L, l, R, r, d);
        keep[i] = k ? 1 : 0;
    }
}

__global__ __launch_bounds__(256) void right_keys_kernel(const int64_t* __restrict__ right, int64_t n,
                                                        uint32_t* __restrict__ key, int32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        key[i] = (uint32_t)right[i];
        idx[i] = (int32_t)i;
    }
}

int64_t refine_workspace_bytes(int64_t n);

int64_t dwithin_workspace_bytes(int64_t n) {
    Carver c(nullptr);
    c.take<char>(refine_workspace_bytes(n));
    c.take<uint32_t>(n);
    c.take<int32_t>(n);
    c.take<uint32_t>(n);
    c.take<int32_t>(n);
    c.take<char>(sort_workspace_bytes_u32(n));
    return c.bytes;
}

// Boxes grown by distance[i] on every side, each bound rounded one ulp
// outward so the grown box holds every point within the distance; a NaN or
// negative distance gives a NaN box (matches nothing).
__global__ __launch_bounds__(256) void expand_boxes_kernel(const double4* __restrict__ b, int64_t n,
                                                          const double* __restrict__ distance,
                                                          double4* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double d = distance[i];
    const double4 v = b[i];
    if (!(d >= 0.0)) {
        out[i] = make_double4(NAN, NAN, NAN, NAN);
        return;
    }
    out[i] = make_double4(nextafter(v.x - d, -INFINITY), nextafter(v.y - d, -INFINITY),
                          nextafter(v.z + d, INFINITY), nextafter(v.w + d, INFINITY));
}

int64_t refine_workspace_bytes(int64_t n) {
    Carver c(nullptr);
    c.take<int32_t>(n);
    c.take<int32_t>(n);
    c.take<long long>(1);
    c.take<char>(scan_workspace_bytes(n));
    return c.bytes;
}

__global__ __launch_bounds__(256) void pack_pairs_kernel(const int64_t* __restrict__ l, const int64_t* __restrict__ r,
                                                        int64_t n, uint64_t* __restrict__ key,
                                                        int32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        key[i] = ((uint64_t)l[i] << 32) | (uint64_t)(uint32_t)r[i];
        idx[i] = (int32_t)i;
    }
}

__global__ __launch_bounds__(256) void unpack_pairs_kernel(const uint64_t* __restrict__ key, int64_t n,
                                                          int64_t* __restrict__ l, int64_t* __restrict__ r) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        l[i] = (int64_t)(key[i] >> 32);
        r[i] = (int64_t)(key[i] & 0xFFFFFFFFull);
    }
}

int64_t lexsort_workspace_bytes(int64_t n) {
    Carver c(nullptr);
    c.take<uint64_t>(n);
    c.take<int32_t>(n);
    c.take<char>(sort_workspace_bytes_u64(n));
    return c.bytes;
}

}  // namespace vs

extern "C" {

int64_t vs_refine_pairs_workspace_bytes(int64_t n_pairs) {
    vs::Carver c(nullptr);
    c.take<char>(vs::refine_workspace_bytes(n_pairs));
    c.take<uint32_t>(n_pairs);
    c.take<int32_t>(n_pairs);
    c.take<uint32_t>(n_pairs);
    c.take<int32_t>(n_pairs);
    c.take<char>(vs::sort_workspace_bytes_u32(n_pairs));
    c.take<int32_t>(n_pairs);
    c.take<int32_t>(n_pairs);
    c.take<unsigned long long>(2);
    return c.bytes;
}

int vs_refine_pairs(int32_t mode, int32_t predicate, const int64_t* left, const int64_t* right, int64_t n_pairs,
                    const double* px, const double* py, const double* ax, const double* ay,
                    const int32_t* a_geom, const int32_t* a_part, const int32_t* a_ring, const double* bx,
                    const double* by, const int32_t* b_geom, const int32_t* b_part, const int32_t* b_ring,
                    int64_t* out_left, int64_t* out_right, int64_t* d_result, void* workspace,
                    int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    const int heavy_log2 = (mode >> 8) & 0xFF;
    mode &= 0xFF;
    VS_REQUIRE(mode >= 0 && mode <= 2, "vs_refine_pairs: mode %d not in {0, 1, 2}", mode);
    VS_REQUIRE(heavy_log2 <= 63, "vs_refine_pairs: heavy-pair threshold 2^%d out of range", heavy_log2);
    const int64_t heavy_pair = heavy_log2 == 0 ? ((int64_t)1 << vs::kHeavyPairLog2)
                               : heavy_log2 >= 63 ? INT64_MAX : ((int64_t)1 << heavy_log2);
    VS_REQUIRE(n_pairs >= 0 && n_pairs < (int64_t)INT32_MAX, "vs_refine_pairs: n_pairs out of range");
    VS_REQUIRE(workspace_bytes >= vs_refine_pairs_workspace_bytes(n_pairs), "vs_refine_pairs: workspace too small");
    VS_REQUIRE(d_result != nullptr, "vs_refine_pairs: null result");
    hipStream_t s = vs::stream_of(stream);
    vs::Carver c0(workspace);
    void* refine_ws = c0.take<char>(vs::refine_workspace_bytes(n_pairs));
    uint32_t* key_in = c0.take<uint32_t>(n_pairs);
    int32_t* idx_in = c0.take<int32_t>(n_pairs);
    uint32_t* key_out = c0.take<uint32_t>(n_pairs);
    int32_t* order = c0.take<int32_t>(n_pairs);
    void* sort_ws = c0.take<char>(vs::sort_workspace_bytes_u32(n_pairs));
    int32_t* heavy = c0.take<int32_t>(n_pairs);
    int32_t* undecided = c0.take<int32_t>(n_pairs);
    unsigned long long* counts = c0.take<unsigned long long>(2);  // heavy, undecided
    unsigned long long* n_heavy = counts;
    unsigned long long* n_undecided = counts + 1;
    vs::Carver c(refine_ws);
    int32_t* keep = c.take<int32_t>(n_pairs);
    int32_t* pos = c.take<int32_t>(n_pairs);
    long long* total = c.take<long long>(1);
    void* scan_ws = c.take<char>(vs::scan_workspace_bytes(n_pairs));
    const vs::Polys A = {ax, ay, a_geom, a_part, a_ring};
    const vs::Polys B = {bx, by, b_geom, b_part, b_ring};
    if (n_pairs > 0) {
        const int32_t* ord = nullptr;
        if (mode == 1) {  // polygon pairs: group by right row
            hipLaunchKernelGGL(vs::right_keys_kernel, dim3((unsigned)vs::ceil_div(n_pairs, 256)), dim3(256), 0, s,
                               right, n_pairs, key_in, idx_in);
            VS_LAUNCH_CHECK();
            VS_TRY(vs::sort_pairs_into_u32(key_in, idx_in, key_out, order, n_pairs, 0, 32, sort_ws,
                                           vs::sort_workspace_bytes_u32(n_pairs), s));
            ord = order;
        }
        if (mode == 1) {
            hipLaunchKernelGGL(vs::zero_count_kernel, dim3(1), dim3(64), 0, s, counts);
            VS_LAUNCH_CHECK();
        }
        const int grid = (int)std::min<int64_t>(vs::ceil_div(n_pairs, 256), 16384);
        hipLaunchKernelGGL(vs::refine_pairs_kernel, dim3(grid), dim3(256), 0, s, mode, predicate, left, right,
                           n_pairs, ord, px, py, A, B, keep, heavy, n_heavy, undecided, n_undecided, heavy_pair);
        VS_LAUNCH_CHECK();
        if (mode == 1) {
            const int hgrid = (int)std::min<int64_t>(vs::ceil_div(n_pairs, 4), 2048);
            hipLaunchKernelGGL(vs::refine_heavy_kernel, dim3(hgrid), dim3(256), 0, s, predicate, left, right, heavy,
                               n_heavy, A, B, keep, undecided, n_undecided);
            VS_LAUNCH_CHECK();
            if (predicate != VS_PRED_INTERSECTS && predicate != VS_PRED_CROSSES) {
                const int rgrid = (int)std::min<int64_t>(vs::ceil_div(n_pairs, 4), 2048);
                hipLaunchKernelGGL(vs::refine_relate_kernel, dim3(rgrid), dim3(256), 0, s, predicate, left, right,
                                   undecided, n_undecided, A, B, keep);
                VS_LAUNCH_CHECK();
            }
        }
    }
    VS_TRY(vs::exclusive_scan_i32(keep, pos, n_pairs, total, scan_ws, s));
    if (n_pairs > 0) {
        hipLaunchKernelGGL(vs::compact_pairs_kernel, dim3((unsigned)vs::ceil_div(n_pairs, 256)), dim3(256), 0, s,
                           left, right, n_pairs, keep, pos, out_left, out_right);
        VS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(vs::total_to_result_kernel, dim3(1), dim3(64), 0, s, total, d_result);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int vs_expand_boxes(const double* boxes, int64_t n, const double* distance, double* out, void* stream) {
    vs::clear_error();
    VS_REQUIRE(n >= 0, "vs_expand_boxes: n < 0");
    if (n == 0) return VS_OK;
    VS_REQUIRE(boxes && distance && out, "vs_expand_boxes: null pointer");
    hipStream_t s = vs::stream_of(stream);
    hipLaunchKernelGGL(vs::expand_boxes_kernel, dim3((unsigned)vs::ceil_div(n, 256)), dim3(256), 0, s,
                       reinterpret_cast<const double4*>(boxes), n, distance, reinterpret_cast<double4*>(out));
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int64_t vs_refine_dwithin_workspace_bytes(int64_t n_pairs) { return vs::dwithin_workspace_bytes(n_pairs); }

int vs_refine_dwithin(int32_t left_kind, int32_t right_kind, const int64_t* left, const int64_t* right,
                      int64_t n_pairs, int64_t n_right, const double* distance, const double* lx, const double* ly,
                      const int32_t* l_geom, const int32_t* l_part, const int32_t* l_ring, const double* rx,
                      const double* ry, const int32_t* r_geom, const int32_t* r_part, const int32_t* r_ring,
                      int64_t* out_left, int64_t* out_right, int64_t* d_result, void* workspace,
                      int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    VS_REQUIRE((left_kind == 0 || left_kind == 1) && (right_kind == 0 || right_kind == 1),
               "vs_refine_dwithin: kinds must be 0 (points) or 1 (polygons)");
    VS_REQUIRE(n_pairs >= 0 && n_pairs < (int64_t)INT32_MAX, "vs_refine_dwithin: n_pairs out of range");
    VS_REQUIRE(n_right >= 0 && n_right <= (int64_t)INT32_MAX, "vs_refine_dwithin: n_right out of range");
    VS_REQUIRE(workspace_bytes >= vs::dwithin_workspace_bytes(n_pairs), "vs_refine_dwithin: workspace too small");
    VS_REQUIRE(d_result != nullptr, "vs_refine_dwithin: null result");
    VS_REQUIRE(n_pairs == 0 || distance != nullptr, "vs_refine_dwithin: null distance");
    hipStream_t s = vs::stream_of(stream);
    vs::Carver c(workspace);
    void* refine_ws = c.take<char>(vs::refine_workspace_bytes(n_pairs));
    uint32_t* key_in = c.take<uint32_t>(n_pairs);
    int32_t* idx_in = c.take<int32_t>(n_pairs);
    uint32_t* key_out = c.take<uint32_t>(n_pairs);
    int32_t* order = c.take<int32_t>(n_pairs);
    void* sort_ws = c.take<char>(vs::sort_workspace_bytes_u32(n_pairs));
    vs::Carver cr(refine_ws);
    int32_t* keep = cr.take<int32_t>(n_pairs);
    int32_t* pos = cr.take<int32_t>(n_pairs);
    long long* total = cr.take<long long>(1);
    void* scan_ws = cr.take<char>(vs::scan_workspace_bytes(n_pairs));
    const vs::Polys L = {lx, ly, l_geom, l_part, l_ring};
    const vs::Polys R = {rx, ry, r_geom, r_part, r_ring};
    if (n_pairs > 0) {
        hipLaunchKernelGGL(vs::right_keys_kernel, dim3((unsigned)vs::ceil_div(n_pairs, 256)), dim3(256), 0, s, right,
                           n_pairs, key_in, idx_in);
        VS_LAUNCH_CHECK();
        int bits = 0;
        while (bits < 31 && ((int64_t)1 << bits) < n_right) ++bits;
        VS_TRY(vs::sort_pairs_into_u32(key_in, idx_in, key_out, order, n_pairs, 0, bits, sort_ws,
                                       vs::sort_workspace_bytes_u32(n_pairs), s));
        const int grid = (int)std::min<int64_t>(vs::ceil_div(n_pairs, 256), 16384);
        hipLaunchKernelGGL(vs::dwithin_pairs_kernel, dim3(grid), dim3(256), 0, s, left_kind, right_kind, left, right,
                           n_pairs, order, distance, L, R, keep);
        VS_LAUNCH_CHECK();
    }
    VS_TRY(vs::exclusive_scan_i32(keep, pos, n_pairs, total, scan_ws, s));
    if (n_pairs > 0) {
        hipLaunchKernelGGL(vs::compact_pairs_kernel, dim3((unsigned)vs::ceil_div(n_pairs, 256)), dim3(256), 0, s,
                           left, right, n_pairs, keep, pos, out_left, out_right);
        VS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(vs::total_to_result_kernel, dim3(1), dim3(64), 0, s, total, d_result);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int64_t vs_lexsort_pairs_workspace_bytes(int64_t n_pairs) { return vs::lexsort_workspace_bytes(n_pairs); }

int vs_lexsort_pairs(int64_t* left, int64_t* right, int64_t n_pairs, void* workspace, int64_t workspace_bytes,
                     void* stream) {
    vs::clear_error();
    VS_REQUIRE(n_pairs >= 0 && n_pairs < (int64_t)INT32_MAX, "vs_lexsort_pairs: n_pairs out of range");
    VS_REQUIRE(workspace_bytes >= vs::lexsort_workspace_bytes(n_pairs), "vs_lexsort_pairs: workspace too small");
    if (n_pairs < 2) return VS_OK;
    hipStream_t s = vs::stream_of(stream);
    vs::Carver c(workspace);
    uint64_t* key = c.take<uint64_t>(n_pairs);
    int32_t* idx = c.take<int32_t>(n_pairs);
    void* sort_ws = c.take<char>(vs::sort_workspace_bytes_u64(n_pairs));
    const unsigned g = (unsigned)vs::ceil_div(n_pairs, 256);
    hipLaunchKernelGGL(vs::pack_pairs_kernel, dim3(g), dim3(256), 0, s, left, right, n_pairs, key, idx);
    VS_LAUNCH_CHECK();
    VS_TRY(vs::sort_pairs_u64(key, idx, n_pairs, 0, 64, sort_ws, vs::sort_workspace_bytes_u64(n_pairs), s));
    hipLaunchKernelGGL(vs::unpack_pairs_kernel, dim3(g), dim3(256), 0, s, key, n_pairs, left, right);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

}  // extern "C"
