// Index build: per-row bounds, centre extent, 64-bit Hilbert keys, packed tree.
//
// Replaces, on the sjoin path:
//   bounds_polygon / bounds_multipolygon   src/vibespatial/kernels/core/geometry_analysis_source.py:31-212
//   _device_total_bounds (D2H fence)       src/vibespatial/spatial/indexing.py:1241-1262
//                                          (kept on the device here: no host round trip)
//   morton_keys_from_bounds                src/vibespatial/spatial/indexing_kernels.py:20-49
//   compute_hilbert_codes (32-bit)         src/vibespatial/io/gpu_parse/indexing_kernels.py:78-176
//   _build_packed_rtree (fan-out 16)       src/vibespatial/io/gpu_parse/indexing.py:262-400
//
// Layout in HBM (all levels in one array of 32-byte boxes, root first, items
// last, like the reference's root-first layout; the children of node j of level
// l are nodes [j*F, j*F+F) of level l+1, so no child-pointer array is stored):
//   node_boxes[level_start[l] + j] = {minx, miny, maxx, maxy}   (fp64)
//   item_ids[k] = original row of the k-th item in Hilbert order
// Empty/null rows have NaN boxes, key UINT64_MAX, sort last and never match.
#include "vs_exact.h"
#include "vs_internal.h"

namespace vs {

int sort_pairs_u64(uint64_t*, int32_t*, int64_t, int, int, void*, int64_t, hipStream_t);
int sort_pairs_u32(uint32_t*, int32_t*, int64_t, int, int, void*, int64_t, hipStream_t);
int64_t sort_workspace_bytes_u64(int64_t n);
int64_t sort_workspace_bytes_u32(int64_t n);

// One wave per 16 consecutive polygonal rows (rows are contiguous, so a wave
// reads one span of memory).  Lanes 0..15 load the rows' coordinate ranges;
// then, by the longest of the 16:
//   <= 16 coordinates: four lanes per row, each loading up to four
//     coordinates at once, reduced with 4-wide shuffles;
//   longer: sixteen lanes per row, four rows at a time, 64 coordinates per
//     pass (four loads in flight per lane), reduced with 16-wide shuffles.
// Every load of a pass is issued before the first use.
constexpr int kBoundsRows = 16;
// With info != nullptr also, per row, info = (c0 when the row is one ring of
// at least one coordinate, else -1; coordinates of the row), and with
// rec != nullptr the row's RowRec (bounds, that ring's first vertex, info):
// the offsets are already in hand.
__global__ __launch_bounds__(256) void polygon_bounds_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const int32_t* __restrict__ geom,
    const int32_t* __restrict__ part, const int32_t* __restrict__ ring, int64_t n_rows,
    double4* __restrict__ out, int2* __restrict__ info, RowRec* __restrict__ rec) {
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * kBoundsRows;
    if (r0 >= n_rows) return;  // wave-uniform
    int32_t c0 = 0, c1 = 0, ci = -1;  // ci: c0 for a single ring, else -1
    if (lane < kBoundsRows && r0 + lane < n_rows) {
        const int32_t p0 = geom[r0 + lane], p1 = geom[r0 + lane + 1];
        const int32_t q0 = part[p0], q1 = part[p1];
        c0 = ring[q0];
        c1 = ring[q1];
        ci = (p1 - p0 == 1 && q1 - q0 == 1 && c1 > c0) ? c0 : -1;
        if (info) info[r0 + lane] = make_int2(ci, c1 - c0);
    }
    // the whole record of a row is written by the lane that holds its first
    // coordinate after the reduction (one 64-byte store per row)
    auto write_rec = [&](int64_t row, double4 box, int32_t a0, int32_t a1, int32_t single_c0, double fx,
                         double fy) {
        RowRec v;
        v.box = box;
        v.first = single_c0 >= 0 ? make_double2(fx, fy) : make_double2(NAN, NAN);
        v.info = make_int2(single_c0, a1 - a0);
        v.pad = make_int2(0, 0);
        rec[row] = v;
    };
    int longest = c1 - c0;
    for (int o = 32; o > 0; o >>= 1) longest = max(longest, __shfl_xor(longest, o, 64));
    if (longest <= 16) {
        const int j = lane >> 2, sub = lane & 3;
        const int32_t a0 = __shfl(c0, j, 64), a1 = __shfl(c1, j, 64), aj = __shfl(ci, j, 64);
        double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
        double xv[4], yv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int32_t c = a0 + sub + 4 * u;
            const bool ok = c < a1;
            xv[u] = ok ? x[c] : INFINITY;
            yv[u] = ok ? y[c] : INFINITY;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (a0 + sub + 4 * u < a1) {
                x0 = fmin(x0, xv[u]); y0 = fmin(y0, yv[u]);
                x1 = fmax(x1, xv[u]); y1 = fmax(y1, yv[u]);
            }
        }
        for (int o = 2; o > 0; o >>= 1) {
            x0 = fmin(x0, __shfl_xor(x0, o, 4));
            y0 = fmin(y0, __shfl_xor(y0, o, 4));
            x1 = fmax(x1, __shfl_xor(x1, o, 4));
            y1 = fmax(y1, __shfl_xor(y1, o, 4));
        }
        const int64_t row = r0 + j;
        if (sub == 0 && row < n_rows) {
            if (a1 <= a0 || !(x0 <= x1) || !(y0 <= y1)) x0 = y0 = x1 = y1 = NAN;
            out[row] = make_double4(x0, y0, x1, y1);
            if (rec) write_rec(row, make_double4(x0, y0, x1, y1), a0, a1, aj, xv[0], yv[0]);  // xv[0] = x[a0]
        }
        return;
    }
    // sixteen lanes per row, four rows at a time, 64 coordinates per pass
    const int g = lane >> 4, sub = lane & 15;
    for (int j0 = 0; j0 < kBoundsRows; j0 += 4) {
        const int j = j0 + g;
        const int64_t row = r0 + j;
        const int32_t a0 = __shfl(c0, j, 64), a1 = __shfl(c1, j, 64), aj = __shfl(ci, j, 64);
        double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY, fx = NAN, fy = NAN;
        for (int32_t cb = a0; cb < a1; cb += 64) {
            double xv[4], yv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int32_t c = cb + sub + 16 * u;
                const bool ok = c < a1;
                xv[u] = ok ? x[c] : INFINITY;
                yv[u] = ok ? y[c] : INFINITY;
            }
            if (cb == a0) {  // sub 0 holds x[a0] in its first pass
                fx = xv[0];
                fy = yv[0];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (cb + sub + 16 * u < a1) {
                    x0 = fmin(x0, xv[u]); y0 = fmin(y0, yv[u]);
                    x1 = fmax(x1, xv[u]); y1 = fmax(y1, yv[u]);
                }
            }
        }
        for (int o = 8; o > 0; o >>= 1) {
            x0 = fmin(x0, __shfl_xor(x0, o, 16));
            y0 = fmin(y0, __shfl_xor(y0, o, 16));
            x1 = fmax(x1, __shfl_xor(x1, o, 16));
            y1 = fmax(y1, __shfl_xor(y1, o, 16));
        }
        if (sub == 0 && row < n_rows) {
            if (a1 <= a0 || !(x0 <= x1) || !(y0 <= y1)) x0 = y0 = x1 = y1 = NAN;
            out[row] = make_double4(x0, y0, x1, y1);
            if (rec) write_rec(row, make_double4(x0, y0, x1, y1), a0, a1, aj, fx, fy);
        }
    }
}

// Extent of box centres (or of points when boxes are degenerate): partial
// reduction per block, then one block folds the partials.
__global__ __launch_bounds__(256) void centre_extent_partial_kernel(const double4* __restrict__ b,
                                                                   int64_t n,
                                                                   double4* __restrict__ part) {
    __shared__ double4 red[4];
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double4 v = b[i];
        if (v.x == v.x) {
            const double cx = 0.5 * (v.x + v.z), cy = 0.5 * (v.y + v.w);
            x0 = fmin(x0, cx); y0 = fmin(y0, cy);
            x1 = fmax(x1, cx); y1 = fmax(y1, cy);
        }
    }
    x0 = wave_min(x0); y0 = wave_min(y0); x1 = wave_max(x1); y1 = wave_max(y1);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = make_double4(x0, y0, x1, y1);
    __syncthreads();
    if (threadIdx.x == 0) {
        double4 r = red[0];
        for (int k = 1; k < 4; ++k) {
            r.x = fmin(r.x, red[k].x); r.y = fmin(r.y, red[k].y);
            r.z = fmax(r.z, red[k].z); r.w = fmax(r.w, red[k].w);
        }
        part[blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(256) void point_extent_partial_kernel(const double* __restrict__ px,
                                                                  const double* __restrict__ py,
                                                                  int64_t n,
                                                                  double4* __restrict__ part) {
    __shared__ double4 red[4];
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    auto fold = [&](double a, double c) {
        if (a == a && c == c) {
            x0 = fmin(x0, a); y0 = fmin(y0, c);
            x1 = fmax(x1, a); y1 = fmax(y1, c);
        }
    };
    const int64_t stride = (int64_t)gridDim.x * 256;
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {  // four points' loads in flight per thread
        double a[4], c[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            a[k] = px[i + k * stride];
            c[k] = py[i + k * stride];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) fold(a[k], c[k]);
    }
    for (; i < n; i += stride) fold(px[i], py[i]);
    x0 = wave_min(x0); y0 = wave_min(y0); x1 = wave_max(x1); y1 = wave_max(y1);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = make_double4(x0, y0, x1, y1);
    __syncthreads();
    if (threadIdx.x == 0) {
        double4 r = red[0];
        for (int k = 1; k < 4; ++k) {
            r.x = fmin(r.x, red[k].x); r.y = fmin(r.y, red[k].y);
            r.z = fmax(r.z, red[k].z); r.w = fmax(r.w, red[k].w);
        }
        part[blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(256) void extent_final_kernel(const double4* __restrict__ part, int nb,
                                                          double4* __restrict__ ext) {
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    for (int i = threadIdx.x; i < nb; i += 256) {
        const double4 v = part[i];
        x0 = fmin(x0, v.x); y0 = fmin(y0, v.y); x1 = fmax(x1, v.z); y1 = fmax(y1, v.w);
    }
    __shared__ double4 red[4];
    x0 = wave_min(x0); y0 = wave_min(y0); x1 = wave_max(x1); y1 = wave_max(y1);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = make_double4(x0, y0, x1, y1);
    __syncthreads();
    if (threadIdx.x == 0) {
        double4 r = red[0];
        for (int k = 1; k < 4; ++k) {
            r.x = fmin(r.x, red[k].x); r.y = fmin(r.y, red[k].y);
            r.z = fmax(r.z, red[k].z); r.w = fmax(r.w, red[k].w);
        }
        ext[0] = r;
    }
}

constexpr int kExtentBlocks = 2048;  // 8 blocks per CU

int64_t extent_workspace_bytes() { return sizeof(double4) * (kExtentBlocks + 1) + 256; }

int centre_extent(const double4* b, int64_t n, double4* ext, void* ws, hipStream_t s) {
    Carver c(ws);
    double4* part = c.take<double4>(kExtentBlocks);
    const int nb = (int)std::min<int64_t>(kExtentBlocks, std::max<int64_t>(1, ceil_div(n, 256)));
    hipLaunchKernelGGL(centre_extent_partial_kernel, dim3(nb), dim3(256), 0, s, b, n, part);
    VS_LAUNCH_CHECK();
    hipLaunchKernelGGL(extent_final_kernel, dim3(1), dim3(256), 0, s, part, nb, ext);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int point_extent(const double* px, const double* py, int64_t n, double4* ext, void* ws, hipStream_t s) {
    Carver c(ws);
    double4* part = c.take<double4>(kExtentBlocks);
    const int nb = (int)std::min<int64_t>(kExtentBlocks, std::max<int64_t>(1, ceil_div(n, 256)));
    hipLaunchKernelGGL(point_extent_partial_kernel, dim3(nb), dim3(256), 0, s, px, py, n, part);
    VS_LAUNCH_CHECK();
    hipLaunchKernelGGL(extent_final_kernel, dim3(1), dim3(256), 0, s, part, nb, ext);
    VS_LAUNCH_CHECK();
    return VS_OK;
}


// Key of bits*2 bits: the Hilbert index of the box centre on a 2^bits grid over
// the extent.  NaN boxes -> all ones (sort last).
__global__ __launch_bounds__(256) void hilbert_keys_boxes_kernel(const double4* __restrict__ b,
                                                                int64_t n,
                                                                const double4* __restrict__ ext,
                                                                int bits,
                                                                uint64_t* __restrict__ keys,
                                                                int32_t* __restrict__ ids) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const KeyScale ks = key_scale(ext, bits);
    const double4 v = b[i];
    uint64_t key = ~0ull;
    if (v.x == v.x) {
        const uint32_t qx = quantize(0.5 * (v.x + v.z), ks.x0, ks.sx, bits) << (32 - bits);
        const uint32_t qy = quantize(0.5 * (v.y + v.w), ks.y0, ks.sy, bits) << (32 - bits);
        key = hilbert64(qx, qy) >> (64 - 2 * bits);
    }
    keys[i] = key;
    ids[i] = (int32_t)i;
}

// Keys of 2*bits bits (right-aligned); the low `zero_bits` bits are cleared, so
// the key is the Hilbert index of a coarser grid, left-aligned.
template <class K>
__global__ __launch_bounds__(256) void hilbert_keys_points_kernel(const double* __restrict__ px,
                                                                 const double* __restrict__ py,
                                                                 int64_t n,
                                                                 const double4* __restrict__ ext,
                                                                 int bits, int zero_bits, K* __restrict__ keys,
                                                                 int32_t* __restrict__ ids,
                                                                 double2* __restrict__ xy) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const KeyScale ks = key_scale(ext, bits);
    const double a = px[i], c = py[i];
    if (xy) xy[i] = make_double2(a, c);  // interleaved copy for the gather (one line per point)
    uint64_t key = ~0ull;
    if (a == a && c == c) {
        const uint32_t qx = quantize(a, ks.x0, ks.sx, bits) << (32 - bits);
        const uint32_t qy = quantize(c, ks.y0, ks.sy, bits) << (32 - bits);
        key = hilbert64(qx, qy) >> (64 - 2 * bits);
    }
    if (zero_bits > 0) key = (key >> zero_bits) << zero_bits;
    keys[i] = (K)key;
    if (ids) ids[i] = (int32_t)i;  // null: the sort makes the indices itself
}

// Items level: boxes gathered into Hilbert order.
__global__ __launch_bounds__(256) void gather_boxes_kernel(const double4* __restrict__ b,
                                                          const int32_t* __restrict__ order, int64_t n,
                                                          double4* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = b[order[i]];
}

// Parent boxes = union of up to F child boxes (NaN children ignored; a parent
// with no valid child is NaN).
__global__ __launch_bounds__(256) void tree_level_kernel(const double4* __restrict__ child,
                                                        int64_t n_child, double4* __restrict__ parent,
                                                        int64_t n_parent, int fan) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n_parent) return;
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    const int64_t k1 = std::min<int64_t>((p + 1) * fan, n_child);
    for (int64_t k = p * fan; k < k1; ++k) {
        const double4 v = child[k];
        if (v.x == v.x) {
            x0 = fmin(x0, v.x); y0 = fmin(y0, v.y);
            x1 = fmax(x1, v.z); y1 = fmax(y1, v.w);
        }
    }
    if (!(x0 <= x1)) x0 = y0 = x1 = y1 = NAN;
    parent[p] = make_double4(x0, y0, x1, y1);
}

// Point boxes (degenerate), NaN for null/empty points.
__global__ __launch_bounds__(256) void point_boxes_kernel(const double* __restrict__ px,
                                                         const double* __restrict__ py, int64_t n,
                                                         double4* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double a = px[i], c = py[i];
    if (!(a == a && c == c)) a = c = NAN;
    out[i] = make_double4(a, c, a, c);
}

int polygon_bounds(const double* x, const double* y, const int32_t* geom, const int32_t* part,
                   const int32_t* ring, int64_t n, double* bounds, hipStream_t s) {
    if (n == 0) return VS_OK;
    hipLaunchKernelGGL(polygon_bounds_kernel, dim3((unsigned)ceil_div(n, 4 * kBoundsRows)), dim3(256), 0, s, x, y,
                       geom, part, ring, n, reinterpret_cast<double4*>(bounds), static_cast<int2*>(nullptr),
                       static_cast<RowRec*>(nullptr));
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int polygon_bounds_info(const double* x, const double* y, const int32_t* geom, const int32_t* part,
                        const int32_t* ring, int64_t n, double* bounds, int2* info, RowRec* rec, hipStream_t s) {
    if (n == 0) return VS_OK;
    hipLaunchKernelGGL(polygon_bounds_kernel, dim3((unsigned)ceil_div(n, 4 * kBoundsRows)), dim3(256), 0, s, x, y,
                       geom, part, ring, n, reinterpret_cast<double4*>(bounds), info, rec);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int point_boxes(const double* px, const double* py, int64_t n, double* boxes, hipStream_t s) {
    if (n == 0) return VS_OK;
    hipLaunchKernelGGL(point_boxes_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, px, py, n,
                       reinterpret_cast<double4*>(boxes));
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int64_t index_workspace_bytes(int64_t n) {
    Carver c(nullptr);
    c.take<uint64_t>(n);
    c.take<double4>(1);
    c.take<char>(extent_workspace_bytes());
    c.take<char>(sort_workspace_bytes_u64(n));
    return c.bytes;
}

int index_build(const double* bounds, int64_t n, const vs_tree_layout* L, double* node_boxes,
                int32_t* item_ids, uint64_t* keys_out, void* ws, int64_t ws_bytes, hipStream_t s) {
    VS_REQUIRE(L && L->n_items == n && L->n_levels >= 1 && L->n_levels <= VS_MAX_LEVELS,
               "vs_index_build: layout does not describe %lld items", (long long)n);
    VS_REQUIRE(ws_bytes >= index_workspace_bytes(n), "vs_index_build: workspace too small");
    if (n == 0) return VS_OK;
    Carver c(ws);
    uint64_t* keys = keys_out ? keys_out : c.take<uint64_t>(n);
    if (keys_out) c.take<uint64_t>(n);
    double4* ext = c.take<double4>(1);
    void* ext_ws = c.take<char>(extent_workspace_bytes());
    void* sort_ws = c.take<char>(sort_workspace_bytes_u64(n));
    const double4* b = reinterpret_cast<const double4*>(bounds);
    VS_TRY(centre_extent(b, n, ext, ext_ws, s));
    hipLaunchKernelGGL(hilbert_keys_boxes_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, b,
                       n, ext, 32, keys, item_ids);
    VS_LAUNCH_CHECK();
    VS_TRY(sort_pairs_u64(keys, item_ids, n, 0, 64, sort_ws, sort_workspace_bytes_u64(n), s));
    double4* nb = reinterpret_cast<double4*>(node_boxes);
    const int last = L->n_levels - 1;
    hipLaunchKernelGGL(gather_boxes_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, b,
                       item_ids, n, nb + L->level_start[last]);
    VS_LAUNCH_CHECK();
    for (int l = last - 1; l >= 0; --l) {
        const int64_t np = L->level_count[l];
        hipLaunchKernelGGL(tree_level_kernel, dim3((unsigned)ceil_div(np, 256)), dim3(256), 0, s,
                           nb + L->level_start[l + 1], L->level_count[l + 1], nb + L->level_start[l],
                           np, L->fanout);
        VS_LAUNCH_CHECK();
    }
    return VS_OK;
}

// Query-box ordering for coherent traversal: 16-level Hilbert keys of the box
// centres over their extent, sorted (NaN boxes last); writes the permutation.
int64_t box_order_workspace_bytes(int64_t n) {
    Carver c(nullptr);
    c.take<uint64_t>(n);
    c.take<uint64_t>(n);
    c.take<int32_t>(n);
    c.take<double4>(1);
    c.take<char>(extent_workspace_bytes());
    c.take<char>(sort_workspace_bytes_u64(n));
    return c.bytes;
}

int box_order(const double4* b, int64_t n, int32_t* perm, void* ws, hipStream_t s) {
    if (n == 0) return VS_OK;
    Carver c(ws);
    uint64_t* keys = c.take<uint64_t>(n);
    uint64_t* keys_in = c.take<uint64_t>(n);
    int32_t* ids_in = c.take<int32_t>(n);
    double4* ext = c.take<double4>(1);
    void* ext_ws = c.take<char>(extent_workspace_bytes());
    void* sort_ws = c.take<char>(sort_workspace_bytes_u64(n));
    VS_TRY(centre_extent(b, n, ext, ext_ws, s));
    hipLaunchKernelGGL(hilbert_keys_boxes_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, b, n, ext, 16,
                       keys_in, ids_in);
    VS_LAUNCH_CHECK();
    // NaN boxes carry all-ones keys: sort 33 bits so they stay after every code
    VS_TRY(sort_pairs_into_u64(keys_in, ids_in, keys, perm, n, 0, 40, sort_ws, sort_workspace_bytes_u64(n), s));
    return VS_OK;
}

// Point ordering for coherent traversal: 16-level (32-bit) Hilbert keys over
// the point extent, sorted; writes the permutation and the gathered coords.
// Four points per thread, every load issued before the first store, so each
// thread keeps eight random reads in flight; blocks stride over the tiles of
// 1024 points, so the grid can be capped to leave compute units free for
// work on another stream.
constexpr int kGatherItems = 4;
constexpr int kGatherTile = 256 * kGatherItems;
__global__ __launch_bounds__(256) void gather_points_kernel(const double* __restrict__ px,
                                                           const double* __restrict__ py,
                                                           const int32_t* __restrict__ perm,
                                                           int64_t n, double* __restrict__ sx,
                                                           double* __restrict__ sy) {
    for (int64_t t0 = (int64_t)blockIdx.x * kGatherTile; t0 < n; t0 += (int64_t)gridDim.x * kGatherTile) {
        const int64_t base = t0 + threadIdx.x;
        int32_t j[kGatherItems];
        double x[kGatherItems], y[kGatherItems];
#pragma unroll
        for (int b = 0; b < kGatherItems; ++b) {
            const int64_t i = base + 256 * b;
            j[b] = i < n ? perm[i] : -1;
        }
#pragma unroll
        for (int b = 0; b < kGatherItems; ++b) {
            x[b] = j[b] >= 0 ? px[j[b]] : 0.0;
            y[b] = j[b] >= 0 ? py[j[b]] : 0.0;
        }
#pragma unroll
        for (int b = 0; b < kGatherItems; ++b) {
            const int64_t i = base + 256 * b;
            if (i < n) {
                sx[i] = x[b];
                sy[i] = y[b];
            }
        }
    }
}

int64_t point_order_workspace_bytes(int64_t n) {
    Carver c(nullptr);
    c.take<uint32_t>(n);
    c.take<uint32_t>(n);
    c.take<double4>(1);
    c.take<char>(extent_workspace_bytes());
    c.take<char>(sort_workspace_bytes_u32(n));
    c.take<double2>(n);
    return c.bytes;
}

// sx[i], sy[i] = xy[perm[i]]: the points interleaved by the key pass, so each
// random read is one 16-byte load within one cache line (two separate x / y
// arrays cost two lines per point).
__global__ __launch_bounds__(256) void gather_xy_kernel(const double2* __restrict__ xy,
                                                       const int32_t* __restrict__ perm, int64_t n,
                                                       double* __restrict__ sx, double* __restrict__ sy) {
    for (int64_t t0 = (int64_t)blockIdx.x * kGatherTile; t0 < n; t0 += (int64_t)gridDim.x * kGatherTile) {
        const int64_t base = t0 + threadIdx.x;
        int32_t j[kGatherItems];
        double2 v[kGatherItems];
#pragma unroll
        for (int b = 0; b < kGatherItems; ++b) {
            const int64_t i = base + 256 * b;
            j[b] = i < n ? perm[i] : -1;
        }
#pragma unroll
        for (int b = 0; b < kGatherItems; ++b) v[b] = j[b] >= 0 ? xy[j[b]] : make_double2(0.0, 0.0);
#pragma unroll
        for (int b = 0; b < kGatherItems; ++b) {
            const int64_t i = base + 256 * b;
            if (i < n) {
                sx[i] = v[b].x;
                sy[i] = v[b].y;
            }
        }
    }
}

int gather_points_xy(const double2* xy, const int32_t* perm, int64_t n, double* sx, double* sy, hipStream_t s,
                     int blocks_per_cu) {
    if (n == 0) return VS_OK;
    hipLaunchKernelGGL(gather_xy_kernel, dim3(stride_grid(n, kGatherTile, blocks_per_cu)), dim3(256), 0, s, xy, perm,
                       n, sx, sy);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

// Point index for the learned form: keys of `bits` per axis (K holds 2*bits),
// sorted, with the permutation, gathered coordinates and the extent the keys
// were quantised over (device double4).
template <class K>
int64_t point_index_workspace_bytes_t(int64_t n) {
    Carver c(nullptr);
    c.take<char>(extent_workspace_bytes());
    c.take<char>(sizeof(K) == 4 ? sort_workspace_bytes_u32(n) : sort_workspace_bytes_u64(n));
    c.take<K>(n);        // unsorted keys (the sort makes the ids)
    c.take<double2>(n);  // interleaved points for the gather
    return c.bytes;
}

// The interleaved points point_index_build_t leaves in its workspace.
template <class K>
const double2* point_index_xy_t(void* ws, int64_t n) {
    Carver c(ws);
    c.take<char>(extent_workspace_bytes());
    c.take<char>(sizeof(K) == 4 ? sort_workspace_bytes_u32(n) : sort_workspace_bytes_u64(n));
    c.take<K>(n);
    return c.take<double2>(n);
}

int gather_points(const double* px, const double* py, const int32_t* perm, int64_t n, double* sx, double* sy,
                  hipStream_t s, int blocks_per_cu) {
    if (n == 0) return VS_OK;
    hipLaunchKernelGGL(gather_points_kernel, dim3(stride_grid(n, kGatherTile, blocks_per_cu)), dim3(256), 0, s, px,
                       py, perm, n, sx, sy);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

// Keys, sort and (when sx is not null) the gathered coordinates.
template <class K>
int point_index_build_t(const double* px, const double* py, int64_t n, K* keys, int32_t* perm, double* sx,
                        double* sy, double4* ext, void* ws, hipStream_t s) {
    Carver c(ws);
    void* ext_ws = c.take<char>(extent_workspace_bytes());
    const int64_t sb = sizeof(K) == 4 ? sort_workspace_bytes_u32(n) : sort_workspace_bytes_u64(n);
    void* sort_ws = c.take<char>(sb);
    K* keys_in = c.take<K>(n);
    double2* xy = c.take<double2>(n);
    VS_TRY(point_extent(px, py, n, ext, ext_ws, s));
    if (n == 0) return VS_OK;
    const int bits = 4 * (int)sizeof(K);
    const int zero_bits = 2 * (bits - point_index_grid_bits<K>());
    hipLaunchKernelGGL(hilbert_keys_points_kernel<K>, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, px, py, n,
                       ext, bits, zero_bits, keys_in, nullptr, xy);
    VS_LAUNCH_CHECK();
    // no id array: the first radix pass takes each key's index as its value
    if (sizeof(K) == 4)
        VS_TRY(sort_pairs_into_u32(reinterpret_cast<const uint32_t*>(keys_in), nullptr,
                                   reinterpret_cast<uint32_t*>(keys), perm, n, zero_bits, 32, sort_ws, sb, s));
    else
        VS_TRY(sort_pairs_into_u64(reinterpret_cast<const uint64_t*>(keys_in), nullptr,
                                   reinterpret_cast<uint64_t*>(keys), perm, n, zero_bits, 64, sort_ws, sb, s));
    if (sx) VS_TRY(gather_points_xy(xy, perm, n, sx, sy, s, 8));
    return VS_OK;
}

int64_t point_index_workspace_bytes_u32(int64_t n) { return point_index_workspace_bytes_t<uint32_t>(n); }
const double2* point_index_xy_u32(void* ws, int64_t n) { return point_index_xy_t<uint32_t>(ws, n); }
const double2* point_index_xy_u64(void* ws, int64_t n) { return point_index_xy_t<uint64_t>(ws, n); }
int64_t point_index_workspace_bytes_u64(int64_t n) { return point_index_workspace_bytes_t<uint64_t>(n); }
int point_index_build_u32(const double* px, const double* py, int64_t n, uint32_t* keys, int32_t* perm, double* sx,
                          double* sy, double4* ext, void* ws, hipStream_t s) {
    return point_index_build_t<uint32_t>(px, py, n, keys, perm, sx, sy, ext, ws, s);
}
int point_index_build_u64(const double* px, const double* py, int64_t n, uint64_t* keys, int32_t* perm, double* sx,
                          double* sy, double4* ext, void* ws, hipStream_t s) {
    return point_index_build_t<uint64_t>(px, py, n, keys, perm, sx, sy, ext, ws, s);
}

int point_order(const double* px, const double* py, int64_t n, int bits, int32_t* perm, double* sx,
                double* sy, void* ws, hipStream_t s) {
    if (n == 0) return VS_OK;
    Carver c(ws);
    uint32_t* keys = c.take<uint32_t>(n);
    uint32_t* keys_in = c.take<uint32_t>(n);
    double4* ext = c.take<double4>(1);
    void* ext_ws = c.take<char>(extent_workspace_bytes());
    void* sort_ws = c.take<char>(sort_workspace_bytes_u32(n));
    double2* xy = c.take<double2>(n);
    VS_TRY(point_extent(px, py, n, ext, ext_ws, s));
    hipLaunchKernelGGL(hilbert_keys_points_kernel<uint32_t>, dim3((unsigned)ceil_div(n, 256)), dim3(256),
                       0, s, px, py, n, ext, 16, 0, keys_in, nullptr, xy);
    VS_LAUNCH_CHECK();
    VS_TRY(sort_pairs_into_u32(keys_in, nullptr, keys, perm, n, 32 - 2 * bits, 32, sort_ws,
                               sort_workspace_bytes_u32(n), s));
    hipLaunchKernelGGL(gather_xy_kernel, dim3((unsigned)ceil_div(n, kGatherTile)), dim3(256), 0, s, xy, perm, n,
                       sx, sy);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

}  // namespace vs

extern "C" {

int vs_tree_plan(int64_t n_items, int32_t fanout, vs_tree_layout* out) {
    vs::clear_error();
    VS_REQUIRE(out != nullptr, "vs_tree_plan: null layout");
    VS_REQUIRE(fanout >= 2 && fanout <= 64, "vs_tree_plan: fanout %d not in [2, 64]", fanout);
    VS_REQUIRE(n_items >= 0 && n_items < (int64_t)INT32_MAX, "vs_tree_plan: n_items out of range");
    memset(out, 0, sizeof(*out));
    int64_t counts[VS_MAX_LEVELS];
    int nl = 0;
    int64_t c = n_items;
    counts[nl++] = c;
    while (c > 1) {
        c = (c + fanout - 1) / fanout;
        VS_REQUIRE(nl < VS_MAX_LEVELS, "vs_tree_plan: too many levels");
        counts[nl++] = c;
    }
    out->fanout = fanout;
    out->n_items = n_items;
    out->n_levels = nl;
    int64_t total = 0;
    for (int l = 0; l < nl; ++l) {
        out->level_count[l] = counts[nl - 1 - l];
        out->level_start[l] = total;
        total += out->level_count[l];
    }
    out->n_nodes = total;
    return VS_OK;
}

int vs_polygon_bounds(const double* x, const double* y, const int32_t* geom_offsets,
                      const int32_t* part_offsets, const int32_t* ring_offsets, int64_t n_rows,
                      double* bounds, void* stream) {
    vs::clear_error();
    VS_REQUIRE(n_rows >= 0, "vs_polygon_bounds: negative n_rows");
    return vs::polygon_bounds(x, y, geom_offsets, part_offsets, ring_offsets, n_rows, bounds,
                              vs::stream_of(stream));
}

int vs_point_boxes(const double* px, const double* py, int64_t n, double* boxes, void* stream) {
    vs::clear_error();
    VS_REQUIRE(n >= 0, "vs_point_boxes: negative n");
    return vs::point_boxes(px, py, n, boxes, vs::stream_of(stream));
}

int64_t vs_index_workspace_bytes(int64_t n_items) { return vs::index_workspace_bytes(n_items); }

int vs_index_build(const double* bounds, int64_t n_items, const vs_tree_layout* layout,
                   double* node_boxes, int32_t* item_ids, uint64_t* item_keys, void* workspace,
                   int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    return vs::index_build(bounds, n_items, layout, node_boxes, item_ids, item_keys, workspace,
                           workspace_bytes, vs::stream_of(stream));
}

int64_t vs_point_order_workspace_bytes(int64_t n_points) {
    return vs::point_order_workspace_bytes(n_points);
}

int vs_point_order(const double* px, const double* py, int64_t n_points, int32_t bits_per_axis,
                   int32_t* perm, double* sorted_x, double* sorted_y, void* workspace,
                   int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    VS_REQUIRE(bits_per_axis >= 1 && bits_per_axis <= 16, "vs_point_order: bits_per_axis %d not in [1,16]",
               bits_per_axis);
    VS_REQUIRE(n_points >= 0 && n_points < (int64_t)INT32_MAX, "vs_point_order: n out of range");
    VS_REQUIRE(workspace_bytes >= vs::point_order_workspace_bytes(n_points),
               "vs_point_order: workspace too small");
    return vs::point_order(px, py, n_points, bits_per_axis, perm, sorted_x, sorted_y, workspace,
                           vs::stream_of(stream));
}

uint64_t vs_host_hilbert64(uint32_t x, uint32_t y) { return vs::hilbert64(x, y); }

}  // extern "C"
