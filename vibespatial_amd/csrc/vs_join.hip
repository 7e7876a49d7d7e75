// Spatial join: packed-tree traversal, exact refine, pair compaction.
//
// Replaces, on the sjoin path:
//   spatial_index_device_query + Morton range count/scatter
//       src/vibespatial/spatial/spatial_index_device.py:228-390, 761-1034
//       src/vibespatial/kernels/core/spatial_query_source.py:1132-1253
//   classify_point_predicates_indexed_device / point_in_polygon_prepared_part_y_index
//       src/vibespatial/predicates/point_relations.py:1040-1191
//       src/vibespatial/predicates/point_location_index_kernels.py:157-315
//   compute_polygonal_intersects_gpu / polygon_polygon_intersects_from_owned
//       src/vibespatial/predicates/polygon.py:849-996, polygon_kernels.py:1150-1197
//   compaction d_left[keep] + count_scatter_total + lexsort
//       src/vibespatial/spatial/query_utils.py:867-868, cuda/_runtime.py:2015-2038,
//       src/vibespatial/spatial/query.py:143-153
//
// Point x polygon (the hot path) is one persistent kernel:
//   * queries are points sorted by 16-level Hilbert key, 64 per wave;
//   * each wave walks the packed tree depth first with one shared stack: a
//     group of up to F sibling boxes is loaded by lanes 0..F-1 (the top levels
//     come from LDS), tested against the union box of the wave's points and the
//     overlapping ones are pushed;
//   * at an item (polygon) the wave streams the polygon's coordinates once,
//     64 at a time with coalesced loads, broadcasts each edge with readlane and
//     every lane whose point is in the item's box runs the exact crossing /
//     boundary step on it (fp64 with the stage-A filter, exact expansion
//     fallback);
//   * refined pairs are buffered per wave in LDS and appended to the stage
//     with one atomic per full buffer.
// The finish pass (vs_finish.hip: radix sort by point id, sort of each
// point's polygon ids) produces pairs lexsorted by (point, polygon) as int64.
// Waves are handed out from 8 queues, one per XCD, each over a contiguous
// range of the Hilbert order, so an XCD's L2 sees a spatially compact slice;
// an idle XCD steals from the others.  Every wave exits when all 8 are empty.
#include <mutex>

#include "vs_exact.h"
#include "vs_internal.h"
#include "vs_join_common.h"

namespace vs {

int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, long long* d_total, void* ws,
                       hipStream_t s);
int64_t scan_workspace_bytes(int64_t n);
int64_t point_order_workspace_bytes(int64_t n);
int64_t box_order_workspace_bytes(int64_t n);
int box_order(const double4* b, int64_t n, int32_t* perm, void* ws, hipStream_t s);
int point_order(const double* px, const double* py, int64_t n, int bits, int32_t* perm, double* sx,
                double* sy, void* ws, hipStream_t s);

constexpr int kJoinThreads = 256;
constexpr int kJoinWaves = kJoinThreads / 64;
constexpr int kStackMax = 256;     // per-wave traversal stack (point queries)
constexpr int kStackMaxBox = 512;  // box queries: one box per wave, so a 64-wide tree fits
constexpr int kTopMax = 640;  // 20 KB of fp64 node boxes staged per block

struct TreeView {
    const double4* box;
    const int32_t* item_ids;
    int64_t level_start[VS_MAX_LEVELS];
    int64_t level_count[VS_MAX_LEVELS];
    int32_t n_levels;
    int32_t fanout;
    int64_t n_top;
};

// Refine layout of the polygons: coordinates in their original order, one
// flag byte per coordinate (bit0: first coordinate of a ring, so no edge ends
// there; bit1: first coordinate of a part, where the even-odd parity of the
// previous part is settled) and, per item slot of the tree (Hilbert order), the
// row's coordinate range.  One load of item_range replaces the
// geom -> part -> ring offset chain.
struct PolyView {
    const double* x;
    const double* y;
    const uint8_t* flags;
    const int2* item_range;
};

constexpr int kOutBuf = 256;  // per-wave LDS buffer of refined pairs
constexpr int kBoxClaim = 4;    // one-box mode: query boxes per queue atomic
constexpr int kOneBoxHits = 32;  // expected hits per box from which box queries run one box per wave


enum { kModePip = 0, kModeBox = 1 };

// Clears the control block and, when given, n_extra further counters.
__global__ void ctl_reset_kernel(JoinCtl* ctl, unsigned long long* extra, int n_extra) {
    if (threadIdx.x == 0) {
        ctl->staged = 0;
        ctl->n_long = 0;
        for (int i = 0; i < 8; ++i) ctl->queue[i] = 0;
        for (int i = 0; i < 4; ++i) ctl->stat[i] = 0;
        ctl->escalated_pairs = 0;
    }
    if (extra && (int)threadIdx.x < n_extra) extra[threadIdx.x] = 0;
}

__global__ void fill_i32_kernel(int32_t* __restrict__ p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

template <int PRED>
__device__ __forceinline__ bool pred_of_location(int loc) {
    if (PRED == VS_PRED_WITHIN || PRED == VS_PRED_CONTAINS) return loc == 2;
    if (PRED == VS_PRED_TOUCHES) return loc == 1;
    return loc != 0;  // intersects, covered_by, covers
}

// Wave-cooperative exact location of each active lane's point in the polygon
// whose coordinates are [c0, c1); returns the predicate result per lane.
// The wave loads 64 coordinates at a time (coalesced), forms the edge ending at
// each lane's coordinate, and keeps only edges that can change the answer of
// some active lane: their y-range meets the region [rx0,rx1]x[ry0,ry1] holding
// every active point and they are not entirely left of it (an edge left of p
// neither crosses p's rightward ray nor contains p).  The kept edges are
// broadcast one by one with readlane and every active lane runs the exact
// crossing / boundary step on them.
template <int PRED>
__device__ bool refine_pip_wave(double px, double py, bool active, int32_t c0, int32_t c1, double rx0,
                                double ry0, double ry1, const PolyView& P) {
    const int lane = lane_id();
    int loc = 0;
    bool done = !active;
    int parity = 0;
    double cx = 0.0, cy = 0.0;  // coordinate base-1 for the chunk's lane 0
    for (int32_t base = c0; base < c1; base += 64) {
        const int32_t c = base + lane;
        const bool inr = c < c1;
        const double bx = inr ? P.x[c] : 0.0;
        const double by = inr ? P.y[c] : 0.0;
        const int f = inr ? (int)P.flags[c] : 0;
        double ax = __shfl_up(bx, 1, 64);
        double ay = __shfl_up(by, 1, 64);
        if (lane == 0) {
            ax = cx;
            ay = cy;
        }
        const bool rel = inr && !(f & 1) && fmax(ay, by) >= ry0 && fmin(ay, by) <= ry1 && fmax(ax, bx) >= rx0;
        const uint64_t m = __ballot(rel);
        const uint64_t pm = __ballot(inr && (f & 2));
        cx = readlane_f64(bx, 63);
        cy = readlane_f64(by, 63);
        uint64_t work = m | pm;
        while (work) {
            const int s = __ffsll((unsigned long long)work) - 1;
            work &= work - 1;
            if ((pm >> s) & 1) {  // a new part starts: settle the previous one
                if (!done && parity) {
                    loc = 2;
                    done = true;
                }
                parity = 0;
            }
            if ((m >> s) & 1) {
                const double eax = readlane_f64(ax, s), eay = readlane_f64(ay, s);
                const double ebx = readlane_f64(bx, s), eby = readlane_f64(by, s);
                if (!done && edge_step_f64(px, py, eax, eay, ebx, eby, parity)) {
                    loc = 1;
                    done = true;
                }
            }
        }
        if (__ballot(!done) == 0) break;
    }
    if (!done && parity) loc = 2;
    return active && pred_of_location<PRED>(loc);
}

// Persistent traversal kernel.  MODE kModePip: queries are points (qa, qb =
// sorted x, y), 64 per wave, with fused refine.  MODE kModeBox: queries are
// boxes given as (qbox, perm) and every bbox pair is emitted; with one_box,
// one box per wave (claimed kBoxClaim at a time): the lanes test a node's
// children and a leaf group's hits are appended with one ballot (boxes with
// many hits); else 64 boxes per wave walk the union of their boxes (many
// boxes with few hits).
template <int MODE, int PRED>
__global__ __launch_bounds__(kJoinThreads) void traverse_kernel(
    const double* __restrict__ qx, const double* __restrict__ qy, const double4* __restrict__ qbox,
    const int32_t* __restrict__ perm, int64_t n, TreeView T, PolyView P, uint32_t* __restrict__ stage_l,
    int32_t* __restrict__ stage_r, int64_t capacity, JoinCtl* __restrict__ ctl, int one_box) {
    __shared__ double4 top[kTopMax];
    __shared__ long long stack_s[kJoinWaves][MODE == kModeBox ? kStackMaxBox : kStackMax];
    __shared__ int2 outbuf_s[kJoinWaves][kOutBuf];
    for (int i = threadIdx.x; i < T.n_top; i += kJoinThreads) top[i] = T.box[i];
    __syncthreads();

    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    long long* stack = stack_s[w];
    int2* outbuf = outbuf_s[w];
    int nbuf = 0;                          // wave-uniform fill of outbuf
    unsigned long long st_cand = 0, st_cverts = 0;
    // Appends the wave's pairs in outbuf to the stage with one atomic.
    auto flush = [&]() {
        if (nbuf == 0) return;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(&ctl->staged, (unsigned long long)nbuf);
        base = __shfl(base, 0, 64);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int i = lane; i < nbuf; i += 64) {
            const unsigned long long slot = base + (unsigned long long)i;
            if (slot < (unsigned long long)capacity) {
                const int2 e = outbuf[i];
                stage_l[slot] = (uint32_t)e.x;
                stage_r[slot] = e.y;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        nbuf = 0;
    };
    // work units: kBoxClaim boxes (one-box mode) or 64 queries
    const bool one = MODE == kModeBox && one_box;
    const int64_t W = one ? (n + kBoxClaim - 1) / kBoxClaim : (n + 63) >> 6;
    const int home = blockIdx.x & 7;
    const int last = T.n_levels - 1;
    const int F = T.fanout;
    const uint64_t lt = lanemask_lt();

    for (int xq = 0; xq < 8;) {
        const int q = (home + xq) & 7;
        const int64_t r0 = W * q / 8, r1 = W * (q + 1) / 8;
        unsigned int t = 0;
        if (lane == 0) t = atomicAdd(&ctl->queue[q], 1u);
        t = __shfl(t, 0, 64);
        if ((int64_t)t >= r1 - r0) {
            ++xq;
            continue;
        }
        if (one) {
            for (int bq = 0; bq < kBoxClaim; ++bq) {
                const int64_t qi = (r0 + t) * kBoxClaim + bq;  // wave-uniform
                if (qi >= n) break;
                const int32_t orig = perm ? perm[qi] : (int32_t)qi;
                const double4 qb = qbox[orig];
                if (!(qb.x == qb.x && qb.y == qb.y && qb.z == qb.z && qb.w == qb.w) || T.level_count[0] <= 0)
                    continue;
                int sp = 1;
                if (lane == 0) stack[0] = 0;  // group 0 of level 0 = the root
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                while (sp > 0) {
                    --sp;
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    const long long e = stack[sp];
                    const int l = (int)(e & 15);
                    const int64_t first = (e >> 4) * F;
                    const int64_t cl = T.level_count[l];
                    const int ng = (int)((cl - first) < F ? (cl - first) : F);
                    const int64_t g = T.level_start[l] + first + lane;
                    bool ov = false;
                    if (lane < ng) {
                        const double4 b = (g < T.n_top) ? top[g] : T.box[g];
                        ov = box_overlap(qb.x, qb.y, qb.z, qb.w, b.x, b.y, b.z, b.w);
                    }
                    uint64_t m = __ballot(ov);
                    if (l == last) {
                        if (m) {  // this leaf group's hits, one lane each
                            const int k = __popcll(m);
                            if (nbuf + k > kOutBuf) flush();
                            if (ov) outbuf[nbuf + __popcll(m & lt)] = make_int2(orig, T.item_ids[first + lane]);
                            nbuf += k;
                        }
                    } else {
                        while (m) {
                            const int s2 = 63 - __clzll((long long)m);
                            m &= ~(1ull << s2);
                            if (lane == 0) stack[sp] = ((first + s2) << 4) | (long long)(l + 1);
                            ++sp;
                        }
                        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    }
                }
            }
            continue;
        }
        const int64_t idx = (r0 + t) * 64 + lane;
        const bool inr = idx < n;
        double x0, y0, x1, y1;
        int32_t orig = 0;
        if (MODE == kModePip) {
            x0 = inr ? qx[idx] : NAN;
            y0 = inr ? qy[idx] : NAN;
            x1 = x0;
            y1 = y0;
            orig = inr ? perm[idx] : 0;
        } else {
            orig = inr ? (perm ? perm[idx] : (int32_t)idx) : 0;
            const double4 b = inr ? qbox[orig] : make_double4(NAN, NAN, NAN, NAN);
            x0 = b.x; y0 = b.y; x1 = b.z; y1 = b.w;
        }
        const bool valid = inr && x0 == x0 && y0 == y0 && x1 == x1 && y1 == y1;
        const double ux0 = wave_min(valid ? x0 : INFINITY), uy0 = wave_min(valid ? y0 : INFINITY);
        const double ux1 = wave_max(valid ? x1 : -INFINITY), uy1 = wave_max(valid ? y1 : -INFINITY);
        if (__ballot(valid) != 0 && T.level_count[0] > 0) {
            int sp = 1;
            if (lane == 0) stack[0] = 0;  // group 0 of level 0 = the root
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            while (sp > 0) {
                --sp;
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const long long e = stack[sp];
                const int l = (int)(e & 15);
                const int64_t first = (e >> 4) * F;
                const int64_t cl = T.level_count[l];
                const int ng = (int)((cl - first) < F ? (cl - first) : F);
                const int64_t g = T.level_start[l] + first + lane;
                double4 b = make_double4(NAN, NAN, NAN, NAN);
                bool ov = false;
                if (lane < ng) {
                    b = (g < T.n_top) ? top[g] : T.box[g];
                    ov = box_overlap(ux0, uy0, ux1, uy1, b.x, b.y, b.z, b.w);
                }
                uint64_t m = __ballot(ov);
                if (l == last) {
                    while (m) {
                        const int s = __ffsll((unsigned long long)m) - 1;
                        m &= m - 1;
                        const double bx0 = readlane_f64(b.x, s), by0 = readlane_f64(b.y, s);
                        const double bx1 = readlane_f64(b.z, s), by1 = readlane_f64(b.w, s);
                        const bool hit = valid && box_overlap(x0, y0, x1, y1, bx0, by0, bx1, by1);
                        const uint64_t hm = __ballot(hit);
                        if (!hm) continue;
                        const int32_t row = T.item_ids[first + s];
                        bool res;
                        if (MODE == kModePip) {
                            const int2 rng = P.item_range[first + s];
                            const unsigned long long h = (unsigned long long)__popcll(hm);
                            st_cand += h;
                            st_cverts += h * (unsigned long long)(rng.y - rng.x);
                            // every active point lies in (wave box) x (item box)
                            res = refine_pip_wave<PRED>(x0, y0, hit, rng.x, rng.y, fmax(ux0, bx0), fmax(uy0, by0),
                                                        fmin(uy1, by1), P);
                        } else {
                            res = hit;
                        }
                        const uint64_t rm = __ballot(res);
                        if (rm) {
                            const int k = __popcll(rm);
                            if (nbuf + k > kOutBuf) flush();
                            if (res) outbuf[nbuf + __popcll(rm & lt)] = make_int2(orig, row);
                            nbuf += k;
                        }
                    }
                } else {
                    while (m) {
                        const int s = 63 - __clzll((long long)m);
                        m &= ~(1ull << s);
                        if (lane == 0) stack[sp] = ((first + s) << 4) | (long long)(l + 1);
                        ++sp;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                }
            }
        }
    }
    flush();
    if (MODE == kModePip && lane == 0 && st_cand) {
        atomicAdd(&ctl->stat[0], st_cand);
        atomicAdd(&ctl->stat[1], st_cverts);
    }
}

// Refine-layout preparation (per join: O(V + R + P + M), tiny next to the join).
__global__ void fill_u8_kernel(uint8_t* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void ring_start_flags_kernel(const int32_t* __restrict__ ring, int64_t n_rings,
                                        uint8_t* __restrict__ flags) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r < n_rings && ring[r + 1] > ring[r]) flags[ring[r]] = 1;
}

__global__ void part_start_flags_kernel(const int32_t* __restrict__ part, const int32_t* __restrict__ ring,
                                        int64_t n_parts, uint8_t* __restrict__ flags) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p < n_parts && part[p + 1] > part[p]) {
        const int32_t c = ring[part[p]];
        if (ring[part[p + 1]] > c) flags[c] = 3;
    }
}

__global__ void item_range_kernel(const int32_t* __restrict__ item_ids, int64_t n_items,
                                  const int32_t* __restrict__ geom, const int32_t* __restrict__ part,
                                  const int32_t* __restrict__ ring, int2* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n_items) return;
    const int32_t row = item_ids[k];
    out[k] = make_int2(ring[part[geom[row]]], ring[part[geom[row + 1]]]);
}

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------
TreeView make_tree_view(const vs_tree_layout* L, const double* node_boxes, const int32_t* item_ids) {
    TreeView T;
    memset(&T, 0, sizeof(T));
    T.box = reinterpret_cast<const double4*>(node_boxes);
    T.item_ids = item_ids;
    T.n_levels = L->n_levels;
    T.fanout = L->fanout;
    for (int l = 0; l < L->n_levels; ++l) {
        T.level_start[l] = L->level_start[l];
        T.level_count[l] = L->level_count[l];
    }
    // stage whole levels (never the items level) while they fit
    int64_t top = 0;
    for (int l = 0; l + 1 < L->n_levels; ++l) {
        const int64_t end = L->level_start[l] + L->level_count[l];
        if (end <= kTopMax) top = end;
    }
    T.n_top = top;
    return T;
}

int check_layout(const vs_tree_layout* L, int stack_max) {
    VS_REQUIRE(L != nullptr, "null tree layout");
    VS_REQUIRE(L->n_levels >= 1 && L->n_levels <= VS_MAX_LEVELS, "tree layout: bad level count %d",
               L->n_levels);
    VS_REQUIRE(L->fanout >= 2 && L->fanout <= 64, "tree layout: bad fanout %d", L->fanout);
    // a depth-first walk keeps at most fanout - 1 siblings per level
    VS_REQUIRE((int64_t)L->n_levels * L->fanout <= stack_max, "tree layout: too deep for the stack");
    VS_REQUIRE(L->level_count[L->n_levels - 1] == L->n_items, "tree layout: item level mismatch");
    return VS_OK;
}

struct JoinWs {
    uint8_t* flags;
    int2* item_range;
    int32_t* perm;
    double* sx;
    double* sy;
    void* order_ws;
    JoinCtl* ctl;
    void* fin_ws;
    uint32_t* stage_l;
    int32_t* stage_r;
};

JoinWs carve_join_ws(void* base, int64_t nq, int64_t n_items, int64_t n_coords, int64_t capacity,
                     bool sorted_points, int64_t* bytes) {
    Carver c(base);
    JoinWs J;
    J.flags = sorted_points ? c.take<uint8_t>(n_coords) : nullptr;
    J.item_range = sorted_points ? c.take<int2>(n_items) : nullptr;
    J.perm = c.take<int32_t>(nq);
    if (sorted_points) {
        J.sx = c.take<double>(nq);
        J.sy = c.take<double>(nq);
        J.order_ws = c.take<char>(point_order_workspace_bytes(nq));
    } else {
        J.sx = J.sy = nullptr;
        J.order_ws = c.take<char>(box_order_workspace_bytes(nq));
    }
    J.ctl = c.take<JoinCtl>(1);
    J.fin_ws = c.take<char>(finish_workspace_bytes(capacity));
    J.stage_l = nullptr;
    J.stage_r = nullptr;
    if (base) finish_stage(J.fin_ws, capacity, &J.stage_l, &J.stage_r);
    if (bytes) *bytes = c.bytes;
    return J;
}

int grid_for(int64_t n_waves) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cus = v;
    }
    const int64_t want = ceil_div(n_waves, kJoinWaves);
    int64_t g = std::min<int64_t>(want, (int64_t)cus * 8);
    g = std::max<int64_t>(g, 8);
    return (int)((g + 7) / 8 * 8);
}

// Optional timing of the dominant kernel of the calling host thread's next
// join or query, with two HIP events recorded on the launch stream
// (instrumentation for benchmarks).  The state is per host thread, so
// threads time their own calls; the events are created by
// vs_set_kernel_timing(1) on the device current there, never inside a join.
struct KernelTiming {
    bool on = false;
    bool recorded = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
};
static thread_local KernelTiming t_timing;

int timing_begin(hipStream_t s) {
    if (!t_timing.on) return VS_OK;
    VS_CHECK_HIP(hipEventRecord(t_timing.ev[0], s));
    return VS_OK;
}

int timing_end(hipStream_t s) {
    if (!t_timing.on) return VS_OK;
    VS_CHECK_HIP(hipEventRecord(t_timing.ev[1], s));
    t_timing.recorded = true;
    return VS_OK;
}

int timing_begin_public(hipStream_t s) { return timing_begin(s); }
int timing_end_public(hipStream_t s) { return timing_end(s); }

template <int MODE>
int launch_traverse(int pred, const double* qx, const double* qy, const double4* qbox,
                    const int32_t* perm, int64_t n, const TreeView& T, const PolyView& P,
                    const JoinWs& J, int64_t capacity, hipStream_t s, bool one_box = false) {
    const int grid = grid_for(one_box ? ceil_div(n, kBoxClaim) : ceil_div(n, 64));
    VS_TRY(timing_begin(s));
#define VS_LAUNCH_T(PR)                                                                          \
    hipLaunchKernelGGL((traverse_kernel<MODE, PR>), dim3(grid), dim3(kJoinThreads), 0, s, qx, qy, \
                       qbox, perm, n, T, P, J.stage_l, J.stage_r, capacity, J.ctl, (int)one_box)
    switch (pred) {
        case VS_PRED_WITHIN: VS_LAUNCH_T(VS_PRED_WITHIN); break;
        case VS_PRED_INTERSECTS: VS_LAUNCH_T(VS_PRED_INTERSECTS); break;
        case VS_PRED_COVERED_BY: VS_LAUNCH_T(VS_PRED_COVERED_BY); break;
        case VS_PRED_TOUCHES: VS_LAUNCH_T(VS_PRED_TOUCHES); break;
        default: VS_REQUIRE(false, "unsupported predicate %d for this join", pred);
    }
#undef VS_LAUNCH_T
    VS_LAUNCH_CHECK();
    VS_TRY(timing_end(s));
    return VS_OK;
}

int build_refine_flags(const int32_t* part, const int32_t* ring, int64_t n_parts, int64_t n_rings, int64_t n_coords,
                       uint8_t* flags, hipStream_t s) {
    if (n_coords > 0) {
        hipLaunchKernelGGL(fill_u8_kernel, dim3((unsigned)ceil_div(n_coords, 256)), dim3(256), 0, s, flags, n_coords);
        VS_LAUNCH_CHECK();
    }
    if (n_rings > 0) {
        hipLaunchKernelGGL(ring_start_flags_kernel, dim3((unsigned)ceil_div(n_rings, 256)), dim3(256), 0, s, ring,
                           n_rings, flags);
        VS_LAUNCH_CHECK();
    }
    if (n_parts > 0) {
        hipLaunchKernelGGL(part_start_flags_kernel, dim3((unsigned)ceil_div(n_parts, 256)), dim3(256), 0, s, part,
                           ring, n_parts, flags);
        VS_LAUNCH_CHECK();
    }
    return VS_OK;
}

void ctl_reset(JoinCtl* ctl, hipStream_t s, unsigned long long* extra, int n_extra) {
    hipLaunchKernelGGL(ctl_reset_kernel, dim3(1), dim3(64), 0, s, ctl, extra, n_extra);
}

void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, p, n, v);
}

int timing_begin_public(hipStream_t s);
int timing_end_public(hipStream_t s);

}  // namespace vs

// Work memory of vs_sjoin_points_polygons_noalloc: a fixed array in device
// memory, one copy per GPU, held from module load to process end.  Host code
// never passes `scratch` itself to a kernel (on the host it names a stand-in,
// not the device address): scratch_ptr() looks its address up once per
// device, with that device current, and every later use goes through that
// pointer.
static __device__ __align__(256) unsigned char scratch[VS_NOALLOC_SCRATCH_BYTES];

namespace vs {

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch[kScratchDevices] = {};

int scratch_ptr(void** out) {
    int dev = 0;
    VS_CHECK_HIP(hipGetDevice(&dev));
    VS_REQUIRE(dev >= 0 && dev < kScratchDevices, "scratch: device %d beyond the %d supported", dev,
               kScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        VS_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        VS_REQUIRE(p != nullptr, "scratch: no device address on device %d", dev);
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return VS_OK;
}

}  // namespace vs

extern "C" {

int vs_set_kernel_timing(int32_t enable) {
    vs::clear_error();
    vs::KernelTiming& T = vs::t_timing;
    if (enable && !T.on) {
        VS_CHECK_HIP(hipEventCreate(&T.ev[0]));
        VS_CHECK_HIP(hipEventCreate(&T.ev[1]));
        T.on = true;
        T.recorded = false;
    } else if (!enable && T.on) {
        VS_CHECK_HIP(hipEventDestroy(T.ev[0]));
        VS_CHECK_HIP(hipEventDestroy(T.ev[1]));
        T = vs::KernelTiming();
    }
    return VS_OK;
}

int vs_last_kernel_ms(float* ms) {
    vs::clear_error();
    VS_REQUIRE(ms != nullptr, "vs_last_kernel_ms: null output");
    const vs::KernelTiming& T = vs::t_timing;
    VS_REQUIRE(T.on && T.recorded, "vs_last_kernel_ms: timing not enabled on this thread or nothing timed yet");
    VS_CHECK_HIP(hipEventSynchronize(T.ev[1]));
    VS_CHECK_HIP(hipEventElapsedTime(ms, T.ev[0], T.ev[1]));
    return VS_OK;
}

int64_t vs_sjoin_points_polygons_workspace_bytes(int64_t n_points, int64_t n_polys, int64_t n_coords,
                                                 int64_t capacity) {
    int64_t b = 0;
    vs::carve_join_ws(nullptr, n_points, n_polys, n_coords, capacity, true, &b);
    return b;
}

int vs_sjoin_points_polygons(const double* px, const double* py, int64_t n_points, const double* x,
                             const double* y, const int32_t* geom_offsets, const int32_t* part_offsets,
                             const int32_t* ring_offsets, int64_t n_polys, int64_t n_parts, int64_t n_rings,
                             int64_t n_coords, const vs_tree_layout* layout, const double* node_boxes,
                             const int32_t* item_ids, int32_t predicate, int64_t capacity, int64_t* out_left,
                             int64_t* out_right, int64_t* d_result, void* workspace, int64_t workspace_bytes,
                             void* stream) {
    vs::clear_error();
    VS_TRY(vs::check_layout(layout, vs::kStackMax));
    VS_REQUIRE(layout->n_items == n_polys, "vs_sjoin_points_polygons: layout has %lld items, %lld polygons",
               (long long)layout->n_items, (long long)n_polys);
    VS_REQUIRE(n_points >= 0 && n_points < (int64_t)INT32_MAX, "vs_sjoin_points_polygons: n_points out of range");
    VS_REQUIRE(n_parts >= 0 && n_rings >= 0 && n_coords >= 0 && n_coords < (int64_t)INT32_MAX,
               "vs_sjoin_points_polygons: bad part/ring/coordinate counts");
    VS_REQUIRE(capacity >= 0 && capacity < (int64_t)INT32_MAX, "vs_sjoin_points_polygons: capacity out of range");
    VS_REQUIRE(predicate == VS_PRED_WITHIN || predicate == VS_PRED_INTERSECTS ||
                   predicate == VS_PRED_COVERED_BY || predicate == VS_PRED_TOUCHES,
               "vs_sjoin_points_polygons: unsupported predicate %d", predicate);
    VS_REQUIRE(d_result != nullptr, "vs_sjoin_points_polygons: null result buffer");
    int64_t need = 0;
    const vs::JoinWs J = vs::carve_join_ws(workspace, n_points, n_polys, n_coords, capacity, true, &need);
    VS_REQUIRE(workspace_bytes >= need, "vs_sjoin_points_polygons: workspace %lld < %lld bytes",
               (long long)workspace_bytes, (long long)need);
    hipStream_t s = vs::stream_of(stream);
    hipLaunchKernelGGL(vs::ctl_reset_kernel, dim3(1), dim3(64), 0, s, J.ctl, nullptr, 0);
    VS_LAUNCH_CHECK();
    VS_TRY(vs::build_refine_flags(part_offsets, ring_offsets, n_parts, n_rings, n_coords, J.flags, s));
    if (n_polys > 0) {
        hipLaunchKernelGGL(vs::item_range_kernel, dim3((unsigned)vs::ceil_div(n_polys, 256)), dim3(256), 0, s, item_ids,
                           n_polys, geom_offsets, part_offsets, ring_offsets, J.item_range);
        VS_LAUNCH_CHECK();
    }
    VS_TRY(vs::point_order(px, py, n_points, 12, J.perm, J.sx, J.sy, J.order_ws, s));  // 3 radix passes
    const vs::TreeView T = vs::make_tree_view(layout, node_boxes, item_ids);
    const vs::PolyView P = {x, y, J.flags, J.item_range};
    if (n_points > 0 && n_polys > 0)
        VS_TRY(vs::launch_traverse<vs::kModePip>(predicate, J.sx, J.sy, nullptr, J.perm, n_points, T, P, J,
                                                 capacity, s));
    return vs::finish_pairs(J.fin_ws, J.ctl, n_points, capacity, true, out_left, out_right, d_result, s);
}

int vs_sjoin_points_polygons_noalloc(const double* px, const double* py, int64_t n_points, const double* x,
                                     const double* y, const int32_t* geom_offsets, const int32_t* part_offsets,
                                     const int32_t* ring_offsets, int64_t n_polys, int64_t n_parts, int64_t n_rings,
                                     int64_t n_coords, const vs_tree_layout* layout, const double* node_boxes,
                                     const int32_t* item_ids, int32_t predicate, int64_t capacity,
                                     int64_t* out_left, int64_t* out_right, int64_t* d_result, void* stream) {
    vs::clear_error();
    VS_REQUIRE(n_points >= 0 && n_polys >= 0 && n_coords >= 0 && capacity >= 0,
               "vs_sjoin_points_polygons_noalloc: negative size");
    const int64_t need = vs_sjoin_points_polygons_workspace_bytes(n_points, n_polys, n_coords, capacity);
    VS_REQUIRE(need <= VS_NOALLOC_SCRATCH_BYTES,
               "vs_sjoin_points_polygons_noalloc: %lld points x %lld polygons (%lld coordinates, capacity %lld) "
               "need %lld bytes of scratch, above the fixed %lld",
               (long long)n_points, (long long)n_polys, (long long)n_coords, (long long)capacity, (long long)need,
               (long long)VS_NOALLOC_SCRATCH_BYTES);
    void* ws = nullptr;
    VS_TRY(vs::scratch_ptr(&ws));
    return vs_sjoin_points_polygons(px, py, n_points, x, y, geom_offsets, part_offsets, ring_offsets, n_polys,
                                    n_parts, n_rings, n_coords, layout, node_boxes, item_ids, predicate, capacity,
                                    out_left, out_right, d_result, ws, VS_NOALLOC_SCRATCH_BYTES, stream);
}

int64_t vs_query_boxes_workspace_bytes(int64_t n_queries, int64_t capacity) {
    int64_t b = 0;
    vs::carve_join_ws(nullptr, n_queries, 0, 0, capacity, false, &b);
    return b;
}

int vs_query_boxes(const double* query_boxes, int64_t n_queries, const vs_tree_layout* layout,
                   const double* node_boxes, const int32_t* item_ids, int64_t capacity, int32_t sort,
                   int64_t* out_left, int64_t* out_right, int64_t* d_result, void* workspace,
                   int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    VS_TRY(vs::check_layout(layout, vs::kStackMaxBox));
    VS_REQUIRE(n_queries >= 0 && n_queries < (int64_t)INT32_MAX, "vs_query_boxes: n_queries out of range");
    VS_REQUIRE(capacity >= 0 && capacity < (int64_t)INT32_MAX, "vs_query_boxes: capacity out of range");
    VS_REQUIRE(d_result != nullptr, "vs_query_boxes: null result buffer");
    int64_t need = 0;
    const vs::JoinWs J = vs::carve_join_ws(workspace, n_queries, 0, 0, capacity, false, &need);
    VS_REQUIRE(workspace_bytes >= need, "vs_query_boxes: workspace %lld < %lld bytes",
               (long long)workspace_bytes, (long long)need);
    hipStream_t s = vs::stream_of(stream);
    hipLaunchKernelGGL(vs::ctl_reset_kernel, dim3(1), dim3(64), 0, s, J.ctl, nullptr, 0);
    VS_LAUNCH_CHECK();
    const vs::TreeView T = vs::make_tree_view(layout, node_boxes, item_ids);
    const vs::PolyView P = {nullptr, nullptr, nullptr, nullptr};
    if (n_queries > 0 && layout->n_items > 0) {
        const double4* qb = reinterpret_cast<const double4*>(query_boxes);
        VS_TRY(vs::box_order(qb, n_queries, J.perm, J.order_ws, s));  // waves of nearby boxes
        // the traversal's shape follows the hits the caller expects per box
        // (capacity / n_queries); both shapes give the same pairs
        const bool one_box = capacity >= (int64_t)vs::kOneBoxHits * n_queries;
        VS_TRY(vs::launch_traverse<vs::kModeBox>(VS_PRED_INTERSECTS, nullptr, nullptr, qb, J.perm, n_queries, T, P,
                                                 J, capacity, s, one_box));
    }
    return vs::finish_pairs(J.fin_ws, J.ctl, n_queries, capacity, sort != 0, out_left, out_right, d_result, s);
}

}  // extern "C"
