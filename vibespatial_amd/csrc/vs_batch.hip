// Batched polygon x polygon intersects: many independent sjoin jobs in one
// fixed sequence of launches, whatever the number of jobs.
//
// Replaces, per job, the reference's sjoin(left, right, predicate='intersects')
// path: bbox candidates from the device index (spatial/spatial_index_device.py:
// 228-390) refined by compute_polygonal_intersects_gpu (predicates/polygon.py:
// 849-996, polygon_kernels.py:163-228), compacted and lexsorted
// (spatial/query.py:143-153).  Job j pairs left rows [left_job[j],
// left_job[j+1]) with right rows [right_job[j], right_job[j+1]) only.
//
//   1. bounds of every left and right row;
//   2. right rows sorted by (job, 8-level Hilbert key of the bbox centre over
//      the extent of the job's right centres); a packed tree per job (fan-out
//      16, root first, items last) is built by one block per job into one
//      forest array;
//   3. left rows sorted the same way; every job is cut into waves of 64
//      consecutive sorted left rows; a persistent kernel deals the waves out
//      round-robin and each wave walks its own job's tree with a shared LDS
//      stack, staging bbox candidates;
//   4. every candidate whose left row is one small ring is first decided
//      from the right row's cell grid (vs_grid.h) where the grid allows:
//      pairs found to intersect go straight to the finish stage, pairs found
//      apart are dropped, the rest are kept for step 5;
//   5. one thread per remaining candidate runs the exact intersects refine
//      (edges of the pair, then one vertex per ring); a right row of one ring
//      with many edges is read through its y-slab index (built once per
//      call), so a candidate visits only the edges in its own y-range;
//      survivors go to the finish pass (radix sort by left row, sort of
//      right rows per left row).
// Tile shape (`tile`): waves per block of the traversal and the refine block
// size, chosen by the host autotuner.
#include "vs_grid.h"
#include "vs_join_common.h"
#include "vs_polys.h"

namespace vs {

int polygon_bounds_info(const double* x, const double* y, const int32_t* geom, const int32_t* part,
                        const int32_t* ring, int64_t n, double* bounds, int2* info, RowRec* rec, hipStream_t s);
int sort_pairs_u64(uint64_t*, int32_t*, int64_t, int, int, void*, int64_t, hipStream_t);
int64_t sort_workspace_bytes_u64(int64_t n);
int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, long long* d_total, void* ws, hipStream_t s);
int64_t scan_workspace_bytes(int64_t n);

constexpr int kFF = 16;          // forest fan-out
constexpr int kFMaxLevels = 10;  // 16^9 > 2^31 items
constexpr int kFStack = kFMaxLevels * kFF;
constexpr int kFOut = 128;       // per-wave LDS candidate buffer

// Level sizes of a packed tree over n items, root first; returns the level count.
__host__ __device__ inline int forest_levels(int64_t n, int64_t* size) {
    if (n <= 0) return 0;
    int64_t tmp[kFMaxLevels];
    int d = 0;
    int64_t m = n;
    for (;;) {
        tmp[d++] = m;
        if (m == 1) break;
        m = (m + kFF - 1) / kFF;
    }
    for (int l = 0; l < d; ++l) size[l] = tmp[d - 1 - l];
    return d;
}

// Job of every row without a search per row: mark[k] counts the jobs j >= 1
// starting at row k (a job starts at left_job[j]; empty jobs share their
// start with the next), the exclusive scan of the marks plus the row's own
// mark is #{j >= 1 : off[j] <= row} = the last job j with off[j] <= row.
__global__ __launch_bounds__(256) void job_marks_kernel(const int64_t* __restrict__ job_off, int64_t n_jobs,
                                                       int64_t n_rows, int32_t* __restrict__ mark) {
    for (int64_t j = 1 + (int64_t)blockIdx.x * 256 + threadIdx.x; j < n_jobs; j += (int64_t)gridDim.x * 256) {
        const int64_t k = job_off[j];
        if (k < n_rows) atomicAdd(&mark[k], 1);
    }
}

__global__ __launch_bounds__(256) void add_marks_kernel(const int32_t* __restrict__ mark, int64_t n,
                                                       int32_t* __restrict__ row_job) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        row_job[i] += mark[i];
}

// Extent of the box centres of each job's rows, as order-preserving 64-bit
// integers (ordered_f64: the integer order of the codes is the order of the
// doubles) so that atomicMin / atomicMax fold it: jext[j] = (min x, min y,
// max x, max y).  NaN boxes are skipped.  jext starts at (~0, ~0, 0, 0)
// (job_extent_init_kernel), which decodes to NaN for a job with no rows.
__device__ __forceinline__ unsigned long long ordered_f64(double d) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double from_ordered_f64(unsigned long long u) {
    return __longlong_as_double((long long)((u >> 63) ? (u & 0x7FFFFFFFFFFFFFFFull) : ~u));
}

__global__ __launch_bounds__(256) void job_extent_init_kernel(int64_t n_jobs, ulonglong4* __restrict__ jext) {
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < n_jobs; j += (int64_t)gridDim.x * 256)
        jext[j] = make_ulonglong4(~0ull, ~0ull, 0ull, 0ull);
}

__global__ __launch_bounds__(256) void job_extent_kernel(const double4* __restrict__ b, int64_t n,
                                                        const int32_t* __restrict__ row_job,
                                                        ulonglong4* __restrict__ jext) {
    // each wave folds a contiguous chunk of rows, carrying the running extent
    // of its current job in registers and folding it into jext when the job
    // changes: a few atomics per wave instead of four per row
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * 4, wv = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t chunk = (n + waves - 1) / waves;
    const int64_t c0 = wv * chunk, c1 = c0 + chunk < n ? c0 + chunk : n;
    int64_t cj = -1;
    double rx0 = INFINITY, ry0 = INFINITY, rx1 = -INFINITY, ry1 = -INFINITY;
    auto fold = [&]() {
        if (cj >= 0 && lane == 0 && rx0 <= rx1) {
            atomicMin(&jext[cj].x, ordered_f64(rx0));
            atomicMin(&jext[cj].y, ordered_f64(ry0));
            atomicMax(&jext[cj].z, ordered_f64(rx1));
            atomicMax(&jext[cj].w, ordered_f64(ry1));
        }
    };
    for (int64_t i0 = c0; i0 < c1; i0 += 64) {
        const int64_t i = i0 + lane;
        const bool in = i < c1;
        int64_t j = -1;
        double cx = NAN, cy = NAN;
        if (in) {
            j = row_job[i];
            const double4 v = b[i];
            cx = 0.5 * (v.x + v.z);
            cy = 0.5 * (v.y + v.w);
        }
        const bool ok = in && cx == cx && cy == cy;
        uint64_t rem = __ballot(in);
        while (rem) {  // one pass per distinct job among the 64 rows (usually one)
            const int sl = __ffsll((unsigned long long)rem) - 1;
            const int64_t js = __shfl(j, sl, 64);
            const bool mine = in && j == js;
            rem &= ~__ballot(mine);
            const double x0 = wave_min(mine && ok ? cx : INFINITY), y0 = wave_min(mine && ok ? cy : INFINITY);
            const double x1 = wave_max(mine && ok ? cx : -INFINITY), y1 = wave_max(mine && ok ? cy : -INFINITY);
            if (js != cj) {
                fold();
                cj = js;
                rx0 = x0;
                ry0 = y0;
                rx1 = x1;
                ry1 = y1;
            } else {
                rx0 = fmin(rx0, x0);
                ry0 = fmin(ry0, y0);
                rx1 = fmax(rx1, x1);
                ry1 = fmax(ry1, y1);
            }
        }
    }
    fold();
}

// key = job << 16 | 8-level Hilbert code of the bbox centre over its job's
// centre extent (NaN -> last): 16 + log2(jobs) bits to sort.
constexpr int kJobKeyBits = 16;
__global__ __launch_bounds__(256) void job_keys_kernel(const double4* __restrict__ b, int64_t n,
                                                      const int32_t* __restrict__ row_job,
                                                      const ulonglong4* __restrict__ jext,
                                                      uint64_t* __restrict__ keys, int32_t* __restrict__ ids) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t j = row_job[i];
    const double4 v = b[i];
    uint64_t h = (1ull << kJobKeyBits) - 1;
    if (v.x == v.x && v.y == v.y && v.z == v.z && v.w == v.w) {
        const ulonglong4 e = jext[j];
        const double4 ext = make_double4(from_ordered_f64(e.x), from_ordered_f64(e.y), from_ordered_f64(e.z),
                                         from_ordered_f64(e.w));
        const KeyScale ks = key_scale(&ext, kJobKeyBits / 2);
        const uint32_t qx = quantize(0.5 * (v.x + v.z), ks.x0, ks.sx, kJobKeyBits / 2) << (32 - kJobKeyBits / 2);
        const uint32_t qy = quantize(0.5 * (v.y + v.w), ks.y0, ks.sy, kJobKeyBits / 2) << (32 - kJobKeyBits / 2);
        h = hilbert64(qx, qy) >> (64 - kJobKeyBits);
    }
    keys[i] = ((uint64_t)j << kJobKeyBits) | h;
    ids[i] = (int32_t)i;
}

// Nodes per job tree and waves per job.
__global__ __launch_bounds__(256) void job_sizes_kernel(const int64_t* __restrict__ left_job,
                                                       const int64_t* __restrict__ right_job, int64_t n_jobs,
                                                       int32_t* __restrict__ nodes, int32_t* __restrict__ waves) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n_jobs) return;
    int64_t size[kFMaxLevels];
    const int d = forest_levels(right_job[j + 1] - right_job[j], size);
    int64_t t = 0;
    for (int l = 0; l < d; ++l) t += size[l];
    nodes[j] = (int32_t)t;
    waves[j] = (int32_t)((left_job[j + 1] - left_job[j] + 63) / 64);
}

// One block per job: item boxes in sorted order, then each parent level.
__global__ __launch_bounds__(256) void forest_build_kernel(const double4* __restrict__ rbox,
                                                          const int32_t* __restrict__ rperm,
                                                          const int64_t* __restrict__ right_job,
                                                          const int32_t* __restrict__ tree_off, int64_t n_jobs,
                                                          double4* __restrict__ fbox) {
    for (int64_t j = blockIdx.x; j < n_jobs; j += gridDim.x) {
        const int64_t r0 = right_job[j], R = right_job[j + 1] - r0;
        int64_t size[kFMaxLevels];
        const int d = forest_levels(R, size);
        if (d == 0) continue;
        int64_t start[kFMaxLevels];
        int64_t acc = tree_off[j];
        for (int l = 0; l < d; ++l) {
            start[l] = acc;
            acc += size[l];
        }
        for (int64_t k = threadIdx.x; k < R; k += blockDim.x) fbox[start[d - 1] + k] = rbox[rperm[r0 + k]];
        __syncthreads();
        for (int l = d - 2; l >= 0; --l) {
            for (int64_t k = threadIdx.x; k < size[l]; k += blockDim.x) {
                double4 u = make_double4(NAN, NAN, NAN, NAN);
                const int64_t c0 = k * kFF, c1 = (c0 + kFF < size[l + 1]) ? c0 + kFF : size[l + 1];
                for (int64_t c = c0; c < c1; ++c) {
                    const double4 v = fbox[start[l + 1] + c];
                    u.x = fmin(u.x, v.x);
                    u.y = fmin(u.y, v.y);
                    u.z = fmax(u.z, v.z);
                    u.w = fmax(u.w, v.w);
                }
                fbox[start[l] + k] = u;
            }
            __syncthreads();
        }
    }
}

// One traversal task (64 consecutive sorted left rows of one job) with every
// per-job word the traversal needs, so a task starts with one 32-byte load
// instead of a binary search over the job offsets.
struct TaskRec {
    int32_t s0;    // first sorted left row
    int32_t cnt;   // rows (1..64)
    int32_t r0;    // first right row of the job
    int32_t nr;    // right rows of the job
    int32_t toff;  // the job's tree in the forest
    int32_t pad[3];
};

// One wave per job (grid-stride): its tasks' records.
__global__ __launch_bounds__(256) void task_recs_kernel(const int64_t* __restrict__ left_job,
                                                       const int64_t* __restrict__ right_job, int64_t n_jobs,
                                                       const int32_t* __restrict__ wave_off,
                                                       const int32_t* __restrict__ tree_off,
                                                       TaskRec* __restrict__ tasks) {
    const int lane = threadIdx.x & 63;
    for (int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); j < n_jobs; j += (int64_t)gridDim.x * 4) {
        const int64_t l0 = left_job[j], nl = left_job[j + 1] - l0;
        const int32_t w0 = wave_off[j], r0 = (int32_t)right_job[j], nr = (int32_t)(right_job[j + 1] - right_job[j]);
        const int32_t toff = tree_off[j];
        for (int64_t k = lane; 64 * k < nl; k += 64) {
            TaskRec t;
            t.s0 = (int32_t)(l0 + 64 * k);
            t.cnt = (int32_t)(nl - 64 * k < 64 ? nl - 64 * k : 64);
            t.r0 = r0;
            t.nr = nr;
            t.toff = toff;
            t.pad[0] = t.pad[1] = t.pad[2] = 0;
            tasks[w0 + k] = t;
        }
    }
}

// Persistent: waves take (job, 64 sorted left rows) tasks round-robin.
__global__ __launch_bounds__(256) void forest_query_kernel(const double4* __restrict__ lbox,
                                                          const int32_t* __restrict__ lperm,
                                                          const TaskRec* __restrict__ tasks,
                                                          const long long* __restrict__ total_waves,
                                                          const double4* __restrict__ fbox,
                                                          const int32_t* __restrict__ rperm,
                                                          uint32_t* __restrict__ cand_l, int32_t* __restrict__ cand_r,
                                                          int64_t capacity, JoinCtl* __restrict__ ctl) {
    __shared__ long long stack_s[4][kFStack];
    __shared__ int2 outbuf_s[4][kFOut];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    long long* stack = stack_s[w];
    int2* outbuf = outbuf_s[w];
    const uint64_t lt = lanemask_lt();
    const long long W = *total_waves;
    int nbuf = 0;
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
                cand_l[slot] = (uint32_t)e.x;
                cand_r[slot] = e.y;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        nbuf = 0;
    };
    // static round-robin of the wave tasks (a shared queue counter would
    // serialise every task start in L2)
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    for (long long tt = (long long)blockIdx.x * (blockDim.x >> 6) + w; tt < W; tt += nw) {
        const TaskRec task = tasks[tt];
        const int64_t idx = (int64_t)task.s0 + lane;
        const bool inr = lane < task.cnt;
        const int32_t orig = inr ? lperm[idx] : 0;
        const double4 b = inr ? lbox[orig] : make_double4(NAN, NAN, NAN, NAN);
        const bool valid = inr && b.x == b.x && b.y == b.y && b.z == b.z && b.w == b.w;
        const double ux0 = wave_min(valid ? b.x : INFINITY), uy0 = wave_min(valid ? b.y : INFINITY);
        const double ux1 = wave_max(valid ? b.z : -INFINITY), uy1 = wave_max(valid ? b.w : -INFINITY);
        const int64_t r0 = task.r0;
        int64_t size[kFMaxLevels];
        const int d = forest_levels(task.nr, size);
        if (d == 0 || __ballot(valid) == 0) continue;
        int64_t start[kFMaxLevels];
        {
            int64_t acc = task.toff;
            for (int l = 0; l < d; ++l) {
                start[l] = acc;
                acc += size[l];
            }
        }
        int sp = 1;
        if (lane == 0) stack[0] = 0;  // group 0 of level 0: the root
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        while (sp > 0) {
            --sp;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const long long e = stack[sp];
            const int l = (int)(e & 15);
            const int64_t first = (e >> 4) * kFF;
            int64_t cl = 0, sl = 0;
            for (int q = 0; q < d; ++q)  // wave-uniform select (no dynamic register indexing)
                if (q == l) {
                    cl = size[q];
                    sl = start[q];
                }
            const int ng = (int)((cl - first) < kFF ? (cl - first) : kFF);
            double4 nb = make_double4(NAN, NAN, NAN, NAN);
            bool ov = false;
            int32_t rid = 0;  // leaf level: the group's right row ids, loaded with the boxes
            if (lane < ng) {
                nb = fbox[sl + first + lane];
                if (l == d - 1) rid = rperm[r0 + first + lane];
                ov = box_overlap(ux0, uy0, ux1, uy1, nb.x, nb.y, nb.z, nb.w);
            }
            uint64_t m = __ballot(ov);
            if (l == d - 1) {
                while (m) {
                    const int s = __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                    const double bx0 = readlane_f64(nb.x, s), by0 = readlane_f64(nb.y, s);
                    const double bx1 = readlane_f64(nb.z, s), by1 = readlane_f64(nb.w, s);
                    const bool hit = valid && box_overlap(b.x, b.y, b.z, b.w, bx0, by0, bx1, by1);
                    const uint64_t hm = __ballot(hit);
                    if (!hm) continue;
                    const int k = __popcll(hm);
                    if (nbuf + k > kFOut) flush();
                    const int32_t rr = __builtin_amdgcn_readlane(rid, s);
                    if (hit) outbuf[nbuf + __popcll(hm & lt)] = make_int2(orig, rr);
                    nbuf += k;
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
    flush();
}

// ---------------------------------------------------------------------------
// Y-slab index of the right rows.  A right row of one ring with at least
// min_edges edges (128, or 16 / 64 / never by the tile; and at most 2^15
// coordinates) gets its y-range
// [ylo, yhi] cut into B equal slabs (B a power of two, about one per four
// edges, at most kRowSlabMax); edge e = (c0 + e - 1, c0 + e) is listed in
// every slab from slab(min y) to slab(max y).  slab() is monotone in y, so a
// point finds in slab(py) every edge whose closed y-range holds py -- every
// edge that can toggle its crossing parity or put it on the boundary -- and a
// box [y0, y1] finds in slabs slab(y0)..slab(y1) every edge whose y-range
// meets it.  Row r's list is 3 n entries from 3 c0 (its first coordinate c0,
// n coordinates; single-ring rows own disjoint coordinate ranges, so disjoint
// lists): B + 1 slab starts, then at most 2 (n - 1) edge numbers (a row that
// would need more takes one slab holding every edge).
// ---------------------------------------------------------------------------
constexpr int kRowSlabMinEdges[4] = {128, 16, 64, 1 << 30};  // by tile bits 26-27 (3: none); smaller rows are streamed whole
constexpr int kRowSlabMax = 64;

struct RowSlabs {
    const double2* geo;    // (ylo, slabs per unit of y)
    const int32_t* nslab;  // B, 0: the row has no index
    const uint16_t* list;
};

__device__ __forceinline__ int row_slab(double y, double ylo, double inv, int nb) {
    double f = (y - ylo) * inv;
    f = fmin(fmax(f, 0.0), (double)(nb - 1));
    return (int)f;
}

__device__ __forceinline__ int64_t row_slab_base(int32_t c0) { return 3 * (int64_t)c0; }

// One 64-thread block per row (grid-stride): counts per slab, starts, entries.
__global__ __launch_bounds__(64) void row_slab_build_kernel(Polys B, int64_t n_rows, const int2* __restrict__ rinfo,
                                                            const double4* __restrict__ rbox,
                                                            double2* __restrict__ geo, int32_t* __restrict__ nslab,
                                                            uint16_t* __restrict__ list, int64_t list_cap,
                                                            int min_edges) {
    __shared__ int cnt[kRowSlabMax + 1];
    __shared__ int cur[kRowSlabMax];
    const int lane = threadIdx.x;
    for (int64_t r = blockIdx.x; r < n_rows; r += gridDim.x) {
        const int2 ri = rinfo[r];
        const int32_t c0 = ri.x, n = ri.y;
        // B + 1 + 2 (n - 1) <= 3 n entries from 3 c0 on, every value below 2^16
        const bool ok = c0 >= 0 && n - 1 >= min_edges && n <= 32768 && row_slab_base(c0) + 3 * (int64_t)n <= list_cap;
        if (!ok) {
            if (lane == 0) nslab[r] = 0;
            continue;  // block-uniform
        }
        const int ne = n - 1;
        const double4 bb = rbox[r];
        int nb = 1;
        while (nb < kRowSlabMax && nb * 4 < ne) nb <<= 1;
        double inv = (bb.w > bb.y) ? (double)nb / (bb.w - bb.y) : 0.0;
        if (!(inv > 0.0) || !(inv < INFINITY)) {
            nb = 1;
            inv = 0.0;
        }
        uint16_t* L = list + row_slab_base(c0);
        if (lane < kRowSlabMax) cnt[lane + 1] = 0;
        if (lane == 0) cnt[0] = 0;
        __syncthreads();
        int total = 0;
        for (int e = 1 + lane; e <= ne; e += 64) {
            const double ay = B.y[c0 + e - 1], by = B.y[c0 + e];
            const int b0 = row_slab(fmin(ay, by), bb.y, inv, nb), b1 = row_slab(fmax(ay, by), bb.y, inv, nb);
            total += b1 - b0 + 1;
            for (int b = b0; b <= b1; ++b) atomicAdd(&cnt[b + 1], 1);
        }
        total = wave_sum_i32(total);
        if (total > 2 * ne) {  // long edges: one slab with every edge
            nb = 1;
            inv = 0.0;
            if (lane == 0) {
                L[0] = 0;
                L[1] = (uint16_t)ne;
            }
            for (int e = 1 + lane; e <= ne; e += 64) L[2 + e - 1] = (uint16_t)e;
        } else {
            int v = lane < nb ? cnt[lane + 1] : 0;
            v = wave_inclusive_scan(v);
            __syncthreads();
            if (lane < nb) {
                cur[lane] = v - cnt[lane + 1];  // exclusive start of slab lane
                L[lane + 1] = (uint16_t)v;
            }
            if (lane == 0) L[0] = 0;
            __syncthreads();
            uint16_t* body = L + nb + 1;
            for (int e = 1 + lane; e <= ne; e += 64) {
                const double ay = B.y[c0 + e - 1], by = B.y[c0 + e];
                const int b0 = row_slab(fmin(ay, by), bb.y, inv, nb), b1 = row_slab(fmax(ay, by), bb.y, inv, nb);
                for (int b = b0; b <= b1; ++b) body[atomicAdd(&cur[b], 1)] = (uint16_t)e;
            }
        }
        if (lane == 0) {
            geo[r] = make_double2(bb.y, inv);
            nslab[r] = nb;
        }
        __syncthreads();  // cnt / cur are reused by the next row
    }
}

// ---------------------------------------------------------------------------
// Cell grids of the right rows (vs_grid.h): 16 lanes per row, 16 rows per
// 256-thread block (grid-stride, block-uniform trip count).  A row gets no
// grid (geo.z = 0) unless it is one closed ring of at least 4 finite
// coordinates with a box of non-empty sides.
// ---------------------------------------------------------------------------
struct RowGrids {
    const double4* geo;  // (x0, y0, wx, wy); wx = 0: no grid
    const double2* inv;  // (1 / wx, 1 / wy)
    const uint2* cells;  // kGrid rows per right row: (boundary bits, inside bits)
};

constexpr int kGridRowsPerBlock = 16;

__global__ __launch_bounds__(256) void row_grid_build_kernel(Polys B, int64_t n_rows, const int2* __restrict__ rinfo,
                                                            const double4* __restrict__ rbox,
                                                            double4* __restrict__ geo, double2* __restrict__ inv,
                                                            uint2* __restrict__ cells) {
    __shared__ uint32_t bnd_s[kGridRowsPerBlock][kGrid];
    __shared__ uint32_t tog_s[kGridRowsPerBlock][kGrid];
    const int grp = threadIdx.x >> 4, sub = threadIdx.x & 15, lane = threadIdx.x & 63;
    uint32_t* bnd = bnd_s[grp];
    uint32_t* tog = tog_s[grp];
    for (int64_t r0 = (int64_t)blockIdx.x * kGridRowsPerBlock; r0 < n_rows;
         r0 += (int64_t)gridDim.x * kGridRowsPerBlock) {
        const int64_t r = r0 + grp;
        int32_t c0 = 0, n = 0;
        GridAxes g{};
        bool ok = false;
        if (r < n_rows) {
            const int2 ri = rinfo[r];
            const double4 bb = rbox[r];
            c0 = ri.x;
            n = ri.y;
            ok = c0 >= 0 && n >= 4 && grid_setup(bb.x, bb.y, bb.z, bb.w, g);
        }
        bnd[sub] = 0u;
        bnd[sub + 16] = 0u;
        tog[sub] = 0u;
        tog[sub + 16] = 0u;
        __syncthreads();
        bool bad = false;
        if (ok) {
            // lane 0 also checks that the ring is closed (first == last), with its first edge's loads
            for (int32_t e = 1 + sub; e < n; e += 16) {
                const double ax = B.x[c0 + e - 1], ay = B.y[c0 + e - 1], bx = B.x[c0 + e], by = B.y[c0 + e];
                if (e == 1) bad = bad || !(ax == B.x[c0 + n - 1] && ay == B.y[c0 + n - 1]);
                if (!(fabs(ax) < INFINITY && fabs(ay) < INFINITY && fabs(bx) < INFINITY && fabs(by) < INFINITY)) {
                    bad = true;
                    continue;
                }
                const GridEdge ge = grid_edge(g, ax, ay, bx, by);
                const uint32_t m = span_mask(ge.i0, ge.i1);
                for (int j = ge.j0; j <= ge.j1; ++j) {
                    atomicOr(&bnd[j], m);
                    if (ge.k < kGrid && grid_crosses(g, j, ay, by)) atomicXor(&tog[j], 1u << ge.k);
                }
            }
        }
        const bool any_bad = ((__ballot(bad) >> (lane & 48)) & 0xFFFFull) != 0;  // this row's 16 lanes
        __syncthreads();
        if (ok) {
            cells[r * kGrid + sub] = make_uint2(bnd[sub], prefix_xor32(tog[sub]));
            cells[r * kGrid + sub + 16] = make_uint2(bnd[sub + 16], prefix_xor32(tog[sub + 16]));
            if (sub == 0) {
                geo[r] = any_bad ? make_double4(0.0, 0.0, 0.0, 0.0) : make_double4(g.x0, g.y0, g.wx, g.wy);
                inv[r] = make_double2(g.ix, g.iy);
            }
        } else if (r < n_rows && sub == 0) {
            geo[r] = make_double4(0.0, 0.0, 0.0, 0.0);
        }
        __syncthreads();  // bnd / tog are cleared for the next rows
    }
}

// The exact intersects of ring A (na coordinates, a lane's LDS column) and
// right row B, a single ring c0b.. of nb coordinates read through its slab
// index: the same three steps as ring_intersects_polys (A's first vertex
// located in B -- inside or on the boundary decides; B's edges near A tested
// against A's edges; B's first vertex located in A), with each step visiting
// only the slab edges that can change its outcome, so the same boolean.
// skip_first: A's first vertex is known to lie outside B, off its boundary
// (the first step's outcome), so that step is not run.
__device__ inline bool ring_intersects_slabbed(const RingRef& A, int32_t na, const double* __restrict__ bx,
                                               const double* __restrict__ by, int32_t c0b,
                                               const uint16_t* __restrict__ L, int nb, double2 g,
                                               bool skip_first) {
    double mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
    for (int32_t k = 0; k < na; ++k) {
        mnx = fmin(mnx, A.X(k));
        mny = fmin(mny, A.Y(k));
        mxx = fmax(mxx, A.X(k));
        mxy = fmax(mxy, A.Y(k));
    }
    const uint16_t* body = L + nb + 1;
    // Slab entries four at a time: the four edge numbers, then the four
    // edges' coordinates, are loaded together (memory-level parallelism).
    constexpr int kU = 4;
    if (!skip_first) {
        const double fx = A.X(0), fy = A.Y(0);
        const int b = row_slab(fy, g.x, g.y, nb);
        int parity = 0;
        const int k0 = L[b], k1 = L[b + 1];
        for (int k = k0; k < k1; k += kU) {
            int32_t c[kU];
            double ax[kU], ay[kU], qx[kU], qy[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) c[u] = c0b + body[k + u < k1 ? k + u : k1 - 1];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                ax[u] = bx[c[u] - 1];
                ay[u] = by[c[u] - 1];
                qx[u] = bx[c[u]];
                qy[u] = by[c[u]];
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                if (k + u >= k1) break;
                if (edge_step_f64(fx, fy, ax[u], ay[u], qx[u], qy[u], parity)) return true;
            }
        }
        if (parity) return true;
    }
    {
        const int k0 = L[row_slab(mny, g.x, g.y, nb)], k1 = L[row_slab(mxy, g.x, g.y, nb) + 1];
        for (int k = k0; k < k1; k += kU) {
            int32_t c[kU];
            double ax[kU], ay[kU], qx[kU], qy[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) c[u] = c0b + body[k + u < k1 ? k + u : k1 - 1];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                ax[u] = bx[c[u] - 1];
                ay[u] = by[c[u] - 1];
                qx[u] = bx[c[u]];
                qy[u] = by[c[u]];
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                if (k + u >= k1) break;
                const double px = ax[u], py = ay[u], ex = qx[u], ey = qy[u];
                if (fmax(px, ex) >= mnx && fmin(px, ex) <= mxx && fmax(py, ey) >= mny && fmin(py, ey) <= mxy) {
                    for (int32_t j = 1; j < na; ++j)
                        if (segments_touch(A.X(j - 1), A.Y(j - 1), A.X(j), A.Y(j), px, py, ex, ey)) return true;
                }
            }
        }
    }
    const double px = bx[c0b], py = by[c0b];
    if (px < mnx || px > mxx || py < mny || py > mxy) return false;  // outside A's bounds: location 0
    int parity = 0;
    for (int32_t j = 1; j < na; ++j)
        if (edge_step_f64(px, py, A.X(j - 1), A.Y(j - 1), A.X(j), A.Y(j), parity)) return true;
    return parity != 0;
}

constexpr int kRingTile = 13;  // left rings of at most this many coordinates are refined from LDS

// One thread per bbox candidate, in the traversal's order.  A pair whose
// left row is one ring of at most kRingTile coordinates (no more than the
// right row's) and whose right row has a cell grid is decided from the grid
// where it can be (grid_decide: the left ring's box first, then its
// vertices): pairs that intersect are appended to the finish stage, pairs
// that do not are dropped.  Every other pair is appended to (key, left id) for the sort that
// orders the refine: key = the right row, or with by_work
// (15 - min(15, floor(log2 coords of the right row))) << (rbits + 1) |
// (first step still to run) << rbits | right row, so the candidates of the
// largest right rows come first, each right row's together (in at most two
// groups), and a wave's lanes refine rows of similar edge counts through the
// same steps.  Bit 31
// of the left id marks a pair whose exact test may skip its first step
// (a left vertex lies in a free outside cell of the right row).  Both lists are gathered per
// block in LDS and appended with one global atomic per kClsBuf - 256 or more
// entries (one atomic per wave on the two counters serialised the kernel).
// ctl->stat[1] accumulates, over all candidates, the coordinates of both rows.
constexpr int kClsBuf = 1024;

__global__ __launch_bounds__(256) void cand_classify_kernel(const uint32_t* __restrict__ cand_l,
                                                           const int32_t* __restrict__ cand_r,
                                                           JoinCtl* __restrict__ cctl, int64_t ccap,
                                                           const RowRec* __restrict__ lrec,
                                                           const int2* __restrict__ rinfo,
                                                           const double* __restrict__ ax,
                                                           const double* __restrict__ ay, RowGrids G,
                                                           int use_grid, int by_work, int rbits,
                                                           uint32_t* __restrict__ ukey, int32_t* __restrict__ uval,
                                                           uint32_t* __restrict__ out_l, int32_t* __restrict__ out_r,
                                                           int64_t capacity, JoinCtl* __restrict__ ctl) {
    __shared__ uint2 tbuf[kClsBuf];  // decided true: (left, right)
    __shared__ uint2 ubuf[kClsBuf];  // undecided: (key, left | flag)
    __shared__ unsigned int tn, un;
    __shared__ unsigned long long gbase;
    const unsigned long long nc = cctl->staged;
    if (nc > (unsigned long long)ccap) return;  // grid-uniform
    if (threadIdx.x == 0) {
        tn = 0u;
        un = 0u;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanemask_lt();
    const int64_t stride = (int64_t)gridDim.x * 256;
    unsigned long long cverts = 0;
    // flush the cnt entries of buf to (a, b) at a slot block from `counter`, bounded by cap
    auto flush = [&](uint2* buf, unsigned int cnt, unsigned int* n, unsigned long long* counter, uint32_t* a,
                     int32_t* b, int64_t cap) {
        if (threadIdx.x == 0) gbase = atomicAdd(counter, (unsigned long long)cnt);
        __syncthreads();
        const unsigned long long base = gbase;
        for (unsigned int k = threadIdx.x; k < cnt; k += 256) {
            const unsigned long long slot = base + k;
            if (slot < (unsigned long long)cap) {
                const uint2 v = buf[k];
                a[slot] = v.x;
                b[slot] = (int32_t)v.y;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) *n = 0u;
        __syncthreads();
    };
    for (int64_t b0 = (int64_t)blockIdx.x * 256; b0 < (int64_t)nc; b0 += stride) {  // block-uniform
        const int64_t i = b0 + threadIdx.x;
        int dec = -1;  // -1 no candidate, 0 apart, 1 intersects, 2 / 3 undecided
        uint32_t l = 0;
        int32_t r = 0, vb = 0;
        if (i < (int64_t)nc) {
            l = cand_l[i];
            r = cand_r[i];
            // every load that depends only on (l, r) issued together
            const RowRec L = lrec[l];
            const int2 ri = rinfo[r];
            const double4 gg = use_grid ? G.geo[r] : make_double4(0.0, 0.0, 0.0, 0.0);
            const double2 gi = use_grid ? G.inv[r] : make_double2(0.0, 0.0);
            const int2 li = L.info;
            vb = ri.y;
            cverts += (unsigned long long)(li.y + vb);
            dec = 3;
            if (use_grid && li.x >= 0 && li.y <= kRingTile && li.y <= vb && gg.z > 0.0)
                dec = grid_decide(GridAxes{gg.x, gg.y, gg.z, gg.w, gi.x, gi.y}, G.cells + (int64_t)r * kGrid,
                                  L.box.x, L.box.y, L.box.z, L.box.w, L.first.x, L.first.y, ax + li.x, ay + li.x,
                                  li.y);
        }
        const uint64_t tm = __ballot(dec == 1), um = __ballot(dec >= 2);
        unsigned int tb = 0u, ub = 0u;
        if (lane == 0) {
            if (tm) tb = atomicAdd(&tn, (unsigned int)__popcll(tm));
            if (um) ub = atomicAdd(&un, (unsigned int)__popcll(um));
        }
        tb = __shfl(tb, 0, 64);
        ub = __shfl(ub, 0, 64);
        if (dec == 1) tbuf[tb + (unsigned int)__popcll(tm & lt)] = make_uint2(l, (uint32_t)r);
        if (dec >= 2) {
            uint32_t key = (uint32_t)r;
            if (by_work) {
                const int lg = vb > 1 ? 31 - __clz(vb) : 0;
                key |= ((uint32_t)(15 - (lg < 15 ? lg : 15)) << (rbits + 1)) | ((dec == 3 ? 1u : 0u) << rbits);
            }
            ubuf[ub + (unsigned int)__popcll(um & lt)] = make_uint2(key, l | (dec == 2 ? 0x80000000u : 0u));
        }
        __syncthreads();
        const unsigned int tcount = tn, ucount = un;  // block-uniform: read between two barriers
        __syncthreads();
        const bool last = b0 + stride >= (int64_t)nc;  // each step adds at most 256 entries per list
        if (tcount > (unsigned int)(kClsBuf - 256) || (last && tcount > 0u))
            flush(tbuf, tcount, &tn, &ctl->staged, out_l, out_r, capacity);
        if (ucount > (unsigned int)(kClsBuf - 256) || (last && ucount > 0u))
            flush(ubuf, ucount, &un, &cctl->stat[0], ukey, uval, ccap);
    }
    for (int o = 32; o > 0; o >>= 1) cverts += __shfl_xor(cverts, o, 64);
    if (lane == 0 && cverts) atomicAdd(&ctl->stat[1], cverts);
}

// The exact intersects of ring A (na coordinates, a lane's LDS column) and
// right ring B (coordinates c0b .. c0b + nb - 1, nb <= 65536): the three
// terms of ring_intersects_polys -- A's first vertex on B's boundary or
// inside B, some edge of B touching some edge of A, B's first vertex in A --
// OR-ed, so the order of evaluation does not change the boolean.  One stream
// over B's edges (four loaded together) runs the first term's crossing steps
// (unless skip_first: then that term is known to be false unless the second
// is true) and lists, in the lane's LDS column `near` (kNearList uint16
// entries, stride 64), the edges whose box meets A's box; the segment tests
// run over that list when it fills and at the end.  Each lane thus tests its
// own near edges: a wave runs the segment loop max(list length) times rather
// than once for every edge index that is near for some lane.
constexpr int kNearList = 16;

__device__ inline bool ring_intersects_listed(const RingRef& A, int32_t na, const double* __restrict__ bx,
                                              const double* __restrict__ by, int32_t c0b, int32_t nb,
                                              bool skip_first, uint16_t* near) {
    double mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
    for (int32_t k = 0; k < na; ++k) {
        mnx = fmin(mnx, A.X(k));
        mny = fmin(mny, A.Y(k));
        mxx = fmax(mxx, A.X(k));
        mxy = fmax(mxy, A.Y(k));
    }
    const double fx = A.X(0), fy = A.Y(0);
    int cnt = 0, parity = 0;
    auto touches = [&]() -> bool {  // segment tests over the listed edges; empties the list
        for (int t = 0; t < cnt; ++t) {
            const int32_t e = c0b + (int32_t)near[t * 64];
            const double px = bx[e - 1], py = by[e - 1], ex = bx[e], ey = by[e];
            for (int32_t k = 1; k < na; ++k)
                if (segments_touch(A.X(k - 1), A.Y(k - 1), A.X(k), A.Y(k), px, py, ex, ey)) return true;
        }
        cnt = 0;
        return false;
    };
    double px = bx[c0b], py = by[c0b];
    for (int32_t c0 = 1; c0 < nb; c0 += 4) {
        double qxs[4], qys[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int32_t c = c0 + u < nb ? c0 + u : nb - 1;
            qxs[u] = bx[c0b + c];
            qys[u] = by[c0b + c];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (c0 + u >= nb) break;
            const double qx = qxs[u], qy = qys[u];
            if (!skip_first && edge_step_f64(fx, fy, px, py, qx, qy, parity)) return true;  // on B's boundary
            if (fmax(px, qx) >= mnx && fmin(px, qx) <= mxx && fmax(py, qy) >= mny && fmin(py, qy) <= mxy) {
                near[cnt * 64] = (uint16_t)(c0 + u);
                if (++cnt == kNearList && touches()) return true;
            }
            px = qx;
            py = qy;
        }
    }
    if (parity) return true;  // A's first vertex inside B
    if (touches()) return true;
    const double qx = bx[c0b], qy = by[c0b];
    if (qx < mnx || qx > mxx || qy < mny || qy > mxy) return false;  // B's first vertex outside A's box
    int par = 0;
    for (int32_t k = 1; k < na; ++k)
        if (edge_step_f64(qx, qy, A.X(k - 1), A.Y(k - 1), A.X(k), A.Y(k), par)) return true;
    return par != 0;
}

// One thread per candidate left by cand_classify_kernel (count in *count,
// at most ccap): exact intersects; survivors appended to the finish stage
// with one atomic per wave.  A left row that is one ring of at most
// kRingTile coordinates (the footprints) is copied into the lane's column of
// an LDS tile (x and y, 64 lanes wide) and tested from there: its edges are
// re-read for every right edge near it.  Other rows, and pairs where the
// right row is the smaller ring, take polygons_intersect_auto; the boolean is
// the same.  Bit 31 of a left id: the first step of the tile path (locating
// the left ring's first vertex in the right row) is skipped (grid_decide's
// outcome 2: its answer is "outside, off the boundary" unless the rings'
// boundaries touch, which the second step finds).
constexpr int64_t kRingTileBytesPerWave =
    2 * kRingTile * 64 * (int64_t)sizeof(double) + kNearList * 64 * (int64_t)sizeof(uint16_t);
__global__ __launch_bounds__(256) void batch_refine_kernel(const uint32_t* __restrict__ cand_l,
                                                          const int32_t* __restrict__ cand_r, uint32_t rmask,
                                                          const unsigned long long* __restrict__ count,
                                                          int64_t ccap, Polys A, Polys B,
                                                          const RowRec* __restrict__ lrec,
                                                          const int2* __restrict__ rinfo, RowSlabs RS,
                                                          uint32_t* __restrict__ out_l, int32_t* __restrict__ out_r,
                                                          int64_t capacity, JoinCtl* __restrict__ ctl) {
    extern __shared__ double ring_tile[];
    const unsigned long long nc = *count;
    if (nc > (unsigned long long)ccap) return;
    const int lane = threadIdx.x & 63;
    // per wave: x and y columns of the ring tile, then the near-edge lists
    char* wbase = reinterpret_cast<char*>(ring_tile) + (threadIdx.x >> 6) * kRingTileBytesPerWave;
    double* tx = reinterpret_cast<double*>(wbase) + lane;
    double* ty = tx + kRingTile * 64;
    uint16_t* near = reinterpret_cast<uint16_t*>(wbase + 2 * kRingTile * 64 * sizeof(double)) + lane;
    const uint64_t lt = lanemask_lt();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); i0 < (int64_t)nc; i0 += stride) {
        const int64_t i = i0 + lane;
        bool keep = false;
        uint32_t l = 0;
        int32_t r = 0;
        if (i < (int64_t)nc) {
            const uint32_t lv = cand_l[i];
            const bool skip_first = (lv >> 31) != 0u;
            l = lv & 0x7FFFFFFFu;
            r = (int32_t)((uint32_t)cand_r[i] & rmask);
            const int2 li = lrec[l].info, ri = rinfo[r];
            const int32_t ca = li.x, na = li.y, cb = ri.x, nb = ri.y;
            const int nslab = RS.nslab[r];
            const bool tile_a = ca >= 0 && na <= kRingTile;
            if (tile_a && (nslab > 0 || !(cb >= 0 && nb < na))) {
                for (int32_t k = 0; k < na; ++k) {  // the lane's own column: no other lane reads it
                    tx[k * 64] = A.x[ca + k];
                    ty[k * 64] = A.y[ca + k];
                }
                if (nslab > 0) {  // one ring of more than kRingTile coordinates, indexed
                    keep = ring_intersects_slabbed(RingRef{tx, ty, 64}, na, B.x, B.y, cb,
                                                   RS.list + row_slab_base(cb), nslab, RS.geo[r], skip_first);
                } else {
                    keep = (cb >= 0 && nb <= 65536)
                               ? ring_intersects_listed(RingRef{tx, ty, 64}, na, B.x, B.y, cb, nb, skip_first, near)
                               : ring_intersects_polys(RingRef{tx, ty, 64}, na, B, r, skip_first);
                }
            } else {
                keep = polygons_intersect_auto(A, l, B, r);
            }
        }
        const uint64_t km = __ballot(keep);
        if (!km) continue;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(&ctl->staged, (unsigned long long)__popcll(km));
        base = __shfl(base, 0, 64);
        if (keep) {
            const unsigned long long slot = base + (unsigned long long)__popcll(km & lt);
            if (slot < (unsigned long long)capacity) {
                out_l[slot] = l;
                out_r[slot] = r;
            }
        }
    }
}

// d_result: [0] pairs K, [1] bbox candidates staged (the stage that bounds
// `capacity`: survivors are a subset), [2] bbox candidates, [3] sum over the
// candidates of both rows' coordinates, [4] candidates left to the exact
// refine ([3] and [4] are 0 when the candidates overflowed).
__global__ void publish_batch_kernel(const JoinCtl* __restrict__ cctl, const JoinCtl* __restrict__ ctl,
                                     int64_t* __restrict__ result) {
    if (threadIdx.x == 0) {
        result[1] = (int64_t)cctl->staged;
        result[2] = (int64_t)cctl->staged;
        result[3] = (int64_t)ctl->stat[1];
        result[4] = (int64_t)cctl->stat[0];
    }
}

struct BatchWs {
    double4* lbox;
    double4* rbox;
    uint64_t* lkey;
    int32_t* lperm;
    uint64_t* rkey;
    int32_t* rperm;
    void* sort_ws;
    ulonglong4* jext;
    int32_t* row_mark;
    int32_t* row_job;
    long long* total_rows;
    void* row_scan_ws;
    int32_t* nodes;
    int32_t* tree_off;
    int32_t* waves;
    int32_t* wave_off;
    TaskRec* tasks;
    long long* total_nodes;
    long long* total_waves;
    void* scan_ws;
    double4* fbox;
    uint32_t* cand_l;
    int32_t* cand_r;
    uint32_t* cs_r;  // undecided candidates' sort keys, sorted
    int32_t* cs_l;   // undecided candidates' left ids (bit 31: first step answered)
    uint32_t* ckey;  // undecided candidates' sort keys, in classify order
    void* csort_ws;
    JoinCtl* cctl;
    JoinCtl* ctl;
    void* fin_ws;
    uint32_t* stage_l;
    int32_t* stage_r;
    double2* slab_geo;
    int32_t* slab_n;
    uint16_t* slab_list;
    double4* grid_geo;
    double2* grid_inv;
    uint2* grid_cells;
    RowRec* lrec;
    int2* rinfo;
};

// Forest nodes are bounded by the items and their parents: < n_right * 16/15 + levels * jobs.
static int64_t forest_node_bound(int64_t n_right, int64_t n_jobs) {
    return n_right + n_right / (kFF - 1) + 1 + (int64_t)kFMaxLevels * n_jobs;
}

static BatchWs carve_batch(void* base, int64_t n_jobs, int64_t n_left, int64_t n_right, int64_t n_right_coords,
                           int64_t capacity, int64_t* bytes) {
    Carver c(base);
    BatchWs W;
    W.lbox = c.take<double4>(n_left);
    W.rbox = c.take<double4>(n_right);
    W.lkey = c.take<uint64_t>(n_left);
    W.lperm = c.take<int32_t>(n_left);
    W.rkey = c.take<uint64_t>(n_right);
    W.rperm = c.take<int32_t>(n_right);
    const int64_t nmax = n_left > n_right ? n_left : n_right;
    W.sort_ws = c.take<char>(sort_workspace_bytes_u64(nmax));
    W.jext = c.take<ulonglong4>(n_jobs);
    W.row_mark = c.take<int32_t>(nmax);
    W.row_job = c.take<int32_t>(nmax);
    W.total_rows = c.take<long long>(1);
    W.row_scan_ws = c.take<char>(scan_workspace_bytes(nmax));
    W.nodes = c.take<int32_t>(n_jobs);
    W.tree_off = c.take<int32_t>(n_jobs);
    W.waves = c.take<int32_t>(n_jobs);
    W.wave_off = c.take<int32_t>(n_jobs);
    W.tasks = c.take<TaskRec>(n_left / 64 + n_jobs + 1);  // sum over jobs of ceil(rows / 64)
    W.total_nodes = c.take<long long>(1);
    W.total_waves = c.take<long long>(1);
    W.scan_ws = c.take<char>(scan_workspace_bytes(n_jobs));
    W.fbox = c.take<double4>(forest_node_bound(n_right, n_jobs));
    W.cand_l = c.take<uint32_t>(capacity);
    W.cand_r = c.take<int32_t>(capacity);
    W.cs_r = c.take<uint32_t>(capacity);
    W.cs_l = c.take<int32_t>(capacity);
    W.ckey = c.take<uint32_t>(capacity);
    W.csort_ws = c.take<char>(sort_workspace_bytes_u32(capacity));
    W.cctl = c.take<JoinCtl>(1);
    W.ctl = c.take<JoinCtl>(1);
    W.fin_ws = c.take<char>(finish_workspace_bytes(capacity));
    W.stage_l = nullptr;
    W.stage_r = nullptr;
    if (base) finish_stage(W.fin_ws, capacity, &W.stage_l, &W.stage_r);
    W.slab_geo = c.take<double2>(n_right);
    W.slab_n = c.take<int32_t>(n_right);
    W.slab_list = c.take<uint16_t>(3 * n_right_coords);
    W.grid_geo = c.take<double4>(n_right);
    W.grid_inv = c.take<double2>(n_right);
    W.grid_cells = c.take<uint2>(n_right * kGrid);
    W.lrec = c.take<RowRec>(n_left);
    W.rinfo = c.take<int2>(n_right);
    if (bytes) *bytes = c.bytes;
    return W;
}

static int bits_for_jobs(int64_t n) {
    int b = 0;
    while (b < 31 && ((int64_t)1 << b) < n) ++b;
    return b;
}

int batched_intersects(const Polys& A, int64_t n_left, const Polys& B, int64_t n_right, int64_t n_right_coords,
                       const int64_t* left_job,
                       const int64_t* right_job, int64_t n_jobs, int tile, int64_t capacity, int64_t* out_l,
                       int64_t* out_r, int64_t* d_result, const BatchWs& W, hipStream_t s) {
    ctl_reset(W.cctl, s);
    ctl_reset(W.ctl, s);
    const int waves_per_block = (tile & 0xFF) ? (tile & 0xFF) : 4;
    const int refine_threads = ((tile >> 8) & 0xFFFF) ? ((tile >> 8) & 0xFFFF) : 256;
    const int rbits = bits_for_jobs(n_right);
    const bool by_work = ((tile >> 24) & 1) != 0 && rbits <= 27;
    const bool use_grid = ((tile >> 25) & 1) == 0;
    VS_REQUIRE(waves_per_block >= 1 && waves_per_block <= 4, "batched join: tile waves per block in [1, 4]");
    VS_REQUIRE(refine_threads % 64 == 0 && refine_threads <= 256, "batched join: refine block in {64,..,256}");
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cus = v;
    }
    if (n_jobs > 0 && n_left > 0 && n_right > 0) {
        // bounds; left rows' records (bounds, first vertex, ring info); right rows' ring info
        VS_TRY(polygon_bounds_info(A.x, A.y, A.geom, A.part, A.ring, n_left, reinterpret_cast<double*>(W.lbox),
                                   static_cast<int2*>(nullptr), W.lrec, s));
        VS_TRY(polygon_bounds_info(B.x, B.y, B.geom, B.part, B.ring, n_right, reinterpret_cast<double*>(W.rbox),
                                   W.rinfo, static_cast<RowRec*>(nullptr), s));
        const int jb = bits_for_jobs(n_jobs);
        // right rows: (job, Hilbert) order, forest
        const unsigned jgrid = (unsigned)std::min<int64_t>(ceil_div(n_jobs, 256), (int64_t)cus * 4);
        // job of every row (marks, scan, own mark), extents of the job centres, keys
        auto row_jobs = [&](const int64_t* off, int64_t n) -> int {
            fill_i32(W.row_mark, n, 0, s);
            hipLaunchKernelGGL(job_marks_kernel, dim3(jgrid), dim3(256), 0, s, off, n_jobs, n, W.row_mark);
            VS_LAUNCH_CHECK();
            VS_TRY(exclusive_scan_i32(W.row_mark, W.row_job, n, W.total_rows, W.row_scan_ws, s));
            hipLaunchKernelGGL(add_marks_kernel, dim3(stride_grid(n, 256, 8)), dim3(256), 0, s, W.row_mark, n,
                               W.row_job);
            VS_LAUNCH_CHECK();
            hipLaunchKernelGGL(job_extent_init_kernel, dim3(jgrid), dim3(256), 0, s, n_jobs, W.jext);
            VS_LAUNCH_CHECK();
            return VS_OK;
        };
        VS_TRY(row_jobs(right_job, n_right));
        hipLaunchKernelGGL(job_extent_kernel, dim3(stride_grid(n_right, 256, 4)), dim3(256), 0, s, W.rbox, n_right,
                           W.row_job, W.jext);
        VS_LAUNCH_CHECK();
        hipLaunchKernelGGL(job_keys_kernel, dim3((unsigned)ceil_div(n_right, 256)), dim3(256), 0, s, W.rbox, n_right,
                           W.row_job, W.jext, W.rkey, W.rperm);
        VS_LAUNCH_CHECK();
        VS_TRY(sort_pairs_u64(W.rkey, W.rperm, n_right, 0, kJobKeyBits + jb, W.sort_ws,
                              sort_workspace_bytes_u64(n_right), s));
        hipLaunchKernelGGL(job_sizes_kernel, dim3((unsigned)ceil_div(n_jobs, 256)), dim3(256), 0, s, left_job,
                           right_job, n_jobs, W.nodes, W.waves);
        VS_LAUNCH_CHECK();
        VS_TRY(exclusive_scan_i32(W.nodes, W.tree_off, n_jobs, W.total_nodes, W.scan_ws, s));
        VS_TRY(exclusive_scan_i32(W.waves, W.wave_off, n_jobs, W.total_waves, W.scan_ws, s));
        hipLaunchKernelGGL(row_slab_build_kernel, dim3((unsigned)std::min<int64_t>(n_right, (int64_t)cus * 32)),
                           dim3(64), 0, s, B, n_right, W.rinfo, W.rbox, W.slab_geo, W.slab_n, W.slab_list,
                           3 * n_right_coords, kRowSlabMinEdges[(tile >> 26) & 3]);
        VS_LAUNCH_CHECK();
        if (use_grid) {
            hipLaunchKernelGGL(row_grid_build_kernel,
                               dim3((unsigned)std::min<int64_t>(ceil_div(n_right, kGridRowsPerBlock), (int64_t)cus * 8)),
                               dim3(256), 0, s, B, n_right, W.rinfo, W.rbox, W.grid_geo, W.grid_inv, W.grid_cells);
            VS_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(forest_build_kernel, dim3((unsigned)std::min<int64_t>(n_jobs, 65535)), dim3(256), 0, s,
                           W.rbox, W.rperm, right_job, W.tree_off, n_jobs, W.fbox);
        VS_LAUNCH_CHECK();
        // left rows: (job, Hilbert) order, then the traversal
        VS_TRY(row_jobs(left_job, n_left));
        hipLaunchKernelGGL(job_extent_kernel, dim3(stride_grid(n_left, 256, 4)), dim3(256), 0, s, W.lbox, n_left,
                           W.row_job, W.jext);
        VS_LAUNCH_CHECK();
        hipLaunchKernelGGL(job_keys_kernel, dim3((unsigned)ceil_div(n_left, 256)), dim3(256), 0, s, W.lbox, n_left,
                           W.row_job, W.jext, W.lkey, W.lperm);
        VS_LAUNCH_CHECK();
        VS_TRY(sort_pairs_u64(W.lkey, W.lperm, n_left, 0, kJobKeyBits + jb, W.sort_ws,
                              sort_workspace_bytes_u64(n_left), s));
        VS_TRY(timing_begin_public(s));
        hipLaunchKernelGGL(task_recs_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(n_jobs, 4), (int64_t)cus * 8)),
                           dim3(256), 0, s, left_job, right_job, n_jobs, W.wave_off, W.tree_off, W.tasks);
        VS_LAUNCH_CHECK();
        hipLaunchKernelGGL(forest_query_kernel, dim3((unsigned)(cus * 8)), dim3(64 * waves_per_block), 0, s, W.lbox,
                           W.lperm, W.tasks, W.total_waves, W.fbox, W.rperm, W.cand_l, W.cand_r, capacity, W.cctl);
        VS_LAUNCH_CHECK();
        // candidates decided from the cell grids where possible; the rest
        // grouped by right row (the lanes of a wave then share parcels: same
        // coordinates, same edge count), with the work order the largest
        // parcels' candidates first
        hipLaunchKernelGGL(cand_classify_kernel, dim3(stride_grid(capacity, 256, 8)), dim3(256), 0, s,
                           reinterpret_cast<const uint32_t*>(W.cand_l), W.cand_r, W.cctl, capacity, W.lrec,
                           W.rinfo, A.x, A.y,
                           RowGrids{W.grid_geo, W.grid_inv, W.grid_cells}, use_grid ? 1 : 0, by_work ? 1 : 0, rbits,
                           W.ckey, W.cs_l, W.stage_l, W.stage_r, capacity, W.ctl);
        VS_LAUNCH_CHECK();
        VS_TRY(sort_pairs_into_u32(W.ckey, W.cs_l, W.cs_r, reinterpret_cast<int32_t*>(W.cand_l), capacity, 0,
                                   by_work ? rbits + 5 : rbits, W.csort_ws, sort_workspace_bytes_u32(capacity), s,
                                   &W.cctl->stat[0]));
        const uint32_t rmask = by_work ? (uint32_t)((1ull << rbits) - 1ull) : 0xFFFFFFFFu;
        const unsigned rgrid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(capacity, refine_threads),
                                                                                  (int64_t)cus * 16));
        hipLaunchKernelGGL(batch_refine_kernel, dim3(rgrid), dim3(refine_threads),
                           (unsigned)(refine_threads / 64 * kRingTileBytesPerWave), s, W.cand_l,
                           reinterpret_cast<const int32_t*>(W.cs_r), rmask, &W.cctl->stat[0], capacity, A, B,
                           W.lrec, W.rinfo, RowSlabs{W.slab_geo, W.slab_n, W.slab_list}, W.stage_l, W.stage_r,
                           capacity, W.ctl);
        VS_LAUNCH_CHECK();
        VS_TRY(timing_end_public(s));
    }
    VS_TRY(finish_pairs(W.fin_ws, W.ctl, n_left, capacity, true, out_l, out_r, d_result, s));
    hipLaunchKernelGGL(publish_batch_kernel, dim3(1), dim3(64), 0, s, W.cctl, W.ctl, d_result);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

}  // namespace vs

extern "C" {

// The cell grid of one ring on the host, built as the device builds it.
int vs_host_cell_grid(const double* x, const double* y, int64_t n, const double* box4, double* geo6,
                      uint32_t* cells64) {
    vs::clear_error();
    VS_REQUIRE(x && y && box4 && geo6 && cells64, "vs_host_cell_grid: null argument");
    VS_REQUIRE(n >= 0, "vs_host_cell_grid: negative size");
    vs::GridAxes g;
    uint2 cells[vs::kGrid];
    if (!vs::grid_build_serial(x, y, n, box4[0], box4[1], box4[2], box4[3], g, cells)) return 0;
    const double v[6] = {g.x0, g.y0, g.wx, g.wy, g.ix, g.iy};
    for (int i = 0; i < 6; ++i) geo6[i] = v[i];
    for (int j = 0; j < vs::kGrid; ++j) {
        cells64[2 * j] = cells[j].x;
        cells64[2 * j + 1] = cells[j].y;
    }
    return 1;
}

// grid_decide on the host for ring (x, y, n), n >= 1, with its box
// computed as the device's bounds pass does (min / max, NaN ignored):
// 1 intersects, 0 apart, 2 undecided with the exact test's first step
// skippable, 3 undecided.
int vs_host_cell_decide(const double* geo6, const uint32_t* cells64, const double* x, const double* y, int32_t n) {
    vs::clear_error();
    VS_REQUIRE(geo6 && cells64 && x && y && n >= 1, "vs_host_cell_decide: bad argument");
    const vs::GridAxes g{geo6[0], geo6[1], geo6[2], geo6[3], geo6[4], geo6[5]};
    uint2 cells[vs::kGrid];
    for (int j = 0; j < vs::kGrid; ++j) cells[j] = make_uint2(cells64[2 * j], cells64[2 * j + 1]);
    double b[4] = {x[0], y[0], x[0], y[0]};
    for (int k = 1; k < n; ++k) {
        b[0] = fmin(b[0], x[k]);
        b[1] = fmin(b[1], y[k]);
        b[2] = fmax(b[2], x[k]);
        b[3] = fmax(b[3], y[k]);
    }
    return vs::grid_decide(g, cells, b[0], b[1], b[2], b[3], x[0], y[0], x, y, n);
}

// cell_lo / cell_hi of v along axis 0 (x) or 1 (y) of a grid (geo6 as above).
int vs_host_cell_span(const double* geo6, int32_t axis, double v, int32_t* lo_hi) {
    vs::clear_error();
    VS_REQUIRE(geo6 && lo_hi && (axis == 0 || axis == 1), "vs_host_cell_span: bad argument");
    const double base = geo6[axis], w = geo6[2 + axis], inv = geo6[4 + axis];
    lo_hi[0] = vs::cell_lo(v, base, w, inv);
    lo_hi[1] = vs::cell_hi(v, base, w, inv);
    return VS_OK;
}

int64_t vs_batched_intersects_workspace_bytes(int64_t n_jobs, int64_t n_left, int64_t n_right, int64_t n_right_coords,
                                              int64_t capacity) {
    int64_t b = 0;
    vs::carve_batch(nullptr, n_jobs, n_left, n_right, n_right_coords, capacity, &b);
    return b;
}

int vs_batched_intersects(const double* ax, const double* ay, const int32_t* a_geom, const int32_t* a_part,
                          const int32_t* a_ring, int64_t n_left, const double* bx, const double* by,
                          const int32_t* b_geom, const int32_t* b_part, const int32_t* b_ring, int64_t n_right,
                          int64_t n_right_coords, const int64_t* left_job_offsets, const int64_t* right_job_offsets,
                          int64_t n_jobs, int32_t tile, int64_t capacity, int64_t* out_left, int64_t* out_right,
                          int64_t* d_result, void* workspace, int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    VS_REQUIRE(n_jobs >= 0 && n_jobs < (int64_t)INT32_MAX, "vs_batched_intersects: n_jobs out of range");
    VS_REQUIRE(n_left >= 0 && n_left < (int64_t)INT32_MAX && n_right >= 0 && n_right < (int64_t)INT32_MAX,
               "vs_batched_intersects: row counts out of range");
    VS_REQUIRE(capacity >= 0 && capacity < (int64_t)INT32_MAX, "vs_batched_intersects: capacity out of range");
    VS_REQUIRE(d_result != nullptr, "vs_batched_intersects: null result");
    VS_REQUIRE(n_jobs == 0 || (left_job_offsets && right_job_offsets), "vs_batched_intersects: null job offsets");
    VS_REQUIRE(n_right_coords >= 0 && n_right_coords < (int64_t)INT32_MAX,
               "vs_batched_intersects: n_right_coords out of range");
    int64_t need = 0;
    const vs::BatchWs W = vs::carve_batch(workspace, n_jobs, n_left, n_right, n_right_coords, capacity, &need);
    VS_REQUIRE(workspace_bytes >= need, "vs_batched_intersects: workspace %lld < %lld bytes",
               (long long)workspace_bytes, (long long)need);
    const vs::Polys A = {ax, ay, a_geom, a_part, a_ring};
    const vs::Polys B = {bx, by, b_geom, b_part, b_ring};
    return vs::batched_intersects(A, n_left, B, n_right, n_right_coords, left_job_offsets, right_job_offsets, n_jobs,
                                  tile, capacity,
                                  out_left, out_right, d_result, W, vs::stream_of(stream));
}

}  // extern "C"
