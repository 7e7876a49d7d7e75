// Finish pass shared by every join: staged (left, right) pairs in arbitrary
// order -> lexsorted int64 (left_index, right_index) output.
//
// Replaces the reference's compaction and lexsort of the refined pairs
// (spatial/query_utils.py:867-868 d_left[keep]; spatial/query.py:143-153
// cp.lexsort((right, left))).
//
//   1. LSD radix sort of the staged pairs by left id (u32 keys over
//      ceil(log2 n_left) bits, stable).  The staged count is read on the device
//      (no host sync); an overflowed stage sorts nothing.  Its last pass
//      writes out_l / out_r as int64 directly.
//   2. runs: with sort_right (the default, lexsorted output) the first slot
//      of each left id's run of two or more sorts that run's right ids in
//      out_r, in registers (from an LDS tile loaded and stored coalesced):
//      runs of at most kNetRun by a 19-comparator network, runs of at most
//      kShortRun by a bitonic network.  Longer runs are listed.  (With no sort pass -- at most one left row -- an emit pass
//      converts and sorts instead.)
//   3. long runs: one block per listed run; runs up to kRunTile are bitonic
//      sorted in LDS, longer ones are sorted in kRunTile chunks and merged in
//      global memory (merge by co-rank: each element's output slot is its
//      index in its own half plus its rank in the other half).
//
// Pairs with the same left id come from one join and have distinct right ids,
// but the merge is stable anyway (lower bound from the left half, upper bound
// from the right half).
#include <climits>

#include "vs_join_common.h"

namespace vs {

constexpr int kShortRun = 32;
constexpr int kNetRun = 8;  // runs sorted in registers by fix_runs_kernel
static_assert(kNetRun < kShortRun && (kShortRun & (kShortRun - 1)) == 0, "bitonic runs: a power of two");
constexpr int kFixTile = 1024;  // slots per fix_runs_kernel tile
constexpr int kRunTile = 4096;  // int32 elements sorted in LDS by one block (16 KB)
constexpr int kRunThreads = 256;

struct FinishBufs {
    uint32_t* stage_l;
    int32_t* stage_r;
    void* sort_ws;
    int32_t* long_runs;
};

static FinishBufs carve_finish(Carver& c, int64_t capacity) {
    FinishBufs F;
    F.stage_l = c.take<uint32_t>(capacity);
    F.stage_r = c.take<int32_t>(capacity);
    F.sort_ws = c.take<char>(sort_workspace_bytes_u32(capacity));
    F.long_runs = c.take<int32_t>(capacity / (kShortRun + 1) + 1);
    return F;
}

int64_t finish_workspace_bytes(int64_t capacity) {
    Carver c(nullptr);
    carve_finish(c, capacity);
    return c.bytes;
}

void finish_stage(void* fin_ws, int64_t capacity, uint32_t** stage_l, int32_t** stage_r) {
    Carver c(fin_ws);
    const FinishBufs F = carve_finish(c, capacity);
    *stage_l = F.stage_l;
    *stage_r = F.stage_r;
}

__global__ __launch_bounds__(256) void emit_pairs_kernel(const uint32_t* __restrict__ keys,
                                                        int32_t* __restrict__ vals,
                                                        const JoinCtl* __restrict__ ctl, int64_t capacity,
                                                        int sort_right, int32_t* __restrict__ long_runs,
                                                        unsigned int* __restrict__ n_long,
                                                        int64_t* __restrict__ out_l, int64_t* __restrict__ out_r) {
    const unsigned long long staged = ctl->staged;
    if (staged > (unsigned long long)capacity) return;
    const int64_t n = (int64_t)staged;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const uint32_t k = keys[i];
        out_l[i] = (int64_t)k;
        if (!sort_right) {
            out_r[i] = vals[i];
            continue;
        }
        if (i > 0 && keys[i - 1] == k) continue;  // not the first slot of its run
        int64_t e = i + 1;
        while (e < n && e - i <= kShortRun && keys[e] == k) ++e;
        const int len = (int)(e - i);
        if (len > kShortRun) {
            long_runs[atomicAdd(n_long, 1u)] = (int32_t)i;
            continue;
        }
        int32_t* r = vals + i;
        for (int a = 1; a < len; ++a) {
            const int32_t v = r[a];
            int b = a - 1;
            while (b >= 0 && r[b] > v) {
                r[b + 1] = r[b];
                --b;
            }
            r[b + 1] = v;
        }
        for (int a = 0; a < len; ++a) out_r[i + a] = r[a];
    }
}

// After the sort's last pass wrote out_l / out_r: sorts each short run of
// equal left ids by right id in out_r and lists the long runs.  One block per
// tile of kFixTile slots: the keys (low words) of slots t0 - 1 .. t1 + kShortRun
// and the right ids of t0 .. t1 + kShortRun - 1 are loaded coalesced into LDS;
// each run starting in the tile is sorted in registers by its first slot's
// thread and written back to LDS (its part past t1 straight to out_r); the
// tile is then stored coalesced, except its leading slots that continue a run
// begun before t0 (the previous tile's block owns that run).
__global__ __launch_bounds__(256) void fix_runs_kernel(const int64_t* __restrict__ out_l,
                                                      int64_t* __restrict__ out_r,
                                                      const JoinCtl* __restrict__ ctl, int64_t capacity,
                                                      int32_t* __restrict__ long_runs,
                                                      unsigned int* __restrict__ n_long) {
    constexpr uint32_t kNoKey = 0xFFFFFFFFu;  // left ids are below it (checked on the host)
    __shared__ uint32_t s_key[kFixTile + kShortRun + 2];  // slot t0 - 1 + x
    __shared__ int32_t s_val[kFixTile + kShortRun];       // slot t0 + x
    const unsigned long long staged = ctl->staged;
    if (staged > (unsigned long long)capacity) return;
    const int64_t n = (int64_t)staged;
    for (int64_t t0 = (int64_t)blockIdx.x * kFixTile; t0 < n; t0 += (int64_t)gridDim.x * kFixTile) {
        // Every load of the tile in flight at once, then the LDS stores.
        constexpr int kKeyIt = (kFixTile + kShortRun + 2 + 255) / 256;
        constexpr int kValIt = (kFixTile + kShortRun + 255) / 256;
        uint32_t kr[kKeyIt];
        int32_t vr[kValIt];
#pragma unroll
        for (int m = 0; m < kKeyIt; ++m) {
            const int x = threadIdx.x + m * 256;
            const int64_t p = t0 - 1 + x;
            kr[m] = x < kFixTile + kShortRun + 2 && p >= 0 && p < n ? (uint32_t)out_l[p] : kNoKey;
        }
#pragma unroll
        for (int m = 0; m < kValIt; ++m) {
            const int x = threadIdx.x + m * 256;
            const int64_t p = t0 + x;
            vr[m] = x < kFixTile + kShortRun && p < n ? (int32_t)out_r[p] : 0;
        }
#pragma unroll
        for (int m = 0; m < kKeyIt; ++m) {
            const int x = threadIdx.x + m * 256;
            if (x < kFixTile + kShortRun + 2) s_key[x] = kr[m];
        }
#pragma unroll
        for (int m = 0; m < kValIt; ++m) {
            const int x = threadIdx.x + m * 256;
            if (x < kFixTile + kShortRun) s_val[x] = vr[m];
        }
        __syncthreads();
        for (int x = threadIdx.x; x < kFixTile && t0 + x < n; x += 256) {
            const uint32_t k = s_key[x + 1];
            if (s_key[x] == k || s_key[x + 2] != k) continue;  // not the first slot of a run of two or more
            int32_t w[kShortRun];
            int len = 2;
#pragma unroll
            for (int j = 2; j <= kNetRun; ++j) len += (len == j && s_key[x + 1 + j] == k);
#pragma unroll
            for (int j = 0; j < kNetRun; ++j) w[j] = s_val[x + j];
            if (len <= kNetRun) {
#pragma unroll
                for (int j = 2; j < kNetRun; ++j) w[j] = j < len ? w[j] : INT32_MAX;
                // Optimal 8-input network, depth 6 (Knuth, TAOCP vol. 3, 5.3.4).
                constexpr int kNet[19][2] = {{0, 2}, {1, 3}, {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6},
                                             {3, 7}, {0, 1}, {2, 3}, {4, 5}, {6, 7}, {2, 4}, {3, 5},
                                             {1, 4}, {3, 6}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
                for (int c = 0; c < 19; ++c) {
                    const int32_t a = w[kNet[c][0]], b = w[kNet[c][1]];
                    w[kNet[c][0]] = min(a, b);
                    w[kNet[c][1]] = max(a, b);
                }
            } else {
                // len == kNetRun + 1 so far: slots x .. x + kNetRun hold k.
#pragma unroll
                for (int j = kNetRun + 1; j <= kShortRun; ++j) len += (len == j && s_key[x + 1 + j] == k);
                if (len > kShortRun) {
                    long_runs[atomicAdd(n_long, 1u)] = (int32_t)(t0 + x);
                    continue;
                }
#pragma unroll
                for (int j = kNetRun; j < kShortRun; ++j) w[j] = j < len ? s_val[x + j] : INT32_MAX;
                // Bitonic network over kShortRun values, every index fixed at compile time.
#pragma unroll
                for (int size = 2; size <= kShortRun; size <<= 1) {
#pragma unroll
                    for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
                        for (int u = 0; u < kShortRun; ++u) {
                            const int v = u ^ stride;
                            if (v > u) {
                                const int32_t a = w[u], b = w[v];
                                const bool up = (u & size) == 0;
                                w[u] = up ? min(a, b) : max(a, b);
                                w[v] = up ? max(a, b) : min(a, b);
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < kShortRun; ++j) {
                if (j < len) {
                    if (x + j < kFixTile) s_val[x + j] = w[j];
                    else out_r[t0 + x + j] = w[j];
                }
            }
        }
        __syncthreads();
        for (int x = threadIdx.x; x < kFixTile && t0 + x < n; x += 256)
            if (s_key[x + 1] != s_key[0]) out_r[t0 + x] = s_val[x];
        __syncthreads();
    }
}

__device__ __forceinline__ int64_t lower_bound_i32(const int32_t* a, int64_t lo, int64_t hi, int32_t x) {
    while (lo < hi) {
        const int64_t m = (lo + hi) >> 1;
        if (a[m] < x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

__device__ __forceinline__ int64_t upper_bound_i32(const int32_t* a, int64_t lo, int64_t hi, int32_t x) {
    while (lo < hi) {
        const int64_t m = (lo + hi) >> 1;
        if (a[m] <= x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// Ascending bitonic sort of t[0, P), P a power of two, by the whole block.
__device__ void block_bitonic_sort(int32_t* t, int P) {
    for (int size = 2; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < P / 2; i += kRunThreads) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool asc = (lo & size) == 0;
                const int32_t a = t[lo], b = t[hi];
                if ((a > b) == asc) {
                    t[lo] = b;
                    t[hi] = a;
                }
            }
            __syncthreads();
        }
    }
}

// WIDE: the sorted pairs are in out_l / out_r (int64; wkeys = out_l) and vals
// is scratch; else in keys / vals (u32 / int32).
template <bool WIDE>
__global__ __launch_bounds__(kRunThreads) void long_runs_kernel(const uint32_t* __restrict__ keys,
                                                               const int64_t* __restrict__ wkeys,
                                                               int32_t* __restrict__ vals,
                                                               int32_t* __restrict__ tmp,
                                                               const JoinCtl* __restrict__ ctl, int64_t capacity,
                                                               const int32_t* __restrict__ long_runs,
                                                               const unsigned int* __restrict__ n_long,
                                                               int64_t* __restrict__ out_r) {
    auto key_at = [&](int64_t i) -> int64_t { return WIDE ? wkeys[i] : (int64_t)keys[i]; };
    auto val_at = [&](int64_t i) -> int32_t { return WIDE ? (int32_t)out_r[i] : vals[i]; };
    __shared__ int32_t t[kRunTile];
    __shared__ long long s_end;
    const unsigned long long staged = ctl->staged;
    if (staged > (unsigned long long)capacity) return;
    const int64_t n = (int64_t)staged;
    const unsigned int nl = *n_long;
    for (unsigned int j = blockIdx.x; j < nl; j += gridDim.x) {
        const int64_t start = long_runs[j];
        if (threadIdx.x == 0) {
            // end of the run: gallop from a slot known to hold the key, then bisect
            const int64_t k = key_at(start);
            int64_t lo = start + kShortRun, step = 64, hi = lo + step;
            while (hi < n && key_at(hi) == k) {
                lo = hi;
                step *= 2;
                hi = lo + step;
            }
            if (hi > n) hi = n;
            while (hi - lo > 1) {  // keys[lo] == k; hi == n or keys[hi] != k
                const int64_t m = (lo + hi) >> 1;
                if (key_at(m) == k) lo = m;
                else hi = m;
            }
            s_end = hi;
        }
        __syncthreads();
        const int64_t L = s_end - start;
        __syncthreads();  // s_end is rewritten by the next run
        if (L <= kRunTile) {
            int P = 64;
            while (P < L) P <<= 1;
            for (int i = threadIdx.x; i < P; i += kRunThreads) t[i] = i < L ? val_at(start + i) : INT_MAX;
            __syncthreads();
            block_bitonic_sort(t, P);
            for (int i = threadIdx.x; i < L; i += kRunThreads) out_r[start + i] = t[i];
            __syncthreads();
            continue;
        }
        // sorted chunks of kRunTile, in place
        for (int64_t c0 = 0; c0 < L; c0 += kRunTile) {
            const int64_t cl = (L - c0) < kRunTile ? (L - c0) : kRunTile;
            for (int i = threadIdx.x; i < kRunTile; i += kRunThreads)
                t[i] = i < cl ? val_at(start + c0 + i) : INT_MAX;
            __syncthreads();
            block_bitonic_sort(t, kRunTile);
            for (int i = threadIdx.x; i < cl; i += kRunThreads) vals[start + c0 + i] = t[i];
            __syncthreads();
        }
        // merge passes: sorted blocks of width w in src -> width 2w in dst
        int32_t* src = vals + start;
        int32_t* dst = tmp + start;
        for (int64_t w = kRunTile; w < L; w <<= 1) {
            for (int64_t x = threadIdx.x; x < L; x += kRunThreads) {
                const int64_t a0 = x / (2 * w) * (2 * w);
                const int64_t a1 = (a0 + w) < L ? (a0 + w) : L;
                const int64_t b1 = (a0 + 2 * w) < L ? (a0 + 2 * w) : L;
                const int32_t v = src[x];
                const int64_t pos = x < a1 ? (x - a0) + (lower_bound_i32(src, a1, b1, v) - a1)
                                           : (x - a1) + (upper_bound_i32(src, a0, a1, v) - a0);
                dst[a0 + pos] = v;
            }
            __syncthreads();
            int32_t* sw = src;
            src = dst;
            dst = sw;
        }
        for (int64_t x = threadIdx.x; x < L; x += kRunThreads) out_r[start + x] = src[x];
        __syncthreads();
    }
}

__global__ void publish_result_kernel(const JoinCtl* __restrict__ ctl, int64_t* __restrict__ result) {
    if (threadIdx.x == 0) {
        result[0] = (int64_t)ctl->staged;
        result[1] = (int64_t)ctl->staged;
        result[2] = (int64_t)ctl->stat[0];
        result[3] = (int64_t)ctl->stat[1];
    }
}

static int bits_for(int64_t n) {
    int b = 0;
    while (b < 32 && ((int64_t)1 << b) < n) ++b;
    return b;
}

int finish_pairs(void* fin_ws, JoinCtl* ctl, int64_t n_left, int64_t capacity, bool sort_right, int64_t* out_l,
                 int64_t* out_r, int64_t* d_result, hipStream_t s) {
    VS_REQUIRE(capacity >= 0 && capacity < (int64_t)INT32_MAX, "join capacity %lld out of range [0, 2^31-1)",
               (long long)capacity);
    VS_REQUIRE(n_left >= 0 && n_left < (int64_t)UINT32_MAX, "left row count %lld out of range [0, 2^32-1)",
               (long long)n_left);
    Carver c(fin_ws);
    const FinishBufs F = carve_finish(c, capacity);
    const int nbits = bits_for(n_left);
    SortResult<uint32_t> r;
    if (capacity > 1 && nbits > 0) {  // the sort's last pass writes the output pairs
        VS_TRY(sort_pairs_u32_dev_wide(F.stage_l, F.stage_r, capacity, &ctl->staged, 0, nbits, F.sort_ws,
                                       sort_workspace_bytes_u32(capacity), s, out_l, out_r, &r));
        if (sort_right) {
            const unsigned grid =
                (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(capacity, kFixTile), 65536));
            hipLaunchKernelGGL(fix_runs_kernel, dim3(grid), dim3(256), 0, s, out_l, out_r, ctl, capacity,
                               F.long_runs, &ctl->n_long);
            VS_LAUNCH_CHECK();
            hipLaunchKernelGGL(long_runs_kernel<true>, dim3(512), dim3(kRunThreads), 0, s, nullptr, out_l, r.vals,
                               r.alt_vals, ctl, capacity, F.long_runs, &ctl->n_long, out_r);
            VS_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(publish_result_kernel, dim3(1), dim3(64), 0, s, ctl, d_result);
        VS_LAUNCH_CHECK();
        return VS_OK;
    }
    VS_TRY(sort_pairs_u32_dev(F.stage_l, F.stage_r, capacity, &ctl->staged, 0, nbits, F.sort_ws,
                              sort_workspace_bytes_u32(capacity), s, &r));
    if (capacity > 0) {
        const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(capacity, 256), 65536));
        hipLaunchKernelGGL(emit_pairs_kernel, dim3(grid), dim3(256), 0, s, r.keys, r.vals, ctl, capacity,
                           (int)sort_right, F.long_runs, &ctl->n_long, out_l, out_r);
        VS_LAUNCH_CHECK();
        if (sort_right) {
            hipLaunchKernelGGL(long_runs_kernel<false>, dim3(512), dim3(kRunThreads), 0, s, r.keys, nullptr, r.vals,
                               r.alt_vals, ctl, capacity, F.long_runs, &ctl->n_long, out_r);
            VS_LAUNCH_CHECK();
        }
    }
    hipLaunchKernelGGL(publish_result_kernel, dim3(1), dim3(64), 0, s, ctl, d_result);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

}  // namespace vs
