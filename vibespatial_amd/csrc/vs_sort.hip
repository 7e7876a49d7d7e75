// Device-wide exclusive scan and LSD radix sort of (key, int32 value) pairs.
//
// Replaces the CCCL primitives the reference JITs for this path:
//   exclusive_sum  src/vibespatial/cuda/cccl_primitives.py:596
//   sort_pairs     src/vibespatial/cuda/cccl_primitives.py:1054 (used by
//                  src/vibespatial/spatial/indexing.py:1296-1322)
//
// Design (gfx950, wave64):
//   scan   tile = 256 threads x 16 items; reduce -> single-block scan of tile
//          sums (int64) -> downsweep.  No decoupled look-back, so no inter-block
//          spin waits.
//   sort   8-bit digits, three launches per pass: LDS histogram per tile,
//          device scan of the digit-major histogram, stable scatter.  The
//          scatter ranks keys inside each wave with 8 ballots per item (peer
//          mask of equal digits), combines waves through LDS counters, writes the
//          tile into LDS in digit order and then streams it out so each digit
//          run is written contiguously.
#include "vs_internal.h"

namespace vs {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

// Block exclusive scan of one int per thread (kScanThreads threads).
__device__ __forceinline__ long long block_exclusive_scan_i64(long long v, long long* lds4,
                                                              long long& total) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    long long inc = wave_inclusive_scan_i64(v);
    if (l == 63) lds4[w] = inc;
    __syncthreads();
    long long off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kScanThreads / 64; ++k) {
        const long long t = lds4[k];
        off += (k < w) ? t : 0;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return off + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const int32_t* __restrict__ in,
                                                                  int64_t n,
                                                                  long long* __restrict__ part) {
    __shared__ long long lds[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    long long s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t idx = base + (int64_t)i * kScanThreads + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    long long tot;
    block_exclusive_scan_i64(s, lds, tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// One block of 1024 threads scans the tile sums in place (exclusive) and
// writes the grand total.
__global__ __launch_bounds__(1024) void scan_partials_kernel(long long* __restrict__ part, int64_t nb,
                                                             long long* __restrict__ total_out) {
    __shared__ long long lds[16];
    __shared__ long long carry_s;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    for (int64_t b0 = 0; b0 < nb; b0 += 1024) {
        const int64_t i = b0 + threadIdx.x;
        const long long v = i < nb ? part[i] : 0;
        const long long inc = wave_inclusive_scan_i64(v);
        if (l == 63) lds[w] = inc;
        __syncthreads();
        long long off = 0, tot = 0;
        for (int k = 0; k < 16; ++k) {
            off += (k < w) ? lds[k] : 0;
            tot += lds[k];
        }
        const long long carry = carry_s;
        if (i < nb) part[i] = carry + off + inc - v;
        __syncthreads();
        if (threadIdx.x == 0) carry_s = carry + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry_s;
}

// Downsweep: thread t owns 16 consecutive elements of the tile.
// A thread's kScanItems consecutive ints, loaded / stored as 16-byte
// accesses where its run is whole and aligned (else element by element).
__device__ __forceinline__ void scan_load(const int32_t* in, int64_t base, int64_t n, int (&v)[kScanItems]) {
    if (base + kScanItems <= n && (reinterpret_cast<uintptr_t>(in + base) & 15) == 0) {
#pragma unroll
        for (int j = 0; j < kScanItems / 4; ++j) {
            const int4 q = reinterpret_cast<const int4*>(in + base)[j];
            v[4 * j] = q.x;
            v[4 * j + 1] = q.y;
            v[4 * j + 2] = q.z;
            v[4 * j + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i) v[i] = (base + i < n) ? in[base + i] : 0;
    }
}

// Writes the exclusive prefix of v starting at run.
__device__ __forceinline__ void scan_store(int32_t* out, int64_t base, int64_t n, const int (&v)[kScanItems],
                                           long long run) {
    int o[kScanItems];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        o[i] = (int32_t)run;
        run += v[i];
    }
    if (base + kScanItems <= n && (reinterpret_cast<uintptr_t>(out + base) & 15) == 0) {
#pragma unroll
        for (int j = 0; j < kScanItems / 4; ++j)
            reinterpret_cast<int4*>(out + base)[j] = make_int4(o[4 * j], o[4 * j + 1], o[4 * j + 2], o[4 * j + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < kScanItems; ++i)
            if (base + i < n) out[base + i] = o[i];
    }
}

__global__ __launch_bounds__(kScanThreads) void scan_down_kernel(const int32_t* __restrict__ in,
                                                                int32_t* __restrict__ out, int64_t n,
                                                                const long long* __restrict__ part) {
    __shared__ long long lds[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int v[kScanItems];
    scan_load(in, base, n, v);
    long long s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) s += v[i];
    long long tot;
    const long long run = block_exclusive_scan_i64(s, lds, tot) + part[blockIdx.x];
    scan_store(out, base, n, v, run);
}

// Downsweep that forms its own tile prefix from the tile sums (no separate
// pass over them): for at most kFusedScanTiles tiles each block sums the
// sums of the tiles before it; the last block writes the grand total.
constexpr int kFusedScanTiles = 1024;

__global__ __launch_bounds__(kScanThreads) void scan_down_fused_kernel(const int32_t* __restrict__ in,
                                                                      int32_t* __restrict__ out, int64_t n,
                                                                      const long long* __restrict__ part,
                                                                      long long* __restrict__ total_out) {
    __shared__ long long lds[4];
    long long p = 0;
    for (int t = threadIdx.x; t < (int)blockIdx.x; t += kScanThreads) p += part[t];
    long long prefix;  // the block total of p, the same in every thread
    block_exclusive_scan_i64(p, lds, prefix);
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int v[kScanItems];
    scan_load(in, base, n, v);
    long long sum = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) sum += v[i];
    long long tot;
    const long long run = block_exclusive_scan_i64(sum, lds, tot) + prefix;
    scan_store(out, base, n, v, run);
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total_out = prefix + tot;
}

__global__ void zero_total_kernel(long long* __restrict__ p) {
    if (threadIdx.x == 0) *p = 0;
}

int64_t scan_workspace_bytes(int64_t n) {
    Carver c(nullptr);
    c.take<long long>(ceil_div(n, kScanTile) + 1);
    return c.bytes;
}

// out[i] = sum(in[0..i)); *d_total (device, int64) = sum(in).  Values must fit
// int32 per element; the running sum is formed in int64 and narrowed on store
// (callers check the total against INT32_MAX before trusting offsets).
int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, long long* d_total, void* ws,
                       hipStream_t s) {
    const int64_t nb = ceil_div(n, kScanTile);
    Carver c(ws);
    long long* part = c.take<long long>(nb + 1);
    if (n == 0) {
        // a kernel, not a memset: callers may capture this in a HIP graph
        if (d_total) {
            hipLaunchKernelGGL(zero_total_kernel, dim3(1), dim3(64), 0, s, d_total);
            VS_LAUNCH_CHECK();
        }
        return VS_OK;
    }
    hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, part);
    VS_LAUNCH_CHECK();
    if (nb <= kFusedScanTiles) {
        hipLaunchKernelGGL(scan_down_fused_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, out, n, part,
                           d_total);
        VS_LAUNCH_CHECK();
        return VS_OK;
    }
    hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(1024), 0, s, part, nb, d_total);
    VS_LAUNCH_CHECK();
    hipLaunchKernelGGL(scan_down_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, out, n, part);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

// ---------------------------------------------------------------------------
// Radix sort
// ---------------------------------------------------------------------------
constexpr int kRsThreads = 256;
constexpr int kRsWaves = kRsThreads / 64;
constexpr int kRsItems = 16;
constexpr int kRsTile = kRsThreads * kRsItems;
constexpr int kRsBins = 256;

// Valid element count: n, or the device count *d_n when given.  A device count
// above n means the producer overflowed its buffer: nothing is sorted.
__device__ __forceinline__ int64_t device_count(int64_t n, const unsigned long long* d_n) {
    if (!d_n) return n;
    const unsigned long long c = *d_n;
    return c <= (unsigned long long)n ? (int64_t)c : 0;
}

template <class K>
__global__ __launch_bounds__(kRsThreads) void rs_hist_kernel(const K* __restrict__ keys, int64_t n,
                                                            const unsigned long long* __restrict__ d_n,
                                                            int shift, int dmask, int32_t* __restrict__ hist,
                                                            int nblocks) {
    n = device_count(n, d_n);
    __shared__ int cnt[kRsWaves][kRsBins];
    const int w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kRsWaves * kRsBins; i += kRsThreads) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    K kk[kRsItems];
#pragma unroll
    for (int i = 0; i < kRsItems; ++i) {
        const int64_t idx = base + (int64_t)i * kRsThreads + threadIdx.x;
        kk[i] = idx < n ? keys[idx] : (K)0;
    }
    // A wave whose keys share one digit adds them with one LDS atomic (a
    // join's staged keys arrive in long runs of one value, whose lanes would
    // serialise on one bin); other waves add per lane.  Valid lanes are a
    // prefix of the wave, so lane 0 is valid whenever any lane is.
#pragma unroll
    for (int i = 0; i < kRsItems; ++i) {
        const int64_t idx = base + (int64_t)i * kRsThreads + threadIdx.x;
        const bool valid = idx < n;
        const int d = (int)((kk[i] >> shift) & dmask);
        const int d0 = __builtin_amdgcn_readfirstlane(d);
        const uint64_t vm = __ballot(valid);
        const uint64_t same = __ballot(valid && d == d0);
        if (same == vm) {
            if ((threadIdx.x & 63) == 0 && vm) atomicAdd(&cnt[w][d0], __popcll(vm));
        } else if (valid) {
            atomicAdd(&cnt[w][d], 1);
        }
    }
    __syncthreads();
    const int d = threadIdx.x;
    int t = 0;
#pragma unroll
    for (int k = 0; k < kRsWaves; ++k) t += cnt[k][d];
    hist[(int64_t)d * nblocks + blockIdx.x] = t;
}

// WIDE: the pass writes its keys and values as int64 (a sort's last pass
// writing a join's output pairs directly).
template <class K, bool WIDE>
__global__ __launch_bounds__(kRsThreads) void rs_scatter_kernel(
    const K* __restrict__ keys_in, const int32_t* __restrict__ vals_in, void* __restrict__ keys_out,
    void* __restrict__ vals_out, int64_t n, const unsigned long long* __restrict__ d_n, int shift,
    int dmask, const int32_t* __restrict__ hist_scan, int nblocks) {
    n = device_count(n, d_n);
    if ((int64_t)blockIdx.x * kRsTile >= n) return;  // whole tile past the valid count (block-uniform)
    __shared__ int wave_cnt[kRsWaves][kRsBins];
    __shared__ int digit_block_off[kRsBins];
    __shared__ long long digit_global[kRsBins];
    __shared__ long long lds4[4];
    __shared__ K skeys[kRsTile];
    __shared__ int32_t svals[kRsTile];

    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    for (int i = tid; i < kRsWaves * kRsBins; i += kRsThreads) (&wave_cnt[0][0])[i] = 0;
    digit_global[tid] = hist_scan[(int64_t)tid * nblocks + blockIdx.x];
    __syncthreads();

    const int64_t base = (int64_t)blockIdx.x * kRsTile + (int64_t)w * (64 * kRsItems);
    const uint64_t lt = lanemask_lt();
    K k[kRsItems];
    int32_t v[kRsItems];
    int rank[kRsItems];
    // every load of the tile in flight before the ranking (its wave fences
    // would otherwise order each load after the previous item's ranking)
#pragma unroll
    for (int i = 0; i < kRsItems; ++i) {
        const int64_t idx = base + (int64_t)i * 64 + lane;
        const bool valid = idx < n;
        k[i] = valid ? keys_in[idx] : (K)0;
        v[i] = valid ? (vals_in ? vals_in[idx] : (int32_t)idx) : 0;
    }
#pragma unroll
    for (int i = 0; i < kRsItems; ++i) {
        const int64_t idx = base + (int64_t)i * 64 + lane;
        const bool valid = idx < n;
        const int d = (int)((k[i] >> shift) & dmask);
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        int cur = 0;
        if (valid) cur = wave_cnt[w][d];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        rank[i] = cur + __popcll(m & lt);
        if (valid && (m & lt) == 0) wave_cnt[w][d] = cur + __popcll(m);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: exclusive over waves, then exclusive over digits
    {
        const int d = tid;
        int s = 0;
#pragma unroll
        for (int ww = 0; ww < kRsWaves; ++ww) {
            const int c = wave_cnt[ww][d];
            wave_cnt[ww][d] = s;
            s += c;
        }
        long long tot;
        const long long off = block_exclusive_scan_i64(s, lds4, tot);
        digit_block_off[d] = (int)off;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kRsItems; ++i) {
        const int64_t idx = base + (int64_t)i * 64 + lane;
        if (idx < n) {
            const int d = (int)((k[i] >> shift) & dmask);
            const int pos = digit_block_off[d] + wave_cnt[w][d] + rank[i];
            skeys[pos] = k[i];
            svals[pos] = v[i];
        }
    }
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * kRsTile;
    const int tile_n = (int)((n - t0) < kRsTile ? (n - t0) : kRsTile);
    for (int p = tid; p < tile_n; p += kRsThreads) {
        const K kk = skeys[p];
        const int d = (int)((kk >> shift) & dmask);
        const int64_t g = digit_global[d] + (p - digit_block_off[d]);
        if (WIDE) {
            static_cast<int64_t*>(keys_out)[g] = (int64_t)kk;
            static_cast<int64_t*>(vals_out)[g] = (int64_t)svals[p];
        } else {
            static_cast<K*>(keys_out)[g] = kk;
            static_cast<int32_t*>(vals_out)[g] = svals[p];
        }
    }
}

template <class K>
int64_t sort_workspace_bytes_t(int64_t n) {
    const int64_t nb = ceil_div(n > 0 ? n : 1, kRsTile);
    Carver c(nullptr);
    c.take<K>(n);
    c.take<int32_t>(n);
    c.take<int32_t>(nb * kRsBins);
    c.bytes += scan_workspace_bytes(nb * kRsBins) + 256;
    return c.bytes;
}

// Sorts (keys, vals) by bits [begin_bit, end_bit) of the key, stable, and
// returns the buffers that hold the result (the inputs or the workspace's
// alternates, by the parity of the pass count) and the other pair.
// n: element capacity; d_n (device, optional): see device_count.
// wide_l / wide_r (both or neither): the last pass writes the sorted keys and
// values there as int64 instead of into a K / int32 buffer; then res->keys is
// null and res->vals, res->alt_vals (int32) and res->alt_keys (K), n elements
// each, are free for the caller.
template <class K>
int sort_pairs_nocopy(K* keys, int32_t* vals, int64_t n, const unsigned long long* d_n, int begin_bit, int end_bit,
                      void* ws, int64_t ws_bytes, hipStream_t s, SortResult<K>* res, int64_t* wide_l = nullptr,
                      int64_t* wide_r = nullptr) {
    VS_REQUIRE(n >= 0 && n < (int64_t)INT32_MAX, "sort_pairs: n=%lld out of range", (long long)n);
    VS_REQUIRE(begin_bit >= 0 && end_bit <= (int)(8 * sizeof(K)) && begin_bit <= end_bit,
               "sort_pairs: bad bit range [%d, %d)", begin_bit, end_bit);
    VS_REQUIRE(ws_bytes >= sort_workspace_bytes_t<K>(n), "sort_pairs: workspace too small");
    const int nb = (int)ceil_div(n > 0 ? n : 1, kRsTile);
    Carver c(ws);
    K* kalt = c.take<K>(n);
    int32_t* valt = c.take<int32_t>(n);
    int32_t* hist = c.take<int32_t>((int64_t)nb * kRsBins);
    void* scan_ws = c.take<char>(scan_workspace_bytes((int64_t)nb * kRsBins));
    K* ki = keys;
    int32_t* vi = vals;
    K* ko = kalt;
    int32_t* vo = valt;
    const bool wide = wide_l != nullptr && n > 1 && begin_bit < end_bit;
    if (n > 1) {
        for (int shift = begin_bit; shift < end_bit; shift += 8) {
            const int width = end_bit - shift < 8 ? end_bit - shift : 8;
            const int dmask = (1 << width) - 1;
            hipLaunchKernelGGL(rs_hist_kernel<K>, dim3(nb), dim3(kRsThreads), 0, s, ki, n, d_n, shift, dmask, hist,
                               nb);
            VS_LAUNCH_CHECK();
            VS_TRY(exclusive_scan_i32(hist, hist, (int64_t)nb * kRsBins, nullptr, scan_ws, s));
            if (wide && shift + 8 >= end_bit) {
                hipLaunchKernelGGL((rs_scatter_kernel<K, true>), dim3(nb), dim3(kRsThreads), 0, s, ki, vi, wide_l,
                                   wide_r, n, d_n, shift, dmask, hist, nb);
                VS_LAUNCH_CHECK();
                res->keys = nullptr;
                res->vals = vi;  // read by this pass, free once it ran
                res->alt_keys = ko;
                res->alt_vals = vo;
                return VS_OK;
            }
            hipLaunchKernelGGL((rs_scatter_kernel<K, false>), dim3(nb), dim3(kRsThreads), 0, s, ki, vi, ko, vo, n,
                               d_n, shift, dmask, hist, nb);
            VS_LAUNCH_CHECK();
            K* tk = ki; ki = ko; ko = tk;
            int32_t* tv = vi; vi = vo; vo = tv;
        }
    }
    res->keys = ki;
    res->vals = vi;
    res->alt_keys = ko;
    res->alt_vals = vo;
    return VS_OK;
}

__global__ void iota_i32_kernel(int32_t* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = (int32_t)i;
}

// Sorts (in_keys, in_vals) into (out_keys, out_vals), stable, by bits
// [begin_bit, end_bit), with no copy: with an odd pass count the first pass
// writes the outputs, with an even one the workspace alternates, so the last
// pass always lands in the outputs.  The inputs are only read; they must not
// alias the outputs; in_vals == nullptr: each key's value is its index.  n:
// element capacity; d_n (device, optional): see device_count (an overflowed
// count sorts nothing).
template <class K>
int sort_pairs_into(const K* in_keys, const int32_t* in_vals, K* out_keys, int32_t* out_vals, int64_t n,
                    const unsigned long long* d_n, int begin_bit, int end_bit, void* ws, int64_t ws_bytes,
                    hipStream_t s) {
    VS_REQUIRE(n >= 0 && n < (int64_t)INT32_MAX, "sort_pairs: n=%lld out of range", (long long)n);
    VS_REQUIRE(begin_bit >= 0 && end_bit <= (int)(8 * sizeof(K)) && begin_bit <= end_bit,
               "sort_pairs: bad bit range [%d, %d)", begin_bit, end_bit);
    VS_REQUIRE(ws_bytes >= sort_workspace_bytes_t<K>(n), "sort_pairs: workspace too small");
    VS_REQUIRE(in_keys != out_keys && in_vals != out_vals, "sort_pairs_into: inputs alias outputs");
    const int passes = (end_bit - begin_bit + 7) / 8;
    if (n <= 1 || passes == 0) {  // nothing to order: copy all n slots (a device count only limits the valid ones)
        if (n > 0) {
            VS_CHECK_HIP(hipMemcpyAsync(out_keys, in_keys, sizeof(K) * n, hipMemcpyDeviceToDevice, s));
            if (in_vals) {
                VS_CHECK_HIP(hipMemcpyAsync(out_vals, in_vals, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
            } else {
                hipLaunchKernelGGL(iota_i32_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, out_vals, n);
                VS_LAUNCH_CHECK();
            }
        }
        return VS_OK;
    }
    const int nb = (int)ceil_div(n, kRsTile);
    Carver c(ws);
    K* kalt = c.take<K>(n);
    int32_t* valt = c.take<int32_t>(n);
    int32_t* hist = c.take<int32_t>((int64_t)nb * kRsBins);
    void* scan_ws = c.take<char>(scan_workspace_bytes((int64_t)nb * kRsBins));
    const K* ki = in_keys;
    const int32_t* vi = in_vals;
    K* ko = (passes & 1) ? out_keys : kalt;
    int32_t* vo = (passes & 1) ? out_vals : valt;
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        const int width = end_bit - shift < 8 ? end_bit - shift : 8;
        const int dmask = (1 << width) - 1;
        hipLaunchKernelGGL(rs_hist_kernel<K>, dim3(nb), dim3(kRsThreads), 0, s, ki, n, d_n, shift, dmask, hist, nb);
        VS_LAUNCH_CHECK();
        VS_TRY(exclusive_scan_i32(hist, hist, (int64_t)nb * kRsBins, nullptr, scan_ws, s));
        hipLaunchKernelGGL((rs_scatter_kernel<K, false>), dim3(nb), dim3(kRsThreads), 0, s, ki, vi, ko, vo, n, d_n,
                           shift, dmask, hist, nb);
        VS_LAUNCH_CHECK();
        ki = ko;
        vi = vo;
        ko = ko == out_keys ? kalt : out_keys;
        vo = vo == out_vals ? valt : out_vals;
    }
    return VS_OK;
}

// In-place form: copies the result back when it ends in the alternates.
template <class K>
int sort_pairs_t(K* keys, int32_t* vals, int64_t n, const unsigned long long* d_n, int begin_bit, int end_bit,
                 void* ws, int64_t ws_bytes, hipStream_t s) {
    SortResult<K> r;
    VS_TRY(sort_pairs_nocopy<K>(keys, vals, n, d_n, begin_bit, end_bit, ws, ws_bytes, s, &r));
    if (r.keys != keys) {
        VS_CHECK_HIP(hipMemcpyAsync(keys, r.keys, sizeof(K) * n, hipMemcpyDeviceToDevice, s));
        VS_CHECK_HIP(hipMemcpyAsync(vals, r.vals, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
    }
    return VS_OK;
}

int64_t sort_workspace_bytes_u64(int64_t n) { return sort_workspace_bytes_t<uint64_t>(n); }
int64_t sort_workspace_bytes_u32(int64_t n) { return sort_workspace_bytes_t<uint32_t>(n); }
int sort_pairs_u64(uint64_t* k, int32_t* v, int64_t n, int b0, int b1, void* ws, int64_t wb, hipStream_t s) {
    return sort_pairs_t<uint64_t>(k, v, n, nullptr, b0, b1, ws, wb, s);
}
int sort_pairs_u32(uint32_t* k, int32_t* v, int64_t n, int b0, int b1, void* ws, int64_t wb, hipStream_t s) {
    return sort_pairs_t<uint32_t>(k, v, n, nullptr, b0, b1, ws, wb, s);
}
int sort_pairs_into_u32(const uint32_t* ik, const int32_t* iv, uint32_t* ok, int32_t* ov, int64_t n, int b0, int b1,
                        void* ws, int64_t wb, hipStream_t s, const unsigned long long* d_n) {
    return sort_pairs_into<uint32_t>(ik, iv, ok, ov, n, d_n, b0, b1, ws, wb, s);
}
int sort_pairs_into_u64(const uint64_t* ik, const int32_t* iv, uint64_t* ok, int32_t* ov, int64_t n, int b0, int b1,
                        void* ws, int64_t wb, hipStream_t s, const unsigned long long* d_n) {
    return sort_pairs_into<uint64_t>(ik, iv, ok, ov, n, d_n, b0, b1, ws, wb, s);
}
int sort_pairs_u32_dev(uint32_t* k, int32_t* v, int64_t n, const unsigned long long* d_n, int b0, int b1, void* ws,
                       int64_t wb, hipStream_t s, SortResult<uint32_t>* res) {
    return sort_pairs_nocopy<uint32_t>(k, v, n, d_n, b0, b1, ws, wb, s, res);
}
int sort_pairs_u32_dev_wide(uint32_t* k, int32_t* v, int64_t n, const unsigned long long* d_n, int b0, int b1,
                            void* ws, int64_t wb, hipStream_t s, int64_t* out_k, int64_t* out_v,
                            SortResult<uint32_t>* res) {
    return sort_pairs_nocopy<uint32_t>(k, v, n, d_n, b0, b1, ws, wb, s, res, out_k, out_v);
}

}  // namespace vs

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int64_t vs_scan_workspace_bytes(int64_t n) { return vs::scan_workspace_bytes(n); }

int vs_exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, int64_t* d_total,
                          void* workspace, int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    VS_REQUIRE(n >= 0, "vs_exclusive_scan_i32: negative n");
    VS_REQUIRE(workspace_bytes >= vs::scan_workspace_bytes(n), "vs_exclusive_scan_i32: workspace too small");
    return vs::exclusive_scan_i32(in, out, n, reinterpret_cast<long long*>(d_total), workspace,
                                  vs::stream_of(stream));
}

int64_t vs_sort_workspace_bytes(int64_t n, int32_t key_bytes) {
    return key_bytes == 4 ? vs::sort_workspace_bytes_u32(n) : vs::sort_workspace_bytes_u64(n);
}

int vs_sort_pairs_u64(uint64_t* keys, int32_t* vals, int64_t n, int32_t begin_bit, int32_t end_bit,
                      void* workspace, int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    return vs::sort_pairs_u64(keys, vals, n, begin_bit, end_bit, workspace, workspace_bytes,
                              vs::stream_of(stream));
}

int vs_sort_pairs_u32(uint32_t* keys, int32_t* vals, int64_t n, int32_t begin_bit, int32_t end_bit,
                      void* workspace, int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    return vs::sort_pairs_u32(keys, vals, n, begin_bit, end_bit, workspace, workspace_bytes,
                              vs::stream_of(stream));
}

}  // extern "C"
