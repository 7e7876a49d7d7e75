// Internal helpers shared by the vsjoin HIP sources (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <cstdarg>
#include <cstdio>

#include "../../include/vsjoin.h"

namespace vs {

// ---------------------------------------------------------------------------
// Errors: every entry point returns VS_OK or a negative code and leaves a
// thread-local message readable through vs_last_error().
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);
void clear_error();

#define VS_CHECK_HIP(expr)                                                        \
    do {                                                                          \
        hipError_t e__ = (expr);                                                  \
        if (e__ != hipSuccess) {                                                  \
            ::vs::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e__), \
                            __FILE__, __LINE__);                                  \
            return VS_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

#define VS_REQUIRE(cond, ...)           \
    do {                                \
        if (!(cond)) {                  \
            ::vs::set_error(__VA_ARGS__); \
            return VS_ERR_INVALID;      \
        }                               \
    } while (0)

#define VS_LAUNCH_CHECK()                                                          \
    do {                                                                           \
        hipError_t e__ = hipGetLastError();                                        \
        if (e__ != hipSuccess) {                                                   \
            ::vs::set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e__), \
                            __FILE__, __LINE__);                                   \
            return VS_ERR_HIP;                                                     \
        }                                                                          \
    } while (0)

#define VS_TRY(expr)             \
    do {                         \
        int rc__ = (expr);       \
        if (rc__ != VS_OK) return rc__; \
    } while (0)

inline hipStream_t stream_of(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------------------
// Workspace carving: the caller hands one device buffer; sub-buffers are
// 256-byte aligned slices of it.  `bytes` accumulates the requirement so the
// same code path answers workspace-size queries (base == nullptr).
// ---------------------------------------------------------------------------
struct Carver {
    char* base;
    int64_t bytes = 0;
    explicit Carver(void* b) : base(static_cast<char*>(b)) {}
    template <class T>
    T* take(int64_t count) {
        int64_t off = (bytes + 255) & ~int64_t(255);
        bytes = off + (count > 0 ? count : 1) * int64_t(sizeof(T));
        return base ? reinterpret_cast<T*>(base + off) : nullptr;
    }
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// One row's bounds, ring info and first vertex in one 64-byte record (one
// cache line): box = bounds (NaN when empty); first = first vertex of a
// single ring (NaN otherwise); info = (first coordinate of a single ring of
// at least one coordinate, else -1; coordinates of the row).
struct alignas(64) RowRec {
    double4 box;
    double2 first;
    int2 info;
    int2 pad;
};


// Grid of the learned point index: keys are Hilbert indices of a 2^g x 2^g
// grid over the point extent, left-aligned in K (low bits zero).  32-bit keys
// use g = 12 (three 8-bit radix passes); 64-bit keys g = 32.
template <class K>
constexpr int point_index_grid_bits() {
    return sizeof(K) == 4 ? 12 : 32;
}

// sx[i] = px[perm[i]], sy[i] = py[perm[i]]; at most blocks_per_cu blocks of
// 256 threads per compute unit (vs_index.hip).
int gather_points(const double* px, const double* py, const int32_t* perm, int64_t n, double* sx, double* sy,
                  hipStream_t s, int blocks_per_cu = 64);

// Radix sort of (key, int32 value) pairs (vs_sort.hip).  The no-copy form
// leaves the result in either the inputs or the workspace's alternates.
template <class K>
struct SortResult {
    K* keys;
    int32_t* vals;
    K* alt_keys;
    int32_t* alt_vals;
};
int64_t sort_workspace_bytes_u32(int64_t n);
int64_t sort_workspace_bytes_u64(int64_t n);
// Sorts (in_keys, in_vals) into (out_keys, out_vals) with no copy pass; the
// inputs are only read and must not alias the outputs.  With d_n (device
// count, optional) only the first *d_n of the n slots are sorted, and a count
// above n sorts nothing.
int sort_pairs_into_u32(const uint32_t* in_keys, const int32_t* in_vals, uint32_t* out_keys, int32_t* out_vals,
                        int64_t n, int begin_bit, int end_bit, void* ws, int64_t ws_bytes, hipStream_t s,
                        const unsigned long long* d_n = nullptr);
int sort_pairs_into_u64(const uint64_t* in_keys, const int32_t* in_vals, uint64_t* out_keys, int32_t* out_vals,
                        int64_t n, int begin_bit, int end_bit, void* ws, int64_t ws_bytes, hipStream_t s,
                        const unsigned long long* d_n = nullptr);
int sort_pairs_u32_dev(uint32_t* keys, int32_t* vals, int64_t n, const unsigned long long* d_n, int begin_bit,
                       int end_bit, void* ws, int64_t ws_bytes, hipStream_t s, SortResult<uint32_t>* res);
// The same with the last pass writing int64 keys and values to out_k / out_v
// (res->keys null; res->vals, res->alt_vals and res->alt_keys free buffers of
// n elements).  Needs begin_bit < end_bit and n > 1 for the wide pass to run.
int sort_pairs_u32_dev_wide(uint32_t* keys, int32_t* vals, int64_t n, const unsigned long long* d_n, int begin_bit,
                            int end_bit, void* ws, int64_t ws_bytes, hipStream_t s, int64_t* out_k, int64_t* out_v,
                            SortResult<uint32_t>* res);

// ---------------------------------------------------------------------------
// Device helpers for 64-wide wavefronts.
// ---------------------------------------------------------------------------
constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    union { double d; int i[2]; } u;
    u.d = v;
    u.i[0] = __builtin_amdgcn_readlane(u.i[0], lane);
    u.i[1] = __builtin_amdgcn_readlane(u.i[1], lane);
    return u.d;
}

__device__ __forceinline__ float readlane_f32(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ int readlane_i32(int v, int lane) {
    return __builtin_amdgcn_readlane(v, lane);
}

// Values every lane of the wave holds alike (block-uniform LDS reads, kernel
// arguments indexed by them): the first active lane's copy, kept in scalar
// registers by the compiler.
__device__ __forceinline__ int uniform_i32(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ float uniform_f32(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ double uniform_f64(double v) {
    union { double d; int i[2]; } u;
    u.d = v;
    u.i[0] = __builtin_amdgcn_readfirstlane(u.i[0]);
    u.i[1] = __builtin_amdgcn_readfirstlane(u.i[1]);
    return u.d;
}

template <class T>
__device__ __forceinline__ T wave_min(T v) {
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, 64);
        v = fmin(v, w);
    }
    return v;
}

template <class T>
__device__ __forceinline__ T wave_max(T v) {
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, 64);
        v = fmax(v, w);
    }
    return v;
}

__device__ __forceinline__ int wave_sum_i32(int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Inclusive scan across the 64 lanes.
__device__ __forceinline__ int wave_inclusive_scan(int v) {
    const int l = lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        int w = __shfl_up(v, o, 64);
        if (l >= o) v += w;
    }
    return v;
}

__device__ __forceinline__ long long wave_inclusive_scan_i64(long long v) {
    const int l = lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        long long w = __shfl_up(v, o, 64);
        if (l >= o) v += w;
    }
    return v;
}

// Block-wide sums of N per-thread counters, then one atomicAdd per counter per
// block (`red` is LDS of at least N * (blockDim.x / 64) entries).  Keeps
// same-address atomics to one per block: per-wave atomics on one counter
// serialise in L2 and cost more than the kernels they count.
template <int N>
__device__ __forceinline__ void block_add_counters(unsigned long long (&v)[N], unsigned long long* red,
                                                   unsigned long long* dst) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        unsigned long long x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (l == 0) red[k * nw + w] = x;
    }
    __syncthreads();
    if ((int)threadIdx.x < N) {
        unsigned long long t = 0;
        for (int ww = 0; ww < nw; ++ww) t += red[threadIdx.x * nw + ww];
        if (t) atomicAdd(&dst[threadIdx.x], t);
    }
}

// Grid for grid-stride kernels: enough blocks to fill the device, few
// enough that per-block atomics stay cheap.
inline unsigned stride_grid(int64_t n, int threads, int per_cu = 4) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cus = v;
    }
    const int64_t want = (n + threads - 1) / threads;
    const int64_t cap = (int64_t)cus * per_cu;
    return (unsigned)(want < 1 ? 1 : (want < cap ? want : cap));
}

// Inclusive bbox overlap; NaN boxes never overlap anything.
__device__ __forceinline__ bool box_overlap(double qx0, double qy0, double qx1, double qy1,
                                            double tx0, double ty0, double tx1, double ty1) {
    return qx0 <= tx1 && qx1 >= tx0 && qy0 <= ty1 && qy1 >= ty0;
}

}  // namespace vs
