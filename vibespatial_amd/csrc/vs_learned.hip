// Learned point index and the polygon-major FP32 join.
//
// No counterpart in the reference: its sjoin index is a Morton range scan
// (src/vibespatial/spatial/spatial_index_device.py:761-1034) and its point refine
// admits only fp64 (src/vibespatial/predicates/point_relations.py:557-579).  This
// file restates the same results with a different organisation:
//
// Point index.  Points sorted by Hilbert key (bits per axis = 4 * sizeof(K)).
// A two-stage rank model over the sorted keys:
//   stage 1  a 2-layer MLP (1 -> 32 ReLU -> 1, fp32) maps the key (as a unit
//            float) to an estimate of its CDF, i.e. a leaf number;
//   stage 2  S equal-count leaves; leaf s spans sorted positions
//            [pos(s), pos(s+1)] with pos(s) = floor(s (n-1) / S) and predicts the
//            position by linear interpolation between its end keys.
// Recorded bounds: E = max over indexed keys of |MLP leaf - true leaf| and
// err[s] = max over the leaf's keys of |prediction - position|.  A lower_bound
// lookup searches leaves only inside [leaf - E, leaf + E] and positions only in
// [pred - err, pred + err + 1]; both windows are verified by bracketing keys and
// widened to the full range when a check fails, so the result is always the
// exact lower_bound (the fallbacks are counted).
// The MLP is initialised as the piecewise-linear interpolant of 33 quantile knots
// and refined on the GPU by full-batch Adam with hand-written forward/backward
// passes (mlp_grad); the best-loss parameters are kept.
//
// Query windows.  A box maps to <= 9 Hilbert cells of one level (<= 3 per axis),
// merged into key ranges; each range becomes a rank range through two lookups.
//
// Polygon-major join (learned+FP32 form).  Work items are (polygon, chunk of 512
// scanned positions) taken from one atomic queue; a 256-thread block scans its
// positions with the exact fp64 point-in-box test, keeps the candidates in LDS,
// stages the polygon's vertices in LDS as fp32 coordinates recentred on the
// polygon's box centre, and runs the crossing-number / boundary rule per
// candidate with these decisions:
//   y ties in fp32 -> escalate; strict fp32 order is the true order (the
//   recentring x -> fl32(fl64(x - c)) is monotone);
//   orientation: |det_fp32| > E_q decides the sign, E_q a forward error bound
//   covering both the input rounding and the fp32 arithmetic; otherwise escalate.
// An escalated edge is evaluated by the exact fp64 step (stage-A filter, then
// the exact expansion) on the original coordinates.
#include <mutex>

#include "vs_exact.h"
#include "vs_internal.h"
#include "vs_join_common.h"

namespace vs {

int64_t point_index_workspace_bytes_u32(int64_t n);
int64_t point_index_workspace_bytes_u64(int64_t n);
const double2* point_index_xy_u32(void* ws, int64_t n);
const double2* point_index_xy_u64(void* ws, int64_t n);
int gather_points_xy(const double2* xy, const int32_t* perm, int64_t n, double* sx, double* sy, hipStream_t s,
                     int blocks_per_cu);
int point_index_build_u32(const double* px, const double* py, int64_t n, uint32_t* keys, int32_t* perm, double* sx,
                          double* sy, double4* ext, void* ws, hipStream_t s);
int point_index_build_u64(const double* px, const double* py, int64_t n, uint64_t* keys, int32_t* perm, double* sx,
                          double* sy, double4* ext, void* ws, hipStream_t s);
int build_refine_flags(const int32_t* part, const int32_t* ring, int64_t n_parts, int64_t n_rings, int64_t n_coords,
                       uint8_t* flags, hipStream_t s);
int sort_pairs_u32(uint32_t*, int32_t*, int64_t, int, int, void*, int64_t, hipStream_t);
int64_t sort_workspace_bytes_u32(int64_t n);

constexpr int kHidden = 32;
constexpr int kNParams = 3 * kHidden + 1;  // w1[32], b1[32], w2[32], b2
constexpr int kSample = 1024;              // training keys (evenly spaced ranks)
constexpr int kMaxRanges = 16;             // cells of a query's cover (merged into at most this many ranges)
constexpr int kChunk = 512;                // positions scanned per round
constexpr int kItemRounds = 8;             // rounds per work item (items span kChunk * kItemRounds positions)
constexpr int kItemPos = kChunk * kItemRounds;
constexpr int kLJThreads = 256;
constexpr int kClaim = 4;                  // work items taken per queue atomic (one per wave in wave mode)
constexpr int kWavePos = 256;              // wave mode: at most this many positions per item (4 per lane)
constexpr int kWaveEdges = 128;            // wave mode: at most this many edges per polygon (2 per lane)
static_assert(kLJThreads / 64 == kClaim, "wave mode runs one claimed item per wave");
constexpr int kCandPerThread = kChunk / kLJThreads;

// ---------------------------------------------------------------------------
// MLP: forward, backward, training
// ---------------------------------------------------------------------------
template <class K>
__device__ __forceinline__ float key_unit(K k) {
    // monotone map of the key to [0, 1]
    return sizeof(K) == 4 ? (float)k * 2.3283064365386963e-10f : (float)k * 5.421010862427522e-20f;
}

__device__ __forceinline__ float mlp_eval(const float* P, float u) {
    float y = P[3 * kHidden];
#pragma unroll 4
    for (int h = 0; h < kHidden; ++h) {
        const float z = P[h] * u + P[kHidden + h];
        y += P[2 * kHidden + h] * fmaxf(z, 0.0f);
    }
    return y;
}

// Gradient of the mean squared rank error, computed by one block in two
// phases so that no thread holds all 97 partial derivatives:
//   phase 1  thread-strided over samples: residual r_j = y_j - t_j into LDS,
//            partial loss and output-bias gradient;
//   phase 2  thread t owns hidden unit h = t mod 32 for the samples
//            j = t div 32 (mod blocks/32): with z = w1_h u_j + b1_h > 0,
//            dy = 2 r_j scale:  dL/dw2_h += dy z,  dL/dw1_h += dy w2_h u_j,
//            dL/db1_h += dy w2_h.
// Parameters: P[h] = w1, P[32 + h] = b1, P[64 + h] = w2, P[96] = b2.
constexpr int kTrainThreads = 1024;
// Adam step (relative to each parameter's scale) of the join's index build:
// larger steps leave the quantile-knot start best for the first tens of steps.
constexpr float kJoinTrainLr = 3.0e-4f;
constexpr int kTrainWaves = kTrainThreads / 64;

struct GradPart {
    float gw1 = 0.0f, gb1 = 0.0f, gw2 = 0.0f, gb2 = 0.0f, loss = 0.0f;
};

// Accumulates the contribution of samples (su, st)[0, S) (LDS) into gp; sr is
// LDS scratch of S floats.  Called by every thread of the block.
__device__ void mlp_grad_accumulate(const float* P, const float* su, const float* st, float* sr, int S, float scale,
                                    GradPart& gp) {
    const int tid = threadIdx.x;
    for (int j = tid; j < S; j += blockDim.x) {
        const float r = mlp_eval(P, su[j]) - st[j];
        sr[j] = r;
        gp.loss += r * r;
        gp.gb2 += 2.0f * r * scale;
    }
    __syncthreads();
    const int h = tid & (kHidden - 1), grp = tid / kHidden, ngrp = blockDim.x / kHidden;
    const float w1 = P[h], b1 = P[kHidden + h], w2 = P[2 * kHidden + h];
    for (int j = grp; j < S; j += ngrp) {
        const float u = su[j];
        const float z = w1 * u + b1;
        if (z > 0.0f) {
            const float dy = 2.0f * sr[j] * scale;
            gp.gw2 += dy * z;
            const float dz = dy * w2;
            gp.gw1 += dz * u;
            gp.gb1 += dz;
        }
    }
    __syncthreads();
}

// Block reduction of the per-thread parts into out[0..kNParams] (gradient,
// then the summed squared residual at out[kNParams]).
__device__ void mlp_grad_reduce(const GradPart& gp, float (*red)[kNParams + 1], float* out) {
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    float gw1 = gp.gw1, gb1 = gp.gb1, gw2 = gp.gw2, gb2 = gp.gb2, loss = gp.loss;
    // lanes l and l ^ 32 hold the same hidden unit (two sample groups per wave)
    gw1 += __shfl_xor(gw1, 32, 64);
    gb1 += __shfl_xor(gb1, 32, 64);
    gw2 += __shfl_xor(gw2, 32, 64);
    for (int o = 32; o > 0; o >>= 1) {
        gb2 += __shfl_xor(gb2, o, 64);
        loss += __shfl_xor(loss, o, 64);
    }
    if (l < kHidden) {
        red[w][l] = gw1;
        red[w][kHidden + l] = gb1;
        red[w][2 * kHidden + l] = gw2;
    }
    if (l == 0) {
        red[w][3 * kHidden] = gb2;
        red[w][kNParams] = loss;
    }
    __syncthreads();
    for (int k = tid; k <= kNParams; k += blockDim.x) {
        float v = 0.0f;
        for (int ww = 0; ww < (int)(blockDim.x >> 6); ++ww) v += red[ww][k];
        out[k] = v;
    }
    __syncthreads();
}

// One block: sample, quantile-knot initialisation, `iters` full-batch Adam
// steps; writes the best-loss parameters and [initial loss, best loss].
template <class K>
__global__ __launch_bounds__(kTrainThreads) void mlp_train_kernel(const K* __restrict__ keys, int64_t n, int iters,
                                                                 float lr, float* __restrict__ params_out,
                                                                 float* __restrict__ losses,
                                                                 int32_t* __restrict__ meta) {
    __shared__ float su[kSample], st[kSample], sr[kSample];
    __shared__ float P[kNParams], M1[kNParams], M2[kNParams], SC[kNParams], BEST[kNParams];
    __shared__ float red[kTrainWaves][kNParams + 1];
    __shared__ float tot[kNParams + 1];
    __shared__ float best_loss, init_loss;
    // one block beside the coordinate gather: raise its waves' issue priority
    __builtin_amdgcn_s_setprio(3);
    const int tid = threadIdx.x;
    if (meta && tid < 4) meta[tid] = 0;  // the leaf error pass that follows takes maxima there
    const int S = n >= kSample ? kSample : (int)n;
    for (int j = tid; j < S; j += blockDim.x) {
        const int64_t i = S > 1 ? (int64_t)((n - 1) * (double)j / (double)(S - 1)) : 0;
        su[j] = key_unit<K>(keys[i]);
        st[j] = (float)((double)i / (double)n);
    }
    __syncthreads();
    if (tid < kHidden) {
        // knots q_h at sample ranks h/H, PL interpolant g(u) = t_0 + sum_h dS_h relu(u - q_h)
        const int h = tid;
        auto slope_of = [&](int hh) {
            const int j0 = (int)((int64_t)hh * (S - 1) / kHidden);
            const int j1 = (int)((int64_t)(hh + 1) * (S - 1) / kHidden);
            return (st[j1] - st[j0]) / fmaxf(su[j1] - su[j0], 1.0e-7f);
        };
        const int j0 = (int)((int64_t)h * (S - 1) / kHidden);
        const int j1 = (int)((int64_t)(h + 1) * (S - 1) / kHidden);
        const float q0 = su[j0], q1 = su[j1];
        const float slope = slope_of(h);
        const float prev_slope = h > 0 ? slope_of(h - 1) : 0.0f;
        P[h] = 1.0f;
        P[kHidden + h] = -q0;
        P[2 * kHidden + h] = slope - prev_slope;
        SC[h] = 0.05f;
        SC[kHidden + h] = fmaxf(q1 - q0, 1.0e-7f);
        SC[2 * kHidden + h] = fmaxf(fabsf(slope), 1.0e-3f);
        if (h == 0) {
            P[3 * kHidden] = st[0];
            SC[3 * kHidden] = 1.0f / kHidden;
            best_loss = INFINITY;
        }
    }
    for (int k = tid; k < kNParams; k += blockDim.x) {
        M1[k] = 0.0f;
        M2[k] = 0.0f;
    }
    __syncthreads();
    const float scale = 1.0f / (float)(S > 0 ? S : 1);
    const float b1 = 0.9f, b2 = 0.999f;
    float b1t = 1.0f, b2t = 1.0f;
    for (int it = 0; it <= iters; ++it) {
        GradPart gp;
        mlp_grad_accumulate(P, su, st, sr, S, scale, gp);
        mlp_grad_reduce(gp, red, tot);
        const float L = tot[kNParams] * scale;
        const bool better = L < best_loss;  // uniform: every thread reads the same LDS values
        if (tid < kNParams && better) BEST[tid] = P[tid];
        __syncthreads();
        if (tid == 0) {
            if (it == 0) init_loss = L;
            if (better) best_loss = L;
        }
        if (it == iters) break;
        b1t *= b1;
        b2t *= b2;
        if (tid < kNParams) {
            const int k = tid;
            const float gk = tot[k];
            M1[k] = b1 * M1[k] + (1.0f - b1) * gk;
            M2[k] = b2 * M2[k] + (1.0f - b2) * gk * gk;
            const float mh = M1[k] / (1.0f - b1t), vh = M2[k] / (1.0f - b2t);
            P[k] -= lr * SC[k] * mh / (sqrtf(vh) + 1.0e-12f);
        }
        __syncthreads();
    }
    __syncthreads();
    for (int k = tid; k < kNParams; k += blockDim.x) params_out[k] = BEST[k];
    if (tid == 0) {
        losses[0] = init_loss;
        losses[1] = best_loss;
    }
}

// Full-batch loss and gradient of given parameters on given (u, t) data
// (the same device code the trainer uses), for checking against a CPU oracle.
__global__ __launch_bounds__(kTrainThreads) void mlp_loss_grad_kernel(const float* __restrict__ params,
                                                                     const float* __restrict__ u,
                                                                     const float* __restrict__ t, int64_t n,
                                                                     float* __restrict__ grad_out,
                                                                     float* __restrict__ loss_out) {
    __shared__ float su[kSample], st[kSample], sr[kSample];
    __shared__ float P[kNParams];
    __shared__ float red[kTrainWaves][kNParams + 1];
    __shared__ float tot[kNParams + 1];
    for (int k = threadIdx.x; k < kNParams; k += blockDim.x) P[k] = params[k];
    const float scale = 1.0f / (float)(n > 0 ? n : 1);
    GradPart gp;
    for (int64_t c0 = 0; c0 < n; c0 += kSample) {
        const int S = (n - c0) < kSample ? (int)(n - c0) : kSample;
        __syncthreads();
        for (int j = threadIdx.x; j < S; j += blockDim.x) {
            su[j] = u[c0 + j];
            st[j] = t[c0 + j];
        }
        __syncthreads();
        mlp_grad_accumulate(P, su, st, sr, S, scale, gp);
    }
    __syncthreads();
    mlp_grad_reduce(gp, red, tot);
    for (int k = threadIdx.x; k < kNParams; k += blockDim.x) grad_out[k] = tot[k];
    if (threadIdx.x == 0) loss_out[0] = tot[kNParams] * scale;
}

// ---------------------------------------------------------------------------
// Leaf table, error bounds, lookups
// ---------------------------------------------------------------------------
template <class K>
struct LearnedView {
    const K* keys;
    int64_t n;
    const float* mlp;
    const K* leaf_key;  // S + 1
    const int32_t* leaf_err;
    const int2* leaf_win;  // per predicted leaf: [first, last] true leaf of the keys it routes
    const int32_t* meta;  // [0] E (MLP leaf error)
    int32_t S;
    int32_t mode;  // 0 learned, 1 leaf table only (binary search over leaves), 2 plain binary search
};

__device__ __forceinline__ int64_t leaf_pos(int64_t s, int64_t n, int64_t S) { return s * (n - 1) / S; }

__device__ __forceinline__ int leaf_of_pred(float g, int S) {
    const float f = g * (float)S;
    if (!(f > 0.0f)) return 0;  // also NaN
    if (f >= (float)(S - 1)) return S - 1;
    return (int)f;
}

template <class K>
__device__ __forceinline__ int mlp_leaf(const float* P, K k, int S) {
    return leaf_of_pred(mlp_eval(P, key_unit<K>(k)), S);
}

// Linear interpolation of the rank of q inside leaf s, which spans sorted
// positions [p0, p1] and keys [k0, k1].
template <class K>
__device__ __forceinline__ double plm_pred(K q, K k0, K k1, int64_t p0, int64_t p1) {
    if (!(k1 > k0)) return (double)p0;
    const double f = (double)(K)(q - k0) / (double)(K)(k1 - k0);
    return (double)p0 + f * (double)(p1 - p0);
}

// One wave per leaf: leaf_key[s] = keys[pos(s)] (and leaf_key[S] = keys[n - 1]),
// err[s] = ceil(max |pred - i|) over i in [pos(s), pos(s+1)]; the MLP leaf
// error over the leaf's first occurrences goes to meta[0] (cleared by the
// trainer, which runs first).
template <class K>
__global__ __launch_bounds__(256) void leaf_err_kernel(const K* __restrict__ keys, int64_t n, int S,
                                                      K* __restrict__ leaf_key, const float* __restrict__ mlp,
                                                      int32_t* __restrict__ leaf_err, int2* __restrict__ leaf_win,
                                                      int32_t* __restrict__ meta) {
    __shared__ float P[kNParams];
    for (int k = threadIdx.x; k < kNParams; k += blockDim.x) P[k] = mlp[k];
    __syncthreads();
    __shared__ int red[4];
    const int lane = threadIdx.x & 63;
    int mlp_e = 0;  // over every leaf this thread visits
    for (int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); s < S; s += (int64_t)gridDim.x * 4) {
        const int64_t p0 = leaf_pos(s, n, S), p1 = leaf_pos(s + 1, n, S);
        const K k0 = keys[p0], k1 = keys[p1];
        if (lane == 0) {
            leaf_key[s] = k0;
            if (s == S - 1) leaf_key[S] = k1;
        }
        double emax = 0.0;
        int sp_lo = INT_MAX, sp_hi = -1;  // predicted leaves of this leaf's first occurrences
        // Positions [p0, p1) in passes of 64 kB keys (one pass for a 512-key
        // leaf), then p1 itself from registers.  The network's weights are read
        // once per hidden unit for kB keys (same operation order as mlp_eval,
        // so the predictions are bit-identical to the lookups').
        constexpr int kB = 8;
        const K kbefore = keys[p1 > p0 ? p1 - 1 : p0];
        for (int64_t base = p0; base < p1; base += 64 * kB) {
            float u[kB], y[kB];
            bool first[kB];
#pragma unroll
            for (int b = 0; b < kB; ++b) {
                const int64_t i = base + lane + 64 * b;
                const bool valid = i < p1;
                const K k = valid ? keys[i] : (K)0;
                first[b] = valid && i > p0 && k != keys[i - 1];
                if (valid) emax = fmax(emax, fabs(plm_pred<K>(k, k0, k1, p0, p1) - (double)i));
                u[b] = key_unit<K>(k);
                y[b] = P[3 * kHidden];
            }
#pragma unroll 4
            for (int h = 0; h < kHidden; ++h) {
                const float w1 = P[h], b1 = P[kHidden + h], w2 = P[2 * kHidden + h];
#pragma unroll
                for (int b = 0; b < kB; ++b) {
                    const float z = w1 * u[b] + b1;
                    y[b] += w2 * fmaxf(z, 0.0f);
                }
            }
#pragma unroll
            for (int b = 0; b < kB; ++b) {
                if (!first[b]) continue;
                // a first occurrence in (p0, p1]: the lookup of its key must land
                // in leaf s (the last leaf whose first key is below it)
                const int sp = leaf_of_pred(y[b], S);
                const int d = sp > (int)s ? sp - (int)s : (int)s - sp;
                mlp_e = d > mlp_e ? d : mlp_e;
                sp_lo = min(sp_lo, sp);
                sp_hi = max(sp_hi, sp);
            }
        }
        // position p1 (the same in every lane)
        emax = fmax(emax, fabs(plm_pred<K>(k1, k0, k1, p0, p1) - (double)p1));
        if (p1 > p0 && k1 != kbefore) {
            const int sp = leaf_of_pred(mlp_eval(P, key_unit<K>(k1)), S);
            const int d = sp > (int)s ? sp - (int)s : (int)s - sp;
            mlp_e = d > mlp_e ? d : mlp_e;
            sp_lo = min(sp_lo, sp);
            sp_hi = max(sp_hi, sp);
        }
        emax = wave_max(emax);
        if (lane == 0) leaf_err[s] = (int32_t)ceil(emax);
        // Routing windows.  A query q with keys[0] < q <= keys[n-1] resolves
        // to the true leaf of the first occurrence k_b of the least key >= q;
        // the distinct key before k_b, k_a, is keys[p0] or a first occurrence
        // of this leaf.  For a monotone network the predicted leaf of q lies
        // between those of k_a and k_b, so leaf s is recorded for every
        // predicted leaf from the least to the greatest over keys[p0] and this
        // leaf's first occurrences.  Lookups verify the window (a non-monotone
        // network only costs a fallback).
        sp_lo = -wave_max(-sp_lo);
        sp_hi = wave_max(sp_hi);
        if (sp_hi >= 0) {
            const int spa = leaf_of_pred(mlp_eval(P, key_unit<K>(k0)), S);
            const int a = min(sp_lo, spa), b = max(sp_hi, spa);
            for (int t = a + lane; t <= b; t += 64) {
                atomicMin(&leaf_win[t].x, (int)s);
                atomicMax(&leaf_win[t].y, (int)s);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int v = __shfl_xor(mlp_e, o, 64);
        mlp_e = v > mlp_e ? v : mlp_e;
    }
    if (lane == 0) red[threadIdx.x >> 6] = mlp_e;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int m = max(max(red[0], red[1]), max(red[2], red[3]));
        if (m > 0) atomicMax(&meta[0], m);
    }
}

struct LookupStats {
    unsigned long long probes;      // key / leaf-key reads in binary searches
    unsigned long long leaf_fb;     // MLP leaf window failed its bracket check
    unsigned long long win_fb;      // leaf error window failed its bracket check
};

// First position j with keys[j] >= q (n if none).
template <class K>
__device__ int64_t learned_lower_bound(K q, const LearnedView<K>& V, LookupStats& st) {
    const int64_t n = V.n;
    if (n == 0) return 0;
    if (q <= V.keys[0]) return 0;
    if (q > V.keys[n - 1]) return n;
    int64_t a, b;
    if (V.mode == 2) {
        a = 1;
        b = n - 1;
    } else {
        const int S = V.S;
        int wl = 0, wh = S - 1;
        if (V.mode == 0) {
            const int sp = mlp_leaf<K>(V.mlp, q, S);
            const int2 w = V.leaf_win[sp];
            int l2, h2;
            if (w.x <= w.y) {  // the leaves this prediction routes to
                l2 = w.x;
                h2 = w.y;
            } else {  // nothing routed here: the global error bound
                const int E = V.meta[0];
                l2 = sp - E < 0 ? 0 : sp - E;
                h2 = sp + E > S - 1 ? S - 1 : sp + E;
            }
            st.probes += 2;
            if (V.leaf_key[l2] < q && (h2 == S - 1 || V.leaf_key[h2 + 1] >= q)) {
                wl = l2;
                wh = h2;
            } else {
                ++st.leaf_fb;
            }
        }
        while (wl < wh) {  // last leaf whose first key is < q
            const int mid = (wl + wh + 1) >> 1;
            ++st.probes;
            if (V.leaf_key[mid] < q) wl = mid;
            else wh = mid - 1;
        }
        const int s = wl;
        const int64_t p0 = leaf_pos(s, n, S), hi = leaf_pos(s + 1, n, S), lo = p0 + 1;
        const double p = plm_pred<K>(q, V.leaf_key[s], V.leaf_key[s + 1], p0, hi);
        const int32_t e = V.leaf_err[s];
        a = (int64_t)ceil(p - (double)e);
        b = (int64_t)floor(p + (double)e) + 1;
        a = a < lo ? lo : a;
        b = b > hi ? hi : b;
        st.probes += 2;
        if (!(a <= b && (a == lo || V.keys[a - 1] < q) && (b == hi || V.keys[b] >= q))) {
            ++st.win_fb;
            a = lo;
            b = hi;
        }
    }
    while (a < b) {
        const int64_t mid = (a + b) >> 1;
        ++st.probes;
        if (V.keys[mid] >= q) b = mid;
        else a = mid + 1;
    }
    return a;
}

// Cover of a box on the key grid: the finest level, no finer than min_l (the
// index grid), whose cells covering the box number at most kMaxRanges (long
// thin boxes get thin covers; by level bits-1 at most 2 x 2 remain), each cell
// a key range [code 4^l, (code+1) 4^l); the codes are sorted (a bitonic
// network on registers, each code carrying its cell).  Then each cell is split
// into its four children one level finer (never finer than the index grid;
// a child's code is its parent's followed by two digits) and the children that
// miss the box are dropped, when the kept children still merge into at most
// kMaxRanges ranges (else the cells themselves are used).  Consecutive codes
// merge.  Writes (lo, hi) pairs to out and returns their count.
__host__ __device__ int cover_ranges(double x0, double y0, double x1, double y1, const KeyScale& ks, int bits,
                                     int min_l, unsigned long long* __restrict__ out) {
    if (!(x0 <= x1 && y0 <= y1)) return 0;
    const uint32_t qx0 = quantize(x0, ks.x0, ks.sx, bits), qx1 = quantize(x1, ks.x0, ks.sx, bits);
    const uint32_t qy0 = quantize(y0, ks.y0, ks.sy, bits), qy1 = quantize(y1, ks.y0, ks.sy, bits);
    // cells per axis in 64 bits (a box over a whole 2^32 axis spans 2^32
    // cells), their product only once both are small
    auto too_many = [&](int lv) {
        const uint64_t cx = (uint64_t)(qx1 >> lv) - (uint64_t)(qx0 >> lv) + 1;
        const uint64_t cy = (uint64_t)(qy1 >> lv) - (uint64_t)(qy0 >> lv) + 1;
        return cx > (uint64_t)kMaxRanges || cy > (uint64_t)kMaxRanges || cx * cy > (uint64_t)kMaxRanges;
    };
    int l = min_l;  // cells no finer than the index grid, so every range is a union of grid cells
    while (l < bits - 1 && too_many(l)) ++l;
    const int lev = bits - l;
    const uint32_t ax = qx0 >> l, ay = qy0 >> l;
    const uint32_t nx = (qx1 >> l) - ax + 1, ncell = nx * ((qy1 >> l) - ay + 1);  // <= 16 here
    uint64_t code[kMaxRanges];
    uint32_t tag[kMaxRanges];
#pragma unroll
    for (int t = 0; t < kMaxRanges; ++t) {
        const uint32_t a = ax + (uint32_t)t % nx, c = ay + (uint32_t)t / nx;
        code[t] = (uint32_t)t < ncell ? hilbert64(a << (32 - lev), c << (32 - lev)) >> (64 - 2 * lev) : ~0ull;
        tag[t] = (uint32_t)t;
    }
#pragma unroll
    for (int k = 2; k <= kMaxRanges; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < kMaxRanges; ++i) {
                const int o = i ^ j;
                if (o > i) {
                    const bool asc = (i & k) == 0;
                    const uint64_t a = code[i], b = code[o];
                    if ((a > b) == asc) {
                        code[i] = b;
                        code[o] = a;
                        const uint32_t ta = tag[i];
                        tag[i] = tag[o];
                        tag[o] = ta;
                    }
                }
            }
    if (l > min_l) {
        // children one level finer: kept ones as a 4-bit mask per cell,
        // indexed by the child's last Hilbert digit
        const int lc = l - 1, levc = bits - lc;
        const uint32_t cx0 = qx0 >> lc, cx1 = qx1 >> lc, cy0 = qy0 >> lc, cy1 = qy1 >> lc;
        uint64_t masks = 0;
#pragma unroll
        for (int t = 0; t < kMaxRanges; ++t) {
            if ((uint32_t)t >= ncell) continue;
            const uint32_t px = 2 * (ax + tag[t] % nx), py = 2 * (ay + tag[t] / nx);
            // last digits of the children (px, py) and (px + 1, py); the curve
            // visits the four children around the 2 x 2 square, so a child's
            // diagonal partner has its digit xor 2
            const uint32_t d00 = (uint32_t)(hilbert64(px << (32 - levc), py << (32 - levc)) >> (64 - 2 * levc)) & 3u;
            const uint32_t d10 =
                (uint32_t)(hilbert64((px + 1) << (32 - levc), py << (32 - levc)) >> (64 - 2 * levc)) & 3u;
            const uint32_t dig[4] = {d00, d10, d10 ^ 2u, d00 ^ 2u};  // (0,0) (1,0) (0,1) (1,1)
            uint32_t m = 0;
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) {
                const uint32_t cx = px + (ch & 1), cy = py + (ch >> 1);
                if (cx < cx0 || cx > cx1 || cy < cy0 || cy > cy1) continue;
                m |= 1u << dig[ch];
            }
            masks |= (uint64_t)m << (4 * t);
        }
        // ranges of the kept children (count first; emit when they fit);
        // loops unrolled so the code array stays in registers
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
            int m = 0;
            uint64_t rlo = 0, rhi = 0;
            bool open = false;
#pragma unroll
            for (int t = 0; t < kMaxRanges; ++t) {
                if ((uint32_t)t >= ncell) continue;
                const uint32_t cm = (uint32_t)(masks >> (4 * t)) & 15u;
#pragma unroll
                for (uint32_t h = 0; h < 4; ++h) {
                    if (!(cm & (1u << h))) continue;
                    const uint64_t c = (code[t] << 2) | h;
                    if (open && c == rhi + 1) {
                        rhi = c;
                        continue;
                    }
                    if (open) {
                        if (pass == 1) {
                            out[2 * m] = rlo << (2 * lc);
                            out[2 * m + 1] = (rhi << (2 * lc)) | ((1ull << (2 * lc)) - 1ull);
                        }
                        ++m;
                    }
                    rlo = rhi = c;
                    open = true;
                }
            }
            if (open) {
                if (pass == 1) {
                    out[2 * m] = rlo << (2 * lc);
                    out[2 * m + 1] = (rhi << (2 * lc)) | ((1ull << (2 * lc)) - 1ull);
                }
                ++m;
            }
            if (pass == 1) return m;
            if (m > kMaxRanges || m == 0) break;  // too many: the cells themselves
        }
    }
    const uint64_t span = (1ull << (2 * l)) - 1ull;
    int m = 0;
    uint64_t rlo = code[0], rhi = code[0];
#pragma unroll
    for (int t = 1; t < kMaxRanges; ++t) {
        if ((uint32_t)t >= ncell) continue;
        if (code[t] == rhi + 1) {
            rhi = code[t];
        } else {
            out[2 * m] = rlo << (2 * l);
            out[2 * m + 1] = (rhi << (2 * l)) | span;
            ++m;
            rlo = rhi = code[t];
        }
    }
    out[2 * m] = rlo << (2 * l);
    out[2 * m + 1] = (rhi << (2 * l)) | span;
    return m + 1;
}

// Cover of each query box (and, for polygon queries, its coordinate range
// and part count): up to kMaxRanges merged key ranges [lo, hi].
template <class K>
__global__ __launch_bounds__(256) void cover_cells_kernel(const double4* __restrict__ qbox, int64_t nq,
                                                         const double4* __restrict__ ext,
                                                         unsigned long long* __restrict__ keyr,
                                                         int32_t* __restrict__ nranges,
                                                         const int32_t* __restrict__ geom,
                                                         const int32_t* __restrict__ part,
                                                         const int32_t* __restrict__ ring, int4* __restrict__ poly_info) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= nq) return;
    if (poly_info) poly_info[q] = make_int4(ring[part[geom[q]]], ring[part[geom[q + 1]]], geom[q + 1] - geom[q], 0);
    const int bits = 4 * (int)sizeof(K);
    const KeyScale ks = key_scale(ext, bits);
    const double4 b = qbox[q];
    nranges[q] = cover_ranges(b.x, b.y, b.z, b.w, ks, bits, bits - point_index_grid_bits<K>(),
                              keyr + q * kMaxRanges * 2);
}

// One thread per range end: rank of the range's first key (x) and of the key
// after its last (y) through the learned lower bound.
template <class K>
__global__ __launch_bounds__(256) void cover_lookup_kernel(int64_t nq, const unsigned long long* __restrict__ keyr,
                                                          const int32_t* __restrict__ nranges, LearnedView<K> V,
                                                          int32_t* __restrict__ ranges,
                                                          unsigned long long* __restrict__ stat) {
    __shared__ float P[kNParams];
    __shared__ unsigned long long red[3 * 4];
    for (int k = threadIdx.x; k < kNParams; k += blockDim.x) P[k] = V.mlp[k];
    __syncthreads();
    LearnedView<K> Vl = V;
    Vl.mlp = P;
    LookupStats st = {0, 0, 0};
    const int64_t n_ends = nq * kMaxRanges * 2;
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n_ends; t += (int64_t)gridDim.x * 256) {
        const int64_t q = t / (kMaxRanges * 2);
        const int k = (int)((t / 2) % kMaxRanges), end = (int)(t & 1);
        if (k < nranges[q]) {
            const uint64_t key = keyr[t];
            int64_t r;
            if (end == 0) r = learned_lower_bound<K>((K)key, Vl, st);
            else if (key >= (sizeof(K) == 4 ? 0xFFFFFFFFull : ~0ull)) r = V.n;
            else r = learned_lower_bound<K>((K)(key + 1), Vl, st);
            ranges[t] = (int32_t)r;  // int2 {x, y} per range
        }
    }
    unsigned long long v[3] = {st.probes, st.leaf_fb, st.win_fb};
    block_add_counters<3>(v, red, stat);
}

// Per query: total scan length and its number of work items (kItemPos
// positions each); stat[3] += positions scanned.  Also the query's cost key
// for the longest-first order, 8 bits, ascending = most expensive first:
// (127 or, for a small item, 255) - min(127, 4 log2(1 + items x vertices)),
// with idx[q] = q.
__global__ __launch_bounds__(256) void cover_total_kernel(int64_t nq, const int2* __restrict__ ranges,
                                                         const int32_t* __restrict__ nranges,
                                                         const int4* __restrict__ poly_info,
                                                         int32_t* __restrict__ n_items, uint32_t* __restrict__ key,
                                                         int32_t* __restrict__ idx,
                                                         unsigned long long* __restrict__ stat) {
    __shared__ unsigned long long red[4];
    unsigned long long scan[1] = {0};
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < nq; q += (int64_t)gridDim.x * 256) {
        int64_t total = 0;
        for (int k = 0; k < nranges[q]; ++k) {
            const int2 r = ranges[q * kMaxRanges + k];
            total += r.y - r.x;
        }
        const int32_t items = (int32_t)((total + kItemPos - 1) / kItemPos);
        n_items[q] = items;
        scan[0] += (unsigned long long)total;
        const int4 info = poly_info ? poly_info[q] : make_int4(0, 0, 1, 0);
        const double v = poly_info ? (double)(info.y - info.x) : 1.0;
        const double cost = (double)items * (v > 1.0 ? v : 1.0);
        int lg = (int)(4.0 * log2(1.0 + cost));
        lg = lg > 127 ? 127 : lg;
        // small items (one wave's positions, a single part of at most
        // kWaveEdges edges) sort after all others: a block that claims four of
        // them runs one per wave
        const bool small = total <= kWavePos && info.z == 1 && info.y - info.x <= kWaveEdges + 1;
        key[q] = (uint32_t)((small ? 255 : 127) - lg);
        idx[q] = (int32_t)q;
    }
    block_add_counters<1>(scan, red, stat + 3);
}

template <class K>
__global__ __launch_bounds__(256) void lookup_batch_kernel(const K* __restrict__ qk, int64_t m, LearnedView<K> V,
                                                          int64_t* __restrict__ out,
                                                          unsigned long long* __restrict__ stat) {
    __shared__ float P[kNParams];
    __shared__ unsigned long long red[3 * 4];
    for (int k = threadIdx.x; k < kNParams; k += blockDim.x) P[k] = V.mlp[k];
    __syncthreads();
    LearnedView<K> Vl = V;
    Vl.mlp = P;
    LookupStats st = {0, 0, 0};
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (int64_t)gridDim.x * 256)
        out[i] = learned_lower_bound<K>(qk[i], Vl, st);
    unsigned long long v[3] = {st.probes, st.leaf_fb, st.win_fb};
    block_add_counters<3>(v, red, stat);
}

// ---------------------------------------------------------------------------
// FP32 edge step with escalation
// ---------------------------------------------------------------------------
// Returns 1 when the edge must be re-evaluated exactly; otherwise the fp32
// decision is exact and `parity` has been updated.
__device__ __forceinline__ int edge_step_f32(float px, float py, float ax, float ay, float bx, float by, float E,
                                             int& parity, unsigned& n_orient) {
    if (ay == py || by == py) return 1;
    const bool above_a = ay > py, above_b = by > py;
    if (above_a == above_b) return 0;
    if (px < fminf(ax, bx)) {
        parity ^= 1;
        return 0;
    }
    if (px > fmaxf(ax, bx)) return 0;
    ++n_orient;
    const float d = (ax - px) * (by - py) - (ay - py) * (bx - px);
    if (d > E) {
        if (above_b) parity ^= 1;
        return 0;
    }
    if (d < -E) {
        if (!above_b) parity ^= 1;
        return 0;
    }
    return 1;
}

// Forward error bound of the fp32 determinant on coordinates recentred on the
// polygon's box centre, all of magnitude <= M (vertices and, for candidates,
// the point lie in the box).  Inputs a' = fl32(fl64(x - c)): |a' - (x - c)| <=
// d*M + eta with d = 2^-24 + 2^-53 + 2^-77, eta = 2^-149 (subnormals).  With
// e = 2(d*M + eta) per coordinate difference and g4 = (1 + 2^-24)^4 - 1 for the
// four roundings on each product path:
//   |det' - det| <= 2 g4 (2M + e)^2 + 8 M e + 2 e^2,
// evaluated in fp64, inflated by 2^-40 and rounded up to fp32.
__device__ __forceinline__ float fp32_det_bound(double M) {
    const double eps = 5.9604644775390625e-08;  // 2^-24
    const double d = eps + 1.1102230246251565e-16 + 6.617444900424222e-24;
    const double eta = 1.401298464324817e-45;
    const double e = 2.0 * (d * M + eta);
    const double g4 = 4.0 * eps + 6.0 * eps * eps + 4.0 * eps * eps * eps + eps * eps * eps * eps;
    double E = 2.0 * g4 * (2.0 * M + e) * (2.0 * M + e) + 8.0 * M * e + 2.0 * e * e;
    E *= 1.0 + 9.094947017729282e-13;
    float Ef = (float)E;
    if ((double)Ef < E) Ef = nextafterf(Ef, INFINITY);
    return Ef;
}

struct RefineView {
    const double* x;
    const double* y;
    const int32_t* geom;
    const int32_t* part;
    const int32_t* ring;
    const uint8_t* flags;
};

// kLModePip: query q is polygon row q, pairs (point, q), the point index is
// built over the left column; kLModePipQ: the same refine with pairs (q,
// point), the polygons being the query column of a prebuilt point index;
// kLModeBox: query q is a box, pairs (q, point).
enum { kLModePip = 0, kLModeBox = 1, kLModePipQ = 2 };

template <int PRED>
__device__ __forceinline__ bool loc_pred(int loc) {
    if (PRED == VS_PRED_WITHIN) return loc == 2;
    if (PRED == VS_PRED_TOUCHES) return loc == 1;
    return loc != 0;
}

constexpr int kEdgeCap = 512;  // edges staged in LDS per batch
constexpr int kPosPerThread = 4;                       // direct path: positions per thread per round
constexpr int kRoundPos = kLJThreads * kPosPerThread;  // direct path: positions per round
constexpr int kOutCap = 2048;  // per-block output buffer (pairs)

// ---------------------------------------------------------------------------
// y-slab lists of the staged edges.  The staged batch's y-range [ylo, yhi]
// (recentred fp32) is cut into B equal slabs; edge e is listed in every slab
// from slab(min(ay, by)) to slab(max(ay, by)).  slab() is monotone in y, so a
// point whose fp32 y lies in an edge's closed fp32 y-range finds that edge in
// its own slab; every other edge has both endpoints strictly above or below
// the point and leaves its parity and boundary state unchanged
// (edge_step_f32 returns 0 for it without escalating).  Parity is an XOR and
// any boundary hit decides, so the order within a slab does not matter: the
// binned refine makes exactly the decisions of the full edge loop.
// ---------------------------------------------------------------------------
constexpr int kMaxBins = 64;
constexpr int kBinCap = 4096;  // slab entries per staged batch (u16 edge indices)

struct Slabs {
    int B;      // slab count (1: one slab holding every edge)
    float ylo;  // fp32 y of the first slab's lower edge
    float inv;  // slabs per unit of y
};

__device__ __forceinline__ int slab_of(float y, const Slabs& S) {
    float f = (y - S.ylo) * S.inv;
    f = fminf(fmaxf(f, 0.0f), (float)(S.B - 1));
    return (int)f;
}

// Block-uniform.  Builds start[0..B] and list[] for the ne staged edges.
__device__ Slabs build_slabs(const float4* __restrict__ Et, int ne, float ylo, float yhi, int* start, int* cur,
                             uint16_t* list) {
    const int tid = threadIdx.x;
    Slabs S;
    S.B = 1;
    while (S.B < kMaxBins && S.B * 2 < ne) S.B <<= 1;
    S.ylo = ylo;
    S.inv = (yhi > ylo) ? (float)S.B / (yhi - ylo) : 0.0f;
    if (!(S.inv > 0.0f) || !(S.inv < INFINITY)) {
        S.B = 1;
        S.inv = 0.0f;
    }
    if (tid <= kMaxBins) start[tid] = 0;
    __syncthreads();
    if (S.B > 1) {
        for (int e = tid; e < ne; e += kLJThreads) {
            const float4 ed = Et[e];
            const int b0 = slab_of(fminf(ed.y, ed.w), S), b1 = slab_of(fmaxf(ed.y, ed.w), S);
            for (int b = b0; b <= b1; ++b) atomicAdd(&start[b + 1], 1);
        }
        __syncthreads();
        if (tid < 64) {
            int v = tid < S.B ? start[tid + 1] : 0;
            v = wave_inclusive_scan(v);
            if (tid < S.B) start[tid + 1] = v;
        }
        __syncthreads();
        const int total = start[S.B];
        __syncthreads();
        if (total <= kBinCap) {
            if (tid < S.B) cur[tid] = start[tid];
            __syncthreads();
            for (int e = tid; e < ne; e += kLJThreads) {
                const float4 ed = Et[e];
                const int b0 = slab_of(fminf(ed.y, ed.w), S), b1 = slab_of(fmaxf(ed.y, ed.w), S);
                for (int b = b0; b <= b1; ++b) list[atomicAdd(&cur[b], 1)] = (uint16_t)e;
            }
            __syncthreads();
            return S;
        }
        S.B = 1;  // too many long edges for the slab lists: one slab
        S.inv = 0.0f;
    }
    if (tid == 0) {
        start[0] = 0;
        start[1] = ne;
    }
    for (int e = tid; e < ne; e += kLJThreads) list[e] = (uint16_t)e;
    __syncthreads();
    return S;
}

// Runs this thread's candidates against the edges of their own slabs.
__device__ __forceinline__ void refine_slabs(const float4* __restrict__ Et, const int32_t* __restrict__ Ec,
                                             const int* __restrict__ start, const uint16_t* __restrict__ list,
                                             const Slabs& S, const RefineView& R, const float* pu, const float* pv,
                                             const double* pxd, const double* pyd, float E, int* loc, int* par,
                                             bool* done, bool* esc, unsigned& n_orient, unsigned long long& n_esc) {
#pragma unroll
    for (int j = 0; j < kCandPerThread; ++j) {
        if (done[j]) continue;
        const int b = slab_of(pv[j], S);
        const int k1 = start[b + 1];
        for (int k = start[b]; k < k1; ++k) {
            const int e = list[k];
            const float4 ed = Et[e];
            if (edge_step_f32(pu[j], pv[j], ed.x, ed.y, ed.z, ed.w, E, par[j], n_orient)) {
                ++n_esc;
                esc[j] = true;
                const int32_t c = Ec[e];
                if (edge_step_f64(pxd[j], pyd[j], R.x[c - 1], R.y[c - 1], R.x[c], R.y[c], par[j])) {
                    loc[j] = 1;
                    done[j] = true;
                    break;
                }
            }
        }
    }
}

// One candidate against the edges of its slab; returns its location code.
__device__ __forceinline__ int refine_one(const float4* __restrict__ Et, const int32_t* __restrict__ Ec,
                                         const int* __restrict__ start, const uint16_t* __restrict__ list,
                                         const Slabs& S, const RefineView& R, float pu, float pv, double pxd,
                                         double pyd, float E, bool& esc, unsigned& n_orient,
                                         unsigned long long& n_esc) {
    int par = 0;
    const int b = slab_of(pv, S);
    const int k1 = start[b + 1];
    int k = start[b];
    // two edges per step: both list entries, then both edges, load together
    for (; k + 1 < k1; k += 2) {
        const int e0 = list[k], e1 = list[k + 1];
        const float4 d0 = Et[e0], d1 = Et[e1];
        if (edge_step_f32(pu, pv, d0.x, d0.y, d0.z, d0.w, E, par, n_orient)) {
            ++n_esc;
            esc = true;
            const int32_t c = Ec[e0];
            if (edge_step_f64(pxd, pyd, R.x[c - 1], R.y[c - 1], R.x[c], R.y[c], par)) return 1;
        }
        if (edge_step_f32(pu, pv, d1.x, d1.y, d1.z, d1.w, E, par, n_orient)) {
            ++n_esc;
            esc = true;
            const int32_t c = Ec[e1];
            if (edge_step_f64(pxd, pyd, R.x[c - 1], R.y[c - 1], R.x[c], R.y[c], par)) return 1;
        }
    }
    if (k < k1) {
        const int e = list[k];
        const float4 ed = Et[e];
        if (edge_step_f32(pu, pv, ed.x, ed.y, ed.z, ed.w, E, par, n_orient)) {
            ++n_esc;
            esc = true;
            const int32_t c = Ec[e];
            if (edge_step_f64(pxd, pyd, R.x[c - 1], R.y[c - 1], R.x[c], R.y[c], par)) return 1;
        }
    }
    return par ? 2 : 0;
}

// Stages every edge (c-1, c), c in (lo, hi), of a single-part polygon with at
// most kEdgeCap edges into Et/Ec (recentred fp32 and coordinate index).
// Block-uniform; returns the number staged.
__device__ int stage_all_edges(int32_t lo, int32_t hi, double cx, double cy, const RefineView& R, float4* Et,
                               int32_t* Ec, int* s_wsum) {
    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    int ne = 0;
    for (int32_t base = lo; base < hi; base += kLJThreads) {
        const int32_t c = base + tid;
        const bool inr = c > lo && c < hi;
        // flags and coordinates in one round of loads
        double ax = 0, ay = 0, bx = 0, by = 0;
        int f = 1;
        if (inr) {
            f = R.flags[c];
            ax = R.x[c - 1];
            ay = R.y[c - 1];
            bx = R.x[c];
            by = R.y[c];
        }
        const bool rel = inr && !(f & 1);
        const int inc = wave_inclusive_scan(rel ? 1 : 0);
        if (lane == 63) s_wsum[wid] = inc;
        __syncthreads();
        int woff = 0, tot = 0;
        for (int w2 = 0; w2 < kLJThreads / 64; ++w2) {
            woff += w2 < wid ? s_wsum[w2] : 0;
            tot += s_wsum[w2];
        }
        if (rel) {
            const int slot = ne + woff + inc - 1;
            Et[slot] = make_float4((float)(ax - cx), (float)(ay - cy), (float)(bx - cx), (float)(by - cy));
            Ec[slot] = c;
        }
        ne += tot;
        __syncthreads();
    }
    return ne;
}

// Orders one wave's LDS accesses around a point where its lanes exchange data
// through LDS (a compiler barrier: a wave's LDS instructions run in order).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// stage_all_edges + build_slabs for a polygon whose edges fit one pass of the
// block (hi - lo <= kLJThreads): each thread stages its own edge and counts
// and lists it in its slabs from registers, so the pair takes four barriers
// instead of eight.  The slabs and lists are those build_slabs makes (list
// order within a slab aside, which no decision depends on).  Block-uniform.
__device__ Slabs stage_and_bin(int32_t lo, int32_t hi, double cx, double cy, float ylo, float yhi,
                               const RefineView& R, float4* Et, int32_t* Ec, int* s_wsum, int* start, int* cur,
                               uint16_t* list) {
    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    const int32_t c = lo + tid;
    const bool inr = c > lo && c < hi;
    double ax = 0, ay = 0, bx = 0, by = 0;
    int f = 1;
    if (inr) {
        f = R.flags[c];
        ax = R.x[c - 1];
        ay = R.y[c - 1];
        bx = R.x[c];
        by = R.y[c];
    }
    const bool rel = inr && !(f & 1);
    if (tid <= kMaxBins) start[tid] = 0;
    const int inc = wave_inclusive_scan(rel ? 1 : 0);
    if (lane == 63) s_wsum[wid] = inc;
    __syncthreads();
    int woff = 0, ne = 0;
    for (int w2 = 0; w2 < kLJThreads / 64; ++w2) {
        woff += w2 < wid ? s_wsum[w2] : 0;
        ne += s_wsum[w2];
    }
    Slabs S;
    S.B = 1;
    while (S.B < kMaxBins && S.B * 2 < ne) S.B <<= 1;
    S.ylo = ylo;
    S.inv = (yhi > ylo) ? (float)S.B / (yhi - ylo) : 0.0f;
    if (!(S.inv > 0.0f) || !(S.inv < INFINITY)) {
        S.B = 1;
        S.inv = 0.0f;
    }
    const int slot = woff + inc - 1;
    int b0 = 0, b1 = -1;
    if (rel) {
        const float4 ed = make_float4((float)(ax - cx), (float)(ay - cy), (float)(bx - cx), (float)(by - cy));
        Et[slot] = ed;
        Ec[slot] = c;
        if (S.B > 1) {
            b0 = slab_of(fminf(ed.y, ed.w), S);
            b1 = slab_of(fmaxf(ed.y, ed.w), S);
            for (int b = b0; b <= b1; ++b) atomicAdd(&start[b + 1], 1);
        }
    }
    __syncthreads();
    if (S.B > 1) {
        if (tid < 64) {  // counts -> slab starts, and the fill cursors beside them
            const int v = tid < S.B ? start[tid + 1] : 0;
            const int incs = wave_inclusive_scan(v);
            if (tid < S.B) {
                start[tid + 1] = incs;
                cur[tid] = incs - v;
            }
        }
        __syncthreads();
        if (start[S.B] <= kBinCap) {
            for (int b = b0; b <= b1; ++b) list[atomicAdd(&cur[b], 1)] = (uint16_t)slot;
            __syncthreads();
            return S;
        }
        S.B = 1;  // too many long edges for the slab lists: one slab (start[0..1] below, read of start[S.B >= 2] above)
        S.inv = 0.0f;
    }
    if (tid == 0) {
        start[0] = 0;
        start[1] = ne;
    }
    for (int e = tid; e < ne; e += kLJThreads) list[e] = (uint16_t)e;
    __syncthreads();
    return S;
}

// LDS buffers of the staged edges and their slab lists.
struct EdgeLds {
    float4* Et;
    int32_t* Ec;
    int* start;
    int* cur;
    uint16_t* list;
    int* wsum;
};

// Stages the edges (c-1, c), c in (lo, hi), of a single-part polygon that can
// affect a point in [xmn, inf) x [ymn, ymx] (all of them when the bounds are
// infinite) into Et/Ec in coordinate order, refining the candidates against a
// full buffer before reusing it.  Block-uniform; returns the number left staged.
__device__ int stage_edges(int32_t lo, int32_t hi, double xmn, double ymn, double ymx, double cx, double cy,
                           const RefineView& R, const EdgeLds& L, int ne, const float* pu, const float* pv,
                           const double* pxd, const double* pyd, float E, int* loc, int* par, bool* done, bool* esc,
                           unsigned& n_orient, unsigned long long& n_esc) {
    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    float4* Et = L.Et;
    int32_t* Ec = L.Ec;
    int* s_wsum = L.wsum;
    for (int32_t base = lo; base < hi; base += kLJThreads) {
        const int32_t c = base + tid;
        bool rel = false;
        double ax = 0, ay = 0, bx = 0, by = 0;
        if (c > lo && c < hi && !(R.flags[c] & 1)) {
            ax = R.x[c - 1];
            ay = R.y[c - 1];
            bx = R.x[c];
            by = R.y[c];
            rel = fmax(ay, by) >= ymn && fmin(ay, by) <= ymx && fmax(ax, bx) >= xmn;
        }
        const int inc = wave_inclusive_scan(rel ? 1 : 0);
        if (lane == 63) s_wsum[wid] = inc;
        __syncthreads();
        int woff = 0, tot = 0;
        for (int w2 = 0; w2 < kLJThreads / 64; ++w2) {
            woff += w2 < wid ? s_wsum[w2] : 0;
            tot += s_wsum[w2];
        }
        if (ne + tot > kEdgeCap) {
            const Slabs S = build_slabs(Et, ne, (float)(ymn - cy), (float)(ymx - cy), L.start, L.cur, L.list);
            refine_slabs(Et, Ec, L.start, L.list, S, R, pu, pv, pxd, pyd, E, loc, par, done, esc, n_orient, n_esc);
            ne = 0;
            __syncthreads();
        }
        if (rel) {
            const int slot = ne + woff + inc - 1;
            Et[slot] = make_float4((float)(ax - cx), (float)(ay - cy), (float)(bx - cx), (float)(by - cy));
            Ec[slot] = c;
        }
        ne += tot;
        __syncthreads();
    }
    return ne;
}

// Wave mode carves each wave's quarter out of the block's buffers: edges and
// their coordinate indices (kEdgeCap / kClaim), slab lists (kBinCap / kClaim
// entries) and slab starts and cursors (kMaxBins / 2 + 1 and kMaxBins / 2
// ints of the candidate buffer).
static_assert(kEdgeCap / kClaim >= kWaveEdges, "wave mode: a wave's edge buffer holds kWaveEdges edges");
static_assert(kChunk / kClaim >= kMaxBins + 1, "wave mode: a wave's slab starts and cursors fit its quarter of cand");

// Persistent polygon-major join.  MODE kLModePip: query q is polygon row q
// (qbox = its bounds), output pairs (point, q).  MODE kLModeBox: query q is a
// box, output pairs (q, point).  Items (up to kItemPos scanned positions of one
// query) are taken in LPT order (queries sorted by decreasing estimated cost)
// and processed in rounds of kChunk positions.  A small single-part polygon's
// edges are staged in LDS once per item; larger ones are staged per round,
// keeping only edges that can affect the round's candidates; multi-part
// polygons stream every edge in order to settle each part's parity.
template <int MODE, int PRED>
// Four waves per SIMD (the register cap costs a few spills; the latency of one
// block's item start-up is hidden by the others on the CU).
__global__ __launch_bounds__(kLJThreads, 4) void learned_join_kernel(
    int64_t nq, const int32_t* __restrict__ item_off, const long long* __restrict__ total_items,
    const int4* __restrict__ item_rec, int64_t item_cap, const int32_t* __restrict__ order,
    const int4* __restrict__ poly_info, const int2* __restrict__ ranges, const int32_t* __restrict__ nranges,
    const double* __restrict__ sx, const double* __restrict__ sy, const int32_t* __restrict__ perm,
    const double4* __restrict__ qbox, RefineView R, uint32_t* __restrict__ stage_l, int32_t* __restrict__ stage_r,
    int64_t capacity, JoinCtl* __restrict__ ctl) {
    __shared__ int s_ncand, s_nout;
    __shared__ int4 s_claim[kClaim];  // records (query, chunk, range count) of the block's claimed items
    __shared__ unsigned long long s_base;
    __shared__ int2 s_rng[kMaxRanges];
    __shared__ int s_pre[kMaxRanges + 1];  // exclusive prefix of the range lengths
    __shared__ int32_t cand[kChunk];
    __shared__ int2 outb[kOutCap];
    __shared__ float4 Et[kEdgeCap];
    __shared__ int32_t Ec[kEdgeCap];
    __shared__ int s_wsum[kLJThreads / 64];
    __shared__ double s_ext[kLJThreads / 64][3];
    __shared__ int s_bstart[kMaxBins + 1], s_bcur[kMaxBins];
    __shared__ uint16_t s_blist[kBinCap];
    const EdgeLds L = {Et, Ec, s_bstart, s_bcur, s_blist, s_wsum};
    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    const long long T = *total_items;
    unsigned long long st_cand = 0, st_cverts = 0, st_orient = 0, st_esc = 0, st_escpairs = 0;
    if (tid == 0) s_nout = 0;
    // Appends the block's buffered pairs to the stage with one atomic.
    auto flush = [&]() {
        __syncthreads();
        if (tid == 0) s_base = s_nout ? atomicAdd(&ctl->staged, (unsigned long long)s_nout) : 0ull;
        __syncthreads();
        const int nout = s_nout;
        for (int k = tid; k < nout; k += kLJThreads) {
            const unsigned long long slot = s_base + (unsigned long long)k;
            if (slot < (unsigned long long)capacity) {
                const int2 e = outb[k];
                stage_l[slot] = (uint32_t)e.x;
                stage_r[slot] = e.y;
            }
        }
        __syncthreads();
        if (tid == 0) s_nout = 0;
        __syncthreads();
    };
    // Wave 0 claims kClaim consecutive items per queue atomic (the items are
    // in descending cost order, so late claims are small: one same-address
    // atomic per item would serialise in L2) and loads their records at once.
    int claim_k = kClaim;  // the next claimed item to run (block-uniform)
    for (;;) {
        if (claim_k == kClaim) {
            if (tid < 64) {
                int cbase = 0;
                if (lane == 0) cbase = (int)atomicAdd(&ctl->queue[0], (unsigned)kClaim);
                cbase = __shfl(cbase, 0, 64);
                if (lane < kClaim) {
                    const int item = cbase + lane;
                    int4 rec = make_int4(-1, 0, 0, 1);  // q < 0: past the last item (no work in wave mode)
                    if ((long long)item < T) {
                        if (item < item_cap) {
                            rec = item_rec[item];
                        } else {  // beyond the map: last sorted index with item_off <= item
                            int64_t lo = 0, hi = nq - 1;
                            while (lo < hi) {
                                const int64_t mid = (lo + hi + 1) >> 1;
                                if (item_off[mid] <= item) lo = mid;
                                else hi = mid - 1;
                            }
                            const int qq = order[lo];
                            rec = make_int4(qq, item - item_off[lo], nranges[qq], 0);
                        }
                    }
                    s_claim[lane] = rec;
                }
            }
            // the buffered pair count, read by every thread before the barrier
            // that precedes any append of the claimed items
            const int nout_claim = uniform_i32(s_nout);
            claim_k = 0;
            __syncthreads();
            const int4 c0 = s_claim[0], c1 = s_claim[1], c2 = s_claim[2], c3 = s_claim[3];
            if (uniform_i32(c0.x) >= 0 && uniform_i32(c0.w & c1.w & c2.w & c3.w)) {
                // Wave mode: four small items, one per wave, each synchronised
                // within its wave only (no block barrier per item).  Each
                // appends at most kWavePos pairs.
                if (nout_claim > kOutCap - kClaim * kWavePos) flush();
                const int4 rec = s_claim[wid];
                const int wq = uniform_i32(rec.x), wnr = uniform_i32(rec.z);
                if (wq >= 0) {
                    // ranges in lanes: lane k < wnr holds range k and the
                    // exclusive prefix of the range lengths
                    const int2 rg = lane < wnr ? ranges[(int64_t)wq * kMaxRanges + lane] : make_int2(0, 0);
                    const double4 wbox = qbox[wq];
                    const int4 winfo = MODE != kLModeBox ? poly_info[wq] : make_int4(0, 0, 0, 0);
                    const int len = rg.y - rg.x;
                    const int winc = wave_inclusive_scan(len);
                    const int wpre = winc - len;
                    const int wtotal = wnr > 0 ? readlane_i32(winc, wnr - 1) : 0;
                    int32_t wpos[kPosPerThread], worig[kPosPerThread];
                    double wpx[kPosPerThread], wpy[kPosPerThread];
#pragma unroll
                    for (int j = 0; j < kPosPerThread; ++j) {
                        const int off = lane + 64 * j;
                        int r = 0;  // last range with prefix <= off (every lane takes part in the shuffles)
                        for (int step = kMaxRanges / 2; step > 0; step >>= 1) {
                            const int cand_r = r + step;
                            const int pv = __shfl(wpre, cand_r, 64);
                            if (cand_r < wnr && pv <= off) r = cand_r;
                        }
                        const int rx = __shfl(rg.x, r, 64), rp = __shfl(wpre, r, 64);
                        wpos[j] = off < wtotal ? rx + (off - rp) : -1;
                    }
#pragma unroll
                    for (int j = 0; j < kPosPerThread; ++j) {
                        wpx[j] = wpos[j] >= 0 ? sx[wpos[j]] : NAN;
                        wpy[j] = wpos[j] >= 0 ? sy[wpos[j]] : NAN;
                        worig[j] = wpos[j] >= 0 ? perm[wpos[j]] : 0;
                    }
                    double wcx = 0, wcy = 0;
                    float wE = 0.0f;
                    Slabs WS = {1, 0.0f, 0.0f};
                    // this wave's quarter of the edge, slab and list buffers
                    float4* wEt = Et + wid * (kEdgeCap / kClaim);
                    int32_t* wEc = Ec + wid * (kEdgeCap / kClaim);
                    uint16_t* wlist = s_blist + wid * (kBinCap / kClaim);
                    int* wstart = cand + wid * (kChunk / kClaim);  // cand is free in wave mode
                    int* wcur = wstart + kMaxBins / 2 + 1;
                    if (MODE != kLModeBox) {
                        wcx = 0.5 * (wbox.x + wbox.z);
                        wcy = 0.5 * (wbox.y + wbox.w);
                        double M = fmax(fmax(wbox.z - wcx, wcx - wbox.x), fmax(wbox.w - wcy, wcy - wbox.y));
                        M *= 1.0 + 4.440892098500626e-16;
                        wE = uniform_f32(fp32_det_bound(M));
                        wcx = uniform_f64(wcx);
                        wcy = uniform_f64(wcy);
                        // edges (c-1, c), c = lo + 1 + lane + 64 k: two per lane
                        const int32_t lo = winfo.x, hi = winfo.y;
                        float4 ed[2];
                        int eslot[2];
                        bool erel[2];
                        int ne = 0;
#pragma unroll
                        for (int k = 0; k < 2; ++k) {
                            const int32_t c = lo + 1 + lane + 64 * k;
                            double ax = 0, ay = 0, bx = 0, by = 0;
                            int f = 1;
                            if (c < hi) {
                                f = R.flags[c];
                                ax = R.x[c - 1];
                                ay = R.y[c - 1];
                                bx = R.x[c];
                                by = R.y[c];
                            }
                            erel[k] = c < hi && !(f & 1);
                            ed[k] = make_float4((float)(ax - wcx), (float)(ay - wcy), (float)(bx - wcx),
                                                (float)(by - wcy));
                            const int einc = wave_inclusive_scan(erel[k] ? 1 : 0);
                            eslot[k] = ne + einc - 1;
                            ne += readlane_i32(einc, 63);
                            if (erel[k]) {
                                wEt[eslot[k]] = ed[k];
                                wEc[eslot[k]] = c;
                            }
                        }
                        // y-slabs as build_slabs makes them, within the wave
                        WS.B = 1;
                        while (WS.B < kMaxBins / 2 && WS.B * 2 < ne) WS.B <<= 1;
                        const float ylo = (float)(wbox.y - wcy), yhi = (float)(wbox.w - wcy);
                        WS.ylo = ylo;
                        WS.inv = (yhi > ylo) ? (float)WS.B / (yhi - ylo) : 0.0f;
                        if (!(WS.inv > 0.0f) || !(WS.inv < INFINITY)) {
                            WS.B = 1;
                            WS.inv = 0.0f;
                        }
                        bool listed = false;
                        if (WS.B > 1) {
                            if (lane <= WS.B) wstart[lane] = 0;
                            wave_lds_sync();
                            int eb0[2], eb1[2];
#pragma unroll
                            for (int k = 0; k < 2; ++k) {
                                eb0[k] = 0;
                                eb1[k] = -1;
                                if (erel[k]) {
                                    eb0[k] = slab_of(fminf(ed[k].y, ed[k].w), WS);
                                    eb1[k] = slab_of(fmaxf(ed[k].y, ed[k].w), WS);
                                    for (int b = eb0[k]; b <= eb1[k]; ++b) atomicAdd(&wstart[b + 1], 1);
                                }
                            }
                            wave_lds_sync();
                            const int v = lane < WS.B ? wstart[lane + 1] : 0;
                            const int incs = wave_inclusive_scan(v);
                            if (lane < WS.B) {
                                wstart[lane + 1] = incs;
                                wcur[lane] = incs - v;
                            }
                            const int wtot = readlane_i32(incs, 63);
                            wave_lds_sync();
                            if (wtot <= kBinCap / kClaim) {
#pragma unroll
                                for (int k = 0; k < 2; ++k)
                                    for (int b = eb0[k]; b <= eb1[k]; ++b)
                                        wlist[atomicAdd(&wcur[b], 1)] = (uint16_t)eslot[k];
                                listed = true;
                            } else {  // too many long edges: one slab
                                WS.B = 1;
                                WS.inv = 0.0f;
                            }
                        }
                        if (!listed) {
                            if (lane == 0) {
                                wstart[0] = 0;
                                wstart[1] = ne;
                            }
#pragma unroll
                            for (int k = 0; k < 2; ++k)
                                if (erel[k]) wlist[eslot[k]] = (uint16_t)eslot[k];
                        }
                        wave_lds_sync();
                    }
                    unsigned n_orient = 0;
                    unsigned long long n_esc = 0;
#pragma unroll
                    for (int j = 0; j < kPosPerThread; ++j) {
                        const bool inbox =
                            wpx[j] >= wbox.x && wpx[j] <= wbox.z && wpy[j] >= wbox.y && wpy[j] <= wbox.w;
                        bool keep = inbox;
                        if (MODE != kLModeBox) {
                            bool esc = false;
                            if (inbox) {
                                const int loc = refine_one(wEt, wEc, wstart, wlist, WS, R, (float)(wpx[j] - wcx),
                                                           (float)(wpy[j] - wcy), wpx[j], wpy[j], wE, esc, n_orient,
                                                           n_esc);
                                keep = loc_pred<PRED>(loc);
                            }
                            const unsigned nb = (unsigned)__popcll(__ballot(inbox));
                            st_cand += nb;
                            st_cverts += (unsigned long long)nb * (unsigned long long)(winfo.y - winfo.x);
                            st_escpairs += (unsigned)__popcll(__ballot(esc));
                        }
                        const uint64_t m = __ballot(keep);
                        if (m) {
                            const int leader = __ffsll((unsigned long long)m) - 1;
                            int base = 0;
                            if (lane == leader) base = atomicAdd(&s_nout, __popcll(m));
                            base = __shfl(base, leader, 64);
                            if (keep)
                                outb[base + __popcll(m & lanemask_lt())] =
                                    MODE == kLModePip ? make_int2(worig[j], wq) : make_int2(wq, worig[j]);
                        }
                    }
                    st_orient += n_orient;
                    st_esc += n_esc;
                }
                __syncthreads();
                claim_k = kClaim;
                continue;
            }
        }
        // item values are block-uniform: kept in scalar registers
        const int q = uniform_i32(s_claim[claim_k].x);
        const int chunk = uniform_i32(s_claim[claim_k].y);
        const int nr = uniform_i32(s_claim[claim_k].z);
        ++claim_k;
        if (q < 0) break;
        // the query's box and polygon info load beside its ranges
        const double4 box = qbox[q];
        const int4 qinfo = MODE != kLModeBox ? poly_info[q] : make_int4(0, 0, 0, 0);
        if (tid < 64) {  // wave 0: the ranges and their exclusive prefix in one pass
            const int2 rg = tid < nr ? ranges[(int64_t)q * kMaxRanges + tid] : make_int2(0, 0);
            const int len = rg.y - rg.x;
            const int inc = wave_inclusive_scan(len);
            if (tid < nr) s_rng[tid] = rg;
            if (tid <= nr) s_pre[tid] = inc - len;
        }
        // the block's buffered pair count, read by every thread before the
        // barrier that precedes the item's first append
        const int nout_item = uniform_i32(s_nout);
        __syncthreads();
        const int total = uniform_i32(s_pre[nr]);
        const int first = chunk * kItemPos;
        const int last = (first + kItemPos < total) ? first + kItemPos : total;
        int4 info = make_int4(0, 0, 0, 0);
        double cx = 0, cy = 0;
        float E = 0.0f;
        int ne_fixed = -1;  // edges staged once for the whole item (small single-part polygons)
        Slabs fixed_slabs = {1, 0.0f, 0.0f};
        if (MODE != kLModeBox) {
            info = qinfo;
            cx = 0.5 * (box.x + box.z);
            cy = 0.5 * (box.y + box.w);
            double M = fmax(fmax(box.z - cx, cx - box.x), fmax(box.w - cy, cy - box.y));
            M *= 1.0 + 4.440892098500626e-16;
            E = uniform_f32(fp32_det_bound(M));
            cx = uniform_f64(cx);
            cy = uniform_f64(cy);
        }
        const bool direct = MODE == kLModeBox || (info.z == 1 && info.y - info.x <= kEdgeCap + 1);
        if (direct) {
            // Box queries and small single-part polygons: every thread scans
            // kPosPerThread positions per round and refines its own candidates
            // at once against the item's slab lists (no candidate compaction).
            int32_t pos[kPosPerThread], orig[kPosPerThread];
            double px[kPosPerThread], py[kPosPerThread];
            // a round's positions, their coordinates and original ids in one
            // round of loads (the ids of the scanned ranges are contiguous, so
            // the extra ones cost little next to a dependent load after the
            // refine)
            auto load_round = [&](int r0) {
#pragma unroll
                for (int j = 0; j < kPosPerThread; ++j) {
                    const int off = r0 + tid + j * kLJThreads;
                    pos[j] = -1;
                    if (off < last) {
                        int r = 0;  // last range with s_pre[r] <= off
                        for (int step = kMaxRanges / 2; step > 0; step >>= 1)
                            if (r + step < nr && s_pre[r + step] <= off) r += step;
                        pos[j] = s_rng[r].x + (off - s_pre[r]);
                    }
                }
#pragma unroll
                for (int j = 0; j < kPosPerThread; ++j) {
                    px[j] = pos[j] >= 0 ? sx[pos[j]] : NAN;
                    py[j] = pos[j] >= 0 ? sy[pos[j]] : NAN;
                    orig[j] = pos[j] >= 0 ? perm[pos[j]] : 0;
                }
            };
            // the first round's loads go out before the edge staging's, so the
            // two round trips overlap
            load_round(first);
            Slabs S = {1, 0.0f, 0.0f};
            if (MODE != kLModeBox) {
                const float ylo = (float)(box.y - cy), yhi = (float)(box.w - cy);
                if (info.y - info.x <= kLJThreads) {
                    S = stage_and_bin(info.x, info.y, cx, cy, ylo, yhi, R, Et, Ec, s_wsum, s_bstart, s_bcur, s_blist);
                } else {
                    const int ne = stage_all_edges(info.x, info.y, cx, cy, R, Et, Ec, s_wsum);
                    S = build_slabs(Et, ne, ylo, yhi, s_bstart, s_bcur, s_blist);
                }
            }
            unsigned n_orient = 0;
            unsigned long long n_esc = 0;
            for (int r0 = first; r0 < last; r0 += kRoundPos) {
                // the flush test must see one value in every wave: a count read
                // between the previous round's appends and a barrier that this
                // round's appends follow (for the first round, the count read
                // before the ranges barrier)
                int nout_now = nout_item;
                if (r0 != first) {
                    __syncthreads();
                    nout_now = uniform_i32(s_nout);
                    load_round(r0);
                    __syncthreads();
                }
                if (nout_now > kOutCap - kRoundPos) flush();
#pragma unroll
                for (int j = 0; j < kPosPerThread; ++j) {
                    const bool inbox = px[j] >= box.x && px[j] <= box.z && py[j] >= box.y && py[j] <= box.w;
                    bool keep = inbox;
                    if (MODE != kLModeBox) {
                        bool esc = false;
                        if (inbox) {
                            const int loc = refine_one(Et, Ec, s_bstart, s_blist, S, R, (float)(px[j] - cx),
                                                       (float)(py[j] - cy), px[j], py[j], E, esc, n_orient, n_esc);
                            keep = loc_pred<PRED>(loc);
                        }
                        // wave-uniform counters (scalar registers)
                        const unsigned nb = (unsigned)__popcll(__ballot(inbox));
                        st_cand += nb;
                        st_cverts += (unsigned long long)nb * (unsigned long long)(info.y - info.x);
                        st_escpairs += (unsigned)__popcll(__ballot(esc));
                    }
                    // wave-aggregated append to the block buffer
                    const uint64_t m = __ballot(keep);
                    if (m) {
                        const int leader = __ffsll((unsigned long long)m) - 1;
                        int base = 0;
                        if (lane == leader) base = atomicAdd(&s_nout, __popcll(m));
                        base = __shfl(base, leader, 64);
                        if (keep)
                            outb[base + __popcll(m & lanemask_lt())] =
                                MODE == kLModePip ? make_int2(orig[j], q) : make_int2(q, orig[j]);
                    }
                }
            }
            st_orient += n_orient;
            st_esc += n_esc;
            __syncthreads();
            continue;
        }
        if (tid == 0) s_ncand = 0;
        __syncthreads();
        for (int r0 = first; r0 < last; r0 += kLJThreads) {
            // 1. scan kLJThreads positions with the exact box test; candidates
            // accumulate until a full block's worth is buffered
            {
                const int off = r0 + tid;
                if (off < last) {
                    int r = 0;  // last range with s_pre[r] <= off (off < total, so r < nr)
                    for (int step = kMaxRanges / 2; step > 0; step >>= 1)
                        if (r + step < nr && s_pre[r + step] <= off) r += step;
                    {
                        const int32_t pos = s_rng[r].x + (off - s_pre[r]);
                        const double px = sx[pos], py = sy[pos];
                        if (px >= box.x && px <= box.z && py >= box.y && py <= box.w) {
                            const int slot = atomicAdd(&s_ncand, 1);
                            cand[slot] = pos;
                        }
                    }
                }
            }
            __syncthreads();
            const int ncand = s_ncand;
            __syncthreads();  // every wave has read s_ncand before it can change again
            if (ncand < kLJThreads && r0 + kLJThreads < last) continue;  // keep accumulating
            if (ncand == 0) continue;
            if (s_nout + ncand > kOutCap) flush();
            if (MODE == kLModeBox) {
                const int base = s_nout;
                for (int k = tid; k < ncand; k += kLJThreads) outb[base + k] = make_int2(q, perm[cand[k]]);
                __syncthreads();
                if (tid == 0) {
                    s_nout = base + ncand;
                    s_ncand = 0;
                }
                __syncthreads();
                continue;
            }
            // 2. refine the round's candidates against the polygon
            const int32_t c0 = info.x, c1 = info.y;
            int loc[kCandPerThread], par[kCandPerThread];
            bool done[kCandPerThread], esc[kCandPerThread];
            float pu[kCandPerThread], pv[kCandPerThread];
            double pxd[kCandPerThread], pyd[kCandPerThread];
            double exmin = INFINITY, eymin = INFINITY, eymax = -INFINITY;
#pragma unroll
            for (int j = 0; j < kCandPerThread; ++j) {
                const int k = tid + j * kLJThreads;
                loc[j] = 0;
                par[j] = 0;
                esc[j] = false;
                done[j] = k >= ncand;
                if (k < ncand) {
                    pxd[j] = sx[cand[k]];
                    pyd[j] = sy[cand[k]];
                    pu[j] = (float)(pxd[j] - cx);
                    pv[j] = (float)(pyd[j] - cy);
                    exmin = fmin(exmin, pxd[j]);
                    eymin = fmin(eymin, pyd[j]);
                    eymax = fmax(eymax, pyd[j]);
                } else {
                    pxd[j] = pyd[j] = 0.0;
                    pu[j] = pv[j] = 0.0f;
                }
            }
            unsigned n_orient = 0;
            unsigned long long n_esc = 0;
            if (info.z == 1 && c1 - c0 <= kEdgeCap + 1) {
                if (ne_fixed < 0) {  // stage every edge and its slabs once for the item
                    ne_fixed = stage_edges(c0, c1, -INFINITY, -INFINITY, INFINITY, cx, cy, R, L, 0, pu, pv, pxd, pyd,
                                           E, loc, par, done, esc, n_orient, n_esc);
                    fixed_slabs = build_slabs(Et, ne_fixed, (float)(box.y - cy), (float)(box.w - cy), s_bstart,
                                              s_bcur, s_blist);
                }
                refine_slabs(Et, Ec, s_bstart, s_blist, fixed_slabs, R, pu, pv, pxd, pyd, E, loc, par, done, esc,
                             n_orient, n_esc);
            } else if (info.z == 1) {
                exmin = wave_min(exmin);
                eymin = wave_min(eymin);
                eymax = wave_max(eymax);
                if (lane == 0) {
                    s_ext[wid][0] = exmin;
                    s_ext[wid][1] = eymin;
                    s_ext[wid][2] = eymax;
                }
                __syncthreads();
                double xmn = s_ext[0][0], ymn = s_ext[0][1], ymx = s_ext[0][2];
                for (int w2 = 1; w2 < kLJThreads / 64; ++w2) {
                    xmn = fmin(xmn, s_ext[w2][0]);
                    ymn = fmin(ymn, s_ext[w2][1]);
                    ymx = fmax(ymx, s_ext[w2][2]);
                }
                __syncthreads();
                const int ne = stage_edges(c0, c1, xmn, ymn, ymx, cx, cy, R, L, 0, pu, pv, pxd, pyd, E, loc, par,
                                           done, esc, n_orient, n_esc);
                const Slabs S = build_slabs(Et, ne, (float)(ymn - cy), (float)(ymx - cy), s_bstart, s_bcur, s_blist);
                refine_slabs(Et, Ec, s_bstart, s_blist, S, R, pu, pv, pxd, pyd, E, loc, par, done, esc, n_orient,
                             n_esc);
            } else {
                float2* Vt = reinterpret_cast<float2*>(Et);
                uint8_t* Ft = reinterpret_cast<uint8_t*>(Ec);
                constexpr int kV = kEdgeCap;  // vertices per batch: 8 B + 1 B each fit the edge buffers
                for (int32_t b0 = c0; b0 < c1; b0 += kV) {
                    const int32_t b1 = (b0 + kV < c1) ? b0 + kV : c1;
                    __syncthreads();
                    for (int32_t c = b0 - 1 + tid; c < b1; c += kLJThreads) {
                        if (c < c0) continue;
                        Vt[c - b0 + 1] = make_float2((float)(R.x[c] - cx), (float)(R.y[c] - cy));
                        Ft[c - b0 + 1] = R.flags[c];
                    }
                    __syncthreads();
                    for (int32_t c = b0; c < b1; ++c) {
                        const int f = Ft[c - b0 + 1];
                        if (f & 2) {
#pragma unroll
                            for (int j = 0; j < kCandPerThread; ++j) {
                                if (!done[j] && par[j]) {
                                    loc[j] = 2;
                                    done[j] = true;
                                }
                                par[j] = 0;
                            }
                        }
                        if (f & 1) continue;
                        const float2 a = Vt[c - b0], b = Vt[c - b0 + 1];
#pragma unroll
                        for (int j = 0; j < kCandPerThread; ++j) {
                            if (done[j]) continue;
                            if (edge_step_f32(pu[j], pv[j], a.x, a.y, b.x, b.y, E, par[j], n_orient)) {
                                ++n_esc;
                                esc[j] = true;
                                if (edge_step_f64(pxd[j], pyd[j], R.x[c - 1], R.y[c - 1], R.x[c], R.y[c], par[j])) {
                                    loc[j] = 1;
                                    done[j] = true;
                                }
                            }
                        }
                    }
                }
                __syncthreads();
                ne_fixed = -1;  // the edge buffers were reused as vertex buffers
            }
            st_orient += n_orient;
            st_esc += n_esc;
#pragma unroll
            for (int j = 0; j < kCandPerThread; ++j) {
                const int k = tid + j * kLJThreads;
                st_escpairs += (unsigned)__popcll(__ballot(k < ncand && esc[j]));
                if (k >= ncand) continue;
                if (!done[j] && par[j]) loc[j] = 2;
                if (loc_pred<PRED>(loc[j])) {
                    const int32_t orig = perm[cand[k]];
                    const int slot = atomicAdd(&s_nout, 1);
                    outb[slot] = MODE == kLModePip ? make_int2(orig, q) : make_int2(q, orig);
                }
            }
            if (wid == 0) {  // block counts, taken once per block
                st_cand += (unsigned long long)ncand;
                st_cverts += (unsigned long long)ncand * (unsigned long long)(c1 - c0);
            }
            __syncthreads();
            if (tid == 0) s_ncand = 0;
            __syncthreads();
        }
        __syncthreads();
    }
    flush();
    // counters: [0] candidates, [1] candidate vertices (wave-uniform), [2] fp32
    // orientation evaluations, [3] escalated edges (per lane); candidate pairs
    // with an escalation (wave-uniform)
    unsigned long long v[2] = {st_orient, st_esc};
    for (int k = 0; k < 2; ++k) {
        unsigned long long x = v[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        v[k] = x;
    }
    if (lane == 0) {
        if (st_cand) atomicAdd(&ctl->stat[0], st_cand);
        if (st_cverts) atomicAdd(&ctl->stat[1], st_cverts);
        if (v[0]) atomicAdd(&ctl->stat[2], v[0]);
        if (v[1]) atomicAdd(&ctl->stat[3], v[1]);
        if (st_escpairs) atomicAdd(&ctl->escalated_pairs, st_escpairs);
    }
}

// LPT key: queries with larger estimated work (items x vertices) sort first.
__global__ __launch_bounds__(256) void gather_counts_kernel(const int32_t* __restrict__ n_items,
                                                           const int32_t* __restrict__ order, int64_t nq,
                                                           int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < nq) out[i] = n_items[order[i]];
}

// item_rec[k] = (query, chunk within the query, the query's range count,
// small-item flag) of item k, for the first `cap` items (items beyond are located by binary
// search over item_off).  q runs over the queries in LPT order.
__global__ __launch_bounds__(256) void item_map_kernel(const int32_t* __restrict__ item_off,
                                                      const int32_t* __restrict__ n_items,
                                                      const int32_t* __restrict__ order,
                                                      const uint32_t* __restrict__ cost_key,
                                                      const int32_t* __restrict__ nranges, int64_t nq,
                                                      int4* __restrict__ item_rec, int64_t cap) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= nq) return;
    const int64_t o = item_off[q];
    const int32_t n = n_items[q];
    if (n == 0 || o >= cap) return;
    const int32_t qq = order[q];
    const int32_t nr = nranges[qq];
    const int32_t small = cost_key[q] >= 128u;  // see cover_total_kernel
    for (int32_t k = 0; k < n && o + k < cap; ++k) item_rec[o + k] = make_int4(qq, k, nr, small);
}

__global__ void publish_learned_kernel(const JoinCtl* __restrict__ ctl, const unsigned long long* __restrict__ cst,
                                       const int32_t* __restrict__ meta, int64_t* __restrict__ result) {
    if (threadIdx.x == 0) {
        result[4] = (int64_t)ctl->stat[2];
        result[5] = (int64_t)ctl->stat[3];
        result[6] = (int64_t)ctl->escalated_pairs;
        result[7] = (int64_t)cst[0];   // lookup probes
        result[8] = (int64_t)cst[1];   // MLP leaf-window fallbacks
        result[9] = (int64_t)cst[2];   // leaf error-window fallbacks
        result[10] = (int64_t)cst[3];  // positions scanned
        result[11] = meta ? (int64_t)meta[0] : 0;
    }
}

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------
template <class K>
struct IndexBufs {
    K* keys;
    int32_t* perm;
    double* sx;
    double* sy;
    double4* ext;
    float* mlp;
    float* losses;
    K* leaf_key;
    int32_t* leaf_err;
    int2* leaf_win;
    int32_t* meta;
    void* build_ws;
};

template <class K>
IndexBufs<K> carve_index(Carver& c, int64_t n, int S) {
    IndexBufs<K> B;
    B.keys = c.take<K>(n);
    B.perm = c.take<int32_t>(n);
    B.sx = c.take<double>(n);
    B.sy = c.take<double>(n);
    B.ext = c.take<double4>(1);
    B.mlp = c.take<float>(kNParams);
    B.losses = c.take<float>(2);
    B.leaf_key = c.take<K>(S + 1);
    B.leaf_err = c.take<int32_t>(S);
    B.leaf_win = c.take<int2>(S);
    B.meta = c.take<int32_t>(4);
    B.build_ws = c.take<char>(sizeof(K) == 4 ? point_index_workspace_bytes_u32(n) : point_index_workspace_bytes_u64(n));
    return B;
}

inline int leaves_for(int64_t n, int leaf_size) {
    const int64_t S = n > 1 ? (n - 1 + leaf_size - 1) / leaf_size : 1;
    return (int)(S < 1 ? 1 : S);
}

__global__ void zero_i32_kernel(int32_t* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// Empty routing windows (first > last) before the error pass records them.
__global__ __launch_bounds__(256) void clear_win_kernel(int2* __restrict__ w, int S) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < S; i += gridDim.x * 256) w[i] = make_int2(INT_MAX, -1);
}

// A second stream (default priority) for work that overlaps the caller's
// stream inside one call, with a fork / join event pair: one per host thread
// and device, created on the thread's first such call on the device and kept
// until the process ends.  Threads never share one, so concurrent calls from
// different host threads do not interfere.
struct SideStream {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

static int side_stream(SideStream** out) {
    static thread_local SideStream per_dev[64];
    int dev = 0;
    VS_CHECK_HIP(hipGetDevice(&dev));
    VS_REQUIRE(dev >= 0 && dev < 64, "device index %d out of range", dev);
    SideStream& S = per_dev[dev];
    if (!S.stream) {
        VS_CHECK_HIP(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
        VS_CHECK_HIP(hipEventCreateWithFlags(&S.fork, hipEventDisableTiming));
        VS_CHECK_HIP(hipEventCreateWithFlags(&S.join, hipEventDisableTiming));
    }
    *out = &S;
    return VS_OK;
}

// Keys and their sort on the caller's stream; then the network's training,
// the leaf table and the error pass (which need only the sorted keys) run on
// the side stream while the caller's stream gathers the sorted coordinates
// (fork), or after them on the caller's stream (no fork: nothing is created).
// The call joins both before returning.
// The side stream of an index build between its fork and its join.
struct IndexFork {
    SideStream* side = nullptr;
    hipStream_t t = nullptr;
};

// Everything up to the join: keys and sort on `s`, then (fork) training and
// the error pass on the side stream beside the coordinate gather on `s`.
// The caller may queue more work on `s` before index_build_join.
template <class K>
int index_build_fork(const double* px, const double* py, int64_t n, int S, int iters, float lr,
                     const IndexBufs<K>& B, hipStream_t s, bool fork, IndexFork* out) {
    if (sizeof(K) == 4)
        VS_TRY(point_index_build_u32(px, py, n, reinterpret_cast<uint32_t*>(B.keys), B.perm, nullptr, nullptr, B.ext,
                                     B.build_ws, s));
    else
        VS_TRY(point_index_build_u64(px, py, n, reinterpret_cast<uint64_t*>(B.keys), B.perm, nullptr, nullptr, B.ext,
                                     B.build_ws, s));
    SideStream* side = nullptr;
    hipStream_t t = s;
    if (fork) {
        VS_TRY(side_stream(&side));
        VS_CHECK_HIP(hipEventRecord(side->fork, s));
        VS_CHECK_HIP(hipStreamWaitEvent(side->stream, side->fork, 0));
        t = side->stream;
    }
    if (n >= 2) {
        hipLaunchKernelGGL(clear_win_kernel, dim3(stride_grid(S, 256, 64)), dim3(256), 0, t, B.leaf_win, S);
        VS_LAUNCH_CHECK();
        hipLaunchKernelGGL(mlp_train_kernel<K>, dim3(1), dim3(vs::kTrainThreads), 0, t, B.keys, n, iters, lr, B.mlp,
                           B.losses, B.meta);
        VS_LAUNCH_CHECK();
        hipLaunchKernelGGL(leaf_err_kernel<K>, dim3(stride_grid(S, 4, 64)), dim3(256), 0, t, B.keys, n, S, B.leaf_key,
                           B.mlp, B.leaf_err, B.leaf_win, B.meta);
        VS_LAUNCH_CHECK();
    } else {
        hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, t, B.meta, 4);
        VS_LAUNCH_CHECK();
    }
    // two gather blocks per CU (enough to keep the random reads in flight)
    // leave room for the side stream's kernels
    const double2* xy = sizeof(K) == 4 ? point_index_xy_u32(B.build_ws, n) : point_index_xy_u64(B.build_ws, n);
    VS_TRY(gather_points_xy(xy, B.perm, n, B.sx, B.sy, s, fork ? 2 : 8));
    out->side = side;
    out->t = t;
    return VS_OK;
}

// `s` waits for the side stream's work.
inline int index_build_join(const IndexFork& f, hipStream_t s) {
    if (f.side) {
        VS_CHECK_HIP(hipEventRecord(f.side->join, f.t));
        VS_CHECK_HIP(hipStreamWaitEvent(s, f.side->join, 0));
    }
    return VS_OK;
}

template <class K>
int index_build_t(const double* px, const double* py, int64_t n, int S, int iters, float lr, const IndexBufs<K>& B,
                  hipStream_t s, bool fork = true) {
    IndexFork f;
    VS_TRY(index_build_fork<K>(px, py, n, S, iters, lr, B, s, fork, &f));
    return index_build_join(f, s);
}

template <class K>
LearnedView<K> view_of(const IndexBufs<K>& B, int64_t n, int S, int mode) {
    LearnedView<K> V;
    V.keys = B.keys;
    V.n = n;
    V.mlp = B.mlp;
    V.leaf_key = B.leaf_key;
    V.leaf_err = B.leaf_err;
    V.leaf_win = B.leaf_win;
    V.meta = B.meta;
    V.S = S;
    V.mode = n < 2 ? 2 : mode;
    return V;
}

struct QueryBufs {
    int4* item_rec;
    int64_t item_cap;
    int4* poly_info;
    uint32_t* cost_in;  // unsorted cost keys and query ids
    int32_t* order_in;
    uint32_t* cost_key;
    int32_t* order;
    int32_t* n_items_sorted;
    void* sort_ws;
    int2* ranges;
    unsigned long long* keyr;
    int32_t* nranges;
    int32_t* n_items;
    int32_t* item_off;
    long long* total_items;
    unsigned long long* cstat;
    void* scan_ws;
    JoinCtl* ctl;
    void* fin_ws;
    uint32_t* stage_l;
    int32_t* stage_r;
};

// Items map: room for 4 items per query plus one per kChunk/4 left rows covers
// typical covers; items beyond the map fall back to a binary search.
inline int64_t item_cap_for(int64_t nq, int64_t n_left) { return 2 * nq + (8 * n_left) / kItemPos + 1024; }

QueryBufs carve_query(Carver& c, int64_t nq, int64_t n_left, int64_t capacity) {
    QueryBufs Q;
    Q.item_cap = item_cap_for(nq, n_left);
    Q.item_rec = c.take<int4>(Q.item_cap);
    Q.poly_info = c.take<int4>(nq);
    Q.cost_in = c.take<uint32_t>(nq);
    Q.order_in = c.take<int32_t>(nq);
    Q.cost_key = c.take<uint32_t>(nq);
    Q.order = c.take<int32_t>(nq);
    Q.n_items_sorted = c.take<int32_t>(nq);
    Q.sort_ws = c.take<char>(sort_workspace_bytes_u32(nq));
    Q.ranges = c.take<int2>(nq * kMaxRanges);
    Q.keyr = c.take<unsigned long long>(nq * kMaxRanges * 2);
    Q.nranges = c.take<int32_t>(nq);
    Q.n_items = c.take<int32_t>(nq);
    Q.item_off = c.take<int32_t>(nq);
    Q.total_items = c.take<long long>(1);
    Q.cstat = c.take<unsigned long long>(4);
    Q.scan_ws = c.take<char>(scan_workspace_bytes(nq));
    Q.ctl = c.take<JoinCtl>(1);
    Q.fin_ws = c.take<char>(finish_workspace_bytes(capacity));
    Q.stage_l = nullptr;
    Q.stage_r = nullptr;
    if (c.base) finish_stage(Q.fin_ws, capacity, &Q.stage_l, &Q.stage_r);
    return Q;
}

// Cover cells of every query box (and, for polygon queries, their coordinate
// ranges and part counts); needs only the point extent of the index.
template <class K>
int cover_cells_launch(int mode, const double4* qbox, int64_t nq, const IndexBufs<K>& B, const RefineView& R,
                       const QueryBufs& Q, hipStream_t s) {
    if (nq <= 0) return VS_OK;
    int4* pinfo = mode != kLModeBox ? Q.poly_info : nullptr;
    hipLaunchKernelGGL(cover_cells_kernel<K>, dim3((unsigned)ceil_div(nq, 256)), dim3(256), 0, s, qbox, nq, B.ext,
                       Q.keyr, Q.nranges, R.geom, R.part, R.ring, pinfo);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

// cells_ready: cover_cells_launch already ran on `s` for these queries.
template <class K>
int learned_query_t(int mode, int pred, const double4* qbox, int64_t nq, const LearnedView<K>& V,
                    const IndexBufs<K>& B, const RefineView& R, int64_t n_left, int64_t capacity, bool sort,
                    const QueryBufs& Q, int64_t* out_l, int64_t* out_r, int64_t* d_result, hipStream_t s,
                    bool cells_ready = false) {
    ctl_reset(Q.ctl, s, Q.cstat, 4);
    VS_LAUNCH_CHECK();
    if (nq > 0) {
        const unsigned gq = (unsigned)ceil_div(nq, 256);
        int4* pinfo = mode != kLModeBox ? Q.poly_info : nullptr;
        if (!cells_ready) VS_TRY(cover_cells_launch<K>(mode, qbox, nq, B, R, Q, s));
        hipLaunchKernelGGL(cover_lookup_kernel<K>, dim3(stride_grid(nq * kMaxRanges * 2, 256, 8)), dim3(256), 0, s,
                           nq, Q.keyr, Q.nranges, V, reinterpret_cast<int32_t*>(Q.ranges), Q.cstat);
        VS_LAUNCH_CHECK();
        hipLaunchKernelGGL(cover_total_kernel, dim3(stride_grid(nq, 256)), dim3(256), 0, s, nq, Q.ranges, Q.nranges,
                           pinfo, Q.n_items, Q.cost_in, Q.order_in, Q.cstat);
        VS_LAUNCH_CHECK();
        // longest-first order of the queries, then item offsets in that order
        VS_TRY(sort_pairs_into_u32(Q.cost_in, Q.order_in, Q.cost_key, Q.order, nq, 0, 8, Q.sort_ws,
                                   sort_workspace_bytes_u32(nq), s));
        hipLaunchKernelGGL(gather_counts_kernel, dim3(gq), dim3(256), 0, s, Q.n_items, Q.order, nq, Q.n_items_sorted);
        VS_LAUNCH_CHECK();
        VS_TRY(exclusive_scan_i32(Q.n_items_sorted, Q.item_off, nq, Q.total_items, Q.scan_ws, s));
        hipLaunchKernelGGL(item_map_kernel, dim3(gq), dim3(256), 0, s, Q.item_off, Q.n_items_sorted, Q.order,
                           Q.cost_key, Q.nranges, nq, Q.item_rec, Q.item_cap);
        VS_LAUNCH_CHECK();
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess) {
            int v = 0;
            if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) cus = v;
        }
        const int grid = cus * 4;
        VS_TRY(timing_begin_public(s));
#define VS_LJ(M, P)                                                                                          \
    hipLaunchKernelGGL((learned_join_kernel<M, P>), dim3(grid), dim3(kLJThreads), 0, s, nq, Q.item_off,       \
                       Q.total_items, Q.item_rec, Q.item_cap, Q.order, Q.poly_info, Q.ranges, Q.nranges, B.sx, B.sy,  \
                       B.perm, qbox, R, Q.stage_l, Q.stage_r, capacity, Q.ctl)
        if (mode == kLModeBox) VS_LJ(kLModeBox, VS_PRED_INTERSECTS);
        else if (mode == kLModePipQ) {
            // polygon (left) over point (right): interior for contains and
            // contains_properly, boundary for touches, else not outside
            if (pred == VS_PRED_CONTAINS || pred == VS_PRED_CONTAINS_PROPERLY) VS_LJ(kLModePipQ, VS_PRED_WITHIN);
            else if (pred == VS_PRED_TOUCHES) VS_LJ(kLModePipQ, VS_PRED_TOUCHES);
            else VS_LJ(kLModePipQ, VS_PRED_INTERSECTS);
        } else if (pred == VS_PRED_WITHIN) VS_LJ(kLModePip, VS_PRED_WITHIN);
        else if (pred == VS_PRED_TOUCHES) VS_LJ(kLModePip, VS_PRED_TOUCHES);
        else VS_LJ(kLModePip, VS_PRED_INTERSECTS);
#undef VS_LJ
        VS_LAUNCH_CHECK();
        VS_TRY(timing_end_public(s));
    }
    VS_TRY(finish_pairs(Q.fin_ws, Q.ctl, n_left, capacity, sort, out_l, out_r, d_result, s));
    hipLaunchKernelGGL(publish_learned_kernel, dim3(1), dim3(64), 0, s, Q.ctl, Q.cstat, B.meta, d_result);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

// The learned join on a carved workspace (fork: the index build's side
// stream; the allocation-free entry point runs everything on `s`).
int learned_join_impl(const double* px, const double* py, int64_t n_points, const double* x, const double* y,
                      const int32_t* geom_offsets, const int32_t* part_offsets, const int32_t* ring_offsets,
                      int64_t n_polys, int64_t n_parts, int64_t n_rings, int64_t n_coords, const double* poly_bounds,
                      int32_t predicate, int32_t train_iters, int32_t lookup_mode, int64_t capacity,
                      int64_t* out_left, int64_t* out_right, int64_t* d_result, void* workspace, hipStream_t s,
                      bool fork) {
    VS_REQUIRE(n_points >= 0 && n_points < (int64_t)INT32_MAX, "learned join: n_points out of range");
    VS_REQUIRE(n_polys >= 0 && n_polys < (int64_t)INT32_MAX, "learned join: n_polys out of range");
    VS_REQUIRE(n_coords >= 0 && n_coords < (int64_t)INT32_MAX, "learned join: n_coords out of range");
    VS_REQUIRE(capacity >= 0 && capacity < (int64_t)INT32_MAX, "learned join: capacity out of range");
    VS_REQUIRE(predicate == VS_PRED_WITHIN || predicate == VS_PRED_INTERSECTS || predicate == VS_PRED_COVERED_BY ||
                   predicate == VS_PRED_TOUCHES,
               "learned join: unsupported predicate %d", predicate);
    VS_REQUIRE(lookup_mode >= 0 && lookup_mode <= 2, "learned join: lookup_mode in {0, 1, 2}");
    VS_REQUIRE(train_iters >= 0 && train_iters <= 100000, "learned join: bad train_iters");
    VS_REQUIRE(d_result != nullptr, "learned join: null result");
    const int S = leaves_for(n_points, 512);
    Carver c(workspace);
    const auto B = carve_index<uint32_t>(c, n_points, S);
    uint8_t* flags = c.take<uint8_t>(n_coords);
    const QueryBufs Q = carve_query(c, n_polys, n_points, capacity);
    // the polygon side (refine flags, cover cells) runs on `s` after the
    // gather, beside the side stream's training
    IndexFork f;
    VS_TRY(index_build_fork<uint32_t>(px, py, n_points, S, train_iters, kJoinTrainLr, B, s, fork, &f));
    VS_TRY(build_refine_flags(part_offsets, ring_offsets, n_parts, n_rings, n_coords, flags, s));
    const RefineView R = {x, y, geom_offsets, part_offsets, ring_offsets, flags};
    const double4* qb = reinterpret_cast<const double4*>(poly_bounds);
    VS_TRY(cover_cells_launch<uint32_t>(kLModePip, qb, n_polys, B, R, Q, s));
    VS_TRY(index_build_join(f, s));
    return learned_query_t<uint32_t>(kLModePip, predicate, qb, n_polys, view_of<uint32_t>(B, n_points, S, lookup_mode),
                                     B, R, n_points, capacity, true, Q, out_left, out_right, d_result, s, true);
}

}  // namespace vs

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int32_t vs_learned_leaf_count(int64_t n_points, int32_t leaf_size) {
    return vs::leaves_for(n_points, leaf_size > 0 ? leaf_size : 512);
}

int64_t vs_learned_index_bytes(int64_t n_points, int32_t key_bytes, int32_t n_leaves) {
    vs::Carver c(nullptr);
    if (key_bytes == 4) vs::carve_index<uint32_t>(c, n_points, n_leaves);
    else vs::carve_index<uint64_t>(c, n_points, n_leaves);
    return c.bytes;
}

int vs_learned_index_build(const double* px, const double* py, int64_t n_points, int32_t key_bytes, int32_t n_leaves,
                           int32_t train_iters, float lr, void* index, int64_t index_bytes, void* stream) {
    vs::clear_error();
    VS_REQUIRE(key_bytes == 4 || key_bytes == 8, "vs_learned_index_build: key_bytes must be 4 or 8");
    VS_REQUIRE(n_points >= 0 && n_points < (int64_t)INT32_MAX, "vs_learned_index_build: n_points out of range");
    VS_REQUIRE(n_leaves >= 1, "vs_learned_index_build: n_leaves must be >= 1");
    VS_REQUIRE(train_iters >= 0 && train_iters <= 100000, "vs_learned_index_build: bad train_iters");
    VS_REQUIRE(index_bytes >= vs_learned_index_bytes(n_points, key_bytes, n_leaves),
               "vs_learned_index_build: index buffer too small");
    vs::Carver c(index);
    hipStream_t s = vs::stream_of(stream);
    if (key_bytes == 4) {
        const auto B = vs::carve_index<uint32_t>(c, n_points, n_leaves);
        return vs::index_build_t<uint32_t>(px, py, n_points, n_leaves, train_iters, lr, B, s);
    }
    const auto B = vs::carve_index<uint64_t>(c, n_points, n_leaves);
    return vs::index_build_t<uint64_t>(px, py, n_points, n_leaves, train_iters, lr, B, s);
}

// Offsets (bytes from the index base) of the index's arrays, for callers that
// read them (tests, persistence): keys, perm, sorted x, sorted y, extent, mlp
// params, losses, leaf keys, leaf errors, meta, routing windows.
int vs_learned_index_layout(int64_t n_points, int32_t key_bytes, int32_t n_leaves, int64_t* offsets11) {
    vs::clear_error();
    VS_REQUIRE(offsets11 != nullptr, "vs_learned_index_layout: null output");
    VS_REQUIRE(key_bytes == 4 || key_bytes == 8, "vs_learned_index_layout: key_bytes must be 4 or 8");
    char* base = reinterpret_cast<char*>(4096);  // any non-null base: offsets are relative
    vs::Carver c(base);
    if (key_bytes == 4) {
        const auto B = vs::carve_index<uint32_t>(c, n_points, n_leaves);
        const void* p[11] = {B.keys, B.perm, B.sx, B.sy, B.ext, B.mlp, B.losses, B.leaf_key, B.leaf_err, B.meta,
                             B.leaf_win};
        for (int k = 0; k < 11; ++k) offsets11[k] = (const char*)p[k] - base;
    } else {
        const auto B = vs::carve_index<uint64_t>(c, n_points, n_leaves);
        const void* p[11] = {B.keys, B.perm, B.sx, B.sy, B.ext, B.mlp, B.losses, B.leaf_key, B.leaf_err, B.meta,
                             B.leaf_win};
        for (int k = 0; k < 11; ++k) offsets11[k] = (const char*)p[k] - base;
    }
    return VS_OK;
}

int vs_learned_lower_bound(const void* index, int64_t n_points, int32_t key_bytes, int32_t n_leaves,
                           const void* query_keys, int64_t m, int32_t lookup_mode, int64_t* out, uint64_t* d_stat3,
                           void* stream) {
    vs::clear_error();
    VS_REQUIRE(key_bytes == 4 || key_bytes == 8, "vs_learned_lower_bound: key_bytes must be 4 or 8");
    VS_REQUIRE(lookup_mode >= 0 && lookup_mode <= 2, "vs_learned_lower_bound: lookup_mode in {0, 1, 2}");
    vs::Carver c(const_cast<void*>(index));
    hipStream_t s = vs::stream_of(stream);
    if (m == 0) return VS_OK;
    if (key_bytes == 4) {
        const auto B = vs::carve_index<uint32_t>(c, n_points, n_leaves);
        const auto V = vs::view_of<uint32_t>(B, n_points, n_leaves, lookup_mode);
        hipLaunchKernelGGL(vs::lookup_batch_kernel<uint32_t>, dim3(vs::stride_grid(m, 256)), dim3(256), 0, s,
                           reinterpret_cast<const uint32_t*>(query_keys), m, V, out,
                           reinterpret_cast<unsigned long long*>(d_stat3));
    } else {
        const auto B = vs::carve_index<uint64_t>(c, n_points, n_leaves);
        const auto V = vs::view_of<uint64_t>(B, n_points, n_leaves, lookup_mode);
        hipLaunchKernelGGL(vs::lookup_batch_kernel<uint64_t>, dim3(vs::stride_grid(m, 256)), dim3(256), 0, s,
                           reinterpret_cast<const uint64_t*>(query_keys), m, V, out,
                           reinterpret_cast<unsigned long long*>(d_stat3));
    }
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int vs_mlp_loss_grad(const float* params, const float* u, const float* t, int64_t n, float* grad_out, float* loss_out,
                     void* stream) {
    vs::clear_error();
    VS_REQUIRE(n >= 0, "vs_mlp_loss_grad: negative n");
    hipLaunchKernelGGL(vs::mlp_loss_grad_kernel, dim3(1), dim3(vs::kTrainThreads), 0, vs::stream_of(stream), params, u, t, n,
                       grad_out, loss_out);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int64_t vs_learned_query_boxes_workspace_bytes(int64_t n_boxes, int64_t capacity) {
    vs::Carver c(nullptr);
    vs::carve_query(c, n_boxes, n_boxes, capacity);
    return c.bytes;
}

int vs_learned_query_boxes(const void* index, int64_t n_points, int32_t key_bytes, int32_t n_leaves,
                           const double* boxes, int64_t n_boxes, int32_t lookup_mode, int64_t capacity,
                           int32_t sort, int64_t* out_left, int64_t* out_right, int64_t* d_result,
                           void* workspace, int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    VS_REQUIRE(key_bytes == 4 || key_bytes == 8, "vs_learned_query_boxes: key_bytes must be 4 or 8");
    VS_REQUIRE(n_boxes >= 0 && n_boxes < (int64_t)INT32_MAX, "vs_learned_query_boxes: n_boxes out of range");
    VS_REQUIRE(capacity >= 0 && capacity < (int64_t)INT32_MAX, "vs_learned_query_boxes: capacity out of range");
    VS_REQUIRE(lookup_mode >= 0 && lookup_mode <= 2, "vs_learned_query_boxes: lookup_mode in {0, 1, 2}");
    VS_REQUIRE(workspace_bytes >= vs_learned_query_boxes_workspace_bytes(n_boxes, capacity),
               "vs_learned_query_boxes: workspace too small");
    VS_REQUIRE(d_result != nullptr, "vs_learned_query_boxes: null result");
    vs::Carver ci(const_cast<void*>(index));
    vs::Carver cw(workspace);
    const vs::QueryBufs Q = vs::carve_query(cw, n_boxes, n_boxes, capacity);
    const vs::RefineView R = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    hipStream_t s = vs::stream_of(stream);
    const double4* qb = reinterpret_cast<const double4*>(boxes);
    if (key_bytes == 4) {
        const auto B = vs::carve_index<uint32_t>(ci, n_points, n_leaves);
        return vs::learned_query_t<uint32_t>(vs::kLModeBox, VS_PRED_INTERSECTS, qb, n_boxes,
                                             vs::view_of<uint32_t>(B, n_points, n_leaves, lookup_mode), B, R, n_boxes,
                                             capacity, sort != 0, Q, out_left, out_right, d_result, s);
    }
    const auto B = vs::carve_index<uint64_t>(ci, n_points, n_leaves);
    return vs::learned_query_t<uint64_t>(vs::kLModeBox, VS_PRED_INTERSECTS, qb, n_boxes,
                                         vs::view_of<uint64_t>(B, n_points, n_leaves, lookup_mode), B, R, n_boxes,
                                         capacity, sort != 0, Q, out_left, out_right, d_result, s);
}

int64_t vs_learned_query_polygons_workspace_bytes(int64_t n_polys, int64_t n_coords, int64_t capacity) {
    vs::Carver c(nullptr);
    c.take<uint8_t>(n_coords);
    vs::carve_query(c, n_polys, n_polys, capacity);
    return c.bytes;
}

int vs_learned_query_polygons(const void* index, int64_t n_points, int32_t key_bytes, int32_t n_leaves,
                              const double* x, const double* y, const int32_t* geom_offsets,
                              const int32_t* part_offsets, const int32_t* ring_offsets, int64_t n_polys,
                              int64_t n_parts, int64_t n_rings, int64_t n_coords, const double* poly_bounds,
                              int32_t predicate, int32_t lookup_mode, int64_t capacity, int32_t sort,
                              int64_t* out_left, int64_t* out_right, int64_t* d_result, void* workspace,
                              int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    VS_REQUIRE(key_bytes == 4 || key_bytes == 8, "vs_learned_query_polygons: key_bytes must be 4 or 8");
    VS_REQUIRE(n_polys >= 0 && n_polys < (int64_t)INT32_MAX, "vs_learned_query_polygons: n_polys out of range");
    VS_REQUIRE(n_coords >= 0 && n_coords < (int64_t)INT32_MAX, "vs_learned_query_polygons: n_coords out of range");
    VS_REQUIRE(capacity >= 0 && capacity < (int64_t)INT32_MAX, "vs_learned_query_polygons: capacity out of range");
    VS_REQUIRE(predicate == VS_PRED_CONTAINS || predicate == VS_PRED_CONTAINS_PROPERLY || predicate == VS_PRED_COVERS ||
                   predicate == VS_PRED_INTERSECTS || predicate == VS_PRED_TOUCHES,
               "vs_learned_query_polygons: unsupported predicate %d", predicate);
    VS_REQUIRE(lookup_mode >= 0 && lookup_mode <= 2, "vs_learned_query_polygons: lookup_mode in {0, 1, 2}");
    VS_REQUIRE(d_result != nullptr, "vs_learned_query_polygons: null result");
    VS_REQUIRE(workspace_bytes >= vs_learned_query_polygons_workspace_bytes(n_polys, n_coords, capacity),
               "vs_learned_query_polygons: workspace too small");
    vs::Carver ci(const_cast<void*>(index));
    vs::Carver cw(workspace);
    uint8_t* flags = cw.take<uint8_t>(n_coords);
    const vs::QueryBufs Q = vs::carve_query(cw, n_polys, n_polys, capacity);
    hipStream_t s = vs::stream_of(stream);
    VS_TRY(vs::build_refine_flags(part_offsets, ring_offsets, n_parts, n_rings, n_coords, flags, s));
    const vs::RefineView R = {x, y, geom_offsets, part_offsets, ring_offsets, flags};
    const double4* qb = reinterpret_cast<const double4*>(poly_bounds);
    if (key_bytes == 4) {
        const auto B = vs::carve_index<uint32_t>(ci, n_points, n_leaves);
        return vs::learned_query_t<uint32_t>(vs::kLModePipQ, predicate, qb, n_polys,
                                             vs::view_of<uint32_t>(B, n_points, n_leaves, lookup_mode), B, R, n_polys,
                                             capacity, sort != 0, Q, out_left, out_right, d_result, s);
    }
    const auto B = vs::carve_index<uint64_t>(ci, n_points, n_leaves);
    return vs::learned_query_t<uint64_t>(vs::kLModePipQ, predicate, qb, n_polys,
                                         vs::view_of<uint64_t>(B, n_points, n_leaves, lookup_mode), B, R, n_polys,
                                         capacity, sort != 0, Q, out_left, out_right, d_result, s);
}

int64_t vs_sjoin_points_polygons_learned_workspace_bytes(int64_t n_points, int64_t n_polys, int64_t n_coords,
                                                         int64_t capacity) {
    const int S = vs::leaves_for(n_points, 512);
    vs::Carver c(nullptr);
    vs::carve_index<uint32_t>(c, n_points, S);
    c.take<uint8_t>(n_coords);
    vs::carve_query(c, n_polys, n_points, capacity);
    return c.bytes;
}

int vs_sjoin_points_polygons_learned(const double* px, const double* py, int64_t n_points, const double* x,
                                     const double* y, const int32_t* geom_offsets, const int32_t* part_offsets,
                                     const int32_t* ring_offsets, int64_t n_polys, int64_t n_parts, int64_t n_rings,
                                     int64_t n_coords, const double* poly_bounds, int32_t predicate,
                                     int32_t train_iters, int32_t lookup_mode, int64_t capacity, int64_t* out_left,
                                     int64_t* out_right, int64_t* d_result, void* workspace, int64_t workspace_bytes,
                                     void* stream) {
    vs::clear_error();
    VS_REQUIRE(workspace_bytes >= vs_sjoin_points_polygons_learned_workspace_bytes(n_points, n_polys, n_coords, capacity),
               "learned join: workspace too small");
    return vs::learned_join_impl(px, py, n_points, x, y, geom_offsets, part_offsets, ring_offsets, n_polys, n_parts,
                                 n_rings, n_coords, poly_bounds, predicate, train_iters, lookup_mode, capacity, out_left,
                                 out_right, d_result, workspace, vs::stream_of(stream), true);
}

// The cover of a box as the learned index computes it, on the host.
int vs_host_cover_ranges(double x0, double y0, double x1, double y1, const double* extent4, int32_t key_bytes,
                         uint64_t* out32) {
    vs::clear_error();
    VS_REQUIRE(key_bytes == 4 || key_bytes == 8, "vs_host_cover_ranges: key_bytes must be 4 or 8");
    VS_REQUIRE(extent4 != nullptr && out32 != nullptr, "vs_host_cover_ranges: null argument");
    const int bits = 4 * key_bytes;
    const double cells = bits >= 32 ? 4294967295.0 : (double)((1ull << bits) - 1);
    vs::KeyScale ks;
    ks.x0 = extent4[0];
    ks.y0 = extent4[1];
    ks.sx = extent4[2] > extent4[0] ? cells / (extent4[2] - extent4[0]) : 0.0;
    ks.sy = extent4[3] > extent4[1] ? cells / (extent4[3] - extent4[1]) : 0.0;
    const int min_l = bits - (key_bytes == 4 ? vs::point_index_grid_bits<uint32_t>() : vs::point_index_grid_bits<uint64_t>());
    return vs::cover_ranges(x0, y0, x1, y1, ks, bits, min_l, reinterpret_cast<unsigned long long*>(out32));
}

}  // extern "C"
