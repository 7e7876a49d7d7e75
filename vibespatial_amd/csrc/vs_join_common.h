// Shared by the join translation units: the control block every join keeps in
// its workspace and the finish pass (per-left counts -> scan -> scatter ->
// per-left sort of right indices) that turns staged pairs into lexsorted
// int64 output.
#pragma once

#include "vs_internal.h"

namespace vs {

struct JoinCtl {
    unsigned long long staged;   // pairs appended (may exceed capacity)
    long long total;             // scan total of per-left counts
    unsigned int queue[8];       // work queues (per XCD for point waves, [0] for items)
    unsigned long long stat[4];  // join-specific counters, published in d_result[2..5]
    unsigned long long escalated_pairs;
};

void ctl_reset(JoinCtl* ctl, hipStream_t s);
void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);

// counts[nq] must hold the number of staged pairs of each left row.
int finish_pairs_raw(const int2* stage, int32_t* counts, int32_t* offsets, int32_t* cursor, void* scan_ws,
                     JoinCtl* ctl, int64_t nq, int64_t capacity, int64_t* out_l, int64_t* out_r, int64_t* d_result,
                     hipStream_t s);

// Per-coordinate flags of the refine layout: bit0 first coordinate of a ring,
// bit1 first coordinate of a part.
int build_refine_flags(const int32_t* part, const int32_t* ring, int64_t n_parts, int64_t n_rings, int64_t n_coords,
                       uint8_t* flags, hipStream_t s);

int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, long long* d_total, void* ws, hipStream_t s);
int64_t scan_workspace_bytes(int64_t n);

// Benchmark timing of a join's dominant kernel (vs_set_kernel_timing).
int timing_begin_public(hipStream_t s);
int timing_end_public(hipStream_t s);

}  // namespace vs
