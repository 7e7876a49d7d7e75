// Shared by the join translation units: the control block every join keeps in
// its workspace and the finish pass (radix sort of the staged pairs by left
// id, then a sort of the right ids of each left id) that turns staged pairs
// into lexsorted int64 output.
#pragma once

#include <algorithm>

#include "vs_internal.h"

namespace vs {

struct JoinCtl {
    unsigned long long staged;   // pairs appended (may exceed capacity)
    unsigned int n_long;         // finish: runs of one left id longer than the short-run limit
    unsigned int pad;
    unsigned int queue[8];       // work queues (per XCD for point waves, [0] for items)
    unsigned long long stat[4];  // join-specific counters, published in d_result[2..5]
    unsigned long long escalated_pairs;
};

// Clears ctl and extra[0, n_extra) (n_extra <= 64) with one launch.
void ctl_reset(JoinCtl* ctl, hipStream_t s, unsigned long long* extra = nullptr, int n_extra = 0);
void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);

// Finish (vs_finish.hip).  A join appends pairs to the two stage arrays that
// finish_stage carves from the finish workspace (left ids as u32, right ids as
// i32, `capacity` slots, ctl->staged the count); finish_pairs turns them into
// int64 output sorted by left id (and by right id within a left id when
// sort_right) and publishes d_result[0..3].
int64_t finish_workspace_bytes(int64_t capacity);
void finish_stage(void* fin_ws, int64_t capacity, uint32_t** stage_l, int32_t** stage_r);
int finish_pairs(void* fin_ws, JoinCtl* ctl, int64_t n_left, int64_t capacity, bool sort_right, int64_t* out_l,
                 int64_t* out_r, int64_t* d_result, hipStream_t s);

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
