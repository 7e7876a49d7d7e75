// Nearest-neighbour join support: fp64 distances of candidate pairs and the
// per-query selection of the nearest ones (ties kept), driven by an
// expanding-radius box search on the host side.
//
// Replaces the reference's nearest refine (spatial/nearest.py:3278
// _nearest_refine_gpu, distances in _compute_pair_distances_gpu :2228 and
// _compute_mixed_distances_gpu :2482, reduction in _nearest_from_distances
// :3237).  The distance of two geometries is the fp64 value GEOS computes:
//   point - point     sqrt(dx*dx + dy*dy)
//   point - segment   GEOS Distance::pointToSegment: the projection parameter
//                     r = ((p - a).(b - a)) / |b - a|^2 decides the end point
//                     (r <= 0, r >= 1) or the perpendicular
//                     |((a.y - p.y)(b.x - a.x) - (a.x - p.x)(b.y - a.y)) / |b - a|^2| * sqrt(|b - a|^2)
//   point - polygon   0 when the point is not outside (exact location), else
//                     the least point - segment distance over the boundary
//   polygon - polygon 0 when they intersect (exact test), else the least
//                     vertex - segment distance either way
// Empty geometries and NaN points are at +inf (never nearest).  Ties are
// equal fp64 distances, as in GEOS's nearest query; the reference's GPU
// reduction instead keeps distances within numpy.isclose of the minimum.
//
// A round of the search (host side): every open query's box grown by its
// radius -> box query over the tree -> vs_pair_distances -> vs_nearest_select,
// which keeps, for each query whose least candidate distance is <= its radius,
// the pairs at that least distance and closes it; other queries get the next
// radius.  Any tree geometry at fp64 distance <= r from a query lies in the
// query box grown by r (1 + 2^-40) (the search radius), so a closed query's
// pairs are complete.
#include "vs_exact.h"
#include "vs_internal.h"
#include "vs_polys.h"

namespace vs {

int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, long long* d_total, void* ws,
                       hipStream_t s);
int64_t scan_workspace_bytes(int64_t n);
int sort_pairs_u64(uint64_t* k, int32_t* v, int64_t n, int b0, int b1, void* ws, int64_t wb, hipStream_t s);
int box_order(const double4* b, int64_t n, int32_t* perm, void* ws, hipStream_t s);
int64_t box_order_workspace_bytes(int64_t n);

constexpr double kSearchSlack = 1.0 + 0x1p-40;  // query boxes are grown by radius * this

__device__ __forceinline__ double point_dist_f64(double px, double py, double qx, double qy) {
    const double dx = px - qx, dy = py - qy;
    return sqrt(dx * dx + dy * dy);
}

__device__ __forceinline__ double segment_dist_f64(double px, double py, double ax, double ay, double bx,
                                                   double by) {
    if (ax == bx && ay == by) return point_dist_f64(px, py, ax, ay);
    const double ux = bx - ax, uy = by - ay;
    const double len2 = ux * ux + uy * uy;
    const double r = ((px - ax) * ux + (py - ay) * uy) / len2;
    if (r <= 0.0) return point_dist_f64(px, py, ax, ay);
    if (r >= 1.0) return point_dist_f64(px, py, bx, by);
    const double s = ((ay - py) * ux - (ax - px) * uy) / len2;
    return fabs(s) * sqrt(len2);
}

__device__ __forceinline__ bool poly_empty(const Polys& P, int64_t row) {
    return P.ring[P.part[P.geom[row + 1]]] <= P.ring[P.part[P.geom[row]]];
}

// Least distance from p to the boundary segments of polygonal row `row`
// (min with `best`).  A segment whose box lies more than best (1 + 2^-30)
// away on an axis is skipped: its computed distance (a few ulps from the
// true one) exceeds best, so the minimum is the one over every segment.
__device__ double boundary_dist_f64(double px, double py, const Polys& P, int64_t row, double best) {
    const int32_t r0 = P.part[P.geom[row]], r1 = P.part[P.geom[row + 1]];
    for (int32_t r = r0; r < r1; ++r)
        for (int32_t c = P.ring[r] + 1; c < P.ring[r + 1]; ++c) {
            const double ax = P.x[c - 1], ay = P.y[c - 1], bx = P.x[c], by = P.y[c];
            const double lim = best * (1.0 + 0x1p-30);
            if (px - fmax(ax, bx) > lim || fmin(ax, bx) - px > lim || py - fmax(ay, by) > lim ||
                fmin(ay, by) - py > lim)
                continue;
            best = fmin(best, segment_dist_f64(px, py, ax, ay, bx, by));
        }
    return best;
}

// One pass over the edges: 0 when p is on the boundary or inside a part (the
// exact location's even-odd steps), else the least segment distance.  With
// `outside` the point is known to lie outside the row and the location steps
// are skipped.  Distances above `cap` are not needed: segments farther than
// cap are skipped and, when no segment comes within cap, the result is +inf.
__device__ double point_polygon_dist_f64(double px, double py, const Polys& P, int64_t row, bool outside,
                                         double cap) {
    if (px != px || py != py || poly_empty(P, row)) return INFINITY;
    double best = cap;
    bool found = false;
    for (int32_t pp = P.geom[row]; pp < P.geom[row + 1]; ++pp) {
        int parity = 0;
        for (int32_t r = P.part[pp]; r < P.part[pp + 1]; ++r)
            for (int32_t c = P.ring[r] + 1; c < P.ring[r + 1]; ++c) {
                const double ax = P.x[c - 1], ay = P.y[c - 1], bx = P.x[c], by = P.y[c];
                if (!outside && edge_step_f64(px, py, ax, ay, bx, by, parity)) return 0.0;  // on the boundary
                const double lim = best * (1.0 + 0x1p-30);
                if (px - fmax(ax, bx) > lim || fmin(ax, bx) - px > lim || py - fmax(ay, by) > lim ||
                    fmin(ay, by) - py > lim)
                    continue;
                const double d = segment_dist_f64(px, py, ax, ay, bx, by);
                if (d <= best) {
                    best = d;
                    found = true;
                }
            }
        if (parity) return 0.0;  // inside this part
    }
    return found ? best : INFINITY;
}

__device__ double polygons_dist_f64(const Polys& A, int64_t ra, const Polys& B, int64_t rb) {
    if (poly_empty(A, ra) || poly_empty(B, rb)) return INFINITY;
    if (polygons_intersect_auto(A, ra, B, rb)) return 0.0;
    double best = INFINITY;
    const int32_t a0 = A.ring[A.part[A.geom[ra]]], a1 = A.ring[A.part[A.geom[ra + 1]]];
    for (int32_t v = a0; v < a1; ++v) best = boundary_dist_f64(A.x[v], A.y[v], B, rb, best);
    const int32_t b0 = B.ring[B.part[B.geom[rb]]], b1 = B.ring[B.part[B.geom[rb + 1]]];
    for (int32_t v = b0; v < b1; ++v) best = boundary_dist_f64(B.x[v], B.y[v], A, ra, best);
    return best;
}

// Kinds: 0 point rows (x[row], y[row]), 1 polygonal rows.  Flags:
// kDistExclusive, a point pair with equal coordinates is at +inf (equal
// geometries excluded); kDistOutside, left points lie outside every right
// polygon (no location steps).  cap (per left row, may be null): a pair
// farther than cap[left] gets +inf.  Pairs are visited in right-row order
// (order[t]) when given.
constexpr int kDistExclusive = 1, kDistOutside = 2;

__global__ __launch_bounds__(256) void pair_distances_kernel(int left_kind, int right_kind,
                                                            const int64_t* __restrict__ left,
                                                            const int64_t* __restrict__ right, int64_t n,
                                                            const int32_t* __restrict__ order, int flags,
                                                            const double* __restrict__ cap, Polys L, Polys R,
                                                            double* __restrict__ out) {
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
        const int64_t i = order ? (int64_t)order[t] : t;
        const int64_t l = left[i], r = right[i];
        const double c = cap && cap[l] == cap[l] ? cap[l] : INFINITY;  // NaN: no cap
        double d;
        if (left_kind == 0 && right_kind == 0) {
            const double px = L.x[l], py = L.y[l], qx = R.x[r], qy = R.y[r];
            if (px != px || py != py || qx != qx || qy != qy) d = INFINITY;
            else if ((flags & kDistExclusive) && px == qx && py == qy) d = INFINITY;
            else d = point_dist_f64(px, py, qx, qy);
        } else if (left_kind == 0) {
            d = point_polygon_dist_f64(L.x[l], L.y[l], R, r, (flags & kDistOutside) != 0, c);
        } else if (right_kind == 0) {
            d = point_polygon_dist_f64(R.x[r], R.y[r], L, l, false, c);
        } else {
            d = polygons_dist_f64(L, l, R, r);
        }
        out[i] = d > c ? INFINITY : d;
    }
}

__global__ __launch_bounds__(256) void right_order_keys_kernel(const int64_t* __restrict__ right, int64_t n,
                                                              uint32_t* __restrict__ key, int32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        key[i] = (uint32_t)right[i];
        idx[i] = (int32_t)i;
    }
}

int64_t pair_distances_workspace_bytes(int64_t n) {
    Carver c(nullptr);
    c.take<uint32_t>(n);
    c.take<int32_t>(n);
    c.take<uint32_t>(n);
    c.take<int32_t>(n);
    c.take<char>(sort_workspace_bytes_u32(n));
    return c.bytes;
}

// Non-negative doubles order as their bit patterns.
__global__ __launch_bounds__(256) void nearest_init_kernel(int64_t nq, unsigned long long* __restrict__ mind,
                                                          unsigned long long* __restrict__ first,
                                                          unsigned long long* __restrict__ open) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q < nq) {
        mind[q] = (unsigned long long)__double_as_longlong(INFINITY);
        first[q] = ~0ull;
    }
    if (q == 0) *open = 0;
}

__global__ __launch_bounds__(256) void nearest_min_kernel(const int64_t* __restrict__ left,
                                                         const double* __restrict__ dist, int64_t n,
                                                         unsigned long long* __restrict__ mind) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const double d = dist[i];
        if (d < INFINITY) atomicMin(&mind[left[i]], (unsigned long long)__double_as_longlong(d));
    }
}

// keep = the pair is at its query's least distance and that distance is
// within the query's radius; first = the least right row among kept pairs.
__global__ __launch_bounds__(256) void nearest_keep_kernel(const int64_t* __restrict__ left,
                                                          const int64_t* __restrict__ right,
                                                          const double* __restrict__ dist, int64_t n,
                                                          const double* __restrict__ radius,
                                                          const unsigned long long* __restrict__ mind,
                                                          unsigned long long* __restrict__ first, int return_all,
                                                          int32_t* __restrict__ keep) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t q = left[i];
    const double d = dist[i];
    const double m = __longlong_as_double((long long)mind[q]);
    const bool k = d < INFINITY && d == m && m <= radius[q];
    keep[i] = k ? 1 : 0;
    if (k && !return_all) atomicMin(&first[q], (unsigned long long)right[i]);
}

__global__ __launch_bounds__(256) void nearest_first_kernel(const int64_t* __restrict__ left,
                                                           const int64_t* __restrict__ right, int64_t n,
                                                           const unsigned long long* __restrict__ first,
                                                           int32_t* __restrict__ keep) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n && keep[i] && (unsigned long long)right[i] != first[left[i]]) keep[i] = 0;
}

__global__ __launch_bounds__(256) void nearest_compact_kernel(const int64_t* __restrict__ left,
                                                             const int64_t* __restrict__ right,
                                                             const double* __restrict__ dist, int64_t n,
                                                             const int32_t* __restrict__ keep,
                                                             const int32_t* __restrict__ pos,
                                                             int64_t* __restrict__ out_l, int64_t* __restrict__ out_r,
                                                             double* __restrict__ out_d) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n && keep[i]) {
        out_l[pos[i]] = left[i];
        out_r[pos[i]] = right[i];
        if (out_d) out_d[pos[i]] = dist[i];
    }
}

// Next radius per query: NaN once closed (nearest found within the radius,
// or the final radius searched without a match), else radius * growth
// (start_radius after a radius-0 round), or final_radius once that reaches
// switch_radius; search = the radius the query box is grown by,
// radius * (1 + 2^-40).
__global__ __launch_bounds__(256) void nearest_radius_kernel(int64_t nq, double* __restrict__ radius,
                                                            double* __restrict__ search,
                                                            const unsigned long long* __restrict__ mind,
                                                            double growth, double start_radius,
                                                            double switch_radius, double final_radius,
                                                            unsigned long long* __restrict__ open) {
    __shared__ int cnt[4];
    int o = 0;
    // grid-stride: one count atomic per block, few blocks
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < nq; q += (int64_t)gridDim.x * 256) {
        const double r = radius[q];
        if (r == r) {
            const double m = __longlong_as_double((long long)mind[q]);
            double nr;
            if (m <= r || r >= final_radius) {
                nr = NAN;
            } else {
                nr = r > 0.0 ? r * growth : start_radius;
                if (!(nr < switch_radius)) nr = final_radius;
                ++o;
            }
            radius[q] = nr;
            search[q] = nr * kSearchSlack;
        }
    }
    const int w = wave_sum_i32(o);
    if ((threadIdx.x & 63) == 0) cnt[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int t = cnt[0] + cnt[1] + cnt[2] + cnt[3];
        if (t) atomicAdd(open, (unsigned long long)t);
    }
}

__global__ void nearest_result_kernel(const long long* __restrict__ total, const unsigned long long* __restrict__ open,
                                      int64_t* __restrict__ result) {
    if (threadIdx.x == 0) {
        result[0] = *total;
        result[1] = (int64_t)*open;
    }
}

// First radius of every query: r0, or NaN (closed, never matches) for a
// query whose box is not finite (empty geometry, NaN point).
__global__ __launch_bounds__(256) void nearest_init_radius_kernel(const double4* __restrict__ boxes, int64_t n,
                                                                 double r0, double* __restrict__ radius,
                                                                 double* __restrict__ search) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const double4 b = boxes[q];
    const bool ok = isfinite(b.x) && isfinite(b.y) && isfinite(b.z) && isfinite(b.w);
    radius[q] = ok ? r0 : NAN;
    search[q] = ok ? r0 * kSearchSlack : NAN;
}

// Open queries (search radius not NaN): their rows, and their boxes grown by
// the search radius with every bound rounded one ulp outward.
__global__ __launch_bounds__(256) void open_flags_kernel(const double* __restrict__ search, int64_t n,
                                                        int32_t* __restrict__ keep) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n) keep[q] = search[q] == search[q] ? 1 : 0;
}

__global__ __launch_bounds__(256) void open_boxes_kernel(const double4* __restrict__ boxes,
                                                        const double* __restrict__ search, int64_t n,
                                                        const int32_t* __restrict__ keep,
                                                        const int32_t* __restrict__ pos,
                                                        int64_t* __restrict__ open_rows,
                                                        double4* __restrict__ open_boxes) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n || !keep[q]) return;
    const double d = search[q];
    const double4 v = boxes[q];
    const int32_t k = pos[q];
    open_rows[k] = q;
    open_boxes[k] = make_double4(nextafter(v.x - d, -INFINITY), nextafter(v.y - d, -INFINITY),
                                 nextafter(v.z + d, INFINITY), nextafter(v.w + d, INFINITY));
}

__global__ __launch_bounds__(256) void map_left_kernel(const int64_t* __restrict__ rows, int64_t* __restrict__ left,
                                                      int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) left[i] = rows[left[i]];
}

int64_t nearest_open_workspace_bytes(int64_t n) {
    Carver c(nullptr);
    c.take<int32_t>(n);
    c.take<int32_t>(n);
    c.take<long long>(1);
    c.take<char>(scan_workspace_bytes(n));
    return c.bytes;
}

__global__ void count_to_result_kernel(const long long* __restrict__ total, int64_t* __restrict__ result) {
    if (threadIdx.x == 0) result[0] = *total;
}

struct NearestTree {
    const double4* box;
    const double2* rep;  // per item slot: a point of the item's geometry (NaN when empty)
    Polys P;                 // polygonal items (exact distances from point queries) when polys
    const int32_t* item_ids;  // item slot -> row of P
    int32_t polys;
    int64_t level_start[VS_MAX_LEVELS];
    int64_t level_count[VS_MAX_LEVELS];
    int32_t n_levels;
    int32_t fanout;
};

// Upper bound of a query's nearest distance: the least, over the tree's
// items, of the farthest-point distance from the query box to the item box or
// to a point of the item (every geometry lies in its box and contains that
// point, so its distance is at most either).  A point query against
// polygonal items also takes the own fp64 distance (the value the pair
// distances compute) of the first kBoundExact items the walk reaches (the
// nearest boxes first): the bound is then usually the nearest distance
// itself, at a bounded cost per query.
// Depth-first over the packed tree, nearest child first, pruning nodes whose
// box is farther (least axis-gap distance) than the bound so far, for at most
// kBoundPops nodes.  A full stack or the visit limit only drops nodes, which
// leaves a valid (looser) bound.  +inf when no
// finite box is met; scaled up by 2^-40 for the rounding of the fp64 values.
constexpr int kBoundStack = 256;
constexpr int kBoundPops = 4096;  // visit limit per query (bounds the time of a pathological query)
constexpr int kBoundExact = 2;    // exact item distances per query

__device__ __forceinline__ double box_gap2(const double4 q, const double4 b) {
    const double gx = fmax(0.0, fmax(b.x - q.z, q.x - b.z));
    const double gy = fmax(0.0, fmax(b.y - q.w, q.y - b.w));
    return gx * gx + gy * gy;
}

__device__ double nearest_upper_bound(const double4 q, const NearestTree& T) {
    int64_t stack[kBoundStack];
    int sp = 0;
    const int last = T.n_levels - 1;
    for (int64_t c = 0; c < T.level_count[0] && sp < kBoundStack; ++c) stack[sp++] = c;  // level 0
    double ub2 = INFINITY;
    int exact = (T.polys && q.x == q.z && q.y == q.w) ? kBoundExact : 0;
    for (int pops = 0; sp > 0 && pops < kBoundPops; ++pops) {
        const int64_t e = stack[--sp];
        const int l = (int)(e >> 48);
        const int64_t c = e & ((1ll << 48) - 1);
        const double4 b = T.box[T.level_start[l] + c];
        if (!(b.x == b.x) || box_gap2(q, b) > ub2) continue;
        if (l == last) {
            // farthest point of the query box from the item box, and from a
            // point of the item itself (never farther, often much nearer)
            const double dx = fmax(q.z - b.x, b.z - q.x), dy = fmax(q.w - b.y, b.w - q.y);
            ub2 = fmin(ub2, dx * dx + dy * dy);
            const double2 p = T.rep[c];
            if (p.x == p.x) {
                const double ex = fmax(fabs(p.x - q.x), fabs(p.x - q.z)), ey = fmax(fabs(p.y - q.y), fabs(p.y - q.w));
                ub2 = fmin(ub2, ex * ex + ey * ey);
            }
            if (exact > 0) {
                --exact;
                // at most the box bound just taken up to rounding; +inf
                // (rounding, empty row) leaves the bound as it is
                const double d = point_polygon_dist_f64(q.x, q.y, T.P, T.item_ids[c], false, sqrt(ub2));
                if (d < INFINITY) ub2 = fmin(ub2, d * d);
            }
            continue;
        }
        const int64_t lo = c * T.fanout;
        const int64_t hi = lo + T.fanout < T.level_count[l + 1] ? lo + T.fanout : T.level_count[l + 1];
        // push the children within the bound, the nearest one last (popped first)
        int64_t best = -1;
        double bestd = INFINITY;
        for (int64_t k = lo; k < hi; ++k) {
            const double4 cb = T.box[T.level_start[l + 1] + k];
            if (!(cb.x == cb.x)) continue;
            const double g = box_gap2(q, cb);
            if (g > ub2) continue;
            if (g < bestd) {
                if (best >= 0 && sp < kBoundStack) stack[sp++] = ((int64_t)(l + 1) << 48) | best;
                best = k;
                bestd = g;
            } else if (sp < kBoundStack) {
                stack[sp++] = ((int64_t)(l + 1) << 48) | k;
            }
        }
        if (best >= 0 && sp < kBoundStack) stack[sp++] = ((int64_t)(l + 1) << 48) | best;
    }
    return sqrt(ub2) * (1.0 + 0x1p-40);
}

// radius[q] = min(bound, cap) for each listed query row q (search follows);
// a query without a finite bound keeps its radius.
// Threads take the rows in the Hilbert order of their boxes (perm), so the
// lanes of a wave walk nearby parts of the tree.
__global__ __launch_bounds__(256) void nearest_bound_kernel(const double4* __restrict__ qboxes,
                                                           const int64_t* __restrict__ rows,
                                                           const int32_t* __restrict__ perm, int64_t m,
                                                           NearestTree T, double cap, double* __restrict__ radius,
                                                           double* __restrict__ search) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const int64_t q = rows[perm[i]];
    const double ub = nearest_upper_bound(qboxes[q], T);
    if (ub < INFINITY) {
        const double r = fmin(ub, cap);
        radius[q] = r;
        search[q] = r * kSearchSlack;
    }
}

__global__ __launch_bounds__(256) void left_keys_kernel(const int64_t* __restrict__ left, int64_t n,
                                                       uint64_t* __restrict__ key, int32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        key[i] = (uint64_t)left[i];
        idx[i] = (int32_t)i;
    }
}

__global__ __launch_bounds__(256) void gather_pairs_kernel(const int32_t* __restrict__ idx, int64_t n,
                                                          const int64_t* __restrict__ left,
                                                          const int64_t* __restrict__ right,
                                                          const double* __restrict__ dist,
                                                          int64_t* __restrict__ out_l, int64_t* __restrict__ out_r,
                                                          double* __restrict__ out_d) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const int32_t k = idx[i];
        out_l[i] = left[k];
        out_r[i] = right[k];
        out_d[i] = dist[k];
    }
}

// rep[slot] = a point of item item_ids[slot]: the point itself (kind 0) or
// the first coordinate of the polygonal row (kind 1; NaN when it has none).
__global__ __launch_bounds__(256) void rep_points_kernel(int kind, const int32_t* __restrict__ item_ids, int64_t n,
                                                        Polys G, double2* __restrict__ rep) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t row = item_ids[i];
    double2 p = make_double2(NAN, NAN);
    if (kind == 0) {
        p = make_double2(G.x[row], G.y[row]);
    } else {
        const int32_t c0 = G.ring[G.part[G.geom[row]]], c1 = G.ring[G.part[G.geom[row + 1]]];
        if (c1 > c0) p = make_double2(G.x[c0], G.y[c0]);
    }
    rep[i] = p;
}

int64_t nearest_finish_workspace_bytes(int64_t n) {
    Carver c(nullptr);
    c.take<uint64_t>(n);
    c.take<int32_t>(n);
    c.take<char>(sort_workspace_bytes_u64(n));
    return c.bytes;
}

struct NearestWs {
    unsigned long long* mind;
    unsigned long long* first;
    unsigned long long* open;
    int32_t* keep;
    int32_t* pos;
    long long* total;
    void* scan_ws;
};

static NearestWs carve_nearest(Carver& c, int64_t n_pairs, int64_t n_query) {
    NearestWs W;
    W.mind = c.take<unsigned long long>(n_query);
    W.first = c.take<unsigned long long>(n_query);
    W.open = c.take<unsigned long long>(1);
    W.keep = c.take<int32_t>(n_pairs);
    W.pos = c.take<int32_t>(n_pairs);
    W.total = c.take<long long>(1);
    W.scan_ws = c.take<char>(scan_workspace_bytes(n_pairs));
    return W;
}

int64_t nearest_select_workspace_bytes(int64_t n_pairs, int64_t n_query) {
    Carver c(nullptr);
    carve_nearest(c, n_pairs, n_query);
    return c.bytes;
}

}  // namespace vs

extern "C" {

int64_t vs_pair_distances_workspace_bytes(int64_t n_pairs) { return vs::pair_distances_workspace_bytes(n_pairs); }

int vs_pair_distances(int32_t left_kind, int32_t right_kind, const int64_t* left, const int64_t* right,
                      int64_t n_pairs, int64_t n_right, int32_t flags, const double* cap, const double* lx,
                      const double* ly, const int32_t* l_geom, const int32_t* l_part, const int32_t* l_ring,
                      const double* rx, const double* ry, const int32_t* r_geom, const int32_t* r_part,
                      const int32_t* r_ring, double* out_distance, void* workspace, int64_t workspace_bytes,
                      void* stream) {
    vs::clear_error();
    VS_REQUIRE((left_kind == 0 || left_kind == 1) && (right_kind == 0 || right_kind == 1),
               "vs_pair_distances: kinds must be 0 (points) or 1 (polygons)");
    VS_REQUIRE(n_pairs >= 0 && n_pairs < (int64_t)INT32_MAX, "vs_pair_distances: n_pairs out of range");
    VS_REQUIRE(n_right >= 0 && n_right <= (int64_t)INT32_MAX, "vs_pair_distances: n_right out of range");
    VS_REQUIRE(workspace_bytes >= vs::pair_distances_workspace_bytes(n_pairs),
               "vs_pair_distances: workspace too small");
    if (n_pairs == 0) return VS_OK;
    VS_REQUIRE(left && right && out_distance, "vs_pair_distances: null pointer");
    hipStream_t s = vs::stream_of(stream);
    vs::Carver c(workspace);
    uint32_t* key_in = c.take<uint32_t>(n_pairs);
    int32_t* idx_in = c.take<int32_t>(n_pairs);
    uint32_t* key_out = c.take<uint32_t>(n_pairs);
    int32_t* order = c.take<int32_t>(n_pairs);
    void* sort_ws = c.take<char>(vs::sort_workspace_bytes_u32(n_pairs));
    const vs::Polys L = {lx, ly, l_geom, l_part, l_ring};
    const vs::Polys R = {rx, ry, r_geom, r_part, r_ring};
    const int32_t* ord = nullptr;
    if (right_kind == 1) {  // lanes of a wave share a polygon
        hipLaunchKernelGGL(vs::right_order_keys_kernel, dim3((unsigned)vs::ceil_div(n_pairs, 256)), dim3(256), 0, s,
                           right, n_pairs, key_in, idx_in);
        VS_LAUNCH_CHECK();
        int bits = 0;
        while (bits < 31 && ((int64_t)1 << bits) < n_right) ++bits;
        VS_TRY(vs::sort_pairs_into_u32(key_in, idx_in, key_out, order, n_pairs, 0, bits, sort_ws,
                                       vs::sort_workspace_bytes_u32(n_pairs), s));
        ord = order;
    }
    const int grid = (int)std::min<int64_t>(vs::ceil_div(n_pairs, 256), 16384);
    hipLaunchKernelGGL(vs::pair_distances_kernel, dim3(grid), dim3(256), 0, s, left_kind, right_kind, left, right,
                       n_pairs, ord, flags & (vs::kDistExclusive | vs::kDistOutside), cap, L, R, out_distance);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int vs_nearest_init(const double* boxes, int64_t n_query, double first_radius, double* radius, double* search,
                    void* stream) {
    vs::clear_error();
    VS_REQUIRE(n_query >= 0, "vs_nearest_init: n_query < 0");
    if (n_query == 0) return VS_OK;
    VS_REQUIRE(boxes && radius && search, "vs_nearest_init: null pointer");
    hipLaunchKernelGGL(vs::nearest_init_radius_kernel, dim3((unsigned)vs::ceil_div(n_query, 256)), dim3(256), 0,
                       vs::stream_of(stream), reinterpret_cast<const double4*>(boxes), n_query, first_radius, radius,
                       search);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int64_t vs_nearest_open_workspace_bytes(int64_t n_query) { return vs::nearest_open_workspace_bytes(n_query); }

int vs_nearest_open(const double* boxes, const double* search, int64_t n_query, int64_t* open_rows,
                    double* open_boxes, int64_t* d_count, void* workspace, int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    VS_REQUIRE(n_query >= 0 && n_query < (int64_t)INT32_MAX, "vs_nearest_open: n_query out of range");
    VS_REQUIRE(workspace_bytes >= vs::nearest_open_workspace_bytes(n_query), "vs_nearest_open: workspace too small");
    VS_REQUIRE(d_count != nullptr, "vs_nearest_open: null count");
    hipStream_t s = vs::stream_of(stream);
    vs::Carver c(workspace);
    int32_t* keep = c.take<int32_t>(n_query);
    int32_t* pos = c.take<int32_t>(n_query);
    long long* total = c.take<long long>(1);
    void* scan_ws = c.take<char>(vs::scan_workspace_bytes(n_query));
    const unsigned g = (unsigned)vs::ceil_div(n_query, 256);
    if (n_query > 0) {
        hipLaunchKernelGGL(vs::open_flags_kernel, dim3(g), dim3(256), 0, s, search, n_query, keep);
        VS_LAUNCH_CHECK();
    }
    VS_TRY(vs::exclusive_scan_i32(keep, pos, n_query, total, scan_ws, s));
    if (n_query > 0) {
        hipLaunchKernelGGL(vs::open_boxes_kernel, dim3(g), dim3(256), 0, s, reinterpret_cast<const double4*>(boxes),
                           search, n_query, keep, pos, open_rows, reinterpret_cast<double4*>(open_boxes));
        VS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(vs::count_to_result_kernel, dim3(1), dim3(64), 0, s, total, d_count);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int vs_nearest_rep_points(int32_t kind, const int32_t* item_ids, int64_t n_items, const double* x, const double* y,
                          const int32_t* geom, const int32_t* part, const int32_t* ring, double* rep, void* stream) {
    vs::clear_error();
    VS_REQUIRE(kind == 0 || kind == 1, "vs_nearest_rep_points: kind must be 0 (points) or 1 (polygons)");
    VS_REQUIRE(n_items >= 0, "vs_nearest_rep_points: n_items < 0");
    if (n_items == 0) return VS_OK;
    VS_REQUIRE(item_ids && x && y && rep && (kind == 0 || (geom && part && ring)), "vs_nearest_rep_points: null pointer");
    const vs::Polys G = {x, y, geom, part, ring};
    hipLaunchKernelGGL(vs::rep_points_kernel, dim3((unsigned)vs::ceil_div(n_items, 256)), dim3(256), 0,
                       vs::stream_of(stream), kind, item_ids, n_items, G, reinterpret_cast<double2*>(rep));
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int64_t vs_nearest_bound_workspace_bytes(int64_t n_rows) {
    vs::Carver c(nullptr);
    c.take<int32_t>(n_rows);
    c.take<char>(vs::box_order_workspace_bytes(n_rows));
    return c.bytes;
}

int vs_nearest_bound(const double* query_boxes, const int64_t* rows, int64_t n_rows, const double* row_boxes,
                     const vs_tree_layout* layout, const double* node_boxes, const double* rep, int32_t tree_kind,
                     const int32_t* item_ids, const double* tx, const double* ty, const int32_t* t_geom,
                     const int32_t* t_part, const int32_t* t_ring, double cap, double* radius, double* search,
                     void* workspace, int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    VS_REQUIRE(n_rows >= 0 && n_rows < (int64_t)INT32_MAX, "vs_nearest_bound: n_rows out of range");
    VS_REQUIRE(workspace_bytes >= vs_nearest_bound_workspace_bytes(n_rows), "vs_nearest_bound: workspace too small");
    VS_REQUIRE(layout && layout->n_levels >= 1 && layout->n_levels <= VS_MAX_LEVELS && layout->fanout >= 2,
               "vs_nearest_bound: bad tree layout");
    if (n_rows == 0 || layout->n_items == 0) return VS_OK;
    VS_REQUIRE(query_boxes && rows && row_boxes && node_boxes && rep && radius && search,
               "vs_nearest_bound: null pointer");
    VS_REQUIRE(tree_kind == 0 || tree_kind == 1, "vs_nearest_bound: tree_kind must be 0 (points) or 1 (polygons)");
    VS_REQUIRE(tree_kind == 0 || (item_ids && tx && ty && t_geom && t_part && t_ring),
               "vs_nearest_bound: polygonal tree needs item_ids and its arrays");
    hipStream_t s = vs::stream_of(stream);
    vs::Carver cw(workspace);
    int32_t* perm = cw.take<int32_t>(n_rows);
    void* order_ws = cw.take<char>(vs::box_order_workspace_bytes(n_rows));
    VS_TRY(vs::box_order(reinterpret_cast<const double4*>(row_boxes), n_rows, perm, order_ws, s));
    vs::NearestTree T;
    memset(&T, 0, sizeof(T));
    T.box = reinterpret_cast<const double4*>(node_boxes);
    T.rep = reinterpret_cast<const double2*>(rep);
    if (tree_kind == 1) {
        T.P = vs::Polys{tx, ty, t_geom, t_part, t_ring};
        T.item_ids = item_ids;
        T.polys = 1;
    }
    T.n_levels = layout->n_levels;
    T.fanout = layout->fanout;
    for (int l = 0; l < layout->n_levels; ++l) {
        T.level_start[l] = layout->level_start[l];
        T.level_count[l] = layout->level_count[l];
    }
    hipLaunchKernelGGL(vs::nearest_bound_kernel, dim3((unsigned)vs::ceil_div(n_rows, 256)), dim3(256), 0, s,
                       reinterpret_cast<const double4*>(query_boxes), rows, perm, n_rows, T, cap, radius, search);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int64_t vs_nearest_finish_workspace_bytes(int64_t n_pairs) { return vs::nearest_finish_workspace_bytes(n_pairs); }

int vs_nearest_finish(const int64_t* left, const int64_t* right, const double* distance, int64_t n_pairs,
                      int64_t n_query, int64_t* out_left, int64_t* out_right, double* out_distance, void* workspace,
                      int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    VS_REQUIRE(n_pairs >= 0 && n_pairs < (int64_t)INT32_MAX, "vs_nearest_finish: n_pairs out of range");
    VS_REQUIRE(n_query >= 0 && n_query <= (int64_t)INT32_MAX, "vs_nearest_finish: n_query out of range");
    VS_REQUIRE(workspace_bytes >= vs::nearest_finish_workspace_bytes(n_pairs),
               "vs_nearest_finish: workspace too small");
    if (n_pairs == 0) return VS_OK;
    VS_REQUIRE(left && right && distance && out_left && out_right && out_distance, "vs_nearest_finish: null pointer");
    hipStream_t s = vs::stream_of(stream);
    vs::Carver c(workspace);
    uint64_t* key = c.take<uint64_t>(n_pairs);
    int32_t* idx = c.take<int32_t>(n_pairs);
    void* sort_ws = c.take<char>(vs::sort_workspace_bytes_u64(n_pairs));
    const unsigned g = (unsigned)vs::ceil_div(n_pairs, 256);
    hipLaunchKernelGGL(vs::left_keys_kernel, dim3(g), dim3(256), 0, s, left, n_pairs, key, idx);
    VS_LAUNCH_CHECK();
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) < n_query) ++bits;
    VS_TRY(vs::sort_pairs_u64(key, idx, n_pairs, 0, bits, sort_ws, vs::sort_workspace_bytes_u64(n_pairs), s));
    hipLaunchKernelGGL(vs::gather_pairs_kernel, dim3(g), dim3(256), 0, s, idx, n_pairs, left, right, distance,
                       out_left, out_right, out_distance);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int vs_map_left(const int64_t* rows, int64_t* left, int64_t n_pairs, void* stream) {
    vs::clear_error();
    VS_REQUIRE(n_pairs >= 0, "vs_map_left: n_pairs < 0");
    if (n_pairs == 0) return VS_OK;
    VS_REQUIRE(rows && left, "vs_map_left: null pointer");
    hipLaunchKernelGGL(vs::map_left_kernel, dim3((unsigned)vs::ceil_div(n_pairs, 256)), dim3(256), 0,
                       vs::stream_of(stream), rows, left, n_pairs);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int64_t vs_nearest_select_workspace_bytes(int64_t n_pairs, int64_t n_query) {
    return vs::nearest_select_workspace_bytes(n_pairs, n_query);
}

int vs_nearest_select(const int64_t* left, const int64_t* right, const double* distance, int64_t n_pairs,
                      int64_t n_query, double* radius, double* search, double growth, double start_radius,
                      double switch_radius, double final_radius, int32_t return_all, int64_t* out_left, int64_t* out_right,
                      double* out_distance, int64_t* d_result, void* workspace, int64_t workspace_bytes,
                      void* stream) {
    vs::clear_error();
    VS_REQUIRE(n_pairs >= 0 && n_pairs < (int64_t)INT32_MAX, "vs_nearest_select: n_pairs out of range");
    VS_REQUIRE(n_query >= 0 && n_query < (int64_t)INT32_MAX, "vs_nearest_select: n_query out of range");
    VS_REQUIRE(workspace_bytes >= vs::nearest_select_workspace_bytes(n_pairs, n_query),
               "vs_nearest_select: workspace too small");
    VS_REQUIRE(d_result != nullptr && (n_query == 0 || (radius != nullptr && search != nullptr)),
               "vs_nearest_select: null pointer");
    VS_REQUIRE(growth > 1.0, "vs_nearest_select: growth must exceed 1");
    hipStream_t s = vs::stream_of(stream);
    vs::Carver c(workspace);
    const vs::NearestWs W = vs::carve_nearest(c, n_pairs, n_query);
    hipLaunchKernelGGL(vs::nearest_init_kernel, dim3((unsigned)std::max<int64_t>(1, vs::ceil_div(n_query, 256))),
                       dim3(256), 0, s, n_query, W.mind, W.first, W.open);
    VS_LAUNCH_CHECK();
    const unsigned gp = (unsigned)vs::ceil_div(n_pairs, 256);
    if (n_pairs > 0) {
        hipLaunchKernelGGL(vs::nearest_min_kernel, dim3(gp), dim3(256), 0, s, left, distance, n_pairs, W.mind);
        VS_LAUNCH_CHECK();
        hipLaunchKernelGGL(vs::nearest_keep_kernel, dim3(gp), dim3(256), 0, s, left, right, distance, n_pairs, radius,
                           W.mind, W.first, return_all ? 1 : 0, W.keep);
        VS_LAUNCH_CHECK();
        if (!return_all) {
            hipLaunchKernelGGL(vs::nearest_first_kernel, dim3(gp), dim3(256), 0, s, left, right, n_pairs, W.first,
                               W.keep);
            VS_LAUNCH_CHECK();
        }
    }
    VS_TRY(vs::exclusive_scan_i32(W.keep, W.pos, n_pairs, W.total, W.scan_ws, s));
    if (n_pairs > 0) {
        hipLaunchKernelGGL(vs::nearest_compact_kernel, dim3(gp), dim3(256), 0, s, left, right, distance, n_pairs,
                           W.keep, W.pos, out_left, out_right, out_distance);
        VS_LAUNCH_CHECK();
    }
    if (n_query > 0) {
        hipLaunchKernelGGL(vs::nearest_radius_kernel, dim3((unsigned)std::min<int64_t>(vs::ceil_div(n_query, 256), 2048)),
                           dim3(256), 0, s,
                           n_query, radius, search, W.mind, growth, start_radius, switch_radius, final_radius,
                           W.open);
        VS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(vs::nearest_result_kernel, dim3(1), dim3(64), 0, s, W.total, W.open, d_result);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

}  // extern "C"
