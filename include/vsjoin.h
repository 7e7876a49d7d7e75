/*
 * vsjoin.h -- C-ABI of libvsjoin.so, the MI355X (gfx950) spatial-join library.
 *
 * The reference (vibeSpatial) has no native library: its hot path is Python
 * that JIT-compiles CUDA strings (NVRTC) and calls CCCL/CuPy.  Each entry point
 * below replaces one step of that path; the comment names the reference code
 * it replaces (paths relative to the reference's src/vibespatial/).  The
 * Python binding that calls these is vibespatial_amd/_lib.py (ctypes).
 *
 * Conventions
 *   - Pointers are DEVICE pointers unless marked "host".  Buffers are owned by
 *     the caller (the Python layer allocates them as torch ROCm tensors).
 *   - Coordinates are fp64 structure-of-arrays.  Polygonal rows use
 *     GeoArrow-style int32 offsets (geometry/buffers.py:140-250):
 *       geom_offsets[row..row+1] -> parts, part_offsets[part..part+1] -> rings,
 *       ring_offsets[ring..ring+1] -> coordinates; rings are closed
 *     (first == last coordinate).  A polygon is a 1-part row; a row with no
 *     parts or no coordinates is empty and never matches (like null).
 *   - Points with a NaN coordinate are null/empty and never match.
 *   - `stream` is a hipStream_t (void* here); all work is enqueued on it and
 *     nothing synchronises the host, except where a function says so.
 *   - Return value: VS_OK (0) or a negative VS_ERR_* code; the message is
 *     available from vs_last_error() on the calling thread.
 *   - No entry point allocates device memory: work buffers come from the
 *     caller as one `workspace` of at least the size the matching
 *     *_workspace_bytes() query reports.
 *   - Output of variable size (pairs) is written into caller buffers of
 *     `capacity` pairs.  d_result (device int64[4]) receives
 *       [0] number of output pairs K, [1] number of pairs produced before the
 *       capacity check, [2] bbox candidate pairs examined (point joins),
 *       [3] sum over those candidates of the polygon's vertex count.
 *     If result[1] > capacity the outputs are not valid: call again with
 *     capacity >= result[1].
 *   - Thread safety: calls are reentrant; concurrent calls must use distinct
 *     workspaces (vs_sjoin_points_polygons_noalloc shares one array per GPU:
 *     its calls must not overlap).
 *   - Side stream: vs_learned_index_build and vs_sjoin_points_polygons_learned
 *     overlap the network's training and the error pass with the coordinate
 *     gather on a second stream (non-blocking, default priority) with a
 *     fork / join event pair.  There is one such stream per host thread and
 *     device, created on that thread's first such call on the device and
 *     kept until the process ends; before returning, the call makes the
 *     caller's stream wait for it, so the work stays ordered on the caller's
 *     stream.  Other entry points use only the caller's stream.
 */
#ifndef VSJOIN_H
#define VSJOIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VS_OK 0
#define VS_ERR_INVALID (-1)
#define VS_ERR_HIP (-2)

#define VS_MAX_LEVELS 16

/* Predicates, as predicate(left, right) with points on the left and
 * polygons on the right; mapping of location codes 0 outside / 1 boundary /
 * 2 interior follows predicates/point_relations.py:100-129. */
#define VS_PRED_WITHIN 0     /* location == 2 */
#define VS_PRED_INTERSECTS 1 /* location != 0 */
#define VS_PRED_COVERED_BY 2 /* location != 0 */
#define VS_PRED_TOUCHES 3    /* location == 1 */
#define VS_PRED_CONTAINS 4   /* polygon (left) contains point (right): location == 2 */
#define VS_PRED_COVERS 5     /* polygon (left) covers point (right): location != 0 */
/* polygon x polygon only (vs_refine_pairs mode 1): */
#define VS_PRED_OVERLAPS 6
#define VS_PRED_CROSSES 7 /* never true for two areal rows */
#define VS_PRED_CONTAINS_PROPERLY 8

/* Library version (major * 10000 + minor * 100 + patch). */
int vs_version(void);

/* Copies the calling thread's last error message (NUL-terminated) into the
 * host buffer; returns the message length. */
int vs_last_error(char* buf, int64_t capacity);

/* ---------------------------------------------------------------------
 * Geometry bounds
 * --------------------------------------------------------------------- */

/* Per-row fp64 bbox {minx, miny, maxx, maxy}; NaN for empty rows.
 * Replaces compute_geometry_bounds_device (kernels/core/geometry_analysis.py:648)
 * and its kernels bounds_polygon/bounds_multipolygon
 * (kernels/core/geometry_analysis_source.py:31-212). */
int vs_polygon_bounds(const double* x, const double* y, const int32_t* geom_offsets,
                      const int32_t* part_offsets, const int32_t* ring_offsets, int64_t n_rows,
                      double* bounds, void* stream);

/* Degenerate boxes of points; NaN for null points.  Replaces bounds_simple
 * (kernels/core/geometry_analysis_source.py:31-72). */
int vs_point_boxes(const double* px, const double* py, int64_t n, double* boxes, void* stream);

/* ---------------------------------------------------------------------
 * Packed Hilbert tree over boxes
 * --------------------------------------------------------------------- */

/* Level layout of a packed tree: level 0 is the root, the last level holds the
 * items (boxes in Hilbert order); the children of node j at level l are nodes
 * [j*fanout, (j+1)*fanout) of level l+1. */
typedef struct {
    int32_t fanout;
    int32_t n_levels;
    int64_t n_items;
    int64_t n_nodes; /* total boxes over all levels, items included */
    int64_t level_start[VS_MAX_LEVELS];
    int64_t level_count[VS_MAX_LEVELS];
} vs_tree_layout;

/* Host-only: fills the layout for n_items and fanout (2..64). */
int vs_tree_plan(int64_t n_items, int32_t fanout, vs_tree_layout* layout);

int64_t vs_index_workspace_bytes(int64_t n_items);

/* Builds the tree from per-item boxes: 64-bit Hilbert keys of box centres over
 * the centre extent (computed on the device, no host fence), stable radix
 * sort, items gathered in key order, parents bottom up.
 *   node_boxes: 4 * layout->n_nodes doubles; item_ids: n_items original rows;
 *   item_keys (optional, may be NULL): sorted keys.
 * Replaces build_flat_spatial_index (spatial/indexing.py:1357-1533),
 * morton_keys_from_bounds (spatial/indexing_kernels.py:20-49), CCCL sort_pairs
 * (cuda/cccl_primitives.py:1054) and _build_packed_rtree
 * (io/gpu_parse/indexing.py:262-400). */
int vs_index_build(const double* bounds, int64_t n_items, const vs_tree_layout* layout,
                   double* node_boxes, int32_t* item_ids, uint64_t* item_keys, void* workspace,
                   int64_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------
 * Joins and queries
 * --------------------------------------------------------------------- */

int64_t vs_sjoin_points_polygons_workspace_bytes(int64_t n_points, int64_t n_polys, int64_t n_coords,
                                                 int64_t capacity);

/* sjoin(points, polygons, predicate): all (point, polygon) pairs with
 * predicate(point, polygon) true, lexsorted, as int64 (left_index, right_index).
 * The polygon tree comes from vs_index_build over the polygons' bounds;
 * n_parts / n_rings / n_coords are the lengths of part_offsets - 1,
 * ring_offsets - 1 and x (host values, used to size the refine layout).
 * predicate: VS_PRED_WITHIN | VS_PRED_INTERSECTS | VS_PRED_COVERED_BY | VS_PRED_TOUCHES.
 * Exact (no tolerance): boundary = the point lies on an edge.
 * Replaces query_spatial_index's point-x-polygon route (spatial/query.py:247-1042):
 * spatial_index_device_query (spatial/spatial_index_device.py:228-390), the
 * point refine classify_point_predicates_indexed_device
 * (predicates/point_relations.py:1040-1191) and the pair compaction
 * (spatial/query_utils.py:867-868, spatial/query.py:143-153). */
int vs_sjoin_points_polygons(const double* px, const double* py, int64_t n_points, const double* x,
                             const double* y, const int32_t* geom_offsets, const int32_t* part_offsets,
                             const int32_t* ring_offsets, int64_t n_polys, int64_t n_parts, int64_t n_rings,
                             int64_t n_coords, const vs_tree_layout* layout, const double* node_boxes,
                             const int32_t* item_ids, int32_t predicate, int64_t capacity, int64_t* out_left,
                             int64_t* out_right, int64_t* d_result, void* workspace, int64_t workspace_bytes,
                             void* stream);

/* Bytes of the fixed device scratch array behind vs_sjoin_points_polygons_noalloc
 * (one copy per GPU, held until the process ends). */
#define VS_NOALLOC_SCRATCH_BYTES (64ll << 20)

/* vs_sjoin_points_polygons without a workspace argument and without any
 * device allocation: its work memory is a fixed VS_NOALLOC_SCRATCH_BYTES
 * array in device memory (one per GPU, its address looked up on the first
 * call on each device).  No counterpart in the reference: it is for callers
 * that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and errors as
 * vs_sjoin_points_polygons; in addition VS_ERR_INVALID when
 * vs_sjoin_points_polygons_workspace_bytes(n_points, n_polys, n_coords,
 * capacity) exceeds VS_NOALLOC_SCRATCH_BYTES (about 1M points x 10k
 * polygons with a capacity of 1M pairs fit).  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, other host
 * threads) race and must be serialised by the caller.  The caller still
 * provides the tree (vs_index_build), outputs and d_result. */
int vs_sjoin_points_polygons_noalloc(const double* px, const double* py, int64_t n_points, const double* x,
                                     const double* y, const int32_t* geom_offsets, const int32_t* part_offsets,
                                     const int32_t* ring_offsets, int64_t n_polys, int64_t n_parts, int64_t n_rings,
                                     int64_t n_coords, const vs_tree_layout* layout, const double* node_boxes,
                                     const int32_t* item_ids, int32_t predicate, int64_t capacity,
                                     int64_t* out_left, int64_t* out_right, int64_t* d_result, void* stream);

int64_t vs_query_boxes_workspace_bytes(int64_t n_queries, int64_t capacity);

/* Bounding-box query: every (query, tree item) pair whose boxes overlap
 * (inclusive; NaN boxes never match), int64.  query_boxes: 4 doubles per
 * query.  sort != 0: lexsorted (query, item); sort == 0: grouped by query in
 * ascending order, items within a query in unspecified order (the reference's
 * sort=False, spatial/query.py:143-153).  capacity also sets the traversal's
 * shape (the pairs are the same either way): at least 32 slots per query box,
 * one box per wave with the lanes over a node's children (boxes with many
 * hits); fewer, 64 boxes per wave over the union of their boxes.  Replaces
 * spatial_index_device_query's candidate output
 * (spatial/spatial_index_device.py:228-390) and the bbox_overlap kernels
 * (kernels/core/spatial_query_source.py:415-555). */
int vs_query_boxes(const double* query_boxes, int64_t n_queries, const vs_tree_layout* layout,
                   const double* node_boxes, const int32_t* item_ids, int64_t capacity, int32_t sort,
                   int64_t* out_left, int64_t* out_right, int64_t* d_result, void* workspace,
                   int64_t workspace_bytes, void* stream);

int64_t vs_refine_pairs_workspace_bytes(int64_t n_pairs);

/* Exact refine of candidate pairs (left[i], right[i]) followed by an
 * order-preserving compaction into out_left/out_right (capacity n_pairs);
 * d_result[0] (device int64) receives the number kept.
 *   mode 0: left rows index points (px, py), right rows polygons B;
 *           keep = predicate(location of the point in the polygon), with
 *           VS_PRED_WITHIN / VS_PRED_CONTAINS: interior, VS_PRED_TOUCHES:
 *           boundary, otherwise: not outside.
 *   mode 1: left rows polygons A, right rows polygons B; keep =
 *           predicate(A, B): intersects, or the exact relations of two areal
 *           rows -- within / covered_by (A inside B, closed), contains /
 *           covers, contains_properly (no boundary contact), touches
 *           (boundaries meet, interiors do not), overlaps (interiors meet,
 *           neither inside the other), crosses (never).
 *   bits 8-15 of mode (mode 1): log2 of the coordinate-count product above
 *           which a pair is refined one wave per pair (0: the default 2^8;
 *           63: never); the results are the same whatever the threshold.
 *   mode 2: left rows polygons B, right rows points (px, py); keep as mode 0
 *           for the right point located in the left polygon (contains,
 *           covers, contains_properly, touches, intersects of polygons over
 *           points, api/sindex.py:144-503 with the swapped mapping of
 *           predicates/point_relations.py:130-146).
 * Replaces _filter_predicate_pairs_owned's device refine routes
 * (spatial/query_utils.py:578-1147), compute_polygonal_intersects_gpu
 * (predicates/polygon.py:849-996) and the boolean compaction
 * (spatial/query_utils.py:867-868). */
int vs_refine_pairs(int32_t mode, int32_t predicate, const int64_t* left, const int64_t* right, int64_t n_pairs,
                    const double* px, const double* py, const double* ax, const double* ay,
                    const int32_t* a_geom, const int32_t* a_part, const int32_t* a_ring, const double* bx,
                    const double* by, const int32_t* b_geom, const int32_t* b_part, const int32_t* b_ring,
                    int64_t* out_left, int64_t* out_right, int64_t* d_result, void* workspace,
                    int64_t workspace_bytes, void* stream);

/* out[i] = boxes[i] grown by distance[i] (per row, device fp64) on every
 * side, each bound rounded one ulp outward so every point within the distance
 * of the box lies in the grown box; NaN or negative distance -> NaN box.
 * The distance-expanded candidate filter of the reference's dwithin query
 * (spatial/spatial_index_device.py:398-430). */
int vs_expand_boxes(const double* boxes, int64_t n, const double* distance, double* out, void* stream);

int64_t vs_refine_dwithin_workspace_bytes(int64_t n_pairs);

/* Exact dwithin refine of candidate pairs (left[i], right[i]) with
 * d = distance[left[i]] (device fp64 per left row), then an order-preserving
 * compaction into out_left/out_right; d_result[0] = pairs kept.  n_right:
 * right row count (right ids < n_right); pairs are visited grouped by right
 * row.  Workspace: vs_refine_dwithin_workspace_bytes(n_pairs).  Kinds: 0 =
 * point rows (x[row],
 * y[row]; offsets unused), 1 = polygonal rows (GeoArrow offsets).  A pair is
 * kept when the exact distance between the geometries is <= d: points by
 * d^2 - |p - q|^2 >= 0; a point and a polygon when the point is not outside
 * it or within d of a boundary segment; two polygons when they intersect or
 * a vertex of one is within d of a boundary segment of the other.  NaN
 * coordinates, empty rows and NaN / negative d never match.  Replaces
 * _dwithin_refine_gpu (spatial/nearest.py:2855-3016), which compares rounded
 * fp64 distances. */
int vs_refine_dwithin(int32_t left_kind, int32_t right_kind, const int64_t* left, const int64_t* right,
                      int64_t n_pairs, int64_t n_right, const double* distance, const double* lx, const double* ly,
                      const int32_t* l_geom, const int32_t* l_part, const int32_t* l_ring, const double* rx,
                      const double* ry, const int32_t* r_geom, const int32_t* r_part, const int32_t* r_ring,
                      int64_t* out_left, int64_t* out_right, int64_t* d_result, void* workspace,
                      int64_t workspace_bytes, void* stream);

int64_t vs_lexsort_pairs_workspace_bytes(int64_t n_pairs);

/* In-place lexsort of int64 pairs by (left, right); values must be in
 * [0, 2^31).  Replaces cp.lexsort (spatial/query.py:105-115, 143-153). */
int vs_lexsort_pairs(int64_t* left, int64_t* right, int64_t n_pairs, void* workspace, int64_t workspace_bytes,
                     void* stream);

/* ---------------------------------------------------------------------
 * Learned point index and the learned + FP32 join form
 * (no reference counterpart; see the comment at the top of vs_learned.hip)
 * --------------------------------------------------------------------- */

/* Number of leaves of the piecewise-linear stage for n points and a leaf size
 * (<= 0: 512 positions per leaf). */
int32_t vs_learned_leaf_count(int64_t n_points, int32_t leaf_size);

/* Bytes of the index buffer (keys, permutation, sorted x/y, extent, MLP
 * parameters, training losses, leaf table, leaf errors, meta). */
int64_t vs_learned_index_bytes(int64_t n_points, int32_t key_bytes, int32_t n_leaves);

/* Byte offsets inside the index buffer of: keys, perm, sorted x, sorted y,
 * extent (4 doubles), MLP params (97 floats: w1[32], b1[32], w2[32], b2),
 * losses (2 floats: before / after training), leaf keys (n_leaves + 1),
 * leaf errors (int32 n_leaves), meta (int32[4]: [0] max MLP leaf error),
 * routing windows (int32 pairs, n_leaves: for each leaf the network predicts,
 * the first and last true leaf of the keys it routes there; first > last
 * when it routes none). */
int vs_learned_index_layout(int64_t n_points, int32_t key_bytes, int32_t n_leaves, int64_t* offsets11);

/* Builds the index over points: Hilbert keys over the point extent, radix-
 * sorted (key_bytes 8: a 2^32 x 2^32 grid; key_bytes 4: a 2^12 x 2^12 grid,
 * left-aligned so the low 8 key bits are zero); the MLP is trained on the GPU for
 * train_iters full-batch Adam steps (lr relative to each parameter's scale);
 * leaf table and error bounds are recorded over every key. */
int vs_learned_index_build(const double* px, const double* py, int64_t n_points, int32_t key_bytes, int32_t n_leaves,
                           int32_t train_iters, float lr, void* index, int64_t index_bytes, void* stream);

/* Batch lower_bound of query keys in the index's sorted keys.  lookup_mode:
 * 0 learned (the network's routing window of leaves, its max leaf error when
 * the window is empty, then the leaf error window), 1 leaf table only (binary
 * search over all leaves, then the leaf error window), 2 plain binary search.
 * d_stat3 (device uint64[3], accumulated): probes, MLP-window fallbacks,
 * error-window fallbacks.  Results are exact in every mode. */
int vs_learned_lower_bound(const void* index, int64_t n_points, int32_t key_bytes, int32_t n_leaves,
                           const void* query_keys, int64_t m, int32_t lookup_mode, int64_t* out, uint64_t* d_stat3,
                           void* stream);

/* Full-batch MSE loss and gradient of MLP parameters on (u, t) data, computed
 * by the device code the trainer uses (for gradient checks). */
int vs_mlp_loss_grad(const float* params, const float* u, const float* t, int64_t n, float* grad_out, float* loss_out,
                     void* stream);

int64_t vs_learned_query_boxes_workspace_bytes(int64_t n_boxes, int64_t capacity);

/* sindex.query(boxes, predicate='intersects') on a point tree: (box, point)
 * pairs with the point inside the box (inclusive) through the learned index
 * (cover cells -> rank ranges -> scan); `sort` as for vs_query_boxes.
 * d_result: int64[16], layout as for the learned join.  Replaces
 * _query_point_tree_box_index (spatial/query_box.py:760-863). */
int vs_learned_query_boxes(const void* index, int64_t n_points, int32_t key_bytes, int32_t n_leaves,
                           const double* boxes, int64_t n_boxes, int32_t lookup_mode, int64_t capacity,
                           int32_t sort, int64_t* out_left, int64_t* out_right, int64_t* d_result,
                           void* workspace, int64_t workspace_bytes, void* stream);

int64_t vs_learned_query_polygons_workspace_bytes(int64_t n_polys, int64_t n_coords, int64_t capacity);

/* sindex.query(polygons, predicate) on a point tree through its learned
 * index: (polygon row, point) pairs with predicate(polygon, point) --
 * VS_PRED_CONTAINS / VS_PRED_CONTAINS_PROPERLY: the point is interior,
 * VS_PRED_TOUCHES: on the boundary, VS_PRED_COVERS / VS_PRED_INTERSECTS: not
 * outside -- decided exactly (the FP32 refine with escalation of the learned
 * join); `sort` as for vs_query_boxes; d_result int64[16] as for the learned
 * join.  poly_bounds from vs_polygon_bounds.  Replaces the polygon-over-point
 * route of query_spatial_index on point trees (spatial/query_box.py:760-863
 * for the candidate boxes, predicates/point_relations.py:130-146 for the
 * swapped predicate mapping). */
int vs_learned_query_polygons(const void* index, int64_t n_points, int32_t key_bytes, int32_t n_leaves,
                              const double* x, const double* y, const int32_t* geom_offsets,
                              const int32_t* part_offsets, const int32_t* ring_offsets, int64_t n_polys,
                              int64_t n_parts, int64_t n_rings, int64_t n_coords, const double* poly_bounds,
                              int32_t predicate, int32_t lookup_mode, int64_t capacity, int32_t sort,
                              int64_t* out_left, int64_t* out_right, int64_t* d_result, void* workspace,
                              int64_t workspace_bytes, void* stream);

int64_t vs_sjoin_points_polygons_learned_workspace_bytes(int64_t n_points, int64_t n_polys, int64_t n_coords,
                                                         int64_t capacity);

/* The learned + FP32 form of vs_sjoin_points_polygons (same pairs, bit for
 * bit): builds the learned index over the points (32-bit keys; the network
 * gets train_iters Adam steps of 3e-4 relative to each parameter's scale
 * after its quantile-knot start), covers each
 * polygon's bounds (poly_bounds, from vs_polygon_bounds) with rank ranges and
 * refines candidates polygon-major in FP32 with escalation to the exact fp64
 * step.  d_result (device int64[16]): [0] K, [1] staged, [2] candidates,
 * [3] candidate vertices, [4] fp32 orientation evaluations, [5] escalated
 * edges, [6] candidate pairs with an escalation, [7] lookup probes,
 * [8] MLP-window fallbacks, [9] error-window fallbacks, [10] positions
 * scanned, [11] max MLP leaf error. */
int vs_sjoin_points_polygons_learned(const double* px, const double* py, int64_t n_points, const double* x,
                                     const double* y, const int32_t* geom_offsets, const int32_t* part_offsets,
                                     const int32_t* ring_offsets, int64_t n_polys, int64_t n_parts, int64_t n_rings,
                                     int64_t n_coords, const double* poly_bounds, int32_t predicate,
                                     int32_t train_iters, int32_t lookup_mode, int64_t capacity, int64_t* out_left,
                                     int64_t* out_right, int64_t* d_result, void* workspace, int64_t workspace_bytes,
                                     void* stream);

/* ---------------------------------------------------------------------
 * Batched polygon x polygon intersects
 * --------------------------------------------------------------------- */

int64_t vs_batched_intersects_workspace_bytes(int64_t n_jobs, int64_t n_left, int64_t n_right, int64_t n_right_coords,
                                              int64_t capacity);

/* Many independent sjoin(left_j, right_j, predicate='intersects') jobs in one
 * fixed launch sequence (a forest of packed trees, one persistent traversal,
 * decisions from 32 x 32 cell grids of the right rows where they settle a
 * pair, one exact refine of the rest, one finish).  Job j owns left rows [left_job_offsets[j],
 * left_job_offsets[j+1]) of A and right rows [right_job_offsets[j],
 * right_job_offsets[j+1]) of B (device int64 arrays of n_jobs + 1 entries,
 * non-decreasing); only pairs inside one job are tested.  n_right_coords:
 * the length of bx / by (sizes the y-slab index of the right rows).  Output:
 * global row pairs lexsorted, so each job's pairs are contiguous and lexsorted.
 * tile: bits 0-7 traversal waves per block (1-4, 0 = 4), bits 8-23 refine
 * block size (multiple of 64 up to 256, 0 = 256), bit 24 candidate order
 * of the refine (0: by right row; 1: by right rows' coordinate-count bins,
 * largest first, then right row), bit 25 set: no cell grids (every bbox
 * candidate refined exactly), bits 26-27: right rings read through a y-slab
 * index from 128 (0), 16 (1) or 64 (2) edges on, or never (3).  Every tile gives the same pairs.  d_result
 * (device int64[5]): [0] pairs, [1] bbox candidates staged (> capacity:
 * retry with that capacity), [2] bbox candidates, [3] sum over the
 * candidates of both rows' coordinate counts, [4] candidates left to the
 * exact refine.  Replaces, per job, the reference's
 * polygon x polygon sjoin path (spatial/spatial_index_device.py:228-390,
 * predicates/polygon.py:849-996, spatial/query.py:143-153). */
int vs_batched_intersects(const double* ax, const double* ay, const int32_t* a_geom, const int32_t* a_part,
                          const int32_t* a_ring, int64_t n_left, const double* bx, const double* by,
                          const int32_t* b_geom, const int32_t* b_part, const int32_t* b_ring, int64_t n_right,
                          int64_t n_right_coords, const int64_t* left_job_offsets, const int64_t* right_job_offsets,
                          int64_t n_jobs, int32_t tile, int64_t capacity, int64_t* out_left, int64_t* out_right,
                          int64_t* d_result, void* workspace, int64_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------
 * WKB ingest (GeoParquet / GeoArrow WKB columns -> SoA on the device)
 * --------------------------------------------------------------------- */

int64_t vs_wkb_workspace_bytes(int64_t n_rows);

/* Pass 1 over n_rows WKB records: record r is payload[offsets[r],
 * offsets[r+1]) (device int64 offsets, n_rows + 1) of the payload_bytes
 * bytes at payload; a record not wholly inside [0, payload_bytes) or whose
 * end offset is below its start is reported as truncated, so no read leaves
 * the payload whatever the offsets hold.  validity (device uint8 per row, or
 * null = all valid): 0 or an empty record = null row.  Supports
 * Point, Polygon and MultiPolygon in 2D, either byte order, ISO or EWKB type
 * words (an SRID is skipped).  d_result (device int64[8]): [0] point rows,
 * [1] polygonal rows, [2] parts, [3] rings, [4] polygon coordinates, [5] first
 * bad record or -1, [6] its error (1 truncated, 2 unsupported type or
 * dimension, 3 bad byte order, 4 multipolygon member not a polygon, 5 count
 * over 2^31, 6 bytes after the geometry), [7] 1 when parts, rings or
 * coordinates of the column exceed INT32_MAX (vs_wkb_fill then writes
 * nothing).  The workspace carries the counts to vs_wkb_fill.  Replaces the
 * reference's device WKB decode (kernels/core/wkb_decode.py:849, used by
 * io/wkb.py:5049 and io/geoparquet.py:3430). */
int vs_wkb_scan(const uint8_t* payload, int64_t payload_bytes, const int64_t* offsets, const uint8_t* validity,
                int64_t n_rows, int64_t* d_result, void* workspace, int64_t workspace_bytes, void* stream);

/* Pass 2 (same records and workspace, after vs_wkb_scan): point rows ->
 * point_x/point_y (result[0] entries) with their record numbers in point_rows;
 * polygonal rows -> x/y (result[4]), ring_offsets (result[3] + 1),
 * part_offsets (result[2] + 1), geom_offsets (result[1] + 1) with their record
 * numbers in polygon_rows; a Polygon is one part (none when it has no rings),
 * a MultiPolygon one part per member. */
int vs_wkb_fill(const uint8_t* payload, int64_t payload_bytes, const int64_t* offsets, const uint8_t* validity,
                int64_t n_rows, const void* workspace, int64_t workspace_bytes, double* point_x, double* point_y,
                int64_t* point_rows, double* x, double* y, int32_t* geom_offsets, int32_t* part_offsets,
                int32_t* ring_offsets, int64_t* polygon_rows, void* stream);

/* ---------------------------------------------------------------------
 * Nearest join (sindex.nearest / sjoin_nearest)
 * --------------------------------------------------------------------- */

int64_t vs_pair_distances_workspace_bytes(int64_t n_pairs);

/* fp64 distance of each candidate pair (kinds 0 points, 1 polygons; geometry
 * arguments as in vs_refine_dwithin; n_right = rows of the right side):
 * point-point sqrt(dx^2 + dy^2), point-segment GEOS pointToSegment, 0 when a
 * point is not outside a polygon or two polygons intersect (exact tests),
 * else the least vertex-segment distance; +inf for empty / NaN geometries.
 * flags: VS_DIST_EXCLUSIVE, point pairs with equal coordinates at +inf;
 * VS_DIST_OUTSIDE, the caller knows every left point lies outside every right
 * polygon it is paired with (the location steps are skipped; a wrong claim
 * gives wrong distances).  cap (device fp64 per left row, or NULL; NaN = no
 * cap): a pair farther than cap[left] gets +inf instead of its distance.
 * Replaces _compute_pair_distances_gpu / _compute_mixed_distances_gpu
 * (spatial/nearest.py:2228, 2482). */
#define VS_DIST_EXCLUSIVE 1
#define VS_DIST_OUTSIDE 2
int vs_pair_distances(int32_t left_kind, int32_t right_kind, const int64_t* left, const int64_t* right,
                      int64_t n_pairs, int64_t n_right, int32_t flags, const double* cap, const double* lx,
                      const double* ly, const int32_t* l_geom, const int32_t* l_part, const int32_t* l_ring,
                      const double* rx, const double* ry, const int32_t* r_geom, const int32_t* r_part,
                      const int32_t* r_ring, double* out_distance, void* workspace, int64_t workspace_bytes,
                      void* stream);

int64_t vs_nearest_select_workspace_bytes(int64_t n_pairs, int64_t n_query);

/* One round of the expanding-radius nearest search.  Pairs (left = query
 * row < n_query) with their distances; radius[q] (device fp64) is the
 * query's radius this round (NaN = closed).  A query whose least finite
 * candidate distance m is <= radius[q] keeps its pairs at distance m (all of
 * them, or with return_all = 0 the one with the least right row) and closes;
 * kept pairs are compacted in input order to out_left / out_right /
 * out_distance (may be null).  radius is updated in place for the next round:
 * NaN when closed or when radius[q] >= final_radius, else radius * growth
 * (start_radius after a radius-0 round; final_radius once that reaches
 * switch_radius); search[q] = the next radius times (1 + 2^-40), the
 * amount to grow the query box by.  d_result (device
 * int64[2]): [0] kept pairs, [1] queries still open.  Replaces
 * _nearest_from_distances (spatial/nearest.py:3237) with ties as equal fp64
 * distances. */
int vs_nearest_select(const int64_t* left, const int64_t* right, const double* distance, int64_t n_pairs,
                      int64_t n_query, double* radius, double* search, double growth, double start_radius,
                      double switch_radius, double final_radius, int32_t return_all, int64_t* out_left,
                      int64_t* out_right, double* out_distance, int64_t* d_result, void* workspace,
                      int64_t workspace_bytes, void* stream);

/* radius[q] = first_radius and search[q] = first_radius times (1 + 2^-40)
 * for every query with a finite box (n_query x 4), NaN (closed) otherwise. */
int vs_nearest_init(const double* boxes, int64_t n_query, double first_radius, double* radius, double* search,
                    void* stream);

int64_t vs_nearest_open_workspace_bytes(int64_t n_query);

/* The open queries of the next round (search[q] not NaN), in row order:
 * open_rows (int64) and their boxes (n x 4 fp64) grown by search[q] with
 * every bound rounded one ulp outward; d_count (device int64[1]) = how
 * many. */
int vs_nearest_open(const double* boxes, const double* search, int64_t n_query, int64_t* open_rows,
                    double* open_boxes, int64_t* d_count, void* workspace, int64_t workspace_bytes, void* stream);

/* For each listed query row q (rows, n_rows; query_boxes n x 4 over all
 * rows): an upper bound of its nearest distance to the tree's geometries --
 * the least, over the items, of the farthest-point distance from the query
 * box to the item box or to the item's rep point, and for a point query
 * against a polygonal tree (tree_kind 1: item_ids and the tree's polygon
 * arrays as in vs_pair_distances) of the item's fp64 distance, by a
 * nearest-first depth-first walk of the tree pruned by the bound so far -- sets
 * radius[q] = min(bound, cap) and search[q] = radius times (1 + 2^-40);
 * rows without a finite bound keep their radius.  rep: vs_nearest_rep_points
 * of the tree's items (tightens the bound).  row_boxes (n_rows x 4, e.g. the
 * open boxes of vs_nearest_open) only order the rows (Hilbert order of their
 * centres, so a wave's lanes walk nearby parts of the tree). */
int64_t vs_nearest_bound_workspace_bytes(int64_t n_rows);
int vs_nearest_bound(const double* query_boxes, const int64_t* rows, int64_t n_rows, const double* row_boxes,
                     const vs_tree_layout* layout, const double* node_boxes, const double* rep, int32_t tree_kind,
                     const int32_t* item_ids, const double* tx, const double* ty, const int32_t* t_geom,
                     const int32_t* t_part, const int32_t* t_ring, double cap, double* radius, double* search,
                     void* workspace, int64_t workspace_bytes, void* stream);

/* rep (n_items x 2 fp64, per item slot of a tree): a point of item
 * item_ids[slot] -- the point (kind 0: x, y per row) or the first coordinate
 * of the polygonal row (kind 1; NaN when empty). */
int vs_nearest_rep_points(int32_t kind, const int32_t* item_ids, int64_t n_items, const double* x, const double* y,
                          const int32_t* geom, const int32_t* part, const int32_t* ring, double* rep, void* stream);

int64_t vs_nearest_finish_workspace_bytes(int64_t n_pairs);

/* The rounds' kept pairs (each query's pairs from one round, sorted by right
 * row there) -> out_* sorted by left row (stable radix sort, left < n_query),
 * distances carried along: the lexsorted result. */
int vs_nearest_finish(const int64_t* left, const int64_t* right, const double* distance, int64_t n_pairs,
                      int64_t n_query, int64_t* out_left, int64_t* out_right, double* out_distance, void* workspace,
                      int64_t workspace_bytes, void* stream);

/* left[i] = rows[left[i]] (pairs of an open-query subset back to query rows;
 * rows ascending keeps the order). */
int vs_map_left(const int64_t* rows, int64_t* left, int64_t n_pairs, void* stream);

/* ---------------------------------------------------------------------
 * Primitives
 * --------------------------------------------------------------------- */

int64_t vs_scan_workspace_bytes(int64_t n);

/* out[i] = sum(in[0..i)), *d_total = sum(in) (device int64).  In-place allowed.
 * Replaces CCCL exclusive_sum (cuda/cccl_primitives.py:596). */
int vs_exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, int64_t* d_total,
                          void* workspace, int64_t workspace_bytes, void* stream);

/* key_bytes: 4 or 8. */
int64_t vs_sort_workspace_bytes(int64_t n, int32_t key_bytes);

/* Stable LSD radix sort of (key, value) in place by bits [begin_bit, end_bit).
 * Replaces CCCL sort_pairs (cuda/cccl_primitives.py:1054). */
int vs_sort_pairs_u64(uint64_t* keys, int32_t* vals, int64_t n, int32_t begin_bit, int32_t end_bit,
                      void* workspace, int64_t workspace_bytes, void* stream);
int vs_sort_pairs_u32(uint32_t* keys, int32_t* vals, int64_t n, int32_t begin_bit, int32_t end_bit,
                      void* workspace, int64_t workspace_bytes, void* stream);

int64_t vs_point_order_workspace_bytes(int64_t n_points);

/* Hilbert order of points (keys over the point extent, bits_per_axis <= 16):
 * perm[i] = original index of the i-th point in order; sorted_x/y gathered. */
int vs_point_order(const double* px, const double* py, int64_t n_points, int32_t bits_per_axis,
                   int32_t* perm, double* sorted_x, double* sorted_y, void* workspace,
                   int64_t workspace_bytes, void* stream);

/* Benchmark instrumentation (no reference counterpart): when enabled, the
 * dominant kernel (traversal / learned join / batched refine) of each
 * join or query the calling host thread makes is bracketed by two HIP events
 * recorded on the call's stream; vs_last_kernel_ms waits for the second
 * event and returns the elapsed milliseconds of the last one.  The state is
 * per host thread (threads time their own calls); vs_set_kernel_timing(1)
 * creates the thread's events on the device current then, and
 * vs_set_kernel_timing(0) destroys them. */
int vs_set_kernel_timing(int32_t enable);
int vs_last_kernel_ms(float* ms);

/* ---------------------------------------------------------------------
 * Host-side functions (no GPU needed): the device arithmetic compiled for the
 * host, so unit tests can check it without a GPU.
 * --------------------------------------------------------------------- */

/* Exact orientation sign of (a, b, c) (orient2d.py:37-84 sign convention). */
int vs_host_orient2d(double ax, double ay, double bx, double by, double cx, double cy);

/* Location code (0/1/2) of (px, py) in polygonal row `row`; host pointers. */
int vs_host_point_location(double px, double py, const double* x, const double* y,
                           const int32_t* geom_offsets, const int32_t* part_offsets,
                           const int32_t* ring_offsets, int64_t row);

/* Exact closed-segment touch-or-cross (predicates/polygon_kernels.py:131-161). */
int vs_host_segments_touch(double ax, double ay, double bx, double by, double cx, double cy,
                           double dx, double dy);

/* Exact |p - q| <= d (1/0); NaN coordinates or a NaN / negative d give 0,
 * d = +inf gives 1.  The device dwithin decision for point pairs. */
int vs_host_point_point_within(double px, double py, double qx, double qy, double d);

/* Exact: the distance from p to the closed segment a-b is <= d (1/0); same
 * NaN / infinity rules.  The device decision for a point and one edge. */
int vs_host_point_segment_within(double px, double py, double ax, double ay, double bx, double by, double d);

/* 64-bit Hilbert index of 32-bit grid coordinates. */
uint64_t vs_host_hilbert64(uint32_t x, uint32_t y);

/* The learned index's cover of the box [x0, x1] x [y0, y1] over a point
 * extent (extent4: x0, y0, x1, y1 of the points) for key_bytes 4 or 8: at
 * most 16 inclusive key ranges (lo, hi) written to out32 (uint64[32]);
 * returns their count.  Every point of the box has its key in one of them. */
int vs_host_cover_ranges(double x0, double y0, double x1, double y1, const double* extent4, int32_t key_bytes,
                         uint64_t* out32);

/* The 32 x 32 cell grid that vs_batched_intersects builds for a right row of
 * one closed ring (x, y: its n coordinates; box4: its bounds): geo6 =
 * (x0, y0, wx, wy, 1/wx, 1/wy), cells64[2 j] = boundary bits and
 * cells64[2 j + 1] = inside bits of grid row j.  Returns 1, or 0 when the
 * ring gets no grid (fewer than 4 coordinates, not closed, a non-finite
 * coordinate, an empty side). */
int vs_host_cell_grid(const double* x, const double* y, int64_t n, const double* box4, double* geo6,
                      uint32_t* cells64);

/* The batched join's decision for a left ring (x, y: its n >= 1
 * coordinates) against such a grid: 1 intersects, 0 does not, 2 undecided with the exact
 * test's first step (the left first vertex located in the right ring)
 * skippable, 3 undecided (also for a non-finite coordinate). */
int vs_host_cell_decide(const double* geo6, const uint32_t* cells64, const double* x, const double* y, int32_t n);

/* The lowest and highest grid cell (0..31) whose closed range holds v along
 * axis 0 (x) or 1 (y) of such a grid, into lo_hi[2]. */
int vs_host_cell_span(const double* geo6, int32_t axis, double v, int32_t* lo_hi);

#ifdef __cplusplus
}
#endif

#endif /* VSJOIN_H */
