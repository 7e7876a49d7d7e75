"""Device pipelines over libvsjoin: index build, point x polygon join, bbox query.

Mirrors the reference's physical operators on this path:
  build_flat_spatial_index      src/vibespatial/spatial/indexing.py:1357-1533
  query_spatial_index           src/vibespatial/spatial/query.py:247-1042
  spatial_index_device_query    src/vibespatial/spatial/spatial_index_device.py:228-390
  classify_point_predicates_indexed_device
                                src/vibespatial/predicates/point_relations.py:1040-1191

Outputs are int64 device tensors (left, right), lexsorted.  The one host
synchronisation per join is the read of the 4-word result header (pair count
and capacity check), the analogue of the reference's count_scatter_total fence
(src/vibespatial/cuda/_runtime.py:2015-2038).  When the caller's capacity is
too small the join is re-run once with the exact capacity, which the
reference's CandidateRelationCapacityError path (spatial/query_types.py:41-93)
turns into an error instead.
"""

from __future__ import annotations

import ctypes
import dataclasses
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import PRED, TreeLayout, check, lib
from .device import DevicePoints, DevicePolygons, ptr, require_gpu, stream

__all__ = ["DeviceTree", "JoinStats", "build_tree", "polygon_bounds", "point_boxes",
           "sjoin_points_polygons", "query_boxes"]

FANOUT = 16


@dataclass
class JoinStats:
    pairs: int
    staged: int
    candidates: int
    capacity: int
    retries: int
    candidate_vertices: int = 0


@dataclass
class BatchStats(JoinStats):
    refined: int = 0  # candidates left to the exact refine after the cell-grid decisions


class DeviceTree:
    """Packed Hilbert tree over item boxes (root first, items last)."""

    def __init__(self, layout: TreeLayout, node_boxes: torch.Tensor, item_ids: torch.Tensor,
                 item_keys: torch.Tensor, item_bounds: torch.Tensor):
        self.layout = layout
        self.node_boxes = node_boxes
        self.item_ids = item_ids
        self.item_keys = item_keys
        self.item_bounds = item_bounds
        # staged pairs per query box of the last box query of each kind
        # ("bbox", "dwithin", ...) over this tree that had no caller
        # capacity: sizes the next such query's output (an undersized one
        # traverses twice)
        self.hits_per_query = {}

    @property
    def n_items(self) -> int:
        return int(self.layout.n_items)


def _ws(nbytes: int, dev) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=dev)


def polygon_bounds(P: DevicePolygons) -> torch.Tensor:
    dev = require_gpu()
    out = torch.empty((len(P), 4), dtype=torch.float64, device=dev)
    check(lib().vs_polygon_bounds(*P.args(), len(P), ptr(out), stream()), "vs_polygon_bounds")
    return out


def point_boxes(Q: DevicePoints) -> torch.Tensor:
    dev = require_gpu()
    out = torch.empty((len(Q), 4), dtype=torch.float64, device=dev)
    check(lib().vs_point_boxes(ptr(Q.x), ptr(Q.y), len(Q), ptr(out), stream()), "vs_point_boxes")
    return out


def build_tree(bounds: torch.Tensor, fanout: int = FANOUT) -> DeviceTree:
    """vs_index_build over (n, 4) fp64 device boxes."""
    dev = require_gpu()
    L = lib()
    n = int(bounds.shape[0])
    layout = TreeLayout()
    check(L.vs_tree_plan(n, fanout, ctypes.byref(layout)), "vs_tree_plan")
    node_boxes = torch.empty(max(int(layout.n_nodes), 1) * 4, dtype=torch.float64, device=dev)
    item_ids = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    item_keys = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
    ws = _ws(L.vs_index_workspace_bytes(n), dev)
    check(
        L.vs_index_build(ptr(bounds), n, ctypes.byref(layout), ptr(node_boxes), ptr(item_ids), ptr(item_keys),
                         ptr(ws), ws.numel(), stream()),
        "vs_index_build",
    )
    return DeviceTree(layout, node_boxes, item_ids, item_keys, bounds)


def expand_boxes(boxes: torch.Tensor, distance: torch.Tensor) -> torch.Tensor:
    """(n, 4) device boxes grown by a per-row fp64 distance on every side, each
    bound rounded one ulp outward (vs_expand_boxes); NaN / negative distance
    gives a NaN box."""
    dev = require_gpu()
    n = int(boxes.shape[0])
    d = distance.to(device=dev, dtype=torch.float64).contiguous()
    if d.numel() != n:
        raise ValueError("one distance per box")
    out = torch.empty_like(boxes)
    check(lib().vs_expand_boxes(ptr(boxes), n, ptr(d), ptr(out), stream()), "vs_expand_boxes")
    return out


def refine_dwithin(left: torch.Tensor, right: torch.Tensor, distance: torch.Tensor, lgeom, rgeom):
    """Exact dwithin of candidate pairs (vs_refine_dwithin); lgeom / rgeom are
    DevicePoints or DevicePolygons; distance is per left row.  Keeps order."""
    dev = require_gpu()
    L = lib()
    n = int(left.numel())
    if n == 0:
        e = torch.empty(0, dtype=torch.int64, device=dev)
        return e, e.clone()

    lk, la = _kind_args(lgeom)
    rk, ra = _kind_args(rgeom)
    d = distance.to(device=dev, dtype=torch.float64).contiguous()
    out_l = torch.empty(n, dtype=torch.int64, device=dev)
    out_r = torch.empty(n, dtype=torch.int64, device=dev)
    res = torch.zeros(4, dtype=torch.int64, device=dev)
    ws = _ws(L.vs_refine_dwithin_workspace_bytes(n), dev)
    check(L.vs_refine_dwithin(lk, rk, ptr(left.contiguous()), ptr(right.contiguous()), n, len(rgeom), ptr(d),
                              *la, *ra, ptr(out_l), ptr(out_r), ptr(res), ptr(ws), ws.numel(), stream()),
          "vs_refine_dwithin")
    k = int(res[0].item())
    return out_l[:k], out_r[:k]


def lexsort_pairs(left: torch.Tensor, right: torch.Tensor):
    """(left, right) int64 pairs sorted by left, then right (vs_lexsort_pairs,
    in place on copies)."""
    dev = require_gpu()
    L = lib()
    l, r = left.clone(), right.clone()
    n = int(l.numel())
    if n > 1:
        ws = _ws(L.vs_lexsort_pairs_workspace_bytes(n), dev)
        check(L.vs_lexsort_pairs(ptr(l), ptr(r), n, ptr(ws), ws.numel(), stream()), "vs_lexsort_pairs")
    return l, r


def _kind_args(g):
    if isinstance(g, DevicePoints):
        return 0, [ptr(g.x), ptr(g.y), ptr(None), ptr(None), ptr(None)]
    return 1, g.args()


def refine_polygon_pairs(predicate: str, left: torch.Tensor, right: torch.Tensor, A: DevicePolygons,
                         B: DevicePolygons):
    """The (left, right) pairs with predicate(A row, B row), exact, in input
    order (vs_refine_pairs mode 1)."""
    dev = require_gpu()
    L = lib()
    n = int(left.numel())
    if n == 0:
        return left[:0], right[:0]
    out_l = torch.empty(n, dtype=torch.int64, device=dev)
    out_r = torch.empty(n, dtype=torch.int64, device=dev)
    res = torch.zeros(4, dtype=torch.int64, device=dev)
    ws = _ws(L.vs_refine_pairs_workspace_bytes(n), dev)
    none = ctypes.c_void_p(0)
    check(L.vs_refine_pairs(1, PRED[predicate], ptr(left.contiguous()), ptr(right.contiguous()), n, none, none,
                            *A.args(), *B.args(), ptr(out_l), ptr(out_r), ptr(res), ptr(ws), ws.numel(), stream()),
          "vs_refine_pairs")
    k = int(res[0].item())
    return out_l[:k], out_r[:k]


def pair_distances(left: torch.Tensor, right: torch.Tensor, lgeom, rgeom, exclusive: bool = False,
                   outside: bool = False, cap: torch.Tensor | None = None):
    """fp64 distance of each (left, right) pair (vs_pair_distances); +inf for
    empty / NaN geometries and, with `exclusive`, for equal geometries: point
    pairs with equal coordinates, and polygon pairs that are topologically
    equal (each covers the other, exact; a pair at distance 0 is tested).
    `outside`: every left point is known to lie outside its right polygons.
    `cap` (fp64 per left row): pairs farther than cap[left] get +inf."""
    dev = require_gpu()
    L = lib()
    n = int(left.numel())
    out = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
    if n:
        lk, la = _kind_args(lgeom)
        rk, ra = _kind_args(rgeom)
        flags = (1 if exclusive else 0) | (2 if outside else 0)
        ws = _ws(L.vs_pair_distances_workspace_bytes(n), dev)
        check(L.vs_pair_distances(lk, rk, ptr(left.contiguous()), ptr(right.contiguous()), n, len(rgeom), flags,
                                  ptr(cap), *la, *ra, ptr(out), ptr(ws), ws.numel(), stream()),
              "vs_pair_distances")
        if exclusive and isinstance(lgeom, DevicePolygons) and isinstance(rgeom, DevicePolygons):
            z = torch.nonzero(out[:n] == 0).flatten()
            if z.numel():
                lz, rz = left[z], right[z]
                l1, r1 = refine_polygon_pairs("covers", lz, rz, lgeom, rgeom)
                l2, r2 = refine_polygon_pairs("covered_by", l1, r1, lgeom, rgeom)
                if l2.numel():
                    nr = max(len(rgeom), 1)
                    eq = torch.isin(lz * nr + rz, l2 * nr + r2)
                    out[z[eq]] = float("inf")
    return out[:n]


def nearest_pairs(qboxes: torch.Tensor, qgeom, tree: DeviceTree, tgeom, *, max_distance=None, return_all=True,
                  exclusive=False, start_radius: float, cover_radius: float, growth: float = 4.0):
    """Lexsorted (left, right, distance) of each query's nearest tree rows
    (ties kept unless not `return_all`), by an expanding-radius search.
    Points against polygons first take a radius-0 round: the point join with
    predicate intersects (learned form) gives every pair at distance 0.  Then each round
    takes the open queries (vs_nearest_open: rows and boxes grown by their
    search radius) -> box query -> rows mapped back (vs_map_left) -> pair
    distances -> vs_nearest_select (closes queries, grows the others' radius).
    `cover_radius` is a radius at which a grown query box covers every tree
    box; without `max_distance` the last round uses it."""
    dev = require_gpu()
    L = lib()
    nq = int(qboxes.shape[0])
    final = float(max_distance) if max_distance is not None else float(cover_radius)
    r0 = min(float(start_radius), final)
    if not (r0 > 0):
        r0 = final
    pip = isinstance(qgeom, DevicePoints) and isinstance(tgeom, DevicePolygons)
    qboxes = qboxes.contiguous()
    radius = torch.empty(max(nq, 1), dtype=torch.float64, device=dev)
    search = torch.empty(max(nq, 1), dtype=torch.float64, device=dev)
    # empty / NaN query rows are closed from the start
    check(L.vs_nearest_init(ptr(qboxes), nq, 0.0 if pip else r0, ptr(radius), ptr(search), stream()),
          "vs_nearest_init")
    outs_l, outs_r = [], []
    res = torch.zeros(2, dtype=torch.int64, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    open_rows = torch.empty(max(nq, 1), dtype=torch.int64, device=dev)
    open_boxes = torch.empty((max(nq, 1), 4), dtype=torch.float64, device=dev)
    open_ws = _ws(L.vs_nearest_open_workspace_bytes(nq), dev)

    outs_d = []

    def select(l, r, d):
        n = int(l.numel())
        ol = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
        orr = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
        od = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
        ws = _ws(L.vs_nearest_select_workspace_bytes(n, nq), dev)
        check(L.vs_nearest_select(ptr(l), ptr(r), ptr(d), n, nq, ptr(radius), ptr(search), float(growth), r0, final,
                                  final, int(bool(return_all)), ptr(ol), ptr(orr), ptr(od), ptr(res), ptr(ws),
                                  ws.numel(), stream()), "vs_nearest_select")
        k, still_open = (int(v) for v in res.cpu().tolist())
        outs_l.append(ol[:k])
        outs_r.append(orr[:k])
        outs_d.append(od[:k])
        return still_open

    def open_queries():
        check(L.vs_nearest_open(ptr(qboxes), ptr(search), nq, ptr(open_rows), ptr(open_boxes), ptr(cnt),
                                ptr(open_ws), open_ws.numel(), stream()), "vs_nearest_open")
        return int(cnt.item())

    still_open = nq
    if pip:
        # radius 0: the pairs with the point not outside the polygon (the
        # learned form of the point join: the same pairs, bit for bit)
        l, r, _ = sjoin_points_polygons_learned(qgeom, tgeom, polygon_bounds(tgeom), "intersects")
        still_open = select(l, r, torch.zeros(max(int(l.numel()), 1), dtype=torch.float64, device=dev))
    if still_open:
        # each open query's radius from a descent of the tree (an upper bound
        # of its nearest distance), so the next round closes nearly all
        m = open_queries()
        rep = torch.empty((max(tree.n_items, 1), 2), dtype=torch.float64, device=dev)
        tk, ta = _kind_args(tgeom)
        check(L.vs_nearest_rep_points(tk, ptr(tree.item_ids), tree.n_items, *ta, ptr(rep), stream()),
              "vs_nearest_rep_points")
        bws = _ws(L.vs_nearest_bound_workspace_bytes(m), dev)
        check(L.vs_nearest_bound(ptr(qboxes), ptr(open_rows), m, ptr(open_boxes), ctypes.byref(tree.layout),
                                 ptr(tree.node_boxes), ptr(rep), tk, ptr(tree.item_ids), *ta, final, ptr(radius),
                                 ptr(search), ptr(bws), bws.numel(), stream()), "vs_nearest_bound")
    # candidate pairs per open query: a box query whose output overflows its
    # capacity traverses the tree again, so the capacity starts generous and
    # follows what the previous round found
    per_query = 16.0
    for _ in range(200):
        if still_open == 0:
            break
        m = open_queries()
        l, r, qs = query_boxes(open_boxes[:m], tree, capacity=max(1024, int(per_query * m)))
        per_query = max(4.0, 1.25 * qs.staged / max(m, 1))
        check(L.vs_map_left(ptr(open_rows), ptr(l), int(l.numel()), stream()), "vs_map_left")
        # only distances up to the query's radius decide this round; after
        # the radius-0 round the open points lie outside every polygon
        still_open = select(l, r, pair_distances(l, r, qgeom, tgeom, exclusive, outside=pip, cap=radius))
    else:
        raise RuntimeError("nearest search did not close every query")
    lo = torch.cat(outs_l) if outs_l else torch.empty(0, dtype=torch.int64, device=dev)
    ro = torch.cat(outs_r) if outs_r else torch.empty(0, dtype=torch.int64, device=dev)
    do = torch.cat(outs_d) if outs_d else torch.empty(0, dtype=torch.float64, device=dev)
    n = int(lo.numel())
    fl = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
    fr = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
    fd = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
    ws = _ws(L.vs_nearest_finish_workspace_bytes(n), dev)
    check(L.vs_nearest_finish(ptr(lo), ptr(ro), ptr(do), n, nq, ptr(fl), ptr(fr), ptr(fd), ptr(ws), ws.numel(),
                              stream()), "vs_nearest_finish")
    return fl[:n], fr[:n], fd[:n]


def _read_result(res: torch.Tensor):
    h = res.cpu().tolist()  # synchronises the current stream
    return int(h[0]), int(h[1]), int(h[2]), int(h[3])


def sjoin_points_polygons(Q: DevicePoints, P: DevicePolygons, tree: DeviceTree, predicate: str = "within",
                          capacity: int | None = None):
    """(left, right, stats): lexsorted int64 pairs with predicate(point, polygon)."""
    dev = require_gpu()
    if predicate not in ("within", "intersects", "covered_by", "touches"):
        raise NotImplementedError(f"point x polygon predicate {predicate!r}")
    if tree.n_items != len(P):
        raise ValueError("tree was not built over these polygons")
    L = lib()
    n = len(Q)
    cap = int(capacity) if capacity is not None else max(1024, 2 * n)
    retries = 0
    while True:
        ws = _ws(L.vs_sjoin_points_polygons_workspace_bytes(n, len(P), P.n_coords, cap), dev)
        out_l = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
        out_r = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
        res = torch.zeros(4, dtype=torch.int64, device=dev)
        check(
            L.vs_sjoin_points_polygons(
                ptr(Q.x), ptr(Q.y), n, *P.args(), len(P), P.n_parts, P.n_rings, P.n_coords,
                ctypes.byref(tree.layout), ptr(tree.node_boxes),
                ptr(tree.item_ids), PRED[predicate], cap, ptr(out_l), ptr(out_r), ptr(res), ptr(ws),
                ws.numel(), stream(),
            ),
            "vs_sjoin_points_polygons",
        )
        k, staged, cand, cverts = _read_result(res)
        if staged <= cap:
            return out_l[:k], out_r[:k], JoinStats(k, staged, cand, cap, retries, cverts)
        if staged >= 2**31 - 1:
            raise OverflowError(f"{staged} pairs exceed the int32 pair limit")
        cap = staged
        retries += 1


PAIR_LIMIT = 2**31 - 2  # pairs one device call can hold (int32 slots)


class _PairLimit(Exception):
    pass


def _split_boxes(run, boxes: torch.Tensor):
    """run(boxes) -> (left, right, stats) or raises _PairLimit; splits the
    query boxes in halves until every call fits PAIR_LIMIT and concatenates
    (left ids offset by each half's start, so grouping and order carry over)."""
    try:
        return run(boxes)
    except _PairLimit:
        n = int(boxes.shape[0])
        if n <= 1:
            raise OverflowError("one query box has more hits than a device call can hold (2^31 - 2)") from None
        h = n // 2
        l0, r0, s0 = _split_boxes(run, boxes[:h])
        l1, r1, s1 = _split_boxes(run, boxes[h:])
        for f in dataclasses.fields(s0):  # counters add up; the model's error bound is a max
            a, b = getattr(s0, f.name), getattr(s1, f.name)
            setattr(s0, f.name, max(a, b) if f.name == "mlp_leaf_error" else a + b)
        return torch.cat([l0, l1 + h]), torch.cat([r0, r1]), s0


# A box query sized from the tree's hint never starts above this many slots
# per query box (at least 1024 in all): a hint from queries with many hits per
# box must not size a large query of few hits (the retry covers the rest).
HINT_SLOTS_PER_QUERY = 32


def query_boxes(qboxes: torch.Tensor, tree: DeviceTree, capacity: int | None = None, sort: bool = True,
                kind: str = "bbox"):
    """(left, right, stats): (query, item) pairs whose boxes overlap; lexsorted
    when `sort`, else grouped by query with items in unspecified order.  Query
    sets whose hits exceed one call's int32 slot limit are split.  Without a
    `capacity` the output is sized from the hits per box of the tree's last
    query of the same `kind`, clamped to HINT_SLOTS_PER_QUERY per box."""
    dev = require_gpu()
    L = lib()

    def run(qb):
        n = int(qb.shape[0])
        if capacity is not None:
            cap = int(capacity)
        else:
            per_query = min(max(2.0, 1.25 * tree.hits_per_query.get(kind, 0.0)), float(HINT_SLOTS_PER_QUERY))
            cap = max(1024, int(per_query * n))
        cap = min(cap, PAIR_LIMIT)
        retries = 0
        while True:
            ws = _ws(L.vs_query_boxes_workspace_bytes(n, cap), dev)
            out_l = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
            out_r = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
            res = torch.zeros(4, dtype=torch.int64, device=dev)
            check(
                L.vs_query_boxes(ptr(qb), n, ctypes.byref(tree.layout), ptr(tree.node_boxes), ptr(tree.item_ids),
                                 cap, int(bool(sort)), ptr(out_l), ptr(out_r), ptr(res), ptr(ws), ws.numel(),
                                 stream()),
                "vs_query_boxes",
            )
            k, staged, cand, cverts = _read_result(res)
            if n and capacity is None:
                tree.hits_per_query[kind] = staged / n
            if staged <= cap:
                return out_l[:k], out_r[:k], JoinStats(k, staged, cand, cap, retries, cverts)
            if staged > PAIR_LIMIT:
                raise _PairLimit()
            del ws, out_l, out_r
            cap = staged
            retries += 1

    return _split_boxes(run, qboxes)


_ = _lib  # re-exported module for callers that need the raw binding


# ---------------------------------------------------------------------------
# learned + FP32 form
# ---------------------------------------------------------------------------
@dataclass
class LearnedStats(JoinStats):
    orient_tests: int = 0
    escalated_edges: int = 0
    escalated_pairs: int = 0
    lookup_probes: int = 0
    mlp_fallbacks: int = 0
    window_fallbacks: int = 0
    scanned: int = 0
    mlp_leaf_error: int = 0

    @property
    def escalation_fraction(self) -> float:
        return self.escalated_edges / max(1, self.orient_tests + self.escalated_edges)


def _learned_stats(h, cap, retries) -> LearnedStats:
    """LearnedStats from the learned join's d_result words (vsjoin.h)."""
    return LearnedStats(pairs=int(h[0]), staged=int(h[1]), candidates=int(h[2]), capacity=cap, retries=retries,
                        candidate_vertices=int(h[3]), orient_tests=int(h[4]), escalated_edges=int(h[5]),
                        escalated_pairs=int(h[6]), lookup_probes=int(h[7]), mlp_fallbacks=int(h[8]),
                        window_fallbacks=int(h[9]), scanned=int(h[10]), mlp_leaf_error=int(h[11]))


LOOKUP_MODES = {"learned": 0, "leaves": 1, "binary": 2}


def sjoin_points_polygons_learned(Q: DevicePoints, P: DevicePolygons, bounds: torch.Tensor, predicate: str = "within",
                                  capacity: int | None = None, train_iters: int = 16, lookup: str = "learned"):
    """Learned + FP32 form of sjoin_points_polygons: same pairs, bit for bit."""
    dev = require_gpu()
    if predicate not in ("within", "intersects", "covered_by", "touches"):
        raise NotImplementedError(f"point x polygon predicate {predicate!r}")
    L = lib()
    n = len(Q)
    cap = int(capacity) if capacity is not None else max(1024, 2 * n)
    retries = 0
    while True:
        ws = _ws(L.vs_sjoin_points_polygons_learned_workspace_bytes(n, len(P), P.n_coords, cap), dev)
        out_l = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
        out_r = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
        res = torch.zeros(16, dtype=torch.int64, device=dev)
        check(
            L.vs_sjoin_points_polygons_learned(
                ptr(Q.x), ptr(Q.y), n, *P.args(), len(P), P.n_parts, P.n_rings, P.n_coords, ptr(bounds),
                PRED[predicate], train_iters, LOOKUP_MODES[lookup], cap, ptr(out_l), ptr(out_r), ptr(res), ptr(ws),
                ws.numel(), stream(),
            ),
            "vs_sjoin_points_polygons_learned",
        )
        h = res.cpu().tolist()
        if h[1] <= cap:
            return out_l[: h[0]], out_r[: h[0]], _learned_stats(h, cap, retries)
        if h[1] >= 2**31 - 1:
            raise OverflowError(f"{h[1]} pairs exceed the int32 pair limit")
        cap = int(h[1])
        retries += 1


class LearnedPointIndex:
    """Learned rank index over Hilbert-sorted points (device buffers)."""

    NAMES = ("keys", "perm", "sx", "sy", "extent", "mlp", "losses", "leaf_keys", "leaf_err", "meta", "leaf_win")

    def __init__(self, Q: DevicePoints, key_bytes: int = 8, leaf_size: int = 512, train_iters: int = 256,
                 lr: float = 3e-4):
        dev = require_gpu()
        L = lib()
        self.n = len(Q)
        self.key_bytes = key_bytes
        self.n_leaves = int(L.vs_learned_leaf_count(self.n, leaf_size))
        nbytes = int(L.vs_learned_index_bytes(self.n, key_bytes, self.n_leaves))
        self.buf = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=dev)
        check(L.vs_learned_index_build(ptr(Q.x), ptr(Q.y), self.n, key_bytes, self.n_leaves, train_iters, lr,
                                       ptr(self.buf), self.buf.numel(), stream()), "vs_learned_index_build")
        off = (ctypes.c_int64 * 11)()
        check(L.vs_learned_index_layout(self.n, key_bytes, self.n_leaves, off), "vs_learned_index_layout")
        self.offsets = dict(zip(self.NAMES, list(off)))

    def _view(self, name, dtype, count):
        o = self.offsets[name]
        itemsize = torch.empty(0, dtype=dtype).element_size()
        return self.buf[o: o + count * itemsize].view(dtype)

    @property
    def keys(self) -> torch.Tensor:
        return self._view("keys", torch.int64 if self.key_bytes == 8 else torch.int32, self.n)

    @property
    def perm(self) -> torch.Tensor:
        return self._view("perm", torch.int32, self.n)

    @property
    def mlp_params(self) -> torch.Tensor:
        return self._view("mlp", torch.float32, 97)

    @property
    def losses(self) -> torch.Tensor:
        return self._view("losses", torch.float32, 2)

    @property
    def mlp_leaf_error(self) -> int:
        return int(self._view("meta", torch.int32, 4)[0].item())

    @property
    def leaf_err(self) -> torch.Tensor:
        return self._view("leaf_err", torch.int32, self.n_leaves)

    @property
    def leaf_win(self) -> torch.Tensor:
        """(n_leaves, 2) int32: per predicted leaf, the first and last true
        leaf of the keys the network routes there (first > last: none)."""
        return self._view("leaf_win", torch.int32, 2 * self.n_leaves).view(self.n_leaves, 2)

    def lower_bound(self, qkeys: torch.Tensor, lookup: str = "learned"):
        dev = require_gpu()
        m = int(qkeys.numel())
        out = torch.empty(max(m, 1), dtype=torch.int64, device=dev)
        st = torch.zeros(3, dtype=torch.int64, device=dev)
        check(lib().vs_learned_lower_bound(ptr(self.buf), self.n, self.key_bytes, self.n_leaves, ptr(qkeys), m,
                                           LOOKUP_MODES[lookup], ptr(out), ptr(st), stream()), "vs_learned_lower_bound")
        s = st.cpu().tolist()
        return out[:m], {"probes": s[0], "mlp_fallbacks": s[1], "window_fallbacks": s[2]}

    def query_boxes(self, boxes: torch.Tensor, lookup: str = "learned", capacity: int | None = None,
                    sort: bool = True):
        """(box, point) pairs with the point in the box (inclusive); lexsorted
        when `sort`, else grouped by box with points in unspecified order."""
        dev = require_gpu()
        L = lib()

        def run(bx):
            nb = int(bx.shape[0])
            cap = int(capacity) if capacity is not None else max(1024, 4 * nb)
            cap = min(cap, PAIR_LIMIT)
            retries = 0
            while True:
                ws = _ws(L.vs_learned_query_boxes_workspace_bytes(nb, cap), dev)
                out_l = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
                out_r = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
                res = torch.zeros(16, dtype=torch.int64, device=dev)
                check(L.vs_learned_query_boxes(ptr(self.buf), self.n, self.key_bytes, self.n_leaves, ptr(bx), nb,
                                               LOOKUP_MODES[lookup], cap, int(bool(sort)), ptr(out_l), ptr(out_r),
                                               ptr(res), ptr(ws), ws.numel(), stream()), "vs_learned_query_boxes")
                h = res.cpu().tolist()
                if h[1] <= cap:
                    return out_l[: h[0]], out_r[: h[0]], _learned_stats(h, cap, retries)
                if h[1] > PAIR_LIMIT:
                    raise _PairLimit()
                del ws, out_l, out_r
                cap = int(h[1])
                retries += 1

        return _split_boxes(run, boxes)

    def query_polygons(self, P: DevicePolygons, bounds: torch.Tensor, predicate: str, lookup: str = "learned",
                       capacity: int | None = None, sort: bool = True):
        """(polygon row, point) pairs with predicate(polygon, point) exact
        (contains / contains_properly: interior, touches: boundary, covers /
        intersects: not outside), through the index and the FP32 refine with
        exact escalation (vs_learned_query_polygons); lexsorted when `sort`."""
        dev = require_gpu()
        L = lib()
        if predicate not in ("contains", "contains_properly", "covers", "intersects", "touches"):
            raise NotImplementedError(f"polygon x point predicate {predicate!r}")
        nq = len(P)
        cap = int(capacity) if capacity is not None else max(1024, 4 * nq)
        cap = min(cap, PAIR_LIMIT)
        retries = 0
        while True:
            ws = _ws(L.vs_learned_query_polygons_workspace_bytes(nq, P.n_coords, cap), dev)
            out_l = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
            out_r = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
            res = torch.zeros(16, dtype=torch.int64, device=dev)
            check(L.vs_learned_query_polygons(ptr(self.buf), self.n, self.key_bytes, self.n_leaves, *P.args(), nq,
                                              P.n_parts, P.n_rings, P.n_coords, ptr(bounds), PRED[predicate],
                                              LOOKUP_MODES[lookup], cap, int(bool(sort)), ptr(out_l), ptr(out_r),
                                              ptr(res), ptr(ws), ws.numel(), stream()), "vs_learned_query_polygons")
            h = res.cpu().tolist()
            if h[1] <= cap:
                return out_l[: h[0]], out_r[: h[0]], _learned_stats(h, cap, retries)
            if h[1] > PAIR_LIMIT:
                raise OverflowError(f"{h[1]} pairs exceed the int32 pair limit of one call")
            del ws, out_l, out_r
            cap = int(h[1])
            retries += 1


def mlp_loss_grad(params: torch.Tensor, u: torch.Tensor, t: torch.Tensor):
    """(loss, grad[97]) of the rank MLP on (u, t), from the device kernels."""
    dev = require_gpu()
    g = torch.empty(97, dtype=torch.float32, device=dev)
    loss = torch.empty(1, dtype=torch.float32, device=dev)
    check(lib().vs_mlp_loss_grad(ptr(params), ptr(u), ptr(t), int(u.numel()), ptr(g), ptr(loss), stream()),
          "vs_mlp_loss_grad")
    return float(loss.item()), g


# ---------------------------------------------------------------------------
# batched polygon x polygon intersects
# ---------------------------------------------------------------------------
def _encode_tile(waves_per_block: int, refine_threads: int, order: int = 0, slab: int = 0) -> int:
    # slab: y-slab index for right rings of 128 (0), 16 (1) or 64 (2) edges and more, or none (3)
    return int(waves_per_block) | (int(refine_threads) << 8) | (int(order) << 24) | (int(slab) << 26)


# (traversal waves per block, refine block size, candidate order: 0 by right
# row, 1 by the right rows' coordinate-count bins, largest first)
TILE_CANDIDATES = [(w, t, o) for o in (0, 1) for w in (1, 2, 4) for t in (64, 128, 256)]


def _tile_key(n_jobs, n_left, n_right):
    # shape class: powers of two of the sizes (an autotune result carries over
    # to batches of similar shape)
    return (int(n_jobs).bit_length(), int(n_left).bit_length(), int(n_right).bit_length())


class TileCache:
    """Autotuned batched-join tiles per (device name, shape class), kept in
    memory and in a JSON file that later processes load (like the
    reference's on-disk kernel cache, cuda/_runtime.py:457-550).  The file is
    $VIBESPATIAL_AMD_CACHE_DIR/batched_tiles.json (default
    ~/.cache/vibespatial_amd); it is rewritten whole through a temporary file
    and an atomic rename, and an unreadable file counts as empty."""

    FILE = "batched_tiles.json"

    def __init__(self):
        self._mem: dict = {}
        self._loaded_from = None

    @staticmethod
    def path() -> str:
        import os

        d = os.environ.get("VIBESPATIAL_AMD_CACHE_DIR") or os.path.join(os.path.expanduser("~"), ".cache",
                                                                         "vibespatial_amd")
        return os.path.join(d, TileCache.FILE)

    def _load(self):
        import json

        p = self.path()
        if self._loaded_from == p:
            return
        self._loaded_from = p
        try:
            with open(p) as f:
                data = json.load(f)
        except (OSError, ValueError):
            return
        for k, v in data.items() if isinstance(data, dict) else ():
            try:
                dev, a, b, c = k.split("|")
                tile = tuple(int(t) for t in v)
            except (ValueError, TypeError, IndexError):
                continue
            if tile in TILE_CANDIDATES:
                self._mem.setdefault((dev, int(a), int(b), int(c)), tile)

    def get(self, device: str, key, default=None):
        self._load()
        return self._mem.get((device, *key), default)

    def put(self, device: str, key, tile) -> None:
        import json
        import os
        import tempfile

        self._load()
        self._mem[(device, *key)] = tuple(int(t) for t in tile)
        p = self.path()
        try:
            os.makedirs(os.path.dirname(p), exist_ok=True)
            data = {"|".join(str(x) for x in k): list(v) for k, v in self._mem.items()}
            fd, tmp = tempfile.mkstemp(dir=os.path.dirname(p), prefix=".tiles.")
            with os.fdopen(fd, "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)
            os.replace(tmp, p)
        except OSError:
            pass  # the in-memory entry still serves this process


_TILE_CACHE = TileCache()


def _device_name() -> str:
    try:
        return torch.cuda.get_device_name(torch.cuda.current_device())
    except Exception:  # pragma: no cover
        return "unknown"


def batched_intersects(A: DevicePolygons, B: DevicePolygons, left_job: torch.Tensor, right_job: torch.Tensor,
                       capacity: int | None = None, tile: tuple[int, int] | None = None, cell_grid: bool = True):
    """Many independent polygon x polygon sjoin(intersects) jobs in one launch
    sequence.  left_job / right_job: device int64 row offsets (n_jobs + 1).
    Returns (left, right, stats) with global row ids, lexsorted; job j's pairs
    are the contiguous run with left in [left_job[j], left_job[j+1]).
    cell_grid=False refines every bbox candidate exactly (same pairs)."""
    dev = require_gpu()
    L = lib()
    n_jobs = int(left_job.numel()) - 1
    if int(right_job.numel()) - 1 != n_jobs:
        raise ValueError("left_job and right_job must have the same number of jobs")
    lj = left_job.to(device=dev, dtype=torch.int64).contiguous()
    rj = right_job.to(device=dev, dtype=torch.int64).contiguous()
    if tile is None:
        tile = _TILE_CACHE.get(_device_name(), _tile_key(n_jobs, len(A), len(B)), (4, 256, 1))
    code = _encode_tile(*tile) | (0 if cell_grid else 1 << 25)
    cap = int(capacity) if capacity is not None else max(1024, 4 * len(A))
    cap = min(cap, PAIR_LIMIT)
    retries = 0
    while True:
        ws = _ws(L.vs_batched_intersects_workspace_bytes(n_jobs, len(A), len(B), B.n_coords, cap), dev)
        out_l = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
        out_r = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
        res = torch.zeros(5, dtype=torch.int64, device=dev)
        check(L.vs_batched_intersects(ptr(A.x), ptr(A.y), ptr(A.geom_offsets), ptr(A.part_offsets),
                                      ptr(A.ring_offsets), len(A), ptr(B.x), ptr(B.y), ptr(B.geom_offsets),
                                      ptr(B.part_offsets), ptr(B.ring_offsets), len(B), B.n_coords, ptr(lj), ptr(rj),
                                      n_jobs, code, cap, ptr(out_l), ptr(out_r), ptr(res), ptr(ws), ws.numel(),
                                      stream()),
              "vs_batched_intersects")
        k, staged, cand, cverts, refined = res.cpu().tolist()
        if staged <= cap:
            return out_l[:k], out_r[:k], BatchStats(k, staged, cand, cap, retries, cverts, refined)
        if staged > PAIR_LIMIT:
            raise OverflowError(f"{staged} bbox candidates exceed the int32 slot limit of one call")
        del ws, out_l, out_r
        cap = staged
        retries += 1


def autotune_batched(A: DevicePolygons, B: DevicePolygons, left_job: torch.Tensor, right_job: torch.Tensor,
                     reps: int = 2, candidates=None):
    """Times every tile shape on this batch (HIP events on the current stream)
    and caches the fastest for batches of the same shape class on this device
    (TileCache: in memory and on disk, so later processes start from it).
    Returns (best tile, {tile: ms})."""
    require_gpu()
    cands = list(candidates or TILE_CANDIDATES)
    _, _, st = batched_intersects(A, B, left_job, right_job, tile=cands[0])
    cap = st.staged + 1024
    times = {}
    for t in cands:
        batched_intersects(A, B, left_job, right_job, capacity=cap, tile=t)  # warm
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            batched_intersects(A, B, left_job, right_job, capacity=cap, tile=t)
        e1.record()
        e1.synchronize()
        times[t] = e0.elapsed_time(e1) / reps
    best = min(times, key=times.get)
    _TILE_CACHE.put(_device_name(), _tile_key(int(left_job.numel()) - 1, len(A), len(B)), best)
    return best, times


class BatchedIntersectsGraph:
    """The batched join captured once in a HIP graph for fixed device inputs
    (same tensors, same sizes): replay() re-runs every launch of the sequence
    with one graph launch.  `capacity` bounds the bbox candidates; a replay
    that overflows raises (capture again with a larger capacity)."""

    def __init__(self, A: DevicePolygons, B: DevicePolygons, left_job: torch.Tensor, right_job: torch.Tensor,
                 capacity: int, tile: tuple[int, int] | None = None, cell_grid: bool = True):
        dev = require_gpu()
        L = lib()
        self.n_jobs = int(left_job.numel()) - 1
        self.lj = left_job.to(device=dev, dtype=torch.int64).contiguous()
        self.rj = right_job.to(device=dev, dtype=torch.int64).contiguous()
        self.A, self.B = A, B
        self.cap = min(int(capacity), PAIR_LIMIT)
        if tile is None:
            tile = _TILE_CACHE.get(_device_name(), _tile_key(self.n_jobs, len(A), len(B)), (4, 256, 1))
        self.code = _encode_tile(*tile) | (0 if cell_grid else 1 << 25)
        self.ws = _ws(L.vs_batched_intersects_workspace_bytes(self.n_jobs, len(A), len(B), B.n_coords, self.cap),
                      dev)
        self.out_l = torch.empty(max(self.cap, 1), dtype=torch.int64, device=dev)
        self.out_r = torch.empty(max(self.cap, 1), dtype=torch.int64, device=dev)
        self.res = torch.zeros(5, dtype=torch.int64, device=dev)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self._launch()  # eager warm-up on the capture stream
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._launch()
        torch.cuda.synchronize()

    def _launch(self):
        A, B, L = self.A, self.B, lib()
        check(L.vs_batched_intersects(ptr(A.x), ptr(A.y), ptr(A.geom_offsets), ptr(A.part_offsets),
                                      ptr(A.ring_offsets), len(A), ptr(B.x), ptr(B.y), ptr(B.geom_offsets),
                                      ptr(B.part_offsets), ptr(B.ring_offsets), len(B), B.n_coords, ptr(self.lj),
                                      ptr(self.rj), self.n_jobs, self.code, self.cap, ptr(self.out_l), ptr(self.out_r),
                                      ptr(self.res), ptr(self.ws), self.ws.numel(), stream()),
              "vs_batched_intersects")

    def replay(self):
        """(left, right, stats) views of the graph's output buffers."""
        self.graph.replay()
        k, staged, cand, cverts, refined = self.res.cpu().tolist()
        if staged > self.cap:
            raise OverflowError(f"{staged} bbox candidates exceed the captured capacity {self.cap}")
        return self.out_l[:k], self.out_r[:k], BatchStats(k, staged, cand, self.cap, 0, cverts, refined)
