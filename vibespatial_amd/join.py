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


class DeviceTree:
    """Packed Hilbert tree over item boxes (root first, items last)."""

    def __init__(self, layout: TreeLayout, node_boxes: torch.Tensor, item_ids: torch.Tensor,
                 item_keys: torch.Tensor, item_bounds: torch.Tensor):
        self.layout = layout
        self.node_boxes = node_boxes
        self.item_ids = item_ids
        self.item_keys = item_keys
        self.item_bounds = item_bounds

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


def query_boxes(qboxes: torch.Tensor, tree: DeviceTree, capacity: int | None = None):
    """(left, right, stats): lexsorted (query, item) pairs whose boxes overlap."""
    dev = require_gpu()
    L = lib()
    n = int(qboxes.shape[0])
    cap = int(capacity) if capacity is not None else max(1024, 2 * n)
    retries = 0
    while True:
        ws = _ws(L.vs_query_boxes_workspace_bytes(n, cap), dev)
        out_l = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
        out_r = torch.empty(max(cap, 1), dtype=torch.int64, device=dev)
        res = torch.zeros(4, dtype=torch.int64, device=dev)
        check(
            L.vs_query_boxes(ptr(qboxes), n, ctypes.byref(tree.layout), ptr(tree.node_boxes), ptr(tree.item_ids),
                             cap, ptr(out_l), ptr(out_r), ptr(res), ptr(ws), ws.numel(), stream()),
            "vs_query_boxes",
        )
        k, staged, cand, cverts = _read_result(res)
        if staged <= cap:
            return out_l[:k], out_r[:k], JoinStats(k, staged, cand, cap, retries, cverts)
        if staged >= 2**31 - 1:
            raise OverflowError(f"{staged} pairs exceed the int32 pair limit")
        cap = staged
        retries += 1


_ = _lib  # re-exported module for callers that need the raw binding
