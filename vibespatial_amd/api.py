"""Public surface of the join path: GeoSeries, GeoDataFrame, SpatialIndex, sjoin.

Signatures and argument checks follow the reference:
  sjoin                src/vibespatial/api/tools/sjoin.py:37-192, checks :268-321
  SpatialIndex.query   src/vibespatial/api/sindex.py:144-503, predicates :31-35,
                       argument validation :122-142
Output order: ``sindex.query`` returns (2, n) [query_idx; tree_idx] and ``sjoin``
joins on (left_index, right_index); both are lexsorted here (the reference
sorts only on request, ``query_utils.py:202-206``; a sorted result satisfies
both contracts).

Routing (all compute on the GPU through libvsjoin; host numpy only maps rows of
mixed point/polygon columns and assembles frames), chosen per family pair by
``_strategy`` (the reference's choice is spatial/spatial_index_device.py:228-390):
  point  x polygon  within / intersects / covered_by / touches -> learned join:
                    rank index over the query points, polygon-major FP32
                    refine with exact escalation (vs_sjoin_points_polygons_learned)
  polygon x point   (point tree) contains / contains_properly / covers /
                    intersects / touches -> the tree's learned index and the
                    same FP32 refine (vs_learned_query_polygons); axis-aligned
                    rectangles with intersects / covers, and predicate None,
                    -> the learned box query alone (vs_learned_query_boxes)
  point  x point    intersects / within / contains / covers / covered_by /
                    contains_properly / None -> learned box query (equality)
  polygon x polygon every predicate -> packed-tree bbox query + exact refine
                    (vs_refine_pairs mode 1: intersects, and the areal
                    relations within / contains / covers / covered_by /
                    contains_properly / touches / overlaps of vs_relate.h,
                    with heavy pairs one wave per pair and undecided pairs
                    through the wave-parallel relation)
  point  x polygon  predicate None -> packed-tree bbox query
  dwithin (every family pair) -> packed-tree query with boxes grown by each
                    query row's distance, exact distance refine
                    (vs_refine_dwithin)
Combinations whose answer is empty by definition (a point cannot contain a
polygon, points cannot cross, two areal rows never cross, ...) return no pairs,
matching the reference's mapping (predicates/point_relations.py:100-146).
"""

from __future__ import annotations

import ctypes
import warnings

import numpy as np
import pandas as pd
import torch

from . import join as J
from ._lib import PRED, check, lib
from .device import device_points, device_polygons, ptr, require_gpu, stream
from .geometry import GeometryArray, PointArray, PolygonArray, _Geom, as_geometry_array

__all__ = ["GeoSeries", "GeoDataFrame", "SpatialIndex", "sjoin", "PREDICATES"]

PREDICATES = {
    None,
    "intersects",
    "within",
    "contains",
    "contains_properly",
    "covers",
    "covered_by",
    "touches",
    "crosses",
    "overlaps",
    "dwithin",
}

_EMPTY_PAIRS = {
    # (query family, tree family): predicates that can never hold
    ("point", "polygon"): {"contains", "contains_properly", "covers", "crosses", "overlaps"},
    ("polygon", "point"): {"within", "covered_by", "crosses", "overlaps"},
    ("point", "point"): {"touches", "crosses", "overlaps"},
    ("polygon", "polygon"): {"crosses"},  # two areal rows never cross
}


# ---------------------------------------------------------------------------
# pair helpers (device int64 tensors)
# ---------------------------------------------------------------------------
def _empty_pairs(dev):
    e = torch.empty(0, dtype=torch.int64, device=dev)
    return e, e.clone()


def _refine(mode: int, predicate: str, left, right, points: PointArray | None, A: PolygonArray | None,
            B: PolygonArray | None, heavy_log2: int = 0):
    """vs_refine_pairs; heavy_log2 moves the one-wave-per-pair threshold of
    polygon pairs (0: the library default, 63: never)."""
    dev = require_gpu()
    L = lib()
    n = int(left.numel())
    if n == 0:
        return _empty_pairs(dev)
    P = device_points(points) if points is not None else None
    DA = device_polygons(A) if A is not None else None
    DB = device_polygons(B) if B is not None else None
    out_l = torch.empty(n, dtype=torch.int64, device=dev)
    out_r = torch.empty(n, dtype=torch.int64, device=dev)
    res = torch.zeros(4, dtype=torch.int64, device=dev)
    ws = torch.empty(max(int(L.vs_refine_pairs_workspace_bytes(n)), 1), dtype=torch.uint8, device=dev)
    none = [ptr(None)] * 5
    a_args = DA.args() if DA is not None else none
    b_args = DB.args() if DB is not None else none
    check(
        L.vs_refine_pairs(mode | (int(heavy_log2) << 8), PRED.get(predicate, 1), ptr(left), ptr(right), n,
                          ptr(P.x if P else None), ptr(P.y if P else None), *a_args, *b_args,
                          ptr(out_l), ptr(out_r), ptr(res), ptr(ws), ws.numel(), stream()),
        "vs_refine_pairs",
    )
    k = int(res[0].item())
    return out_l[:k], out_r[:k]


def _family_bounds_device(ga: GeometryArray):
    """(n_rows, 4) fp64 device bounds of a column (NaN for null rows)."""
    dev = require_gpu()
    if ga.is_points:
        return J.point_boxes(device_points(ga.points))
    if ga.is_polygons:
        return J.polygon_bounds(device_polygons(ga.polygons))
    out = torch.full((len(ga), 4), float("nan"), dtype=torch.float64, device=dev)
    if ga.points is not None:
        out[torch.as_tensor(ga.point_rows, device=dev)] = J.point_boxes(device_points(ga.points))
    if ga.polygons is not None:
        out[torch.as_tensor(ga.polygon_rows, device=dev)] = J.polygon_bounds(device_polygons(ga.polygons))
    return out


class _FamilyIndex:
    """Indexes over the tree column, one set per family, built lazily and kept
    with the column: the packed Hilbert tree of each family, the polygon
    family's device bounds, and a learned rank index over the point family
    (64-bit Hilbert keys)."""

    def __init__(self, ga: GeometryArray):
        self.ga = ga
        self._trees = {}
        self._bounds = {}
        self._learned = None

    def bounds(self, family: str) -> torch.Tensor:
        b = self._bounds.get(family)
        if b is None:
            if family == "point":
                b = J.point_boxes(device_points(self.ga.points))
            elif family == "polygon":
                b = J.polygon_bounds(device_polygons(self.ga.polygons))
            else:
                b = _family_bounds_device(self.ga)
            self._bounds[family] = b
        return b

    def tree(self, family: str) -> J.DeviceTree:
        t = self._trees.get(family)
        if t is None:
            t = J.build_tree(self.bounds(family))
            self._trees[family] = t
        return t

    def learned(self) -> J.LearnedPointIndex:
        if self._learned is None:
            self._learned = J.LearnedPointIndex(device_points(self.ga.points), key_bytes=8)
        return self._learned


def _dev_geom(fam: str, arr):
    return device_points(arr) if fam == "point" else device_polygons(arr)


# Point x polygon predicates the fused learned join decides (location codes).
_POINT_POLY_PREDS = ("within", "intersects", "covered_by", "touches")
# Polygon-over-point predicates of the learned polygon query.
_POLY_POINT_PREDS = ("contains", "contains_properly", "covers", "intersects", "touches")


def all_rectangles(pa: PolygonArray) -> bool:
    """Every non-empty row is one axis-aligned rectangle with positive width
    and height (one part, one closed ring of five coordinates, no holes).  For
    such rows `intersects` / `covers` of a point equals the inclusive point-in-
    box test.  Cached on the array."""
    cached = getattr(pa, "_all_rect", None)
    if cached is not None:
        return cached
    g, p, r = pa.geom_offsets, pa.part_offsets, pa.ring_offsets
    nparts = np.diff(g)
    ok = bool(np.all(nparts <= 1))
    if ok:
        rows = np.flatnonzero(nparts == 1)
        parts = g[rows]
        ok = bool(np.all(p[parts + 1] - p[parts] == 1))
        if ok and len(rows):
            rings = p[parts]
            ok = bool(np.all(r[rings + 1] - r[rings] == 5))
            if ok:
                c = r[rings][:, None] + np.arange(5)
                x, y = pa.x[c], pa.y[c]
                closed = (x[:, 0] == x[:, 4]) & (y[:, 0] == y[:, 4])
                # edges alternate vertical / horizontal, starting either way
                a = (x[:, 0] == x[:, 1]) & (y[:, 1] == y[:, 2]) & (x[:, 2] == x[:, 3]) & (y[:, 3] == y[:, 4])
                b = (y[:, 0] == y[:, 1]) & (x[:, 1] == x[:, 2]) & (y[:, 2] == y[:, 3]) & (x[:, 3] == x[:, 4])
                area = (x[:, 0] != x[:, 2]) & (y[:, 0] != y[:, 2])
                ok = bool(np.all(closed & (a | b) & area))
    pa._all_rect = ok
    return ok


def _strategy(q_fam: str, t_fam: str, predicate, qarr) -> str:
    """The device route of one family pair (the reference chooses among
    point-grid, Morton-range and brute-force candidates in
    spatial/spatial_index_device.py:228-390; here):
      learned_join     point x polygon predicates: learned rank index over the
                       query points built per call, polygon-major FP32 refine
                       with exact escalation (vs_sjoin_points_polygons_learned);
      learned_boxes    point trees queried by boxes whose test is exact:
                       predicate None (bbox), points (equality), or
                       axis-aligned rectangles with intersects / covers
                       (spatial/query_box.py:760-863: the pure box path);
      learned_polygons other polygon-over-point predicates on point trees:
                       the FP32 point-in-polygon refine over the tree's
                       learned index (vs_learned_query_polygons);
      tree             polygon trees queried by polygons, bbox-only point
                       queries of polygon trees, and dwithin: packed Hilbert
                       tree + exact refine."""
    if predicate == "dwithin":
        return "tree"
    if t_fam == "point":
        if q_fam == "point" or predicate is None:
            return "learned_boxes"
        if predicate in ("intersects", "covers") and all_rectangles(qarr):
            return "learned_boxes"
        return "learned_polygons"
    if q_fam == "point" and predicate in _POINT_POLY_PREDS:
        return "learned_join"
    return "tree"


def _pairs_family(q_fam: str, qarr, t_fam: str, tarr, tindex: _FamilyIndex, predicate, dist=None, log=None):
    """Pairs (query family idx, tree family idx), lexsorted, device int64.
    dist: per query-family row distances (fp64) for dwithin.  log: a list the
    chosen strategy is appended to."""
    dev = require_gpu()
    if predicate in _EMPTY_PAIRS.get((q_fam, t_fam), ()):
        return _empty_pairs(dev)
    how = _strategy(q_fam, t_fam, predicate, qarr)
    if log is not None:
        log.append((q_fam, t_fam, how))
    if how == "learned_join":
        l, r, _ = J.sjoin_points_polygons_learned(device_points(qarr), device_polygons(tarr),
                                                  tindex.bounds("polygon"), predicate)
        return l, r
    if how == "learned_boxes":
        qb = J.point_boxes(device_points(qarr)) if q_fam == "point" else J.polygon_bounds(device_polygons(qarr))
        l, r, _ = tindex.learned().query_boxes(qb)
        return l, r
    if how == "learned_polygons":
        P = device_polygons(qarr)
        l, r, _ = tindex.learned().query_polygons(P, J.polygon_bounds(P), predicate)
        return l, r
    tree = tindex.tree(t_fam)
    if predicate == "dwithin":
        # boxes grown by each row's distance (outward-rounded), then the exact
        # distance refine (spatial/query.py:663-826 route)
        d = torch.as_tensor(np.ascontiguousarray(dist, dtype=np.float64), device=dev)
        qb = J.point_boxes(device_points(qarr)) if q_fam == "point" else J.polygon_bounds(device_polygons(qarr))
        l, r, _ = J.query_boxes(J.expand_boxes(qb, d), tree, kind="dwithin")
        return J.refine_dwithin(l, r, d, _dev_geom(q_fam, qarr), _dev_geom(t_fam, tarr))
    qb = J.point_boxes(device_points(qarr)) if q_fam == "point" else J.polygon_bounds(device_polygons(qarr))
    l, r, _ = J.query_boxes(qb, tree)
    if predicate is None:
        return l, r
    if q_fam == "polygon" and t_fam == "polygon":
        if l.numel() == 0:
            return l, r  # no bbox overlap: every predicate but disjoint is false
        return _refine(1, predicate, l, r, None, qarr, tarr)
    raise NotImplementedError(f"{q_fam} x {t_fam} predicate {predicate!r}")


def _map_rows(idx: torch.Tensor, rows: np.ndarray) -> torch.Tensor:
    """Family index -> column row (rows ascending, so order is kept)."""
    if len(rows) == 0 or (rows[0] == 0 and rows[-1] == len(rows) - 1):
        return idx  # identity
    return torch.as_tensor(rows, device=idx.device)[idx]


def query_pairs(query: GeometryArray, tree_ga: GeometryArray, tindex: _FamilyIndex, predicate, distance=None,
                log=None):
    """Lexsorted (query_row, tree_row) pairs as device int64 tensors.
    distance: per query row fp64 array (dwithin only); log: a list that
    receives (query family, tree family, strategy) per family pair."""
    dev = require_gpu()
    fams_q = [(f, a, rows) for f, a, rows in (("point", query.points, query.point_rows),
                                               ("polygon", query.polygons, query.polygon_rows)) if a is not None]
    fams_t = [(f, a, rows) for f, a, rows in (("point", tree_ga.points, tree_ga.point_rows),
                                               ("polygon", tree_ga.polygons, tree_ga.polygon_rows)) if a is not None]
    if len(fams_q) == 1 and len(fams_t) == 1:
        (qf, qa, qrows), (tf, ta, trows) = fams_q[0], fams_t[0]
        dist = None if distance is None else distance[qrows]
        l, r = _pairs_family(qf, qa, tf, ta, tindex, predicate, dist, log)
        return _map_rows(l, qrows), _map_rows(r, trows)
    parts_l, parts_r = [], []
    for qf, qa, qrows in fams_q:
        for tf, ta, trows in fams_t:
            dist = None if distance is None else distance[qrows]
            l, r = _pairs_family(qf, qa, tf, ta, tindex, predicate, dist, log)
            parts_l.append(qrows[l.cpu().numpy()])
            parts_r.append(trows[r.cpu().numpy()])
    if not parts_l:
        return _empty_pairs(dev)
    lh = np.concatenate(parts_l)
    rh = np.concatenate(parts_r)
    o = np.lexsort((rh, lh))
    return torch.as_tensor(lh[o], device=dev), torch.as_tensor(rh[o], device=dev)


# ---------------------------------------------------------------------------
# SpatialIndex
# ---------------------------------------------------------------------------
class SpatialIndex:
    """Spatial index of a geometry column (packed Hilbert trees on the GPU)."""

    def __init__(self, geometry):
        self._ga = as_geometry_array(geometry)
        self._index = _FamilyIndex(self._ga)
        # (query family, tree family, strategy) of each family pair of the
        # last query (see _strategy)
        self.last_strategy = []

    @property
    def valid_query_predicates(self):
        return set(PREDICATES)

    @property
    def size(self) -> int:
        return int((self._ga.tags >= 0).sum())

    def __len__(self) -> int:
        return self.size

    @property
    def is_empty(self) -> bool:
        return self.size == 0

    def _check(self, predicate, distance):
        if predicate not in self.valid_query_predicates:
            raise ValueError(
                f"Got predicate='{predicate}'; `predicate` must be one of {self.valid_query_predicates}"
            )
        if predicate == "dwithin":
            if distance is None:
                raise ValueError("'distance' parameter is required for 'dwithin' predicate")
        elif distance is not None:
            raise ValueError("'distance' parameter is only supported in combination with 'dwithin' predicate")

    def query(self, geometry, predicate=None, sort=False, distance=None, output_format="indices",
              return_device=False):
        """predicate(geometry, tree_geometry) pairs; see api/sindex.py:144-503."""
        self._check(predicate, distance)
        if output_format not in ("indices", "dense", "sparse"):
            raise ValueError(f"Invalid output_format: '{output_format}'. Use one of 'indices', 'sparse', 'dense'.")
        if return_device and output_format != "indices":
            raise ValueError("return_device=True requires output_format='indices'")
        scalar = geometry is None or isinstance(geometry, _Geom)
        if not scalar and not isinstance(geometry, (list, tuple, np.ndarray, GeometryArray, PointArray,
                                                    PolygonArray, GeoSeries)):
            raise TypeError(f"invalid geometry argument: {type(geometry).__name__}")
        q = as_geometry_array(geometry)
        if len(q) == 0 or self._ga.points is None and self._ga.polygons is None or (
                q.points is None and q.polygons is None):
            l = r = torch.empty(0, dtype=torch.int64, device=require_gpu())
        else:
            dist = None
            if predicate == "dwithin":
                try:
                    dist = np.ascontiguousarray(np.broadcast_to(np.asarray(distance, dtype=np.float64), (len(q),)))
                except ValueError as exc:
                    raise ValueError("distance array must be broadcastable to the geometry input") from exc
            self.last_strategy = []
            l, r = query_pairs(q, self._ga, self._index, predicate, dist, self.last_strategy)
        if return_device:
            return r if scalar else torch.stack([l, r])
        lh, rh = l.cpu().numpy(), r.cpu().numpy()
        if output_format == "indices":
            return rh if scalar else np.vstack([lh, rh])
        n_tree = len(self._ga)
        if scalar:
            dense = np.zeros(n_tree, dtype=bool)
            dense[rh] = True
        else:
            dense = np.zeros((n_tree, len(q)), dtype=bool)
            dense[rh, lh] = True
        if output_format == "dense":
            return dense
        from scipy import sparse

        return sparse.coo_array(dense)

    def _nearest_family(self, qf, qa, tf, ta, max_distance, return_all, exclusive):
        """(query family idx, tree family idx, distance) device tensors."""
        qb = J.point_boxes(device_points(qa)) if qf == "point" else J.polygon_bounds(device_polygons(qa))
        tb_h, qb_h = ta.bounds(), qa.bounds()
        fin = np.isfinite(tb_h).all(axis=1)
        if not fin.any():
            return None
        both = np.concatenate([tb_h[fin], qb_h[np.isfinite(qb_h).all(axis=1)]])
        x0, y0 = both[:, 0].min(), both[:, 1].min()
        x1, y1 = both[:, 2].max(), both[:, 3].max()
        # a radius at which every grown query box covers every tree box
        cover = float(np.hypot(x1 - x0, y1 - y0)) * 1.01 + 1e-300
        if not np.isfinite(cover):
            cover = float(np.finfo(np.float64).max)
        t0, t1 = tb_h[fin].min(axis=0), tb_h[fin].max(axis=0)
        w, h = t1[2] - t0[0], t1[3] - t0[1]
        n_t = int(fin.sum())
        # the mean spacing of the tree rows as the first radius
        start = float(np.sqrt(w * h / n_t)) if w > 0 and h > 0 else max(w, h) / n_t
        if not (start > 0) or not np.isfinite(start):
            start = cover / 1024
        return J.nearest_pairs(qb, _dev_geom(qf, qa), self._index.tree(tf), _dev_geom(tf, ta),
                               max_distance=max_distance, return_all=return_all, exclusive=exclusive,
                               start_radius=start, cover_radius=cover)

    def nearest(self, geometry, return_all=True, max_distance=None, return_distance=False, exclusive=False):
        """Nearest tree rows for each input geometry (api/sindex.py:2018-2277):
        (2, n) indices lexsorted by input row, then tree row, and the fp64
        distances when `return_distance`.  Ties (equal distances) are all
        returned unless not `return_all` (then the least tree row).  Columns
        holding both points and polygons are searched per family pair and the
        per-row minimum taken over the families.  `exclusive` leaves out tree
        rows equal to the input geometry (topological equality, as GEOS
        equals: equal coordinates for points, mutual covers for polygons)."""
        q = as_geometry_array(geometry)
        if max_distance is not None and not (float(max_distance) >= 0):
            raise ValueError("max_distance must be a non-negative number")
        fams_q = [(f, a, rows) for f, a, rows in (("point", q.points, q.point_rows),
                                                   ("polygon", q.polygons, q.polygon_rows)) if a is not None]
        fams_t = [(f, a, rows) for f, a, rows in (("point", self._ga.points, self._ga.point_rows),
                                                   ("polygon", self._ga.polygons, self._ga.polygon_rows))
                  if a is not None]
        parts = []
        for qf, qa, qrows in fams_q:
            for tf, ta, trows in fams_t:
                # per family pair all ties are kept; the selection across
                # families below applies return_all
                res = self._nearest_family(qf, qa, tf, ta, max_distance, return_all or len(fams_t) > 1, exclusive)
                if res is None:
                    continue
                l, r, d = (t.cpu().numpy() for t in res)
                parts.append((qrows[l], trows[r], d))
        if parts:
            lh = np.concatenate([p[0] for p in parts]).astype(np.int64)
            rh = np.concatenate([p[1] for p in parts]).astype(np.int64)
            dh = np.concatenate([p[2] for p in parts])
        else:
            lh = rh = np.empty(0, dtype=np.int64)
            dh = np.empty(0, dtype=np.float64)
        if len(fams_t) > 1 and len(lh):
            o = np.lexsort((rh, dh, lh))
            lh, rh, dh = lh[o], rh[o], dh[o]
            first = np.r_[True, lh[1:] != lh[:-1]]
            m = np.maximum.accumulate(np.where(first, np.arange(len(lh)), 0))  # index of each row's first pair
            keep = dh == dh[m]
            if not return_all:
                keep &= first
            lh, rh, dh = lh[keep], rh[keep], dh[keep]
        o = np.lexsort((rh, lh))
        idx = np.vstack([lh[o], rh[o]])
        if return_distance:
            return idx, dh[o]
        return idx

    def intersection(self, coordinates):
        """Tree rows whose bbox intersects a bounds tuple or a point (x, y)."""
        if not isinstance(coordinates, tuple) or len(coordinates) not in (2, 4):
            raise TypeError("Invalid coordinates, must be iterable in format (minx, miny, maxx, maxy) or (x, y)")
        from .geometry import box as _box, Point as _Point

        g = _Point(*coordinates) if len(coordinates) == 2 else _box(*coordinates)
        return self.query(g)


# ---------------------------------------------------------------------------
# GeoSeries / GeoDataFrame (thin: attributes in pandas, geometry in SoA arrays)
# ---------------------------------------------------------------------------
class GeoSeries:
    def __init__(self, data=None, index=None, crs=None, name=None):
        self.values = as_geometry_array([] if data is None else data)
        self.index = pd.RangeIndex(len(self.values)) if index is None else pd.Index(index)
        self.crs = crs
        self.name = name

    def __len__(self) -> int:
        return len(self.values)

    def __getitem__(self, i):
        return self.values[int(i)]

    def __iter__(self):
        for i in range(len(self)):
            yield self.values[i]

    @property
    def sindex(self) -> SpatialIndex:
        if self.values._sindex is None:
            self.values._sindex = SpatialIndex(self.values)
        return self.values._sindex

    @property
    def bounds(self) -> pd.DataFrame:
        return pd.DataFrame(self.values.bounds(), columns=["minx", "miny", "maxx", "maxy"], index=self.index)

    @property
    def geom_type(self):
        names = {0: "Point", 1: "Polygon", -1: None}
        return pd.Series([names[int(t)] for t in self.values.tags], index=self.index)

    def take(self, idx) -> "GeoSeries":
        """Rows idx (positions; -1 gives a null row), buffers gathered without
        per-row objects."""
        return GeoSeries(self.values.take(np.asarray(idx, dtype=np.int64)), crs=self.crs, name=self.name)

    def isna(self) -> pd.Series:
        return pd.Series(self.values.tags < 0, index=self.index)

    def notna(self) -> pd.Series:
        return pd.Series(self.values.tags >= 0, index=self.index)


class GeoDataFrame:
    """Attribute columns in a pandas DataFrame plus one active geometry column."""

    def __init__(self, data=None, geometry=None, crs=None, index=None):
        df = pd.DataFrame(data, index=index) if not isinstance(data, pd.DataFrame) else data.copy()
        name = "geometry"
        if isinstance(geometry, str):
            name = geometry
            geoms = df.pop(geometry)
            geoms = list(geoms) if not isinstance(geoms, GeoSeries) else geoms
        elif geometry is not None:
            geoms = geometry
        elif "geometry" in df.columns:
            geoms = list(df.pop("geometry"))
        else:
            raise ValueError("GeoDataFrame needs a geometry column")
        self._df = df
        gs = geoms if isinstance(geoms, GeoSeries) else GeoSeries(geoms, crs=crs)
        gs.index = df.index
        self._geometry = gs
        self._geom_name = name
        self.crs = crs if crs is not None else gs.crs

    # pandas-like surface
    @property
    def geometry(self) -> GeoSeries:
        return self._geometry

    @property
    def sindex(self) -> SpatialIndex:
        return self._geometry.sindex

    @property
    def index(self):
        return self._df.index

    @property
    def columns(self):
        cols = list(self._df.columns)
        return pd.Index(cols[: self._geom_pos()] + [self._geom_name] + cols[self._geom_pos():])

    def _geom_pos(self) -> int:
        return getattr(self, "_geom_at", len(self._df.columns))

    @property
    def shape(self):
        return (len(self._df), len(self._df.columns) + 1)

    def __len__(self) -> int:
        return len(self._df)

    def __contains__(self, col) -> bool:
        return col == self._geom_name or col in self._df.columns

    def __getitem__(self, col):
        if col == self._geom_name:
            return self._geometry
        return self._df[col]

    def sjoin_nearest(self, right, how="inner", max_distance=None, lsuffix="left", rsuffix="right",
                      distance_col=None, exclusive=False):
        return sjoin_nearest(self, right, how=how, max_distance=max_distance, lsuffix=lsuffix, rsuffix=rsuffix,
                             distance_col=distance_col, exclusive=exclusive)

    def to_parquet(self, path, little_endian: bool = True) -> None:
        """GeoParquet 1.0 with a WKB geometry column."""
        from .io import to_parquet

        to_parquet(self, path, little_endian)

    def to_pandas(self) -> pd.DataFrame:
        out = self._df.copy()
        out.insert(self._geom_pos(), self._geom_name, list(self._geometry))
        return out

    @property
    def loc(self):
        return self.to_pandas().loc

    @property
    def iloc(self):
        return self.to_pandas().iloc

    def iterrows(self):
        return self.to_pandas().iterrows()

    def sjoin(self, df, *args, **kwargs):
        return sjoin(self, df, *args, **kwargs)

    def __repr__(self):
        return repr(self.to_pandas())


# ---------------------------------------------------------------------------
# sjoin
# ---------------------------------------------------------------------------
def _maybe_make_list(obj):
    if isinstance(obj, tuple):
        return list(obj)
    if obj is not None and not isinstance(obj, list):
        return [obj]
    return obj


def _basic_checks(left_df, right_df, how, lsuffix, rsuffix, on_attribute=None, allowed_hows=None):
    """api/tools/sjoin.py:268-321: sjoin allows outer, sjoin_nearest does not."""
    if not isinstance(left_df, GeoDataFrame):
        raise ValueError(f"'left_df' should be GeoDataFrame, got {type(left_df)}")
    if not isinstance(right_df, GeoDataFrame):
        raise ValueError(f"'right_df' should be GeoDataFrame, got {type(right_df)}")
    allowed = allowed_hows if allowed_hows is not None else ["left", "right", "inner"]
    if how not in allowed:
        raise ValueError(f'`how` was "{how}" but is expected to be in {allowed}')
    if left_df.crs != right_df.crs:
        warnings.warn("CRS mismatch between the CRS of left geometries and the CRS of right geometries.",
                      UserWarning, stacklevel=3)
    if on_attribute:
        for attr in on_attribute:
            if attr not in left_df and attr not in right_df:
                raise ValueError(f"Expected column {attr} is missing from both of the dataframes.")
            if attr not in left_df:
                raise ValueError(f"Expected column {attr} is missing from the left dataframe.")
            if attr not in right_df:
                raise ValueError(f"Expected column {attr} is missing from the right dataframe.")
            if attr in (left_df.geometry.name, right_df.geometry.name, left_df._geom_name, right_df._geom_name):
                raise ValueError("Active geometry column cannot be used as an input for on_attribute parameter.")


def sjoin_indices(left, right, predicate="intersects", return_device=False):
    """(left_index, right_index) lexsorted pairs with predicate(left, right)."""
    lg = left.geometry if isinstance(left, GeoDataFrame) else left
    rg = right.geometry if isinstance(right, GeoDataFrame) else right
    rg_series = rg if isinstance(rg, GeoSeries) else GeoSeries(rg)
    res = rg_series.sindex.query(lg.values if isinstance(lg, GeoSeries) else lg, predicate=predicate,
                                 return_device=return_device)
    return res[0], res[1]


def sjoin(left_df, right_df, how="inner", predicate="intersects", lsuffix="left", rsuffix="right",
          distance=None, on_attribute=None, **kwargs):
    """Spatial join of two GeoDataFrames (api/tools/sjoin.py:37-192)."""
    if kwargs:
        first = next(iter(kwargs.keys()))
        raise TypeError(f"sjoin() got an unexpected keyword argument '{first}'")
    on_attribute = _maybe_make_list(on_attribute)
    _basic_checks(left_df, right_df, how, lsuffix, rsuffix, on_attribute,
                  allowed_hows=["left", "right", "inner", "outer"])
    tree = right_df.sindex
    tree._check(predicate, distance)
    res = tree.query(left_df.geometry.values, predicate=predicate, sort=True, distance=distance)
    l_idx, r_idx = res[0], res[1]
    if on_attribute:
        keep = np.ones(len(l_idx), dtype=bool)
        for attr in on_attribute:
            keep &= left_df[attr].to_numpy()[l_idx] == right_df[attr].to_numpy()[r_idx]
        l_idx, r_idx = l_idx[keep], r_idx[keep]
    return _assemble(left_df, right_df, l_idx, r_idx, how, lsuffix, rsuffix, on_attribute)


def sjoin_nearest(left_df, right_df, how="inner", max_distance=None, lsuffix="left", rsuffix="right",
                  distance_col=None, exclusive=False):
    """Join on the nearest geometry (api/tools/sjoin.py:1398-1560): every
    equidistant nearest right row; how='right' queries the right geometries
    against the left index (api/tools/sjoin.py:1173-1212)."""
    _basic_checks(left_df, right_df, how, lsuffix, rsuffix)
    if how == "right":
        res = left_df.sindex.nearest(right_df.geometry.values, return_all=True, max_distance=max_distance,
                                     return_distance=True, exclusive=exclusive)
        (input_idx, tree_idx), dist = res
        l_idx, r_idx = tree_idx, input_idx
        o = np.argsort(l_idx, kind="stable")
        l_idx, r_idx, dist = l_idx[o], r_idx[o], dist[o]
    else:
        (l_idx, r_idx), dist = right_df.sindex.nearest(left_df.geometry.values, return_all=True,
                                                        max_distance=max_distance, return_distance=True,
                                                        exclusive=exclusive)
    extra = {distance_col: dist} if distance_col is not None else None
    return _assemble(left_df, right_df, l_idx, r_idx, how, lsuffix, rsuffix, None, extra)


def _assemble(left_df, right_df, l_idx, r_idx, how, lsuffix, rsuffix, on_attribute, extra=None):
    """GeoPandas-shaped frame for the join (api/_native_results.py:301-472)."""
    ldf, rdf = left_df._df, right_df._df
    on_attribute = on_attribute or []
    index_left = f"index_{lsuffix}"
    index_right = f"index_{rsuffix}"
    lcols = list(ldf.columns)
    rcols = [c for c in rdf.columns if c not in on_attribute]
    overlap = (set(lcols) | {left_df._geom_name}) & set(rcols)
    lren = {c: f"{c}_{lsuffix}" for c in overlap if c in lcols}
    rren = {c: f"{c}_{rsuffix}" for c in overlap}
    if how == "right":
        l_idx0, r_idx0 = l_idx, r_idx
        l_idx, r_idx = l_idx.copy(), r_idx.copy()
        o = np.lexsort((l_idx, r_idx))
        l_idx, r_idx = l_idx[o], r_idx[o]
        matched = np.zeros(len(rdf), dtype=bool)
        matched[r_idx] = True
        miss = np.flatnonzero(~matched)
        r_all = np.concatenate([r_idx, miss])
        l_all = np.concatenate([l_idx, np.full(len(miss), -1)])
        o = np.argsort(r_all, kind="stable")
        r_all, l_all = r_all[o], l_all[o]
        left_part = ldf.rename(columns=lren).reset_index(drop=True).reindex(np.where(l_all >= 0, l_all, -1))
        left_part.index = rdf.index[r_all]
        out = pd.DataFrame(index=rdf.index[r_all])
        out[index_left] = np.where(l_all >= 0, np.asarray(ldf.index)[np.maximum(l_all, 0)], np.nan) \
            if len(l_all) else []
        for c in lcols:
            out[lren.get(c, c)] = left_part[lren.get(c, c)].to_numpy()
        for c in rcols:
            out[rren.get(c, c)] = rdf[c].to_numpy()[r_all]
        for name, vals in (extra or {}).items():
            ov = np.full(len(r_all), np.nan)
            ov[: len(l_idx)] = np.asarray(vals, dtype=np.float64)[np.lexsort((l_idx0, r_idx0))]
            out[name] = ov[o]
        geoms = right_df.geometry.take(r_all)
        g = GeoDataFrame(out, geometry=geoms, crs=right_df.crs)
        g._geom_name = right_df._geom_name
        return g
    if how == "outer":
        return _assemble_outer(left_df, right_df, l_idx, r_idx, index_left, index_right, lcols, rcols, lren, rren,
                               extra)
    extra_cols = {}
    if how == "left":
        matched = np.zeros(len(ldf), dtype=bool)
        matched[l_idx] = True
        miss = np.flatnonzero(~matched)
        l_all = np.concatenate([l_idx, miss])
        r_all = np.concatenate([r_idx, np.full(len(miss), -1)])
        o = np.argsort(l_all, kind="stable")
        l_all, r_all = l_all[o], r_all[o]
        for name, vals in (extra or {}).items():
            extra_cols[name] = np.concatenate([np.asarray(vals, dtype=np.float64), np.full(len(miss), np.nan)])[o]
    else:
        l_all, r_all = l_idx, r_idx
        for name, vals in (extra or {}).items():
            extra_cols[name] = np.asarray(vals, dtype=np.float64)
    out = pd.DataFrame(index=ldf.index[l_all] if len(ldf.index) else None)
    for c in lcols:
        out[lren.get(c, c)] = ldf[c].to_numpy()[l_all]
    valid = r_all >= 0
    ridx = np.asarray(rdf.index)
    out[index_right] = np.where(valid, ridx[np.maximum(r_all, 0)] if len(ridx) else 0, np.nan) \
        if how != "inner" else (ridx[r_all] if len(r_all) else np.empty(0, dtype=ridx.dtype))
    for c in rcols:
        out[rren.get(c, c)] = _column_take(rdf[c].to_numpy(), r_all)
    for name, vals in extra_cols.items():
        out[name] = vals
    geoms = left_df.geometry.take(l_all)
    g = GeoDataFrame(out, geometry=geoms, crs=left_df.crs)
    g._geom_name = left_df._geom_name
    g._geom_at = len(lcols)
    return g


def outer_indexers(l_idx, r_idx, n_left: int, n_right: int):
    """Row indexers of an outer join (api/_native_results.py:402-448): every
    left row in order, repeated once per match (once, with right -1, when it
    has none), then the right rows no left row matched (left -1).  Pairs are
    taken in lexsorted order.  Returns (left, right, position of each input
    pair in the output)."""
    l_idx = np.asarray(l_idx, dtype=np.int64)
    r_idx = np.asarray(r_idx, dtype=np.int64)
    order = np.lexsort((r_idx, l_idx))
    l_s, r_s = l_idx[order], r_idx[order]
    counts = np.bincount(l_s, minlength=n_left) if len(l_s) else np.zeros(n_left, np.int64)
    repeats = np.maximum(counts, 1)
    out_l = np.repeat(np.arange(n_left, dtype=np.int64), repeats)
    out_r = np.full(len(out_l), -1, dtype=np.int64)
    target = np.empty(0, dtype=np.int64)
    if len(l_s):
        first_match = np.cumsum(counts) - counts
        rank = np.arange(len(l_s)) - np.repeat(first_match, counts)
        row_start = np.cumsum(repeats) - repeats
        target = row_start[l_s] + rank
        out_r[target] = r_s
    matched_right = np.zeros(n_right, dtype=bool)
    matched_right[r_idx] = True
    miss = np.flatnonzero(~matched_right)
    pos = np.empty(len(l_idx), dtype=np.int64)
    pos[order] = target
    return (np.concatenate([out_l, np.full(len(miss), -1, np.int64)]), np.concatenate([out_r, miss]), pos)


def _column_take(values: np.ndarray, idx: np.ndarray) -> np.ndarray:
    """values[idx] with NaN where idx < 0 (object or float columns then)."""
    valid = idx >= 0
    out = values[np.maximum(idx, 0)] if len(values) else np.empty(len(idx))
    if (~valid).any():
        # as pandas reindexing fills: integers and floats become float64,
        # everything else object
        out = out.astype(float) if out.dtype.kind in "iufc" else out.astype(object)
        out[~valid] = np.nan
    return out


def _assemble_outer(left_df, right_df, l_idx, r_idx, index_left, index_right, lcols, rcols, lren, rren, extra):
    """Outer join frame: a fresh RangeIndex, both frames' index labels as the
    index_<suffix> columns (NaN where the side is missing), left columns,
    right columns, and the left geometry filled from the right geometry where
    the left row is missing (api/_native_results.py:476-500)."""
    ldf, rdf = left_df._df, right_df._df
    lo, ro, pos = outer_indexers(l_idx, r_idx, len(ldf), len(rdf))
    out = pd.DataFrame(index=pd.RangeIndex(len(lo)))
    out[index_left] = _column_take(np.asarray(ldf.index), lo)
    for c in lcols:
        out[lren.get(c, c)] = _column_take(ldf[c].to_numpy(), lo)
    out[index_right] = _column_take(np.asarray(rdf.index), ro)
    for c in rcols:
        out[rren.get(c, c)] = _column_take(rdf[c].to_numpy(), ro)
    for name, vals in (extra or {}).items():
        col = np.full(len(lo), np.nan)
        col[pos] = np.asarray(vals, dtype=np.float64)
        out[name] = col
    # left rows come first (every lo >= 0), the unmatched right rows last
    n_first = int(np.count_nonzero(lo >= 0))
    geoms = GeometryArray.concat([left_df.geometry.values.take(lo[:n_first]),
                                  right_df.geometry.values.take(ro[n_first:])])
    g = GeoDataFrame(out, geometry=GeoSeries(geoms, crs=left_df.crs), crs=left_df.crs)
    g._geom_name = left_df._geom_name
    return g
