"""ctypes wrapper over liboracle.so (test infrastructure only).

See oracle.c for what each function restates and how it is pinned.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
import threading

import numpy as np

__all__ = [
    "lib",
    "build",
    "orient2d",
    "orient2d_exact",
    "point_location",
    "point_location_pairs",
    "segments_touch",
    "polygon_intersects",
    "polygon_bounds",
    "bbox_pairs",
    "hilbert16",
    "hilbert_textbook",
    "sjoin_points_polygons",
    "sjoin_polygons_polygons",
    "box_query_points",
    "dwithin_pair",
    "dwithin_pairs",
    "segment_dist_f64",
    "distance_f64",
    "nearest_pairs",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lock = threading.RLock()
_lib = None

_d = ctypes.c_double
_i64 = ctypes.c_int64
_p = ctypes.c_void_p


def build() -> str:
    """Compile liboracle.so with make (gcc); returns its path."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)
    return _SO


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(_SO):
                build()
            L = ctypes.CDLL(_SO)
            L.or_orient2d.argtypes = [_d] * 6
            L.or_orient2d.restype = ctypes.c_int
            L.or_orient2d_exact.argtypes = [_d] * 6
            L.or_orient2d_exact.restype = ctypes.c_int
            L.or_point_location.argtypes = [_d, _d, _p, _p, _p, _p, _p, _i64]
            L.or_point_location.restype = ctypes.c_int
            L.or_point_location_pairs.argtypes = [_p, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _p, ctypes.c_int]
            L.or_point_location_pairs.restype = None
            L.or_segments_touch.argtypes = [_d] * 8
            L.or_segments_touch.restype = ctypes.c_int
            L.or_polygon_intersects.argtypes = [_p, _p, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _i64]
            L.or_polygon_intersects.restype = ctypes.c_int
            L.or_polygon_bounds.argtypes = [_p, _p, _p, _p, _p, _i64, _p]
            L.or_polygon_bounds.restype = None
            L.or_bbox_pairs.argtypes = [_p, _i64, _p, _i64, _p, _p]
            L.or_bbox_pairs.restype = _i64
            L.or_hilbert16.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
            L.or_hilbert16.restype = ctypes.c_uint32
            L.or_hilbert16_batch.argtypes = [_p, _p, _i64, _p]
            L.or_hilbert16_batch.restype = None
            L.or_hilbert_textbook.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
            L.or_hilbert_textbook.restype = ctypes.c_uint64
            L.or_sjoin_points_polygons.argtypes = [_p, _p, _i64, _p, _p, _p, _p, _p, _i64, ctypes.c_int, ctypes.c_int]
            L.or_sjoin_points_polygons.restype = _i64
            L.or_sjoin_polygons_polygons.argtypes = [_p, _p, _p, _p, _p, _i64, _p, _p, _p, _p, _p, _i64, ctypes.c_int]
            L.or_sjoin_polygons_polygons.restype = _i64
            L.or_box_query_points.argtypes = [_p, _p, _i64, _p, _i64, ctypes.c_int]
            L.or_box_query_points.restype = _i64
            L.or_join_take.argtypes = [_p, _p]
            L.or_join_take.restype = None
            _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def orient2d(ax, ay, bx, by, cx, cy) -> int:
    return lib().or_orient2d(ax, ay, bx, by, cx, cy)


def orient2d_exact(ax, ay, bx, by, cx, cy) -> int:
    return lib().or_orient2d_exact(ax, ay, bx, by, cx, cy)


class _Polys:
    """Holds contiguous polygon buffers alive for a ctypes call."""

    def __init__(self, polys):
        self.x = _f64(polys.x)
        self.y = _f64(polys.y)
        self.g = _i32(polys.geom_offsets)
        self.p = _i32(polys.part_offsets)
        self.r = _i32(polys.ring_offsets)
        self.n = len(self.g) - 1

    def args(self):
        return [_ptr(self.x), _ptr(self.y), _ptr(self.g), _ptr(self.p), _ptr(self.r)]


def point_location(px: float, py: float, polys, row: int) -> int:
    P = _Polys(polys)
    return lib().or_point_location(px, py, *P.args(), row)


def point_location_pairs(px, py, left, right, polys, nthreads: int = 0) -> np.ndarray:
    P = _Polys(polys)
    px, py = _f64(px), _f64(py)
    li = np.ascontiguousarray(left, dtype=np.int64)
    ri = np.ascontiguousarray(right, dtype=np.int64)
    out = np.zeros(len(li), dtype=np.uint8)
    lib().or_point_location_pairs(_ptr(px), _ptr(py), _ptr(li), _ptr(ri), len(li), *P.args(), _ptr(out), nthreads)
    return out


def segments_touch(a, b, c, d) -> bool:
    return bool(lib().or_segments_touch(*a, *b, *c, *d))


def polygon_intersects(A, ra: int, B, rb: int) -> bool:
    PA, PB = _Polys(A), _Polys(B)
    return bool(lib().or_polygon_intersects(*PA.args(), ra, *PB.args(), rb))


def polygon_bounds(polys) -> np.ndarray:
    P = _Polys(polys)
    out = np.empty((P.n, 4), dtype=np.float64)
    lib().or_polygon_bounds(*P.args(), P.n, _ptr(out))
    return out


def bbox_pairs(qbounds, tbounds):
    qb = _f64(qbounds).reshape(-1, 4)
    tb = _f64(tbounds).reshape(-1, 4)
    n = lib().or_bbox_pairs(_ptr(qb), len(qb), _ptr(tb), len(tb), None, None)
    lq = np.empty(n, dtype=np.int64)
    lt = np.empty(n, dtype=np.int64)
    lib().or_bbox_pairs(_ptr(qb), len(qb), _ptr(tb), len(tb), _ptr(lq), _ptr(lt))
    return lq, lt


def hilbert16(x, y) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.uint32)
    y = np.ascontiguousarray(y, dtype=np.uint32)
    out = np.empty(len(x), dtype=np.uint32)
    lib().or_hilbert16_batch(_ptr(x), _ptr(y), len(x), _ptr(out))
    return out


def hilbert_textbook(order: int, x: int, y: int) -> int:
    return lib().or_hilbert_textbook(order, x, y)


def _take(n: int):
    left = np.empty(n, dtype=np.int64)
    right = np.empty(n, dtype=np.int64)
    lib().or_join_take(_ptr(left), _ptr(right))
    return left, right


_PRED = {"within": 0, "intersects": 1, "covered_by": 2, "touches": 3}


def sjoin_points_polygons(px, py, polys, predicate: str = "within", nthreads: int = 0):
    """Lexsorted (point_idx, poly_idx) int64 pairs satisfying predicate(point, polygon)."""
    P = _Polys(polys)
    px, py = _f64(px), _f64(py)
    with _lock:
        n = lib().or_sjoin_points_polygons(_ptr(px), _ptr(py), len(px), *P.args(), P.n, _PRED[predicate], nthreads)
        return _take(n)


def sjoin_polygons_polygons(A, B, nthreads: int = 0):
    """Lexsorted (left_idx, right_idx) int64 pairs with intersects(A[l], B[r])."""
    PA, PB = _Polys(A), _Polys(B)
    with _lock:
        n = lib().or_sjoin_polygons_polygons(*PA.args(), PA.n, *PB.args(), PB.n, nthreads)
        return _take(n)


def box_query_points(px, py, boxes, nthreads: int = 0):
    """Lexsorted (box_idx, point_idx) pairs with the point inside the box (inclusive)."""
    px, py = _f64(px), _f64(py)
    b = _f64(boxes).reshape(-1, 4)
    with _lock:
        n = lib().or_box_query_points(_ptr(px), _ptr(py), len(px), _ptr(b), len(b), nthreads)
        return _take(n)


def _rings_of(polys, row: int):
    g, pt, r = polys.geom_offsets, polys.part_offsets, polys.ring_offsets
    for ring in range(pt[g[row]], pt[g[row + 1]]):
        yield int(r[ring]), int(r[ring + 1])


def dwithin_pair(left_kind: int, right_kind: int, L, R, li: int, ri: int, d: float) -> bool:
    """Exact dwithin of one pair (kinds: 0 points (x, y arrays), 1 polygons):
    the exact distance between the geometries is <= d.  A point and a polygon:
    location != 0 (C restatement) or a boundary segment within d; two
    polygons: intersects (C restatement) or a vertex of one within d of a
    boundary segment of the other.  Rationals; small inputs only."""
    from . import fraction_ref as FR

    if not (d >= 0.0):
        return False
    if left_kind == 1 and right_kind == 0:
        return dwithin_pair(0, 1, R, L, ri, li, d)
    if left_kind == 0:
        px, py = float(L.x[li]), float(L.y[li])
        if px != px or py != py:
            return False
        if right_kind == 0:
            qx, qy = float(R.x[ri]), float(R.y[ri])
            if qx != qx or qy != qy:
                return False
            return FR.within_distance(FR.dist2_point_point(px, py, qx, qy), d)
        rings = list(_rings_of(R, ri))
        if not any(e > b for b, e in rings):
            return False
        if d == float("inf") or point_location(px, py, R, ri) != 0:
            return True
        for b, e in rings:
            for c in range(b + 1, e):
                if FR.within_distance(FR.dist2_point_segment(px, py, R.x[c - 1], R.y[c - 1], R.x[c], R.y[c]), d):
                    return True
        return False
    ra, rb = list(_rings_of(L, li)), list(_rings_of(R, ri))
    if not any(e > b for b, e in ra) or not any(e > b for b, e in rb):
        return False
    if d == float("inf") or polygon_intersects(L, li, R, ri):
        return True
    for P, rp, Q, rq in ((L, ra, R, rb), (R, rb, L, ra)):
        for b0, e0 in rp:
            for v in range(b0, e0):
                for b, e in rq:
                    for c in range(b + 1, e):
                        if FR.within_distance(FR.dist2_point_segment(P.x[v], P.y[v], Q.x[c - 1], Q.y[c - 1],
                                                                     Q.x[c], Q.y[c]), d):
                            return True
    return False


def dwithin_pairs(left_kind: int, right_kind: int, L, R, distance, left_bounds, right_bounds):
    """Lexsorted (left, right) pairs within each left row's distance: boxes
    grown by the distance (bbox brute force, inclusive), then dwithin_pair."""
    qb = _f64(left_bounds).reshape(-1, 4).copy()
    d = np.asarray(distance, dtype=np.float64)
    with np.errstate(invalid="ignore"):
        qb[:, 0] -= d
        qb[:, 1] -= d
        qb[:, 2] += d
        qb[:, 3] += d
    qb[~(d >= 0)] = np.nan
    lq, lt = bbox_pairs(np.nextafter(qb, np.array([-np.inf, -np.inf, np.inf, np.inf])), right_bounds)
    keep = np.array([dwithin_pair(left_kind, right_kind, L, R, int(a), int(b), float(d[a]))
                     for a, b in zip(lq, lt)], dtype=bool)
    return lq[keep], lt[keep]


# ---------------------------------------------------------------------------
# nearest (fp64 distances as GEOS computes them; brute force)
# ---------------------------------------------------------------------------
def segment_dist_f64(px, py, ax, ay, bx, by):
    """Distance from (px, py) to segments a-b, elementwise over arrays, by
    GEOS Distance::pointToSegment (the reference's CPU path through shapely;
    its GPU kernels, spatial/nearest.py:2228-2482, use the same projection):
    r = ((p - a).(b - a)) / |b - a|^2; r <= 0 -> |p - a|, r >= 1 -> |p - b|,
    else |((a.y - p.y)(b.x - a.x) - (a.x - p.x)(b.y - a.y)) / |b - a|^2| *
    sqrt(|b - a|^2); a degenerate segment is its end point."""
    px, py = np.float64(px), np.float64(py)
    ax, ay, bx, by = (np.asarray(v, dtype=np.float64) for v in (ax, ay, bx, by))
    with np.errstate(invalid="ignore", divide="ignore"):
        da = np.sqrt((px - ax) * (px - ax) + (py - ay) * (py - ay))
        db = np.sqrt((px - bx) * (px - bx) + (py - by) * (py - by))
        ux, uy = bx - ax, by - ay
        len2 = ux * ux + uy * uy
        r = ((px - ax) * ux + (py - ay) * uy) / len2
        s = ((ay - py) * ux - (ax - px) * uy) / len2
        perp = np.abs(s) * np.sqrt(len2)
    out = np.where(r <= 0.0, da, np.where(r >= 1.0, db, perp))
    return np.where((ax == bx) & (ay == by), da, out)


def _boundary_dist(px, py, P, row):
    xs, ys = [], []
    for b, e in _rings_of(P, row):
        if e - b >= 2:
            xs.append((P.x[b:e - 1], P.x[b + 1:e]))
            ys.append((P.y[b:e - 1], P.y[b + 1:e]))
    if not xs:
        return np.inf
    ax = np.concatenate([a for a, _ in xs])
    bx = np.concatenate([b for _, b in xs])
    ay = np.concatenate([a for a, _ in ys])
    by = np.concatenate([b for _, b in ys])
    d = segment_dist_f64(px, py, ax, ay, bx, by)
    d = d[~np.isnan(d)]
    return float(d.min()) if len(d) else np.inf


def _poly_empty(P, row):
    return not any(e > b for b, e in _rings_of(P, row))


def distance_f64(left_kind, right_kind, L, R, li, ri, exclusive=False) -> float:
    """fp64 distance of one pair (kinds: 0 points, 1 polygons), +inf for
    empty / NaN geometries and, with `exclusive`, equal geometries (equal
    points; polygons that cover each other, relate_ref); 0 when a point
    is not outside a polygon (C restatement of the exact location) or two
    polygons intersect (C restatement), else the least vertex - segment
    distance over both directions."""
    if left_kind == 1 and right_kind == 0:
        return distance_f64(0, 1, R, L, ri, li, exclusive)
    if left_kind == 0:
        px, py = float(L.x[li]), float(L.y[li])
        if px != px or py != py:
            return np.inf
        if right_kind == 0:
            qx, qy = float(R.x[ri]), float(R.y[ri])
            if qx != qx or qy != qy or (exclusive and px == qx and py == qy):
                return np.inf
            dx, dy = np.float64(px) - np.float64(qx), np.float64(py) - np.float64(qy)
            return float(np.sqrt(dx * dx + dy * dy))
        if _poly_empty(R, ri):
            return np.inf
        if point_location(px, py, R, ri) != 0:
            return 0.0
        return _boundary_dist(px, py, R, ri)
    if _poly_empty(L, li) or _poly_empty(R, ri):
        return np.inf
    if polygon_intersects(L, li, R, ri):
        if exclusive:
            from . import relate_ref

            # topological equality (GEOS equals): each covers the other
            if relate_ref.predicate("covers", L, li, R, ri) and relate_ref.predicate("covered_by", L, li, R, ri):
                return np.inf
        return 0.0
    best = np.inf
    for P, rp, Q, rq in ((L, li, R, ri), (R, ri, L, li)):
        for b, e in _rings_of(P, rp):
            for v in range(b, e):
                best = min(best, _boundary_dist(float(P.x[v]), float(P.y[v]), Q, rq))
    return best


def nearest_pairs(left_kind, right_kind, L, R, max_distance=None, return_all=True, exclusive=False):
    """Brute-force nearest: for every left row, the right rows at the least
    finite fp64 distance (<= max_distance when given); ties all kept, or the
    least right row when not return_all.  (left, right, distance) lexsorted."""
    nl = len(L.x) if left_kind == 0 else len(L)
    nr = len(R.x) if right_kind == 0 else len(R)
    out_l, out_r, out_d = [], [], []
    for i in range(nl):
        d = np.array([distance_f64(left_kind, right_kind, L, R, i, j, exclusive) for j in range(nr)])
        fin = np.isfinite(d)
        if not fin.any():
            continue
        m = d[fin].min()
        if max_distance is not None and not (m <= max_distance):
            continue
        js = np.flatnonzero(d == m)
        if not return_all:
            js = js[:1]
        out_l += [i] * len(js)
        out_r += js.tolist()
        out_d += [m] * len(js)
    return (np.array(out_l, dtype=np.int64), np.array(out_r, dtype=np.int64), np.array(out_d, dtype=np.float64))
