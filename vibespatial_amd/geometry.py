"""Owned geometry storage: fp64 structure-of-arrays with int32 offsets.

Restates the layout contract of the reference's OwnedGeometryArray /
FamilyGeometryBuffer (src/vibespatial/geometry/owned.py:256-267, 646-664;
schema src/vibespatial/geometry/buffers.py:140-250) for the two families the
join path uses:

* ``PointArray``: x, y fp64; a null or empty point has NaN coordinates.
* ``PolygonArray``: polygons and multipolygons in one layout,
  ``geom_offsets`` (row -> part), ``part_offsets`` (part -> ring),
  ``ring_offsets`` (ring -> coordinate), rings closed (first == last).
  A polygon row has one part; an empty or null row has none.

Small scalar types (``Point``, ``Polygon``, ``MultiPolygon``, ``box``) stand in
for the Shapely objects the public API accepts.  ``GeometryArray`` holds a
possibly mixed column as per-family arrays plus row maps, the way the
reference keeps ``tags`` and ``family_row_offsets`` (owned.py:41-45).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Iterable, Sequence

import numpy as np

__all__ = [
    "Point",
    "Polygon",
    "MultiPolygon",
    "box",
    "PointArray",
    "PolygonArray",
    "GeometryArray",
    "points_from_xy",
    "box_array",
    "as_geometry_array",
]


# ---------------------------------------------------------------------------
# Scalars
# ---------------------------------------------------------------------------
class _Geom:
    geom_type = "Geometry"

    @property
    def is_empty(self) -> bool:
        raise NotImplementedError

    @property
    def bounds(self):
        raise NotImplementedError


class Point(_Geom):
    geom_type = "Point"

    def __init__(self, *args):
        if len(args) == 0:
            self.x = self.y = float("nan")
        elif len(args) == 1:
            self.x, self.y = (float(v) for v in args[0])
        else:
            self.x, self.y = float(args[0]), float(args[1])

    @property
    def is_empty(self) -> bool:
        return not (self.x == self.x and self.y == self.y)

    @property
    def bounds(self):
        return (self.x, self.y, self.x, self.y)

    def __repr__(self):
        return "POINT EMPTY" if self.is_empty else f"POINT ({self.x!r} {self.y!r})"

    def __eq__(self, other):
        return isinstance(other, Point) and (self.x, self.y) == (other.x, other.y)

    def __hash__(self):
        return hash((self.x, self.y))


def _ring(coords) -> np.ndarray:
    a = np.asarray(coords, dtype=np.float64).reshape(-1, 2)
    if len(a) and not np.array_equal(a[0], a[-1]):
        a = np.vstack([a, a[:1]])
    return a


class Polygon(_Geom):
    geom_type = "Polygon"

    def __init__(self, shell=None, holes=None):
        self.rings = []
        if shell is not None and len(shell):
            self.rings.append(_ring(shell))
            for h in holes or ():
                self.rings.append(_ring(h))

    @property
    def is_empty(self) -> bool:
        return not self.rings

    @property
    def exterior(self):
        return self.rings[0] if self.rings else np.empty((0, 2))

    @property
    def interiors(self):
        return self.rings[1:]

    @property
    def bounds(self):
        if not self.rings:
            return (float("nan"),) * 4
        a = np.vstack(self.rings)
        return (a[:, 0].min(), a[:, 1].min(), a[:, 0].max(), a[:, 1].max())

    @property
    def parts(self):
        return [self.rings] if self.rings else []

    def __repr__(self):
        return "POLYGON EMPTY" if self.is_empty else f"POLYGON ({len(self.rings)} rings)"


class MultiPolygon(_Geom):
    geom_type = "MultiPolygon"

    def __init__(self, polygons: Iterable = ()):
        self.polygons = []
        for p in polygons:
            if isinstance(p, Polygon):
                if not p.is_empty:
                    self.polygons.append(p)
            else:  # (shell, holes) tuple
                shell, holes = (p[0], p[1]) if len(p) == 2 and not np.isscalar(p[0][0]) else (p, None)
                self.polygons.append(Polygon(shell, holes))

    @property
    def is_empty(self) -> bool:
        return not self.polygons

    @property
    def geoms(self):
        return list(self.polygons)

    @property
    def bounds(self):
        if not self.polygons:
            return (float("nan"),) * 4
        a = np.vstack([r for p in self.polygons for r in p.rings])
        return (a[:, 0].min(), a[:, 1].min(), a[:, 0].max(), a[:, 1].max())

    @property
    def parts(self):
        return [p.rings for p in self.polygons]

    def __repr__(self):
        return "MULTIPOLYGON EMPTY" if self.is_empty else f"MULTIPOLYGON ({len(self.polygons)} parts)"


def box(minx, miny, maxx, maxy, ccw: bool = True) -> Polygon:
    """Axis-aligned rectangle polygon (Shapely's ``box`` vertex order)."""
    if ccw:
        shell = [(maxx, miny), (maxx, maxy), (minx, maxy), (minx, miny)]
    else:
        shell = [(minx, miny), (minx, maxy), (maxx, maxy), (maxx, miny)]
    return Polygon(shell)


# ---------------------------------------------------------------------------
# Arrays
# ---------------------------------------------------------------------------
@dataclass
class PointArray:
    x: np.ndarray
    y: np.ndarray

    def __post_init__(self):
        self.x = np.ascontiguousarray(self.x, dtype=np.float64)
        self.y = np.ascontiguousarray(self.y, dtype=np.float64)
        if self.x.shape != self.y.shape or self.x.ndim != 1:
            raise ValueError("x and y must be 1-d arrays of equal length")
        self._device = None

    family = "point"

    def __len__(self) -> int:
        return len(self.x)

    def bounds(self) -> np.ndarray:
        b = np.stack([self.x, self.y, self.x, self.y], axis=1)
        b[np.isnan(self.x) | np.isnan(self.y)] = np.nan
        return b

    def take(self, idx) -> "PointArray":
        idx = np.asarray(idx)
        return PointArray(self.x[idx], self.y[idx])

    def scalar(self, i: int) -> Point:
        return Point(self.x[i], self.y[i])

    @property
    def nbytes(self) -> int:
        return self.x.nbytes + self.y.nbytes


def _ranges(starts: np.ndarray, counts: np.ndarray) -> np.ndarray:
    """Concatenation of arange(starts[i], starts[i] + counts[i]) over i."""
    counts = np.asarray(counts, dtype=np.int64)
    total = int(counts.sum())
    if total == 0:
        return np.empty(0, dtype=np.int64)
    first = np.cumsum(counts) - counts
    return np.repeat(np.asarray(starts, dtype=np.int64) - first, counts) + np.arange(total, dtype=np.int64)


def _offsets(counts: np.ndarray) -> np.ndarray:
    return np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)


def concat_polygons(arrays: Sequence["PolygonArray"]) -> "PolygonArray":
    """Rows of the arrays one after another."""
    xs, ys, ro, po, go = [], [], [np.zeros(1, np.int64)], [np.zeros(1, np.int64)], [np.zeros(1, np.int64)]
    nc = nr = npart = 0
    for a in arrays:
        xs.append(a.x)
        ys.append(a.y)
        ro.append(a.ring_offsets[1:].astype(np.int64) + nc)
        po.append(a.part_offsets[1:].astype(np.int64) + nr)
        go.append(a.geom_offsets[1:].astype(np.int64) + npart)
        nc += len(a.x)
        nr += len(a.ring_offsets) - 1
        npart += len(a.part_offsets) - 1
    x = np.concatenate(xs) if xs else np.empty(0)
    y = np.concatenate(ys) if ys else np.empty(0)
    return PolygonArray(x, y, np.concatenate(ro), np.concatenate(po), np.concatenate(go))


def box_array(bounds: np.ndarray, ccw: bool = True) -> "PolygonArray":
    """Rectangles from an (n, 4) [minx, miny, maxx, maxy] array, one closed
    ring of five coordinates each (the vectorised form of ``box``)."""
    b = np.asarray(bounds, dtype=np.float64).reshape(-1, 4)
    n = len(b)
    x0, y0, x1, y1 = b[:, 0], b[:, 1], b[:, 2], b[:, 3]
    if ccw:
        xs = np.stack([x1, x1, x0, x0, x1], axis=1)
        ys = np.stack([y0, y1, y1, y0, y0], axis=1)
    else:
        xs = np.stack([x0, x0, x1, x1, x0], axis=1)
        ys = np.stack([y0, y1, y1, y0, y0], axis=1)
    k = np.arange(n + 1, dtype=np.int64)
    return PolygonArray(xs.ravel(), ys.ravel(), 5 * k, k, k)


def points_from_xy(x, y) -> PointArray:
    return PointArray(np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64))


@dataclass
class PolygonArray:
    x: np.ndarray
    y: np.ndarray
    ring_offsets: np.ndarray
    part_offsets: np.ndarray
    geom_offsets: np.ndarray

    family = "polygon"

    def __post_init__(self):
        self.x = np.ascontiguousarray(self.x, dtype=np.float64)
        self.y = np.ascontiguousarray(self.y, dtype=np.float64)
        self.ring_offsets = np.ascontiguousarray(self.ring_offsets, dtype=np.int32)
        self.part_offsets = np.ascontiguousarray(self.part_offsets, dtype=np.int32)
        self.geom_offsets = np.ascontiguousarray(self.geom_offsets, dtype=np.int32)
        if len(self.x) >= 2**31 - 1:
            raise OverflowError("more than 2^31-1 coordinates in one polygon array")
        self._device = None

    def __len__(self) -> int:
        return len(self.geom_offsets) - 1

    @classmethod
    def from_geometries(cls, geoms: Sequence) -> "PolygonArray":
        """From Polygon / MultiPolygon / None scalars (None and empty rows have no parts)."""
        xs, ys = [], []
        ring_off, part_off, geom_off = [0], [0], [0]
        n_coords = 0
        for g in geoms:
            parts = [] if g is None else g.parts
            for rings in parts:
                for r in rings:
                    xs.append(r[:, 0])
                    ys.append(r[:, 1])
                    n_coords += len(r)
                    ring_off.append(n_coords)
                part_off.append(len(ring_off) - 1)
            geom_off.append(len(part_off) - 1)
        x = np.concatenate(xs) if xs else np.empty(0)
        y = np.concatenate(ys) if ys else np.empty(0)
        return cls(x, y, np.asarray(ring_off), np.asarray(part_off), np.asarray(geom_off))

    def bounds(self) -> np.ndarray:
        n = len(self)
        out = np.full((n, 4), np.nan)
        c0 = self.ring_offsets[self.part_offsets[self.geom_offsets[:-1]]]
        c1 = self.ring_offsets[self.part_offsets[self.geom_offsets[1:]]]
        nonempty = c1 > c0
        if nonempty.any():
            # coordinates of consecutive non-empty rows are contiguous, so one
            # reduceat per column covers every row (x[c1 of the last row:] is empty)
            idx = np.flatnonzero(nonempty)
            starts = c0[idx]
            end = int(c1[idx[-1]])
            x, y = self.x[:end], self.y[:end]
            out[idx, 0] = np.minimum.reduceat(x, starts)
            out[idx, 1] = np.minimum.reduceat(y, starts)
            out[idx, 2] = np.maximum.reduceat(x, starts)
            out[idx, 3] = np.maximum.reduceat(y, starts)
        return out

    def n_coords(self) -> int:
        return len(self.x)

    def take(self, idx) -> "PolygonArray":
        """Rows idx (in that order), by offset arithmetic on the buffers."""
        idx = np.asarray(idx, dtype=np.int64)
        g, p, r = self.geom_offsets.astype(np.int64), self.part_offsets.astype(np.int64), \
            self.ring_offsets.astype(np.int64)
        n_parts = g[idx + 1] - g[idx]
        parts = _ranges(g[idx], n_parts)
        n_rings = p[parts + 1] - p[parts]
        rings = _ranges(p[parts], n_rings)
        n_coords = r[rings + 1] - r[rings]
        coords = _ranges(r[rings], n_coords)
        return PolygonArray(self.x[coords], self.y[coords], _offsets(n_coords), _offsets(n_rings),
                            _offsets(n_parts))

    def scalar(self, i: int):
        parts = []
        for p in range(self.geom_offsets[i], self.geom_offsets[i + 1]):
            rings = []
            for r in range(self.part_offsets[p], self.part_offsets[p + 1]):
                a, b = self.ring_offsets[r], self.ring_offsets[r + 1]
                rings.append(np.stack([self.x[a:b], self.y[a:b]], axis=1))
            parts.append(rings)
        if not parts:
            return Polygon()
        if len(parts) == 1:
            return Polygon(parts[0][0], parts[0][1:])
        return MultiPolygon([Polygon(r[0], r[1:]) for r in parts])

    @property
    def nbytes(self) -> int:
        return (
            self.x.nbytes + self.y.nbytes + self.ring_offsets.nbytes + self.part_offsets.nbytes
            + self.geom_offsets.nbytes
        )


# ---------------------------------------------------------------------------
# Mixed column
# ---------------------------------------------------------------------------
TAG_NULL = -1
TAG_POINT = 0
TAG_POLYGON = 1


class GeometryArray:
    """A geometry column: one PointArray and/or one PolygonArray plus row maps.

    ``point_rows[i]`` is the column row of ``points[i]``; likewise for
    polygons.  Null/empty rows belong to no family and never match.
    """

    def __init__(self, tags, points: PointArray | None, point_rows, polygons: PolygonArray | None,
                 polygon_rows):
        self.tags = np.asarray(tags, dtype=np.int8)
        self.points = points
        self.point_rows = None if point_rows is None else np.asarray(point_rows, dtype=np.int64)
        self.polygons = polygons
        self.polygon_rows = None if polygon_rows is None else np.asarray(polygon_rows, dtype=np.int64)
        self._sindex = None

    def __len__(self) -> int:
        return len(self.tags)

    @classmethod
    def from_points(cls, pa: PointArray) -> "GeometryArray":
        n = len(pa)
        return cls(np.zeros(n, np.int8), pa, np.arange(n), None, None)

    @classmethod
    def from_polygons(cls, pa: PolygonArray) -> "GeometryArray":
        n = len(pa)
        return cls(np.ones(n, np.int8), None, None, pa, np.arange(n))

    @classmethod
    def from_scalars(cls, geoms: Sequence) -> "GeometryArray":
        geoms = list(geoms)
        tags = np.full(len(geoms), TAG_NULL, np.int8)
        pts, prow, polys, qrow = [], [], [], []
        for i, g in enumerate(geoms):
            if g is None:
                continue
            if isinstance(g, Point):
                tags[i] = TAG_POINT
                pts.append((g.x, g.y))
                prow.append(i)
            elif isinstance(g, (Polygon, MultiPolygon)):
                tags[i] = TAG_POLYGON
                polys.append(g)
                qrow.append(i)
            else:
                raise TypeError(f"unsupported geometry type {type(g).__name__}")
        points = None
        if pts:
            a = np.asarray(pts, dtype=np.float64).reshape(-1, 2)
            points = PointArray(a[:, 0], a[:, 1])
        polygons = PolygonArray.from_geometries(polys) if polys else None
        return cls(tags, points, prow if pts else None, polygons, qrow if polys else None)

    @property
    def is_points(self) -> bool:
        return self.polygons is None and self.points is not None

    @property
    def is_polygons(self) -> bool:
        return self.points is None and self.polygons is not None

    def bounds(self) -> np.ndarray:
        out = np.full((len(self), 4), np.nan)
        if self.points is not None:
            out[self.point_rows] = self.points.bounds()
        if self.polygons is not None:
            out[self.polygon_rows] = self.polygons.bounds()
        return out

    def _family_index(self) -> np.ndarray:
        """Row -> index of the row inside its family array (-1 for null rows)."""
        fi = getattr(self, "_fam_idx", None)
        if fi is None:
            fi = np.full(len(self), -1, dtype=np.int64)
            if self.points is not None:
                fi[self.point_rows] = np.arange(len(self.point_rows))
            if self.polygons is not None:
                fi[self.polygon_rows] = np.arange(len(self.polygon_rows))
            self._fam_idx = fi
        return fi

    def take(self, idx) -> "GeometryArray":
        """Rows idx in that order; idx < 0 gives a null row."""
        idx = np.asarray(idx, dtype=np.int64)
        valid = idx >= 0
        src = np.maximum(idx, 0)
        tags = np.where(valid, self.tags[src] if len(self.tags) else TAG_NULL, TAG_NULL).astype(np.int8)
        fi = self._family_index()
        pts = prow = polys = qrow = None
        m = tags == TAG_POINT
        if m.any():
            pts = self.points.take(fi[idx[m]])
            prow = np.flatnonzero(m)
        m = tags == TAG_POLYGON
        if m.any():
            polys = self.polygons.take(fi[idx[m]])
            qrow = np.flatnonzero(m)
        return GeometryArray(tags, pts, prow, polys, qrow)

    @staticmethod
    def concat(arrays: Sequence["GeometryArray"]) -> "GeometryArray":
        """Rows of the columns one after another."""
        tags = np.concatenate([a.tags for a in arrays]) if arrays else np.empty(0, np.int8)
        base, prow, qrow, pts, polys = 0, [], [], [], []
        for a in arrays:
            if a.points is not None:
                pts.append(a.points)
                prow.append(a.point_rows + base)
            if a.polygons is not None:
                polys.append(a.polygons)
                qrow.append(a.polygon_rows + base)
            base += len(a)
        P = PointArray(np.concatenate([p.x for p in pts]), np.concatenate([p.y for p in pts])) if pts else None
        Q = concat_polygons(polys) if polys else None
        return GeometryArray(tags, P, np.concatenate(prow) if pts else None, Q, np.concatenate(qrow) if polys else None)

    def __getitem__(self, i: int):
        t = self.tags[i]
        if t == TAG_POINT:
            j = int(np.searchsorted(self.point_rows, i))
            return self.points.scalar(j)
        if t == TAG_POLYGON:
            j = int(np.searchsorted(self.polygon_rows, i))
            return self.polygons.scalar(j)
        return None


def as_geometry_array(obj) -> GeometryArray:
    """Coerce a GeoSeries / GeometryArray / PointArray / PolygonArray / scalar list."""
    if isinstance(obj, GeometryArray):
        return obj
    if hasattr(obj, "values") and isinstance(getattr(obj, "values"), GeometryArray):
        return obj.values
    if isinstance(obj, (PointArray, PolygonArray)):
        # one wrapper per array (its row tables are O(n) to build), like the
        # column of a GeoSeries
        ga = getattr(obj, "_geometry_array", None)
        if ga is None:
            ga = GeometryArray.from_points(obj) if isinstance(obj, PointArray) else GeometryArray.from_polygons(obj)
            obj._geometry_array = ga
        return ga
    if isinstance(obj, _Geom) or obj is None:
        return GeometryArray.from_scalars([obj])
    if isinstance(obj, (list, tuple, np.ndarray)):
        return GeometryArray.from_scalars(list(obj))
    raise TypeError(f"cannot interpret {type(obj).__name__} as geometries")
