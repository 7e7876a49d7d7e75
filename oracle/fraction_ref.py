"""Exact-rational restatements used to pin the C oracle (test infrastructure).

Every float is converted with ``Fraction(float)``, which is exact, so these
functions compute the mathematically exact answer of the reference algorithm
on the given fp64 inputs.  They are slow and meant for small inputs only.

Semantics followed (reference file:line):
  orientation sign   src/vibespatial/cuda/device_functions/orient2d.py:37-84
  ring classify      src/vibespatial/cuda/device_functions/point_in_ring.py:148-188
  polygon location   src/vibespatial/predicates/point_relations_kernels.py:37-95
  segment touch      src/vibespatial/predicates/polygon_kernels.py:131-161
  point-segment dist src/vibespatial/spatial/point_distance_kernels.py:17-39
                     (closest point by the clamped projection parameter);
                     dwithin = distance <= d (spatial/nearest.py:2855-3016),
                     decided here on exact squared distances
"""

from __future__ import annotations

from fractions import Fraction

__all__ = [
    "orient",
    "point_on_segment",
    "location_by_geometry",
    "segments_touch",
    "dist2_point_point",
    "dist2_point_segment",
    "within_distance",
]


def orient(ax, ay, bx, by, cx, cy) -> int:
    """Exact sign of (ax-cx)(by-cy) - (ay-cy)(bx-cx)."""
    F = Fraction
    d = (F(ax) - F(cx)) * (F(by) - F(cy)) - (F(ay) - F(cy)) * (F(bx) - F(cx))
    return (d > 0) - (d < 0)


def point_on_segment(px, py, ax, ay, bx, by) -> bool:
    """Exact: p lies on the closed segment a-b."""
    if orient(ax, ay, bx, by, px, py) != 0:
        return False
    return min(ax, bx) <= px <= max(ax, bx) and min(ay, by) <= py <= max(ay, by)


def _ring_parity_and_boundary(px, py, ring):
    """Independent formulation: boundary = p on any edge (exact), interior
    parity = number of edges crossing the rightward ray under the half-open
    rule (a.y > p.y) != (b.y > p.y) with the crossing x strictly right of p.

    This does not reuse the reference's shortcut structure: the crossing x is
    computed as an exact rational and compared with px."""
    n = len(ring)
    on_boundary = False
    crossings = 0
    for k in range(n):
        ax, ay = ring[k - 1]
        bx, by = ring[k]
        if point_on_segment(px, py, ax, ay, bx, by):
            on_boundary = True
        if (ay > py) != (by > py):
            F = Fraction
            t = (F(py) - F(ay)) / (F(by) - F(ay))
            xc = F(ax) + t * (F(bx) - F(ax))
            if xc > F(px):
                crossings += 1
    return on_boundary, crossings & 1


def location_by_geometry(px, py, parts) -> int:
    """Location code of p for a (multi)polygon given as a list of parts, each a
    list of rings, each a list of (x, y) tuples (closed or not).

    0 outside, 1 boundary, 2 interior.  Parts are visited in order and the first
    part that yields boundary or interior decides (point_relations_kernels.py:63-95).
    """
    for rings in parts:
        inside = 0
        for ring in rings:
            if len(ring) < 2:
                continue
            b, parity = _ring_parity_and_boundary(px, py, ring)
            if b:
                return 1
            inside ^= parity
        if inside:
            return 2
    return 0


def segments_touch(a, b, c, d) -> bool:
    """Exact: closed segments a-b and c-d share at least one point.

    Formulated independently of the reference's orientation case split: solve
    the 2x2 system in rationals; for parallel segments test collinear overlap by
    projecting onto the dominant axis."""
    F = Fraction
    ax, ay, bx, by = map(F, (*a, *b))
    cx, cy, dx, dy = map(F, (*c, *d))
    rx, ry = bx - ax, by - ay
    sx, sy = dx - cx, dy - cy
    if rx == 0 and ry == 0:
        return point_on_segment(a[0], a[1], c[0], c[1], d[0], d[1])
    if sx == 0 and sy == 0:
        return point_on_segment(c[0], c[1], a[0], a[1], b[0], b[1])
    den = rx * sy - ry * sx
    qpx, qpy = cx - ax, cy - ay
    if den != 0:
        t = (qpx * sy - qpy * sx) / den
        u = (qpx * ry - qpy * rx) / den
        return 0 <= t <= 1 and 0 <= u <= 1
    # parallel: collinear?
    if qpx * ry - qpy * rx != 0:
        return False
    def overlap(p0, p1, q0, q1):
        return max(min(p0, p1), min(q0, q1)) <= min(max(p0, p1), max(q0, q1))

    return overlap(ax, bx, cx, dx) and overlap(ay, by, cy, dy)


def dist2_point_point(px, py, qx, qy) -> Fraction:
    """Exact squared distance between two points."""
    F = Fraction
    dx, dy = F(px) - F(qx), F(py) - F(qy)
    return dx * dx + dy * dy


def dist2_point_segment(px, py, ax, ay, bx, by) -> Fraction:
    """Exact squared distance from p to the closed segment a-b: the projection
    parameter t = (p - a).(b - a) / |b - a|^2 clamped to [0, 1]."""
    F = Fraction
    ux, uy = F(bx) - F(ax), F(by) - F(ay)
    wx, wy = F(px) - F(ax), F(py) - F(ay)
    l2 = ux * ux + uy * uy
    t = F(0) if l2 == 0 else min(max((ux * wx + uy * wy) / l2, F(0)), F(1))
    ex, ey = wx - t * ux, wy - t * uy
    return ex * ex + ey * ey


def within_distance(d2: Fraction, d: float) -> bool:
    """d2 <= d^2 for d >= 0 (NaN / negative d never matches)."""
    if not (d >= 0.0):
        return False
    if d == float("inf"):
        return True
    return d2 <= Fraction(d) * Fraction(d)
