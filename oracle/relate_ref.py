"""Polygon x polygon relations in exact rationals -- TEST INFRASTRUCTURE ONLY
(the checker for the GPU refine of within / contains / covers / covered_by /
contains_properly / touches / overlaps / crosses between polygonal rows).

Restates the reference's DE-9IM predicate semantics for two areal rows
(predicates/polygon.py:19-40 masks; polygon_kernels.py:630-830) through the
boundary-piece classification the device code uses, in fractions.Fraction
arithmetic with independent sub-steps: cut points sorted by their exact
parameter along the edge, ring orientation from the shoelace area (the device
code uses the turn at the lowest vertex), and point location by the even-odd
rule on rational points.  Small inputs only.
"""

from __future__ import annotations

from fractions import Fraction as F


def _rows(P, row):
    """Parts of polygonal row `row`: [[ring, ...], ...], ring = [(x, y), ...] closed, Fractions."""
    out = []
    g, pt, r = P.geom_offsets, P.part_offsets, P.ring_offsets
    for p in range(g[row], g[row + 1]):
        rings = []
        for k in range(pt[p], pt[p + 1]):
            rings.append([(F(float(P.x[c])), F(float(P.y[c]))) for c in range(r[k], r[k + 1])])
        out.append(rings)
    return out


def _orient(a, b, c):
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def _sgn(v):
    return (v > 0) - (v < 0)


def _on_segment(a, b, p):
    return _orient(a, b, p) == 0 and min(a[0], b[0]) <= p[0] <= max(a[0], b[0]) and \
        min(a[1], b[1]) <= p[1] <= max(a[1], b[1])


def locate(p, parts):
    """0 outside, 1 boundary, 2 interior (even-odd per part) for a rational point."""
    for rings in parts:
        parity = 0
        for ring in rings:
            for a, b in zip(ring[:-1], ring[1:]):
                if _on_segment(a, b, p):
                    return 1
                if (a[1] > p[1]) != (b[1] > p[1]):
                    # x of the edge at height p.y, compared with p.x
                    xi = a[0] + (p[1] - a[1]) * (b[0] - a[0]) / (b[1] - a[1])
                    if xi > p[0]:
                        parity ^= 1
        if parity:
            return 2
    return 0


def _area2(ring):
    return sum(a[0] * b[1] - b[0] * a[1] for a, b in zip(ring[:-1], ring[1:]))


def _interior_left(ring, shell):
    ccw = _area2(ring) > 0
    return ccw if shell else not ccw


def _touch(a, b, c, d):
    o1, o2 = _sgn(_orient(a, b, c)), _sgn(_orient(a, b, d))
    o3, o4 = _sgn(_orient(c, d, a)), _sgn(_orient(c, d, b))
    if o1 * o2 < 0 and o3 * o4 < 0:
        return True
    return _on_segment(a, b, c) or _on_segment(a, b, d) or _on_segment(c, d, a) or _on_segment(c, d, b)


def _side(X, Y, first):
    """Boundary pieces of X against Y: dict of the flags of the device code."""
    S = dict(cross=False, touch=False, in_int=False, in_ext=False, same=False, all_in=True)
    y_edges = [(a, b, ring, k == 0) for rings in Y for k, ring in enumerate(rings)
               for a, b in zip(ring[:-1], ring[1:]) if a != b]
    y_verts = [v for rings in Y for ring in rings for v in ring]
    for rings in X:
        part_in = False
        for k, ring in enumerate(rings):
            left_x = _interior_left(ring, k == 0)
            for a, b in zip(ring[:-1], ring[1:]):
                if a == b:
                    continue
                for c, d, _, _ in y_edges:
                    if first and _touch(a, b, c, d):
                        S["touch"] = True
                    if _sgn(_orient(a, b, c)) * _sgn(_orient(a, b, d)) < 0 and \
                            _sgn(_orient(c, d, a)) * _sgn(_orient(c, d, b)) < 0:
                        S["cross"] = True
                # cut points: parameters of the Y vertices strictly inside the edge
                u = (b[0] - a[0], b[1] - a[1])
                L2 = u[0] * u[0] + u[1] * u[1]
                ts = sorted({((v[0] - a[0]) * u[0] + (v[1] - a[1]) * u[1]) / L2 for v in y_verts
                             if _orient(a, b, v) == 0})
                ts = [F(0)] + [t for t in ts if 0 < t < 1] + [F(1)]
                for t0, t1 in zip(ts[:-1], ts[1:]):
                    tm = (t0 + t1) / 2
                    m = (a[0] + tm * u[0], a[1] + tm * u[1])
                    loc = locate(m, Y)
                    if loc == 2:
                        S["in_int"] = True
                        part_in = True
                    elif loc == 0:
                        S["in_ext"] = True
                    else:
                        for c, d, yring, yshell in y_edges:
                            if _on_segment(c, d, m):
                                ly = _interior_left(yring, yshell)
                                same_dir = u[0] * (d[0] - c[0]) + u[1] * (d[1] - c[1]) > 0
                                if (ly == left_x) if same_dir else (ly != left_x):
                                    S["same"] = True
                                    part_in = True
                                break
        if not part_in:
            S["all_in"] = False
    return S


def relate(A, ra, B, rb):
    """(interiors_meet, a_in_b, b_in_a, touch) for polygonal rows ra of A and rb of B."""
    X, Y = _rows(A, ra), _rows(B, rb)
    sa, sb = _side(X, Y, True), _side(Y, X, False)
    cross = sa["cross"] or sb["cross"]
    meet = cross or sa["in_int"] or sb["in_int"] or sa["same"] or sb["same"]
    a_in_b = not cross and not sa["in_ext"] and not sb["in_int"] and sa["all_in"]
    b_in_a = not cross and not sb["in_ext"] and not sa["in_int"] and sb["all_in"]
    return meet, a_in_b, b_in_a, sa["touch"]


def intersects(A, ra, B, rb):
    X, Y = _rows(A, ra), _rows(B, rb)
    for rings in X:
        for ring in rings:
            for a, b in zip(ring[:-1], ring[1:]):
                for rings2 in Y:
                    for ring2 in rings2:
                        for c, d in zip(ring2[:-1], ring2[1:]):
                            if _touch(a, b, c, d):
                                return True
    return any(locate(ring[0], Y) for rings in X for ring in rings[:1]) or \
        any(locate(ring[0], X) for rings in Y for ring in rings[:1])


def predicate(pred, A, ra, B, rb) -> bool:
    """pred(A row, B row) for two polygonal rows."""
    if pred == "intersects":
        return intersects(A, ra, B, rb)
    if pred == "crosses":
        return False
    meet, a_in_b, b_in_a, touch = relate(A, ra, B, rb)
    if pred in ("within", "covered_by"):
        return a_in_b
    if pred in ("contains", "covers"):
        return b_in_a
    if pred == "contains_properly":
        return b_in_a and not touch
    if pred == "touches":
        return not meet and intersects(A, ra, B, rb)
    if pred == "overlaps":
        return meet and not a_in_b and not b_in_a
    raise ValueError(pred)
