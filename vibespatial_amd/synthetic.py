"""Seeded synthetic workloads (NumPy PCG64 ``default_rng``), like the
reference's generators (src/vibespatial/testing/synthetic.py:36-326).

Every polygon produced here is valid (simple rings, holes strictly inside the
shell and pairwise disjoint), so even-odd location with exact boundary
detection is exactly GEOS's ``within`` / ``intersects``:

* star polygons: vertex angles are sorted uniform draws on [0, 2*pi), redrawn
  for any polygon whose largest angular gap (wrap-around included) is >= pi;
  every edge then subtends less than pi at the centre, so the ring is
  star-shaped about the centre and simple.
* holes fit in 0.9 x the shell's inscribed radius about the centre: one hole
  is concentric; 2-3 holes sit on a circle of half that radius in disjoint
  angular sectors.

Workloads:
  c1: 100k points U[0,1000]^2 (seed 0); 1000 stars (seed 1): centre
      U[0,1000]^2, R ~ U(5,30), n ~ U{8..64}, radius R*(1+U(-0.3,0.3)).
  c2: 10M points from a 64-component isotropic Gaussian mixture (means
      U[0,1e5]^2, sigma ~ U(200,3000), equal weights; seed 2); 100k stars
      (seed 3): centre U[0,1e5]^2, R ~ U(50,300), vertices
      clip(round(LogNormal(ln 24, 1.0)), 4, 10000), holes ~ U{0..3} with
      radius U(0.1,0.4)*R (shrunk to fit), hole vertices U{4..16}.
  c4: 50M points from 1024 clusters with Zipf(s=1.1) weights, N(0, 500 m)
      about centres U[0,1e5]^2 (seed 4); 1M axis-aligned boxes (seed 5)
      centred on fresh draws from the same mixture, width and height
      independent LogNormal(ln 50, 1.0).
  c5: 4096 polygon x polygon intersects jobs (job sizes seed 6, rows of
      each run of 64 jobs seed (6, run)): per job
      L = clip(LogNormal(ln 2000, 1.0), 1e3, 2e5) left footprints (stars with
      4-12 vertices, radius U(2.5, 20) m) and R = clip(LogNormal(ln 300, 1.0),
      10, 5000) right parcels (stars, radius U(40, 70) m, vertices
      clip(LogNormal(ln 40, 0.5), 4, 400)) on a job square of side
      120 sqrt(R) m; jobs on a 64-wide grid of 2e4 m cells (or all at the
      origin with overlap=True, so that only the job offsets separate them).
Sizes can be scaled down with n_points / n_polys for tests.
"""

from __future__ import annotations

import numpy as np

from .geometry import PointArray, PolygonArray

__all__ = ["star_polygons", "c1", "c2", "c3", "c4", "c5", "gmm_points", "zipf_points", "shared_edge_triangles",
           "Workload"]


class Workload:
    def __init__(self, name, points, polygons, predicate, **meta):
        self.name = name
        self.points = points
        self.polygons = polygons
        self.predicate = predicate
        self.meta = meta


def _sorted_angles(rng, counts: np.ndarray) -> np.ndarray:
    """Concatenated per-polygon sorted angles; polygons whose largest angular
    gap (wrap-around included) is >= pi are redrawn until none is left."""
    counts = np.asarray(counts, dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(counts)[:-1]])
    ang = np.empty(int(counts.sum()))
    todo = np.arange(len(counts))
    while len(todo):
        c = counts[todo]
        seg = np.repeat(np.arange(len(todo)), c)
        a = rng.uniform(0.0, 2 * np.pi, int(c.sum()))
        a = a[np.lexsort((a, seg))]
        o = np.concatenate([[0], np.cumsum(c)[:-1]])
        last = o + c - 1
        gaps = np.empty_like(a)
        gaps[:-1] = a[1:] - a[:-1]
        gaps[last] = a[o] + 2 * np.pi - a[last]
        ok = np.maximum.reduceat(gaps, o) < np.pi
        dst = np.repeat(off[todo], c) + (np.arange(len(a)) - np.repeat(o, c))
        keep = np.repeat(ok, c)
        ang[dst[keep]] = a[keep]
        todo = todo[~ok]
    return ang


def _sorted_angles_spacing(rng, counts: np.ndarray) -> np.ndarray:
    """Same distribution as _sorted_angles without a sort: the order statistics
    of n uniforms are the normalised partial sums of n + 1 exponential
    spacings; the ring's largest angular gap is its largest spacing, the
    wrap-around gap being E_1 + E_{n+1}.  Rings with a gap >= pi are redrawn.
    Rings are drawn in groups of equal vertex count (one (rings, n + 1) array
    of spacings per group and redraw)."""
    counts = np.asarray(counts, dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(counts)[:-1]])
    ang = np.empty(int(counts.sum()))
    order = np.argsort(counts, kind="stable")
    sc = counts[order]
    starts = np.flatnonzero(np.concatenate([[True], sc[1:] != sc[:-1]])) if len(sc) else np.empty(0, np.int64)
    ends = np.concatenate([starts[1:], [len(sc)]])
    for g0, g1 in zip(starts, ends):
        n = int(sc[g0])
        todo = order[g0:g1]
        while len(todo):
            e = rng.exponential(1.0, (len(todo), n + 1))
            cs = np.cumsum(e, axis=1)
            tot = cs[:, n]
            gap = e[:, 0] + e[:, n]
            if n > 1:
                gap = np.maximum(gap, e[:, 1:n].max(axis=1))
            ok = gap < 0.5 * tot
            rows = todo[ok]
            ang[off[rows][:, None] + np.arange(n)] = (2 * np.pi) * (cs[ok, :n] / tot[ok, None])
            todo = todo[~ok]
    return ang


def _inscribed_radius(cx, cy, x, y, counts):
    """Per ring: min distance from its centre to its (closed) edges."""
    off = np.concatenate([[0], np.cumsum(counts)[:-1]])
    seg = np.repeat(np.arange(len(counts)), counts)
    nxt = np.arange(len(x)) + 1
    last = off + counts - 1
    nxt[last] = off
    ax, ay = x - cx[seg], y - cy[seg]
    bx, by = x[nxt] - cx[seg], y[nxt] - cy[seg]
    dx, dy = bx - ax, by - ay
    L2 = dx * dx + dy * dy
    t = np.clip(-(ax * dx + ay * dy) / np.where(L2 > 0, L2, 1.0), 0.0, 1.0)
    px, py = ax + t * dx, ay + t * dy
    d = np.sqrt(px * px + py * py)
    return np.minimum.reduceat(d, off)


def star_rings(rng, cx, cy, base_r, counts, jitter=0.3, spacing=False):
    """Vertices of star rings (not closed), concatenated; counts per ring."""
    counts = np.asarray(counts, dtype=np.int64)
    ang = _sorted_angles_spacing(rng, counts) if spacing else _sorted_angles(rng, counts)
    seg = np.repeat(np.arange(len(counts)), counts)
    rad = base_r[seg] * (1.0 + rng.uniform(-jitter, jitter, len(ang)))
    return cx[seg] + rad * np.cos(ang), cy[seg] + rad * np.sin(ang)


def _assemble(rings_x, rings_y, ring_counts, rings_per_poly, reverse_mask=None) -> PolygonArray:
    """Close each ring and build offsets (one part per polygon)."""
    ring_counts = np.asarray(ring_counts, dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(ring_counts)])
    nring = len(ring_counts)
    closed_counts = ring_counts + 1
    coff = np.concatenate([[0], np.cumsum(closed_counts)])
    X = np.empty(int(coff[-1]))
    Y = np.empty(int(coff[-1]))
    seg = np.repeat(np.arange(nring), ring_counts)
    local = np.arange(int(off[-1])) - off[seg]
    if reverse_mask is not None:
        rev = reverse_mask[seg]
        local = np.where(rev, ring_counts[seg] - 1 - local, local)
    X[coff[seg] + local] = rings_x
    Y[coff[seg] + local] = rings_y
    # closing vertex = first vertex of the (possibly reversed) ring
    X[coff[1:] - 1] = X[coff[:-1]]
    Y[coff[1:] - 1] = Y[coff[:-1]]
    part_offsets = np.concatenate([[0], np.cumsum(rings_per_poly)])
    geom_offsets = np.arange(len(rings_per_poly) + 1)
    return PolygonArray(X, Y, coff, part_offsets, geom_offsets)


def star_polygons(rng, centres_x, centres_y, R, n_vertices, holes=None, hole_vertices=None,
                  hole_frac=None, jitter=0.3, spacing=False) -> PolygonArray:
    """Star polygons with optional holes; returns a PolygonArray (shell CCW, holes CW)."""
    m = len(R)
    n_vertices = np.asarray(n_vertices, dtype=np.int64)
    sx, sy = star_rings(rng, centres_x, centres_y, R, n_vertices, jitter, spacing)
    if holes is None:
        return _assemble(sx, sy, n_vertices, np.ones(m, dtype=np.int64))
    holes = np.asarray(holes, dtype=np.int64)
    r_in = _inscribed_radius(centres_x, centres_y, sx, sy, n_vertices)
    hole_poly = np.repeat(np.arange(m), holes)
    hk = np.concatenate([np.arange(h) for h in holes]) if len(hole_poly) else np.empty(0, np.int64)
    hcount = holes[hole_poly]
    phi0 = rng.uniform(0, 2 * np.pi, m)[hole_poly]
    d = np.where(hcount == 1, 0.0, 0.5 * r_in[hole_poly])
    fit = np.where(hcount == 1, 0.9 * r_in[hole_poly],
                   0.9 * np.minimum(0.5 * r_in[hole_poly], d * np.sin(np.pi / np.maximum(hcount, 1))))
    want = hole_frac * R[hole_poly]
    base = np.minimum(want, fit / (1.0 + jitter))
    ang = phi0 + 2 * np.pi * hk / np.maximum(hcount, 1)
    hcx = centres_x[hole_poly] + d * np.cos(ang)
    hcy = centres_y[hole_poly] + d * np.sin(ang)
    hv = np.asarray(hole_vertices, dtype=np.int64)
    hx, hy = star_rings(rng, hcx, hcy, base, hv, jitter)
    # interleave rings: polygon i = shell i then its holes
    ring_poly = np.concatenate([np.arange(m), hole_poly])
    ring_is_hole = np.concatenate([np.zeros(m, bool), np.ones(len(hole_poly), bool)])
    order = np.lexsort((ring_is_hole, ring_poly))
    counts_all = np.concatenate([n_vertices, hv])
    xs_rings = np.split(np.concatenate([sx, hx]), np.cumsum(counts_all)[:-1])
    ys_rings = np.split(np.concatenate([sy, hy]), np.cumsum(counts_all)[:-1])
    rx = np.concatenate([xs_rings[i] for i in order]) if len(order) else np.empty(0)
    ry = np.concatenate([ys_rings[i] for i in order]) if len(order) else np.empty(0)
    return _assemble(rx, ry, counts_all[order], 1 + holes, reverse_mask=ring_is_hole[order])


def c1(n_points: int = 100_000, n_polys: int = 1_000) -> Workload:
    rp = np.random.default_rng(0)
    px = rp.uniform(0, 1000, n_points)
    py = rp.uniform(0, 1000, n_points)
    rg = np.random.default_rng(1)
    cx = rg.uniform(0, 1000, n_polys)
    cy = rg.uniform(0, 1000, n_polys)
    R = rg.uniform(5, 30, n_polys)
    nv = rg.integers(8, 65, n_polys)
    polys = star_polygons(rg, cx, cy, R, nv)
    return Workload("c1", PointArray(px, py), polys, "within")


def gmm_points(rng, n, n_comp=64, extent=1e5, sigma=(200.0, 3000.0)):
    mx = rng.uniform(0, extent, n_comp)
    my = rng.uniform(0, extent, n_comp)
    sg = rng.uniform(sigma[0], sigma[1], n_comp)
    comp = rng.integers(0, n_comp, n)
    x = mx[comp] + sg[comp] * rng.standard_normal(n)
    y = my[comp] + sg[comp] * rng.standard_normal(n)
    return x, y


def c2(n_points: int = 10_000_000, n_polys: int = 100_000, extent: float = 1e5) -> Workload:
    rp = np.random.default_rng(2)
    x, y = gmm_points(rp, n_points, extent=extent)
    rg = np.random.default_rng(3)
    cx = rg.uniform(0, extent, n_polys)
    cy = rg.uniform(0, extent, n_polys)
    R = rg.uniform(50, 300, n_polys)
    nv = np.clip(np.round(rg.lognormal(np.log(24.0), 1.0, n_polys)), 4, 10_000).astype(np.int64)
    holes = rg.integers(0, 4, n_polys)
    n_holes = int(holes.sum())
    hole_vertices = rg.integers(4, 17, n_holes)
    hole_frac = rg.uniform(0.1, 0.4, n_holes)
    polys = star_polygons(rg, cx, cy, R, nv, holes=holes, hole_vertices=hole_vertices, hole_frac=hole_frac)
    return Workload("c2", PointArray(x, y), polys, "within")


def _zipf_mixture(n_clusters=1024, s=1.1, extent=1e5, seed=4):
    rng = np.random.default_rng(seed)
    w = 1.0 / np.arange(1, n_clusters + 1) ** s
    w /= w.sum()
    return rng, w, rng.uniform(0, extent, n_clusters), rng.uniform(0, extent, n_clusters)


def zipf_points(rng, n, w, cx, cy, sigma=500.0):
    c = rng.choice(len(w), n, p=w)
    return cx[c] + sigma * rng.standard_normal(n), cy[c] + sigma * rng.standard_normal(n)


def c4(n_points: int = 50_000_000, n_boxes: int = 1_000_000):
    """(points, boxes): the point-tree box query workload; boxes (n, 4) as
    (xmin, ymin, xmax, ymax)."""
    rng, w, cx, cy = _zipf_mixture()
    x, y = zipf_points(rng, n_points, w, cx, cy)
    rb = np.random.default_rng(5)
    bx, by = zipf_points(rb, n_boxes, w, cx, cy)
    bw = rb.lognormal(np.log(50.0), 1.0, n_boxes)
    bh = rb.lognormal(np.log(50.0), 1.0, n_boxes)
    boxes = np.column_stack([bx - bw / 2, by - bh / 2, bx + bw / 2, by + bh / 2])
    return Workload("c4", PointArray(x, y), None, "intersects", boxes=boxes)


def _c5_chunk(k, jobs, nl, nr, side, ox, oy):
    """Right parcels and left footprints of jobs[0] .. jobs[-1], drawn from
    the chunk's own generator (seed (6, k)), so the result does not depend on
    how chunks are spread over threads."""
    rng = np.random.default_rng([6, k])
    jl = np.repeat(jobs, nl[jobs])
    jr = np.repeat(jobs, nr[jobs])
    rcx = ox[jr] + rng.uniform(0, 1, len(jr)) * side[jr]
    rcy = oy[jr] + rng.uniform(0, 1, len(jr)) * side[jr]
    rR = rng.uniform(40.0, 70.0, len(jr))
    rnv = np.clip(np.round(rng.lognormal(np.log(40.0), 0.5, len(jr))), 4, 400).astype(np.int64)
    right = star_polygons(rng, rcx, rcy, rR, rnv, spacing=True)
    lcx = ox[jl] + rng.uniform(0, 1, len(jl)) * side[jl]
    lcy = oy[jl] + rng.uniform(0, 1, len(jl)) * side[jl]
    lR = rng.uniform(2.5, 20.0, len(jl))
    lnv = rng.integers(4, 13, len(jl))
    left = star_polygons(rng, lcx, lcy, lR, lnv, spacing=True)
    return left, right


def _concat_polygons(parts) -> PolygonArray:
    """Row-wise concatenation of PolygonArrays (offsets rebased)."""
    xs, ys, rings, partso, geoms = [], [], [np.zeros(1, np.int64)], [np.zeros(1, np.int64)], [np.zeros(1, np.int64)]
    nc = nr_ = npt = 0
    for P in parts:
        xs.append(P.x)
        ys.append(P.y)
        rings.append(np.asarray(P.ring_offsets[1:], np.int64) + nc)
        partso.append(np.asarray(P.part_offsets[1:], np.int64) + nr_)
        geoms.append(np.asarray(P.geom_offsets[1:], np.int64) + npt)
        nc += len(P.x)
        nr_ += len(P.ring_offsets) - 1
        npt += len(P.part_offsets) - 1
    return PolygonArray(np.concatenate(xs), np.concatenate(ys), np.concatenate(rings), np.concatenate(partso),
                        np.concatenate(geoms))


def c5(n_jobs: int = 4096, left_median: float = 2000.0, right_median: float = 300.0, overlap: bool = False,
       left_clip=(1_000, 200_000), right_clip=(10, 5_000), threads: int | None = None):
    """(left, right, left_job, right_job): PolygonArrays of every job's rows
    and int64 row offsets per job (n_jobs + 1).  Job sizes come from seed 6;
    the rows of each run of 64 jobs from the generator seeded (6, run), built
    on `threads` threads (default: the usable cores, at most 16)."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(6)
    nl = np.clip(np.round(rng.lognormal(np.log(left_median), 1.0, n_jobs)), *left_clip).astype(np.int64)
    nr = np.clip(np.round(rng.lognormal(np.log(right_median), 1.0, n_jobs)), *right_clip).astype(np.int64)
    side = 120.0 * np.sqrt(nr)
    ox = np.zeros(n_jobs) if overlap else (np.arange(n_jobs) % 64) * 2.0e4
    oy = np.zeros(n_jobs) if overlap else (np.arange(n_jobs) // 64) * 2.0e4
    chunks = [np.arange(j, min(j + 64, n_jobs)) for j in range(0, n_jobs, 64)]
    threads = threads or min(16, len(os.sched_getaffinity(0)))
    with ThreadPoolExecutor(max(1, threads)) as ex:
        out = list(ex.map(lambda kj: _c5_chunk(kj[0], kj[1], nl, nr, side, ox, oy), enumerate(chunks)))
    left = _concat_polygons([o[0] for o in out])
    right = _concat_polygons([o[1] for o in out])
    lj = np.concatenate([[0], np.cumsum(nl)]).astype(np.int64)
    rj = np.concatenate([[0], np.cumsum(nr)]).astype(np.int64)
    return left, right, lj, rj


def shared_edge_triangles(rng, polys: PolygonArray, owners: np.ndarray, height: float = 0.5):
    """For each owner polygon, a triangle sharing one of its shell edges
    exactly (same two vertices, reversed) with its apex on the outer side."""
    tris = []
    for q in owners:
        c0 = int(polys.ring_offsets[polys.part_offsets[polys.geom_offsets[q]]])
        c1 = int(polys.ring_offsets[polys.part_offsets[polys.geom_offsets[q]] + 1])
        k = int(rng.integers(c0 + 1, c1))
        ax, ay, bx, by = polys.x[k - 1], polys.y[k - 1], polys.x[k], polys.y[k]
        # shell rings are counter-clockwise: the outer side is to the right of a->b
        nx, ny = (by - ay), -(bx - ax)
        mx, my = 0.5 * (ax + bx), 0.5 * (ay + by)
        tris.append([(bx, by), (ax, ay), (mx + height * nx, my + height * ny), (bx, by)])
    return tris


def c3(n_points: int = 10_000_000, n_polys: int = 100_000, shift=(5.0e5, 3.0e6), on_edge_frac: float = 0.05,
       shared_frac: float = 0.01, verify=None) -> Workload:
    """C2's layout shifted to UTM-scale offsets with exact degeneracies:
    on_edge_frac of the points sit on polygon vertices or on edge midpoints
    that lie exactly on the edge, and shared_frac of the polygons are replaced
    by triangles sharing an edge exactly with another polygon.

    `verify(px, py, polys, rows) -> location codes` (the test oracle) keeps
    only midpoints that are exactly on their edge; without it only vertices
    are used."""
    from .geometry import Polygon

    base = c2(n_points=n_points, n_polys=n_polys)
    P = base.polygons
    X = P.x + shift[0]
    Y = P.y + shift[1]
    P = PolygonArray(X, Y, P.ring_offsets, P.part_offsets, P.geom_offsets)
    rng = np.random.default_rng(4)
    n_shared = int(round(shared_frac * n_polys))
    if n_shared:
        owners = rng.choice(n_polys - n_shared, n_shared, replace=False)
        tris = shared_edge_triangles(rng, P, owners)
        tail = PolygonArray.from_geometries([Polygon(t) for t in tris])
        head_coords = int(P.ring_offsets[P.part_offsets[P.geom_offsets[n_polys - n_shared]]])
        head_rings = int(P.part_offsets[P.geom_offsets[n_polys - n_shared]])
        head_parts = int(P.geom_offsets[n_polys - n_shared])
        x = np.concatenate([P.x[:head_coords], tail.x])
        y = np.concatenate([P.y[:head_coords], tail.y])
        ring = np.concatenate([P.ring_offsets[:head_rings + 1], tail.ring_offsets[1:] + head_coords])
        part = np.concatenate([P.part_offsets[:head_parts + 1], tail.part_offsets[1:] + head_rings])
        geom = np.concatenate([P.geom_offsets[:n_polys - n_shared + 1], tail.geom_offsets[1:] + head_parts])
        P = PolygonArray(x, y, ring, part, geom)
    px = base.points.x + shift[0]
    py = base.points.y + shift[1]
    n_on = int(round(on_edge_frac * n_points))
    if n_on:
        idx = rng.choice(n_points, n_on, replace=False)
        rows = rng.integers(0, len(P), n_on)
        c0 = P.ring_offsets[P.part_offsets[P.geom_offsets[rows]]]
        c1 = P.ring_offsets[P.part_offsets[P.geom_offsets[rows + 1]]]
        k = c0 + 1 + (rng.random(n_on) * (c1 - c0 - 1)).astype(np.int64)
        k = np.minimum(k, c1 - 1)
        vx, vy = P.x[k], P.y[k]
        mx, my = 0.5 * (P.x[k - 1] + P.x[k]), 0.5 * (P.y[k - 1] + P.y[k])
        use_mid = rng.random(n_on) < 0.5
        if verify is not None:
            codes = verify(np.where(use_mid, mx, vx), np.where(use_mid, my, vy), P, rows)
            use_mid &= codes == 1
        else:
            use_mid[:] = False
        px[idx] = np.where(use_mid, mx, vx)
        py[idx] = np.where(use_mid, my, vy)
    return Workload("c3", PointArray(px, py), P, "within", on_edge=n_on, shared=n_shared)
