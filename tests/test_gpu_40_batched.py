"""GPU tests of the batched polygon x polygon intersects join (many
independent jobs in one launch sequence) against the CPU oracle, job by job."""

from __future__ import annotations

import os

import numpy as np
import pytest
import torch

import oracle
from vibespatial_amd import join as J
from vibespatial_amd import synthetic as S
from vibespatial_amd.device import DevicePolygons

pytestmark = pytest.mark.gpu

NTHREADS = min(16, len(os.sched_getaffinity(0)))


def _expected(left, right, lj, rj):
    el, er = [], []
    for j in range(len(lj) - 1):
        if lj[j + 1] == lj[j] or rj[j + 1] == rj[j]:
            continue
        A = left.take(np.arange(lj[j], lj[j + 1]))
        B = right.take(np.arange(rj[j], rj[j + 1]))
        l, r = oracle.sjoin_polygons_polygons(A, B, nthreads=NTHREADS)
        el.append(l + lj[j])
        er.append(r + rj[j])
    if not el:
        return np.empty(0, np.int64), np.empty(0, np.int64)
    return np.concatenate(el), np.concatenate(er)


def _run(left, right, lj, rj, **kw):
    A, B = DevicePolygons.from_host(left), DevicePolygons.from_host(right)
    l, r, st = J.batched_intersects(A, B, torch.from_numpy(lj).cuda(), torch.from_numpy(rj).cuda(), **kw)
    return l.cpu().numpy(), r.cpu().numpy(), st


def _small_batch(n_jobs=48):
    return S.c5(n_jobs=n_jobs, left_median=600, right_median=60, overlap=True, left_clip=(1, 4000),
                right_clip=(1, 400))


@pytest.mark.parametrize("tile", [(4, 256, 0), (4, 256, 1)])
@pytest.mark.parametrize("cell_grid", [True, False])
def test_batched_matches_oracle_per_job_with_overlapping_jobs(tile, cell_grid):
    print(f"batched 48 overlapping jobs, tile {tile}, cell grid {cell_grid}", flush=True)
    left, right, lj, rj = _small_batch()
    l, r, st = _run(left, right, lj, rj, tile=tile, cell_grid=cell_grid)
    el, er = _expected(left, right, lj, rj)
    assert np.array_equal(l, el) and np.array_equal(r, er)
    assert st.candidates >= st.pairs > 0
    if cell_grid:
        assert 0 < st.refined < st.candidates
    else:
        assert st.refined == st.candidates


def _rings_array(rings):
    from vibespatial_amd.geometry import PolygonArray

    xs = [np.asarray(x, np.float64) for x, _ in rings]
    ys = [np.asarray(y, np.float64) for _, y in rings]
    ro = np.concatenate([[0], np.cumsum([len(x) for x in xs])]).astype(np.int64)
    k = np.arange(len(rings) + 1, dtype=np.int64)
    return PolygonArray(np.concatenate(xs), np.concatenate(ys), ro, k, k)


def _touching_batch():
    """Job 0: parcels (stars of 40-200 vertices) and footprints placed on
    them at random, on their vertices (triangles pointing in and out), along
    their edges (triangles sharing an edge), inside and outside; job 1:
    parcels that get no cell grid (not closed, zero height, three
    coordinates) with footprints over them."""
    rng = np.random.default_rng(77)
    n_par = 12
    P = S.star_polygons(rng, rng.uniform(0, 400, n_par), rng.uniform(0, 400, n_par), rng.uniform(40, 70, n_par),
                        rng.integers(40, 200, n_par), spacing=True)
    parcels = [(P.x[P.ring_offsets[k]:P.ring_offsets[k + 1]], P.y[P.ring_offsets[k]:P.ring_offsets[k + 1]])
               for k in range(n_par)]
    feet = []
    F = S.star_polygons(rng, rng.uniform(-20, 420, 600), rng.uniform(-20, 420, 600), rng.uniform(2.5, 20, 600),
                        rng.integers(4, 13, 600))
    feet += [(F.x[F.ring_offsets[k]:F.ring_offsets[k + 1]], F.y[F.ring_offsets[k]:F.ring_offsets[k + 1]])
             for k in range(600)]
    for x, y in parcels:
        for k in rng.choice(len(x) - 1, 8, replace=False):
            vx, vy, d = x[k], y[k], 3.0
            for sgn in (1.0, -1.0):
                feet.append(([vx, vx + sgn * d, vx + sgn * d, vx], [vy, vy - d, vy + d, vy]))
                feet.append(([vx + sgn * d, vx, vx + sgn * d, vx + sgn * d], [vy - d, vy, vy + d, vy - d]))
            ax, ay, bx, by = x[k], y[k], x[k + 1], y[k + 1]
            for sgn in (1.0, -1.0):
                feet.append(([ax, bx, 0.5 * (ax + bx) + sgn * (by - ay), ax],
                             [ay, by, 0.5 * (ay + by) - sgn * (bx - ax), ay]))
    # job 1: rows without a grid
    odd = [([0, 50, 50, 0], [0, 0, 50, 50]),                     # not closed
           ([100, 150, 170, 100], [10, 10, 10, 10]),             # zero height
           ([300, 340, 300], [0, 20, 0])]                        # three coordinates
    odd_feet = [([x0 + 5, x0 + 15, x0 + 10, x0 + 5], [5, 5, 12, 5]) for x0 in (0, 100, 140, 300, 310)]
    odd_feet += [([x0 + 30, x0 + 60, x0 + 45, x0 + 30], [-5, -5, 30, -5]) for x0 in (0, 100, 300)]
    left = _rings_array(feet + odd_feet)
    right = _rings_array(parcels + odd)
    lj = np.array([0, len(feet), len(feet) + len(odd_feet)], np.int64)
    rj = np.array([0, len(parcels), len(parcels) + len(odd)], np.int64)
    return left, right, lj, rj


@pytest.mark.parametrize("cell_grid", [True, False])
def test_batched_footprints_touching_parcels_and_rows_without_grid(cell_grid):
    print(f"batched touching footprints, cell grid {cell_grid}", flush=True)
    left, right, lj, rj = _touching_batch()
    el, er = _expected(left, right, lj, rj)
    for tile in [(4, 256, 0), (4, 256, 1)]:
        l, r, st = _run(left, right, lj, rj, tile=tile, cell_grid=cell_grid)
        assert np.array_equal(l, el) and np.array_equal(r, er), tile
    if cell_grid:
        assert st.refined < st.candidates


def test_batched_refine_on_holes_multiparts_and_long_rings():
    """General rows in the batched refine: left rows with holes and several
    parts, right rings longer than 64 edges (their y-slab index), left rings
    longer than the LDS tile, against the oracle in both candidate orders."""
    from vibespatial_amd.geometry import MultiPolygon, PolygonArray

    rng = np.random.default_rng(12)
    m = 60
    holes = rng.integers(0, 2, m)
    nh = int(holes.sum())
    A = S.star_polygons(rng, rng.uniform(0, 200, m), rng.uniform(0, 200, m), rng.uniform(3, 15, m),
                        rng.integers(4, 90, m), holes=holes, hole_vertices=rng.integers(4, 8, nh),
                        hole_frac=rng.uniform(0.1, 0.3, nh))
    B = S.star_polygons(rng, rng.uniform(0, 200, 40), rng.uniform(0, 200, 40), rng.uniform(10, 40, 40),
                        rng.integers(4, 300, 40))
    extra = PolygonArray.from_geometries([MultiPolygon([A.scalar(0), A.scalar(1)]), A.scalar(2)])
    A2 = PolygonArray.from_geometries([A.scalar(i) for i in range(m)] + [extra.scalar(0), extra.scalar(1)])
    lj = np.array([0, 30, len(A2)], dtype=np.int64)
    rj = np.array([0, 25, 40], dtype=np.int64)
    el, er = _expected(A2, B, lj, rj)
    for tile in [(4, 256, 0), (2, 128, 1)]:
        l, r, _ = _run(A2, B, lj, rj, tile=tile)
        assert np.array_equal(l, el) and np.array_equal(r, er), tile


def test_batched_edge_cases_empty_and_single_row_jobs():
    left, right, lj, rj = _small_batch(12)
    # re-cut the job offsets: an empty-left job, an empty-right job, a one-parcel job
    nl, nr = np.diff(lj), np.diff(rj)
    nl[2] = 0
    nr[5] = 0
    nr[7] = 1
    lj2 = np.concatenate([[0], np.cumsum(nl)]).astype(np.int64)
    rj2 = np.concatenate([[0], np.cumsum(nr)]).astype(np.int64)
    L2 = left.take(np.arange(lj2[-1]))
    R2 = right.take(np.arange(rj2[-1]))
    l, r, _ = _run(L2, R2, lj2, rj2)
    el, er = _expected(L2, R2, lj2, rj2)
    assert np.array_equal(l, el) and np.array_equal(r, er)
    # no jobs at all
    l, r, st = _run(L2, R2, np.zeros(1, np.int64), np.zeros(1, np.int64))
    assert len(l) == 0 and st.pairs == 0


def test_batched_capacity_retry_and_every_tile_shape_agree():
    left, right, lj, rj = _small_batch(16)
    l0, r0, st = _run(left, right, lj, rj, capacity=8)
    assert st.retries == 1
    for tile in J.TILE_CANDIDATES:
        l, r, _ = _run(left, right, lj, rj, tile=tile)
        assert np.array_equal(l, l0) and np.array_equal(r, r0), tile
    A, B = DevicePolygons.from_host(left), DevicePolygons.from_host(right)
    best, times = J.autotune_batched(A, B, torch.from_numpy(lj).cuda(), torch.from_numpy(rj).cuda(), reps=1)
    assert best in J.TILE_CANDIDATES and len(times) == len(J.TILE_CANDIDATES)
    print(f"autotune: best {best}; " + ", ".join(f"{k}: {v:.3f} ms" for k, v in times.items()))


def test_batched_graph_replays_match_eager():
    """Captured once, replayed three times with eager work and host syncs in
    between; every replay equals the eager result."""
    left, right, lj, rj = _small_batch(16)
    A, B = DevicePolygons.from_host(left), DevicePolygons.from_host(right)
    tl, tr = torch.from_numpy(lj).cuda(), torch.from_numpy(rj).cuda()
    l0, r0, st = J.batched_intersects(A, B, tl, tr)
    g = J.BatchedIntersectsGraph(A, B, tl, tr, capacity=st.staged + 64)
    junk = torch.zeros(1 << 20, device="cuda")
    for i in range(3):
        junk.add_(1.0)  # eager work between replays
        l, r, _ = g.replay()
        torch.cuda.synchronize()
        assert torch.equal(l, l0) and torch.equal(r, r0), i
