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
def test_batched_matches_oracle_per_job_with_overlapping_jobs(tile):
    print(f"batched 48 overlapping jobs, tile {tile}", flush=True)
    left, right, lj, rj = _small_batch()
    l, r, st = _run(left, right, lj, rj, tile=tile)
    el, er = _expected(left, right, lj, rj)
    assert np.array_equal(l, el) and np.array_equal(r, er)
    assert st.candidates >= st.pairs > 0


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
