"""vs_sjoin_points_polygons_noalloc (work memory in a fixed device array)
against vs_sjoin_points_polygons (caller workspace) through the C-ABI: two
calls in a row, the second smaller, bit for bit, and both against the
oracle."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
from vibespatial_amd import _lib
from vibespatial_amd import join as J
from vibespatial_amd import synthetic as S
from vibespatial_amd.device import ptr, stream

pytestmark = pytest.mark.gpu


def _noalloc(Q, P, tree, predicate, cap):
    import torch

    L = _lib.lib()
    dev = Q.x.device
    out_l = torch.full((max(cap, 1),), -7, dtype=torch.int64, device=dev)
    out_r = torch.full((max(cap, 1),), -7, dtype=torch.int64, device=dev)
    res = torch.zeros(4, dtype=torch.int64, device=dev)
    rc = L.vs_sjoin_points_polygons_noalloc(
        ptr(Q.x), ptr(Q.y), len(Q), *P.args(), len(P), P.n_parts, P.n_rings, P.n_coords,
        ctypes.byref(tree.layout), ptr(tree.node_boxes), ptr(tree.item_ids), _lib.PRED[predicate], cap,
        ptr(out_l), ptr(out_r), ptr(res), stream())
    return rc, out_l, out_r, res


@pytest.mark.parametrize("predicate", ["within", "intersects"])
def test_noalloc_equals_workspace_entry_point_twice(predicate):
    import torch

    from vibespatial_amd.device import DevicePoints, DevicePolygons

    torch.cuda.set_device(0)
    cases = [S.c1(n_points=200_000, n_polys=2_000), S.c2(n_points=40_000, n_polys=500, extent=5e3)]
    for w in cases:  # the second call is smaller: the scratch still holds the first call's contents
        print(f"noalloc {predicate}: {len(w.points)} points x {len(w.polygons)} polygons", flush=True)
        Q = DevicePoints.from_host(w.points)
        P = DevicePolygons.from_host(w.polygons)
        tree = J.build_tree(J.polygon_bounds(P))
        el, er, st = J.sjoin_points_polygons(Q, P, tree, predicate)
        cap = st.pairs + 16
        need = _lib.lib().vs_sjoin_points_polygons_workspace_bytes(len(Q), len(P), P.n_coords, cap)
        assert need <= _lib.NOALLOC_SCRATCH_BYTES
        rc, ol, orr, res = _noalloc(Q, P, tree, predicate, cap)
        _lib.check(rc, "vs_sjoin_points_polygons_noalloc")
        r = res.cpu().numpy()
        k = int(r[0])
        assert k == st.pairs and int(r[1]) == st.staged and int(r[2]) == st.candidates
        assert torch.equal(ol[:k], el) and torch.equal(orr[:k], er)
        ol_h, or_h = oracle.sjoin_points_polygons(w.points.x, w.points.y, w.polygons, predicate)
        assert np.array_equal(ol[:k].cpu().numpy(), ol_h) and np.array_equal(orr[:k].cpu().numpy(), or_h)


def test_noalloc_capacity_overflow_and_size_limit():
    import torch

    from vibespatial_amd.device import DevicePoints, DevicePolygons

    torch.cuda.set_device(0)
    w = S.c1(n_points=50_000, n_polys=500)
    Q = DevicePoints.from_host(w.points)
    P = DevicePolygons.from_host(w.polygons)
    tree = J.build_tree(J.polygon_bounds(P))
    el, er, st = J.sjoin_points_polygons(Q, P, tree, "within")
    # too small a capacity: the count says how many, as for the workspace entry point
    rc, _, _, res = _noalloc(Q, P, tree, "within", 100)
    _lib.check(rc, "vs_sjoin_points_polygons_noalloc")
    assert int(res[1]) == st.staged > 100
    rc, ol, orr, res = _noalloc(Q, P, tree, "within", int(res[1]))
    _lib.check(rc, "vs_sjoin_points_polygons_noalloc")
    k = int(res[0])
    assert torch.equal(ol[:k], el) and torch.equal(orr[:k], er)
    # beyond the scratch: refused before any launch, outputs untouched
    big = _lib.NOALLOC_SCRATCH_BYTES // 16
    rc, ol, _, res = _noalloc(Q, P, tree, "within", big)
    assert rc == -1 and "scratch" in _lib.last_error()
    assert int(res.abs().sum()) == 0 and int(ol[0]) == -7
