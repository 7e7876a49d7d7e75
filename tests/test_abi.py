"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/vsjoin.h declares, and its host-compiled arithmetic agrees with the
oracle; layout and workspace queries behave."""

from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import oracle
from oracle import fraction_ref as FR
from vibespatial_amd import _lib
from vibespatial_amd import synthetic as S


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "vsjoin.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vs_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _declared_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing
    # and the ctypes table binds exactly the declared set
    assert sorted(_lib.SIGNATURES) == declared


def test_version_and_error_channel():
    L = _lib.lib()
    assert L.vs_version() >= 100
    lay = _lib.TreeLayout()
    rc = L.vs_tree_plan(10, 1, ctypes.byref(lay))
    assert rc == -1
    assert "fanout" in _lib.last_error()
    with pytest.raises(_lib.VsjoinError):
        _lib.check(rc, "vs_tree_plan")


def test_host_orient_matches_rationals():
    L = _lib.lib()
    rng = np.random.default_rng(11)
    for _ in range(3000):
        a = rng.uniform(-50, 50, 4)
        t = rng.uniform(-1, 2)
        c = a[:2] + t * (a[2:] - a[:2])
        c[0] = np.nextafter(c[0], np.inf * rng.choice([-1, 1])) if rng.random() < 0.5 else c[0]
        args = (*a, *c)
        assert L.vs_host_orient2d(*args) == FR.orient(*args)


def test_host_point_location_matches_oracle():
    L = _lib.lib()
    rng = np.random.default_rng(12)
    m = 30
    holes = rng.integers(0, 4, m)
    nh = int(holes.sum())
    P = S.star_polygons(rng, rng.uniform(0, 100, m), rng.uniform(0, 100, m), rng.uniform(5, 20, m),
                        rng.integers(4, 40, m), holes=holes, hole_vertices=rng.integers(4, 9, nh),
                        hole_frac=rng.uniform(0.1, 0.4, nh))
    x, y = P.x, P.y
    g, p, r = P.geom_offsets, P.part_offsets, P.ring_offsets
    ptrs = [a.ctypes.data_as(ctypes.c_void_p) for a in (x, y, g, p, r)]
    for q in range(m):
        b = P.scalar(q).bounds
        pts = list(zip(rng.uniform(b[0], b[2], 40), rng.uniform(b[1], b[3], 40)))
        c0 = P.ring_offsets[P.part_offsets[P.geom_offsets[q]]]
        pts += [(P.x[c0 + k], P.y[c0 + k]) for k in range(3)]
        for px_, py_ in pts:
            assert L.vs_host_point_location(px_, py_, *ptrs, q) == oracle.point_location(px_, py_, P, q)


def test_host_segments_touch_matches_oracle():
    L = _lib.lib()
    rng = np.random.default_rng(13)
    for _ in range(3000):
        v = rng.integers(0, 4, 8).astype(float)
        assert L.vs_host_segments_touch(*v) == int(oracle.segments_touch(v[0:2], v[2:4], v[4:6], v[6:8]))


def test_host_hilbert64_extends_the_16_level_code():
    L = _lib.lib()
    rng = np.random.default_rng(14)
    xs = rng.integers(0, 2**32, 2000, dtype=np.uint64)
    ys = rng.integers(0, 2**32, 2000, dtype=np.uint64)
    k64 = np.array([L.vs_host_hilbert64(int(a), int(b)) for a, b in zip(xs, ys)], dtype=np.uint64)
    k16 = oracle.hilbert16((xs >> np.uint64(16)).astype(np.uint32), (ys >> np.uint64(16)).astype(np.uint32))
    assert np.array_equal((k64 >> np.uint64(32)).astype(np.uint32), k16)
    # a full 2^6 grid placed at the low bits: bijective and 4-adjacent
    n = 64
    gx, gy = np.meshgrid(np.arange(n), np.arange(n))
    gx, gy = gx.ravel(), gy.ravel()
    k = np.array([L.vs_host_hilbert64(int(a), int(b)) for a, b in zip(gx, gy)], dtype=np.uint64)
    assert np.array_equal(np.sort(k), np.arange(n * n, dtype=np.uint64))
    o = np.argsort(k)
    assert np.all(np.abs(np.diff(gx[o])) + np.abs(np.diff(gy[o])) == 1)


@pytest.mark.parametrize("n,fan", [(0, 16), (1, 16), (15, 16), (16, 16), (17, 16), (100_000, 16), (1000, 2)])
def test_tree_plan_layout(n, fan):
    L = _lib.lib()
    lay = _lib.TreeLayout()
    _lib.check(L.vs_tree_plan(n, fan, ctypes.byref(lay)), "vs_tree_plan")
    nl = lay.n_levels
    counts = [lay.level_count[i] for i in range(nl)]
    starts = [lay.level_start[i] for i in range(nl)]
    assert counts[-1] == n and counts[0] == (1 if n else 0)
    for l in range(nl - 1):
        assert counts[l] == -(-counts[l + 1] // fan)
        assert starts[l + 1] == starts[l] + counts[l]
    assert lay.n_nodes == sum(counts)


def test_workspace_queries_are_positive_and_monotone():
    L = _lib.lib()
    for fn in ("vs_index_workspace_bytes", "vs_scan_workspace_bytes", "vs_point_order_workspace_bytes",
               "vs_refine_pairs_workspace_bytes", "vs_lexsort_pairs_workspace_bytes"):
        f = getattr(L, fn)
        a, b = f(1000), f(10_000_000)
        assert 0 < a <= b, fn
    assert 0 < L.vs_sjoin_points_polygons_workspace_bytes(1000, 10, 100, 10) < \
        L.vs_sjoin_points_polygons_workspace_bytes(1000, 10, 100, 10**6)
    assert L.vs_sort_workspace_bytes(10**6, 4) < L.vs_sort_workspace_bytes(10**6, 8)


def test_noalloc_scratch_limit_and_rejection_without_device_work():
    text = open(os.path.join(ROOT, "include", "vsjoin.h")).read()
    m = re.search(r"#define VS_NOALLOC_SCRATCH_BYTES \((\d+)ll << (\d+)\)", text)
    assert m and int(m.group(1)) << int(m.group(2)) == _lib.NOALLOC_SCRATCH_BYTES
    L = _lib.lib()
    # 10M points x 100k polygons needs ~535 MiB of work memory: refused by the
    # size check, which runs before any device call (null pointers are never read)
    n, m_, v, cap = 10_000_000, 100_000, 5_000_000, 10_000_000
    assert L.vs_sjoin_points_polygons_workspace_bytes(n, m_, v, cap) > _lib.NOALLOC_SCRATCH_BYTES
    lay = _lib.TreeLayout()
    _lib.check(L.vs_tree_plan(m_, 16, ctypes.byref(lay)), "vs_tree_plan")
    res = L.vs_sjoin_points_polygons_noalloc(None, None, n, None, None, None, None, None, m_, m_, m_, v,
                                             ctypes.byref(lay), None, None, _lib.PRED["within"], cap, None, None,
                                             None, None)
    assert res == -1
    msg = _lib.last_error()
    assert "scratch" in msg and str(_lib.NOALLOC_SCRATCH_BYTES) in msg
    assert L.vs_sjoin_points_polygons_noalloc(None, None, -1, None, None, None, None, None, 0, 0, 0, 0,
                                              ctypes.byref(lay), None, None, 0, 0, None, None, None, None) == -1


@pytest.mark.parametrize("key_bytes", [4, 8])
def test_host_cover_ranges_hold_every_box_point(key_bytes):
    """The learned index's box cover: at most 16 sorted, disjoint key ranges
    that hold the key of every point inside the box (keys computed with the
    host Hilbert encoder on the grid the index uses), and that hold fewer
    points outside the box than the unrefined cell cover would."""
    L = _lib.lib()
    rng = np.random.default_rng(30 + key_bytes)
    ext = np.array([0.0, 0.0, 1e5, 1e5])
    px, py = rng.uniform(0, 1e5, 200_000), rng.uniform(0, 1e5, 200_000)
    bits = 4 * key_bytes
    cells = 4294967295.0 if bits == 32 else float((1 << bits) - 1)
    qx = np.minimum(np.maximum((px - ext[0]) * (cells / (ext[2] - ext[0])), 0), cells).astype(np.uint64)
    qy = np.minimum(np.maximum((py - ext[1]) * (cells / (ext[3] - ext[1])), 0), cells).astype(np.uint64)
    sh = 0 if key_bytes == 8 else 16
    keys = np.array([L.vs_host_hilbert64(int(a) << sh, int(b) << sh) for a, b in zip(qx, qy)], dtype=np.uint64)
    if key_bytes == 4:  # 16-bit grid key, low 8 bits cleared (the 2^12 index grid)
        keys = (keys >> np.uint64(32)) & np.uint64(0xFFFFFF00)
    order = np.argsort(keys)
    ks = keys[order]
    out = np.zeros(32, dtype=np.uint64)
    scanned = inside = 0
    for _ in range(300):
        c = rng.uniform(0, 1e5, 2)
        w, h = rng.lognormal(np.log(300), 1.0, 2)
        x0, y0, x1, y1 = c[0] - w / 2, c[1] - h / 2, c[0] + w / 2, c[1] + h / 2
        m = L.vs_host_cover_ranges(x0, y0, x1, y1, ext.ctypes.data_as(ctypes.c_void_p), key_bytes,
                                   out.ctypes.data_as(ctypes.c_void_p))
        assert 1 <= m <= 16
        lo, hi = out[0:2 * m:2], out[1:2 * m:2]
        assert np.all(lo <= hi) and np.all(hi[:-1] < lo[1:])  # sorted, disjoint
        inbox = (px >= x0) & (px <= x1) & (py >= y0) & (py <= y1)
        k = keys[inbox]
        j = np.searchsorted(lo, k, side="right") - 1
        assert np.all(j >= 0) and np.all(k <= hi[j])
        a, b = np.searchsorted(ks, lo, "left"), np.searchsorted(ks, hi, "right")
        scanned += int((b - a).sum())
        inside += int(inbox.sum())
    print(f"cover: {scanned} positions scanned for {inside} points in the boxes ({scanned / max(inside, 1):.3f})")
    assert scanned >= inside
    # degenerate boxes: a point box on a data point, zero width / height, the
    # whole extent, beyond the extent, edges on grid-cell boundaries, inverted
    cell = (ext[2] - ext[0]) / (2 ** (12 if key_bytes == 4 else 32))
    special = [(px[7], py[7], px[7], py[7]), (px[9], 0.0, px[9], 1e5), (0.0, py[11], 1e5, py[11]),
               (-1.0, -1.0, 1e5 + 1, 1e5 + 1), (2e5, 2e5, 3e5, 3e5), (-5e4, -5e4, -1.0, -1.0),
               (64 * cell, 64 * cell, 128 * cell, 96 * cell), (0.0, 0.0, 0.0, 0.0), (1e5, 1e5, 1e5, 1e5)]
    for x0, y0, x1, y1 in special:
        m = L.vs_host_cover_ranges(x0, y0, x1, y1, ext.ctypes.data_as(ctypes.c_void_p), key_bytes,
                                   out.ctypes.data_as(ctypes.c_void_p))
        assert 1 <= m <= 16
        lo, hi = out[0:2 * m:2], out[1:2 * m:2]
        assert np.all(lo <= hi) and np.all(hi[:-1] < lo[1:])
        inbox = (px >= x0) & (px <= x1) & (py >= y0) & (py <= y1)
        k = keys[inbox]
        j = np.searchsorted(lo, k, side="right") - 1
        assert np.all(j >= 0) and np.all(k <= hi[j]), (x0, y0, x1, y1)
    assert L.vs_host_cover_ranges(5.0, 5.0, 4.0, 6.0, ext.ctypes.data_as(ctypes.c_void_p), key_bytes,
                                  out.ctypes.data_as(ctypes.c_void_p)) == 0  # inverted: empty
    assert L.vs_host_cover_ranges(np.nan, 5.0, 4.0, 6.0, ext.ctypes.data_as(ctypes.c_void_p), key_bytes,
                                  out.ctypes.data_as(ctypes.c_void_p)) == 0  # NaN: empty


def test_hilbert_children_digits_diagonal_rule():
    """The four children of a Hilbert cell are visited around the 2 x 2
    square: a child's diagonal partner has its last digit xor 2 (the rule the
    box covers use to get all four digits from two encodings)."""
    L = _lib.lib()
    rng = np.random.default_rng(40)
    for lev in (2, 5, 12, 20, 31):
        for _ in range(300):
            px = int(rng.integers(0, 2 ** (lev - 1))) * 2
            py = int(rng.integers(0, 2 ** (lev - 1))) * 2
            sh = 32 - lev

            def digit(x, y):
                return (L.vs_host_hilbert64(x << sh, y << sh) >> (64 - 2 * lev)) & 3

            d = {(i, j): digit(px + i, py + j) for i in (0, 1) for j in (0, 1)}
            assert sorted(d.values()) == [0, 1, 2, 3]
            assert d[(1, 1)] == d[(0, 0)] ^ 2 and d[(0, 1)] == d[(1, 0)] ^ 2
