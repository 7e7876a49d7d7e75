"""GPU tests of the primitives behind the join: radix sort, scan, bounds,
packed-tree build, bbox query (all through the C-ABI)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import oracle
from vibespatial_amd import _lib
from vibespatial_amd import join as J
from vibespatial_amd import synthetic as S
from vibespatial_amd.device import DevicePoints, DevicePolygons, ptr, stream

pytestmark = pytest.mark.gpu


def _dev():
    return torch.device("cuda", 0)


@pytest.mark.parametrize("n", [0, 1, 2, 1000, 4096, 4097, 300_000])
@pytest.mark.parametrize("key_bytes", [4, 8])
def test_radix_sort_matches_stable_argsort(n, key_bytes):
    print(f"radix sort n={n} key_bytes={key_bytes}", flush=True)
    L = _lib.lib()
    rng = np.random.default_rng(n + key_bytes)
    if key_bytes == 8:
        keys = rng.integers(0, 2**63, n, dtype=np.int64).astype(np.uint64)
        keys[: n // 3] = keys[: n // 3] % 7  # heavy duplicates
        tk = torch.from_numpy(keys.view(np.int64)).to(_dev())
        fn = L.vs_sort_pairs_u64
    else:
        keys = rng.integers(0, 2**32, n, dtype=np.int64).astype(np.uint32)
        keys[: n // 3] = keys[: n // 3] % 5
        tk = torch.from_numpy(keys.view(np.int32)).to(_dev())
        fn = L.vs_sort_pairs_u32
    vals = torch.arange(n, dtype=torch.int32, device=_dev())
    ws = torch.empty(max(L.vs_sort_workspace_bytes(n, key_bytes), 1), dtype=torch.uint8, device=_dev())
    _lib.check(fn(ptr(tk), ptr(vals), n, 0, 8 * key_bytes, ptr(ws), ws.numel(), stream()), "sort")
    got_v = vals.cpu().numpy()
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(got_v, order)
    got_k = tk.cpu().numpy().view(np.uint64 if key_bytes == 8 else np.uint32)
    assert np.array_equal(got_k, keys[order])


def test_radix_sort_partial_bits_is_stable():
    L = _lib.lib()
    rng = np.random.default_rng(3)
    n = 50_000
    keys = rng.integers(0, 2**32, n, dtype=np.int64).astype(np.uint32)
    tk = torch.from_numpy(keys.view(np.int32)).to(_dev())
    vals = torch.arange(n, dtype=torch.int32, device=_dev())
    ws = torch.empty(L.vs_sort_workspace_bytes(n, 4), dtype=torch.uint8, device=_dev())
    _lib.check(L.vs_sort_pairs_u32(ptr(tk), ptr(vals), n, 16, 32, ptr(ws), ws.numel(), stream()), "sort")
    order = np.argsort(keys >> 16, kind="stable")
    assert np.array_equal(vals.cpu().numpy(), order)


@pytest.mark.parametrize("n", [0, 1, 4095, 4096, 1_000_003, 4_194_304, 5_000_001])
def test_exclusive_scan(n):
    L = _lib.lib()
    rng = np.random.default_rng(n)
    a = rng.integers(0, 5, n).astype(np.int32)
    t = torch.from_numpy(a).to(_dev())
    out = torch.empty(max(n, 1), dtype=torch.int32, device=_dev())
    tot = torch.zeros(1, dtype=torch.int64, device=_dev())
    ws = torch.empty(L.vs_scan_workspace_bytes(n), dtype=torch.uint8, device=_dev())
    _lib.check(L.vs_exclusive_scan_i32(ptr(t), ptr(out), n, ptr(tot), ptr(ws), ws.numel(), stream()), "scan")
    exp = np.concatenate([[0], np.cumsum(a)[:-1]]) if n else np.empty(0)
    assert np.array_equal(out.cpu().numpy()[:n], exp)
    assert int(tot.item()) == int(a.sum())


def test_polygon_bounds_match_oracle():
    w = S.c2(n_points=10, n_polys=5000)
    P = DevicePolygons.from_host(w.polygons)
    b = J.polygon_bounds(P).cpu().numpy()
    assert np.array_equal(b, oracle.polygon_bounds(w.polygons))


def test_tree_build_invariants():
    w = S.c2(n_points=10, n_polys=20_000)
    P = DevicePolygons.from_host(w.polygons)
    bounds = J.polygon_bounds(P)
    bounds[7] = float("nan")  # an empty row
    tree = J.build_tree(bounds)
    lay = tree.layout
    nb = tree.node_boxes.cpu().numpy().reshape(-1, 4)
    ids = tree.item_ids.cpu().numpy()[: lay.n_items]
    keys = tree.item_keys.cpu().numpy().view(np.uint64)[: lay.n_items]
    hb = bounds.cpu().numpy()
    assert np.array_equal(np.sort(ids), np.arange(lay.n_items))
    assert np.all(keys[:-1] <= keys[1:])
    assert ids[-1] == 7 and keys[-1] == np.uint64(2**64 - 1)  # NaN row sorts last
    last = lay.n_levels - 1
    items = nb[lay.level_start[last]: lay.level_start[last] + lay.n_items]
    assert np.array_equal(items, hb[ids], equal_nan=True)
    F = lay.fanout
    for l in range(last):
        par = nb[lay.level_start[l]: lay.level_start[l] + lay.level_count[l]]
        ch = nb[lay.level_start[l + 1]: lay.level_start[l + 1] + lay.level_count[l + 1]]
        for j in range(0, len(par), max(1, len(par) // 50)):
            c = ch[j * F: (j + 1) * F]
            c = c[~np.isnan(c[:, 0])]
            assert par[j, 0] == c[:, 0].min() and par[j, 1] == c[:, 1].min()
            assert par[j, 2] == c[:, 2].max() and par[j, 3] == c[:, 3].max()


@pytest.mark.parametrize("n_q,n_t", [(1, 1), (500, 3000), (20_000, 50_000)])
def test_query_boxes_matches_bruteforce(n_q, n_t):
    print(f"query_boxes {n_q}x{n_t}", flush=True)
    rng = np.random.default_rng(n_q)
    q = rng.uniform(0, 1000, (n_q, 2))
    qb = np.hstack([q, q + rng.uniform(0, 30, (n_q, 2))])
    t = rng.uniform(0, 1000, (n_t, 2))
    tb = np.hstack([t, t + rng.uniform(0, 30, (n_t, 2))])
    if n_t > 10:
        tb[3] = np.nan
        qb[min(2, n_q - 1)] = np.nan
        tb[4] = tb[5]  # duplicate boxes
    tree = J.build_tree(torch.from_numpy(tb).to(_dev()))
    l, r, st = J.query_boxes(torch.from_numpy(qb).to(_dev()), tree)
    el, er = oracle.bbox_pairs(qb, tb)
    assert np.array_equal(l.cpu().numpy(), el) and np.array_equal(r.cpu().numpy(), er)
    assert st.pairs == len(el)


def test_query_boxes_capacity_retry():
    rng = np.random.default_rng(9)
    t = rng.uniform(0, 100, (2000, 2))
    tb = np.hstack([t, t + 5])
    tree = J.build_tree(torch.from_numpy(tb).to(_dev()))
    qb = np.array([[0, 0, 100, 100.0]] * 3)
    l, r, st = J.query_boxes(torch.from_numpy(qb).to(_dev()), tree, capacity=10)
    assert st.retries == 1 and st.pairs == 6000
    el, er = oracle.bbox_pairs(qb, tb)
    assert np.array_equal(l.cpu().numpy(), el) and np.array_equal(r.cpu().numpy(), er)


def test_query_boxes_every_run_length_is_sorted():
    """Runs of one left id of every length 1..40 (sorted in registers up to 8,
    in place up to 32, by the long-run pass beyond), the right ids arriving
    in tree order rather than id order: unit boxes on a line at shuffled
    positions, each query box covering a stretch of exactly L of them."""
    print("query_boxes run lengths 1..40", flush=True)
    rng = np.random.default_rng(7)
    n_t = 2000
    pos = rng.permutation(n_t).astype(np.float64)
    tb = np.stack([pos * 10, np.zeros(n_t), pos * 10 + 1, np.ones(n_t)], axis=1)
    tree = J.build_tree(torch.from_numpy(tb).to(_dev()))
    lens = np.repeat(np.arange(1, 41), 3)
    rng.shuffle(lens)
    starts = rng.integers(0, n_t - 40, lens.size)
    qb = np.stack([starts * 10 + 0.5, np.full(lens.size, 0.25), (starts + lens - 1) * 10 + 0.5,
                   np.full(lens.size, 0.75)], axis=1)
    l, r, st = J.query_boxes(torch.from_numpy(qb).to(_dev()), tree)
    l, r = l.cpu().numpy(), r.cpu().numpy()
    assert np.array_equal(np.bincount(l, minlength=lens.size), lens)
    order = np.argsort(pos)  # box id at each position
    for q in range(lens.size):
        s0 = starts[q]
        assert np.array_equal(r[l == q], np.sort(order[s0:s0 + lens[q]]))
    el, er = oracle.bbox_pairs(qb, tb)
    assert np.array_equal(l, el) and np.array_equal(r, er)


@pytest.mark.parametrize("n_t", [30_000, 70_001])
def test_query_boxes_long_runs_are_sorted(n_t):
    """Queries with many hits: runs of one left id longer than the short-run
    insertion sort (LDS bitonic sort) and longer than one LDS tile (chunk sort +
    merge passes), next to short runs; and a single query (one run holds every
    pair)."""
    print(f"query_boxes long runs n_t={n_t}", flush=True)
    rng = np.random.default_rng(n_t)
    t = rng.uniform(0, 1000, (n_t, 2))
    tb = np.hstack([t, t + rng.uniform(0, 3, (n_t, 2))])
    tree = J.build_tree(torch.from_numpy(tb).to(_dev()))
    q = rng.uniform(0, 1000, (300, 2))
    qb = np.hstack([q, q + rng.uniform(0, 20, (300, 2))])
    qb[10] = [-1, -1, 2000, 2000]  # every tree box
    qb[11] = [0, 0, 300, 300]  # ~9% of them
    qb[12] = [0, 0, 60, 60]  # a few hundred
    qb[13] = [0, 0, 20, 40]  # tens
    for sel in (slice(None), slice(10, 11)):
        l, r, st = J.query_boxes(torch.from_numpy(qb[sel]).to(_dev()), tree)
        el, er = oracle.bbox_pairs(qb[sel], tb)
        assert np.array_equal(l.cpu().numpy(), el) and np.array_equal(r.cpu().numpy(), er)
