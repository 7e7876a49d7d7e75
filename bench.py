"""Benchmark: sjoin(points, polygons, predicate='within') refined pairs/sec,
learned + FP32 form vs the plain fp64 form.

Workload (one GPU): 10,000,000 points from a 64-component Gaussian mixture
over [0, 1e5]^2 and 100,000 star polygons with 0-3 holes and lognormal vertex
counts (vibespatial_amd.synthetic.c2, seeded).  A step is one full join from
device-resident SoA inputs to device-resident lexsorted int64
(left_index, right_index) pairs: polygon bounds, packed Hilbert tree build,
point Hilbert ordering, traversal with fused exact refine, radix sort of
the staged pairs by left id and the sorted emit.  Input H2D and output D2H are outside the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Both forms run the same step boundaries; `value` is the learned + FP32 form,
`config.plain_fp64` the plain form, `config.learned_vs_plain` their ratio.
Prints one JSON line.  `roofline` is for the learned form's dominant kernel
(learned_join_kernel): algorithmic bytes per launch / its average duration
measured with HIP events on the launch stream.  Algorithmic bytes per launch:
    16 N + sum_q (16 V_q + 32) + sum_{(p,q) in C} (8 + 8 V_q) + 32 E + 8 K
and for the plain form's traversal kernel
    16 N + sum_q (16 V_q + 32) + sum_{(p,q) in C} (16 + 16 V_q) + 8 K
(N points, V_q vertices of polygon q over all rings, C the exact inclusive
bbox candidate pairs, E escalated edges, K the refined pairs).  `cpu_baseline` times the CPU
oracle's join (packed tree + exact location, OpenMP) on a bounded sample of
the same workload on this host: median of 3 runs on all usable cores (up to
16) after a warm-up, plus `one_thread`, the same join on one thread over a
tenth of the sample.

Other workloads (`--workload`): c4 (learned index vs tree, 50M points x 1M
boxes), c5 (4096 batched polygon intersects jobs), dwithin (the c2 data with
predicate='dwithin'), wkb (GPU WKB ingest of the c2 data), nearest
(sjoin_nearest of the c2 points to the c2 polygons), relate (polygon x polygon
predicates, --predicate).
"""

from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "sjoin refined pairs/sec @10M pts×100k polys, bit-exact; learned+FP32 vs plain"
HBM_PEAK_GBS = 8000.0


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-points", type=int, default=10_000_000)
    ap.add_argument("--n-polys", type=int, default=100_000)
    ap.add_argument("--cpu-sample", type=int, default=10_000_000,
                    help="points in the CPU-baseline sample (0 disables it)")
    ap.add_argument("--workload", choices=("c2", "c4", "c5", "dwithin", "wkb", "nearest", "relate"), default="c2",
                    help="c2: the sjoin metric (default); c4: point-tree sindex.query, learned index vs tree; "
                         "c5: 4096 batched polygon x polygon intersects jobs; dwithin: exact point x polygon "
                         "dwithin join on the c2 data")
    ap.add_argument("--distance", type=float, default=25.0, help="dwithin distance (workload dwithin)")
    ap.add_argument("--predicate", default="overlaps", help="polygon x polygon predicate (workload relate)")
    ap.add_argument("--extras", type=int, default=1, choices=(0, 1),
                    help="c2: also time the lookup ablation and the public API (0: only the two forms, for "
                         "profiler runs)")
    ap.add_argument("--tile", default="", metavar="W,T,O",
                    help="c5: use this tile (traversal waves per block, refine block, candidate order) instead "
                         "of autotuning, and skip the ablation (for profiler runs)")
    ap.add_argument("--tree-fanout", type=int, default=64,
                    help="c4: fanout of the point tree of the tree form (one query box per wave tests a node's "
                         "children on its lanes, so 64 fills the wave)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB device-to-device once (a kernel of known traffic for "
                         "scaling hardware byte counters)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the join shards by point ranges across GPUs by design; only the 1-GPU path is implemented)",
              file=sys.stderr)
        return 2

    from __graft_entry__ import _ensure_built

    _ensure_built()
    import numpy as np
    import torch

    from vibespatial_amd import _lib
    from vibespatial_amd import join as J
    from vibespatial_amd import synthetic as S
    from vibespatial_amd.device import DevicePoints, DevicePolygons

    torch.cuda.set_device(0)
    if args.workload == "c4":
        return _with_copy_probe(args, run_c4)
    if args.workload == "c5":
        return _with_copy_probe(args, run_c5)
    if args.workload == "dwithin":
        return _with_copy_probe(args, run_dwithin)
    if args.workload == "wkb":
        return _with_copy_probe(args, run_wkb)
    if args.workload == "nearest":
        return _with_copy_probe(args, run_nearest)
    if args.workload == "relate":
        return _with_copy_probe(args, run_relate)
    _log("generating workload")
    w = S.c2(n_points=args.n_points, n_polys=args.n_polys)
    Q = DevicePoints.from_host(w.points)
    P = DevicePolygons.from_host(w.polygons)
    torch.cuda.synchronize()
    L = _lib.lib()
    P_h = w.polygons
    V = np.diff(P_h.ring_offsets[P_h.part_offsets[P_h.geom_offsets]])
    n_pts = len(w.points)
    poly_bytes = int((16 * V + 32).sum())

    def step_plain(cap):
        tree = J.build_tree(J.polygon_bounds(P))
        return J.sjoin_points_polygons(Q, P, tree, "within", capacity=cap)

    def step_learned(cap):
        return J.sjoin_points_polygons_learned(Q, P, J.polygon_bounds(P), "within", capacity=cap)

    def measure(step, name):
        _log(f"{name}: warmup")
        cap, st = None, None
        for _ in range(max(args.warmup, 1)):
            _, _, st = step(cap)
            cap = st.pairs + 1024
        _lib.check(L.vs_set_kernel_timing(1), "vs_set_kernel_timing")
        kernel_ms = []
        torch.cuda.synchronize()
        _log(f"{name}: timing {args.steps} steps")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _, _, st = step(cap)  # reads the result header: synchronises the stream
            ms = _lib.ctypes.c_float()
            _lib.check(L.vs_last_kernel_ms(_lib.ctypes.byref(ms)), "vs_last_kernel_ms")
            kernel_ms.append(ms.value)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _lib.check(L.vs_set_kernel_timing(0), "vs_set_kernel_timing")
        return st, (t1 - t0) * 1e3 / args.steps, statistics.mean(kernel_ms)

    st_p, ms_p, kms_p = measure(step_plain, "plain")
    st_l, ms_l, kms_l = measure(step_learned, "learned+fp32")
    assert st_p.pairs == st_l.pairs and st_p.candidates == st_l.candidates, "forms disagree"
    K = st_l.pairs

    # the same join with the network left out of the lookups (leaf table
    # alone; plain binary search over the keys) and with other training
    # lengths: the network's share of the learned form's time
    ablation = {}
    for name, kw in (("leaves", {"lookup": "leaves"}), ("binary", {"lookup": "binary"}),
                     ("untrained", {"train_iters": 0}), ("train_64", {"train_iters": 64})) if args.extras else ():
        st_a, ms_a, kms_a = measure(lambda cap, kw=kw: J.sjoin_points_polygons_learned(
            Q, P, J.polygon_bounds(P), "within", capacity=cap, **kw), f"learned lookup={name}")
        assert st_a.pairs == K
        ablation[name] = {"ms_per_step": ms_a, "kernel_ms": kms_a, "lookup_probes": st_a.lookup_probes,
                          "mlp_leaf_error": st_a.mlp_leaf_error}

    # the public API on the same data: GeoSeries(polygons).sindex.query(points,
    # predicate='within', return_device=True), the route sjoin takes (device
    # pairs; frame assembly on the host excluded)
    ms_pub, strategy = None, []
    if args.extras:
        from vibespatial_amd.api import GeoSeries

        gs = GeoSeries(w.polygons)
        w.points._device = Q  # the column's device copy is resident, as after the first query
        w.polygons._device = P
        res = gs.sindex.query(w.points, predicate="within", return_device=True)
        assert torch.equal(res[0], step_learned(None)[0]) and int(res[0].numel()) == K
        del res
        torch.cuda.synchronize()
        _log(f"public api: timing {args.steps} steps")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = gs.sindex.query(w.points, predicate="within", return_device=True)
            del res
        torch.cuda.synchronize()
        ms_pub = (time.perf_counter() - t0) * 1e3 / args.steps
        strategy = gs.sindex.last_strategy
    # algorithmic bytes per launch of each form's dominant kernel
    alg_p = 16 * n_pts + poly_bytes + 16 * st_p.candidates + 16 * st_p.candidate_vertices + 8 * K
    alg_l = 16 * n_pts + poly_bytes + 8 * st_l.candidates + 8 * st_l.candidate_vertices + 32 * st_l.escalated_edges \
        + 8 * K
    ach_l = alg_l / (kms_l * 1e-3) / 1e9
    ach_p = alg_p / (kms_p * 1e-3) / 1e9
    value = K / (ms_l * 1e-3)
    value_p = K / (ms_p * 1e-3)
    ms_per_step = ms_l
    traffic = _measured_traffic(n_pts, len(P_h))

    cpu = None
    if args.cpu_sample > 0:
        import oracle

        ns = min(args.cpu_sample, n_pts)
        px, py = w.points.x[:ns], w.points.y[:ns]
        cpu = _cpu_timing(lambda nt: len(oracle.sjoin_points_polygons(px, py, P_h, "within", nthreads=nt)[0]),
                          lambda: len(oracle.sjoin_points_polygons(px[: ns // 10], py[: ns // 10], P_h, "within",
                                                                   nthreads=1)[0]), "pairs/s")
        cpu["sample"] = (f"first {ns:,} of the {n_pts:,} points x all {len(P_h):,} polygons, within "
                         f"({cpu.pop('count'):,} pairs; {cpu.pop('runs_s')}; tree build included)")
        cpu["one_thread"]["sample"] = (f"first {ns // 10:,} points x all polygons ({cpu['one_thread'].pop('count'):,} "
                                       f"pairs; {cpu['one_thread'].pop('runs_s')})")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"{n_pts:,} GMM points x {len(P_h):,} star polygons (0-3 holes, lognormal vertices), "
                        "sjoin predicate=within",
            "form": "learned+fp32 (learned rank index over Hilbert-sorted points, polygon-major FP32 refine "
                    "with exact escalation)",
            "pairs": K,
            "bbox_candidates": st_l.candidates,
            "candidate_rate_per_s": st_l.candidates / (ms_l * 1e-3),
            "polygon_vertices": int(V.sum()),
            "escalated_edges": st_l.escalated_edges,
            "escalation_fraction": st_l.escalation_fraction,
            "lookup_probes": st_l.lookup_probes,
            "positions_scanned": st_l.scanned,
            "mlp_leaf_error": st_l.mlp_leaf_error,
            "plain_fp64": {"value": value_p, "ms_per_step": ms_p, "kernel_ms": kms_p,
                           "roofline_frac": ach_p / HBM_PEAK_GBS},
            "learned_vs_plain": value / value_p,
            "public_api": {"call": "GeoSeries(polygons).sindex.query(points, predicate='within', "
                                   "return_device=True)", "strategy": [list(t) for t in strategy],
                           "ms_per_step": ms_pub, "value": K / (ms_pub * 1e-3) if ms_pub else None},
            "lookup_ablation": ablation,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "learned_join_kernel<0,0> (scan of rank ranges + FP32 refine)",
            "achieved": ach_l,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": ach_l / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel_ms": kms_l,
            "alg_bytes_per_launch": alg_l,
            # the same algorithmic bytes over the whole step (index build,
            # covers, scan + refine, finish), and the measured DRAM bytes of
            # the kernel over its duration
            "step_frac": alg_l / (ms_l * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "dram_frac": None if traffic is None else traffic / (kms_l * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "traffic_source": None if traffic is None else _TRAFFIC_SOURCE,
            "note": "frac: the byte model's bytes over the kernel's time; the model charges every bbox candidate "
                    "for streaming its polygon's edges (8 + 8 V_q), which the kernel reads once per polygon "
                    "into LDS, so frac can pass 1; step_frac: the same bytes over the whole step; dram_frac: "
                    "the measured DRAM bytes over the kernel's time",
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    _copy_probe(args)
    return 0


# HBM bytes per learned_join_kernel launch on the default workload, from two
# rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE; KiB per dispatch, mean of the
# warm-up and the timed dispatch) scaled by what the same counters read for a
# 1 GiB device-to-device copy in the same run (FETCH_SIZE 524,316.5 KiB ->
# x1.99989, WRITE_SIZE 1,048,576 KiB -> x1.0):
#   (1.99989 * 353,285.6 + 1.0 * 97,870.5) KiB = 823,708,818 bytes.
# The summary with both commands is profiles/learned_join_kernel_10Mpts_100kpolys_pmc_traffic.md.
_TRAFFIC_BYTES = {(10_000_000, 100_000): 823_708_818}


_TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of learned_join_kernel on this workload, scaled by a "
                   "1 GiB copy (profiles/learned_join_kernel_10Mpts_100kpolys_pmc_traffic.md); not measured in "
                   "this run")


def _measured_traffic(n_points, n_polys):
    return _TRAFFIC_BYTES.get((n_points, n_polys))


def _cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    visible = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS")
    cores = min(visible, int(share)) if share and share.isdigit() and int(share) > 0 else visible
    return model, visible, cores


def _cpu_timing(run_all, run_one, unit, reps=5):
    """CPU baseline protocol: one warm-up, then the median of `reps` runs, on
    every usable core (the affinity set, within OMP_NUM_THREADS when the host
    caps this process's share), and on one thread over a tenth of the
    sample.  run_all(nthreads) and run_one() return the result count."""
    model, visible, cores = _cpu_info()
    _log(f"cpu baseline: {cores} threads ({visible} visible, {model})")
    run_all(cores)  # warm-up
    runs = []
    for _ in range(reps):
        t0 = time.perf_counter()
        count = run_all(cores)
        runs.append(time.perf_counter() - t0)
    tc = statistics.median(runs)
    run_one()
    runs1 = []
    for _ in range(reps):
        t0 = time.perf_counter()
        c1 = run_one()
        runs1.append(time.perf_counter() - t0)
    t1 = statistics.median(runs1)
    return {
        "value": count / tc, "unit": unit, "cores": cores, "kind": "port", "cpu_model": model,
        "cores_visible": visible, "count": count, "runs_s": f"median of {reps} runs after a warm-up {tc:.3f} s",
        "one_thread": {"value": c1 / t1, "unit": unit, "cores": 1, "count": c1,
                       "runs_s": f"median of {reps} runs {t1:.3f} s"},
    }


def _copy_probe(args):
    if args.copy_probe <= 0:
        return
    import torch

    n = args.copy_probe * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    _log(f"copy probe: {args.copy_probe} MiB read + {args.copy_probe} MiB written")


def _with_copy_probe(args, fn):
    rc = fn(args)
    _copy_probe(args)
    return rc


def run_c4(args):
    """sindex.query(boxes, predicate='intersects') over a point tree: 50M
    Zipf-clustered points x 1M boxes (vibespatial_amd.synthetic.c4), learned
    index (64-bit keys) vs packed-tree traversal.  A step = index build + the
    query of every box (sort=False, the sindex.query default: pairs grouped by
    box).  The hits exceed one call's int32 slot limit, so the boxes go in
    equal batches.  Algorithmic bytes per step: 16 N + 32 Q + 24 |H|."""
    import numpy as np
    import torch

    from vibespatial_amd import _lib
    from vibespatial_amd import join as J
    from vibespatial_amd import synthetic as S
    from vibespatial_amd.device import DevicePoints

    n_pts = args.n_points if args.n_points != 10_000_000 else 50_000_000
    n_box = 1_000_000
    _log(f"c4: generating {n_pts:,} points x {n_box:,} boxes")
    w = S.c4(n_points=n_pts, n_boxes=n_box)
    Q = DevicePoints.from_host(w.points)
    boxes = torch.from_numpy(w.meta["boxes"]).cuda()
    torch.cuda.synchronize()
    L = _lib.lib()
    n_batches = 4
    edges = np.linspace(0, n_box, n_batches + 1).astype(np.int64)
    batches = [boxes[edges[i]:edges[i + 1]] for i in range(n_batches)]

    def learned_step(caps):
        idx = J.LearnedPointIndex(Q, key_bytes=8)
        hits, scanned, out = 0, 0, []
        for i, b in enumerate(batches):
            l, r, st = idx.query_boxes(b, "learned", capacity=caps[i] if caps else None, sort=False)
            hits += st.pairs
            scanned += st.scanned
            out.append(st.pairs + 1024)
            del l, r
        return hits, out, {"scanned": scanned, "mlp_leaf_error": idx.mlp_leaf_error}

    def tree_step(caps):
        pts = torch.stack([Q.x, Q.y, Q.x, Q.y], dim=1).contiguous()
        tree = J.build_tree(pts, fanout=args.tree_fanout)
        del pts
        hits, out = 0, []
        for i, b in enumerate(batches):
            l, r, st = J.query_boxes(b, tree, capacity=caps[i] if caps else None, sort=False)
            hits += st.pairs
            out.append(st.pairs + 1024)
            del l, r
        return hits, out, {}

    def measure(step, name):
        _log(f"c4 {name}: warmup")
        hits, caps, extra = step(None)
        for _ in range(max(args.warmup - 1, 0)):
            step(caps)
        torch.cuda.synchronize()
        _log(f"c4 {name}: timing {args.steps} steps")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            h, _, extra = step(caps)
            assert h == hits
        torch.cuda.synchronize()
        return hits, (time.perf_counter() - t0) * 1e3 / args.steps, extra

    _ = L
    hl, ms_l, ex_l = measure(learned_step, "learned")
    ht, ms_t, _ = measure(tree_step, "tree")
    assert hl == ht, "forms disagree"
    alg = 16 * n_pts + 32 * n_box + 24 * hl
    cpu = None
    if args.cpu_sample > 0:
        import oracle

        ns, nq = min(n_pts, 5_000_000), min(n_box, 100_000)
        px, py, bx = w.points.x[:ns], w.points.y[:ns], w.meta["boxes"][:nq]
        cpu = _cpu_timing(lambda nt: len(oracle.box_query_points(px, py, bx, nthreads=nt)[0]),
                          lambda: len(oracle.box_query_points(px[: ns // 10], py[: ns // 10], bx, nthreads=1)[0]),
                          "hits/s", reps=3)
        cpu["sample"] = (f"first {ns:,} of the {n_pts:,} points (tree build included) x first {nq:,} boxes "
                         f"({cpu.pop('count'):,} hits; {cpu.pop('runs_s')})")
        cpu["one_thread"]["sample"] = (f"first {ns // 10:,} points x the same boxes "
                                       f"({cpu['one_thread'].pop('count'):,} hits; {cpu['one_thread'].pop('runs_s')})")
    out = {
        "metric": "sindex.query hits/sec @50M pts x 1M boxes (point tree, intersects); learned vs tree",
        "value": hl / (ms_l * 1e-3),
        "unit": "hits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_l,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{n_pts:,} Zipf(1.1)-clustered points (1024 clusters, sigma 500 m) x {n_box:,} boxes "
                        "(LogNormal(ln 50, 1) sides), 64-bit Hilbert keys, sort=False, 4 batches",
            "hits": hl,
            "positions_scanned": ex_l.get("scanned"),
            "mlp_leaf_error": ex_l.get("mlp_leaf_error"),
            "tree": {"value": ht / (ms_t * 1e-3), "ms_per_step": ms_t, "fanout": args.tree_fanout},
            "learned_vs_tree": ms_t / ms_l,
        },
        "roofline": {"bound": "hbm", "kernel": "whole step (build + query)", "achieved": alg / (ms_l * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": alg / (ms_l * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "alg_bytes_per_step": alg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def run_c5(args):
    """4,096 independent polygon x polygon sjoin(intersects) jobs
    (vibespatial_amd.synthetic.c5) in one batched launch sequence: forest
    build, persistent traversal, cell-grid decisions, exact refine of the
    rest, finish.  A step = one batched
    call over every job, inputs resident; the tile shape is autotuned first.
    Algorithmic bytes per step: sum_q 16 V_q + sum_C 16 (V_a + V_b) + 8 K."""
    import numpy as np
    import torch

    from vibespatial_amd import _lib
    from vibespatial_amd import join as J
    from vibespatial_amd.device import DevicePolygons
    from vibespatial_amd import synthetic as S

    _log("c5: generating 4096 jobs")
    left, right, lj, rj = S.c5()
    A, B = DevicePolygons.from_host(left), DevicePolygons.from_host(right)
    tl, tr = torch.from_numpy(lj).cuda(), torch.from_numpy(rj).cuda()
    torch.cuda.synchronize()
    L = _lib.lib()
    _log("c5: autotune")
    if args.tile:
        best, times = tuple(int(v) for v in args.tile.split(",")), {}
    else:
        best, times = J.autotune_batched(A, B, tl, tr, reps=1)
    _, _, st = J.batched_intersects(A, B, tl, tr, tile=best)
    cap = st.staged + 1024
    for _ in range(max(args.warmup, 1)):
        J.batched_intersects(A, B, tl, tr, capacity=cap, tile=best)
    _lib.check(L.vs_set_kernel_timing(1), "vs_set_kernel_timing")
    kms = []
    torch.cuda.synchronize()
    _log(f"c5: timing {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, _, st = J.batched_intersects(A, B, tl, tr, capacity=cap, tile=best)
        ms = _lib.ctypes.c_float()
        _lib.check(L.vs_last_kernel_ms(_lib.ctypes.byref(ms)), "vs_last_kernel_ms")
        kms.append(ms.value)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    _lib.check(L.vs_set_kernel_timing(0), "vs_set_kernel_timing")
    # graph replay of the same sequence
    g = J.BatchedIntersectsGraph(A, B, tl, tr, capacity=cap, tile=best)
    g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, _, stg = g.replay()
    torch.cuda.synchronize()
    ms_graph = (time.perf_counter() - t0) * 1e3 / args.steps
    assert stg.pairs == st.pairs
    # ablation: every bbox candidate refined exactly (no cell-grid decisions)
    ablation = None
    if not args.tile:
        J.batched_intersects(A, B, tl, tr, capacity=cap, tile=best, cell_grid=False)
        _lib.check(L.vs_set_kernel_timing(1), "vs_set_kernel_timing")
        kms0 = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _, _, st0 = J.batched_intersects(A, B, tl, tr, capacity=cap, tile=best, cell_grid=False)
            ms = _lib.ctypes.c_float()
            _lib.check(L.vs_last_kernel_ms(_lib.ctypes.byref(ms)), "vs_last_kernel_ms")
            kms0.append(ms.value)
        torch.cuda.synchronize()
        ms_nogrid = (time.perf_counter() - t0) * 1e3 / args.steps
        _lib.check(L.vs_set_kernel_timing(0), "vs_set_kernel_timing")
        assert st0.pairs == st.pairs and st0.refined == st0.candidates
        ablation = {"ms_per_step": ms_nogrid, "traverse_refine_kernel_ms": statistics.mean(kms0)}
    VA = np.diff(left.ring_offsets[left.part_offsets[left.geom_offsets]])
    VB = np.diff(right.ring_offsets[right.part_offsets[right.geom_offsets]])
    # sum_q 16 V_q + sum_C 16 (V_a + V_b) + 8 K (the candidate term counted on the device)
    alg = 16 * int(VA.sum() + VB.sum()) + 16 * st.candidate_vertices + 8 * st.pairs
    cpu = None
    if args.cpu_sample > 0:
        import oracle

        nj = min(512, len(lj) - 1)  # a bounded sample: the first jobs
        Ls, Rs = left.take(np.arange(lj[nj])), right.take(np.arange(rj[nj]))
        jl = np.repeat(np.arange(nj), np.diff(lj[: nj + 1]))
        jr = np.repeat(np.arange(nj), np.diff(rj[: nj + 1]))

        def per_job(nt):  # every job's join: the oracle over the rows, pairs inside one job
            a, b = oracle.sjoin_polygons_polygons(Ls, Rs, nthreads=nt)
            return int((jl[a] == jr[b]).sum())

        nj1 = max(nj // 10, 1)
        L1, R1 = left.take(np.arange(lj[nj1])), right.take(np.arange(rj[nj1]))
        cpu = _cpu_timing(per_job, lambda: len(oracle.sjoin_polygons_polygons(L1, R1, nthreads=1)[0]), "pairs/s",
                          reps=3)
        cpu["sample"] = (f"first {nj} of the {len(lj) - 1} jobs ({lj[nj]:,} x {rj[nj]:,} rows; "
                         f"{cpu.pop('count'):,} pairs; {cpu.pop('runs_s')})")
        cpu["one_thread"]["sample"] = (f"first {nj1} jobs ({cpu['one_thread'].pop('count'):,} pairs; "
                                       f"{cpu['one_thread'].pop('runs_s')})")
    out = {
        "metric": "batched sjoin(intersects) refined pairs/sec, 4096 polygon x polygon jobs",
        "value": st.pairs / (ms_step * 1e-3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{len(lj) - 1} jobs, {len(left):,} left footprints x {len(right):,} right parcels "
                        "(vibespatial_amd.synthetic.c5)",
            "pairs": st.pairs,
            "bbox_candidates": st.candidates,
            "refined_exactly": st.refined,
            "tile": list(best),
            "autotune_ms": {f"{k[0]}x{k[1]}/order{k[2]}": v for k, v in times.items()},
            "traverse_refine_kernel_ms": statistics.mean(kms),
            "graph_replay_ms_per_step": ms_graph,
            "without_cell_grids": ablation,
        },
        "roofline": {"bound": "hbm", "kernel": "whole step", "achieved": alg / (ms_step * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": alg / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "alg_bytes_per_step": alg,
                     "candidate_vertices": st.candidate_vertices},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def run_dwithin(args):
    """sjoin(points, polygons, predicate='dwithin', distance=d) on the c2 data
    (10M GMM points x 100k star polygons), d = --distance for every point.  A
    step = point boxes grown by d (outward-rounded) -> box query over the
    packed polygon tree -> exact distance refine -> compaction; tree built
    once (the reference builds it once per sindex).  Algorithmic bytes per
    step (lower bound): 32 N for the grown boxes + 16 V + the candidate pairs
    (16 each) + 16 per refined pair."""
    import torch

    from vibespatial_amd import join as J
    from vibespatial_amd import synthetic as S
    from vibespatial_amd.device import DevicePoints, DevicePolygons

    _log("dwithin: generating c2 workload")
    w = S.c2(n_points=args.n_points, n_polys=args.n_polys)
    Q = DevicePoints.from_host(w.points)
    P = DevicePolygons.from_host(w.polygons)
    tree = J.build_tree(J.polygon_bounds(P))
    n = len(w.points)
    d = torch.full((n,), float(args.distance), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()

    def step():
        qb = J.expand_boxes(J.point_boxes(Q), d)
        # output sized by the library (the tree's last hits per query box)
        l, r, st = J.query_boxes(qb, tree)
        kl, kr = J.refine_dwithin(l, r, d, Q, P)
        return int(l.numel()), int(kl.numel())

    for _ in range(max(args.warmup, 1)):
        cands, pairs = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        c2, p2 = step()
        assert (c2, p2) == (cands, pairs)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    V = int(len(w.polygons.x))
    alg = 32 * n + 16 * V + 16 * cands + 16 * pairs
    out = {
        "metric": "sjoin(dwithin) refined pairs/sec, 10M points x 100k polygons, exact distance decisions",
        "value": pairs / (ms * 1e-3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{n:,} GMM points x {len(w.polygons):,} star polygons (0-3 holes), "
                               f"dwithin distance {args.distance}",
                   "pairs": pairs, "bbox_candidates": cands},
        "roofline": {"bound": "hbm", "kernel": "whole step", "achieved": alg / (ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "alg_bytes_per_step_lower_bound": alg},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def run_nearest(args):
    """sjoin_nearest(points, polygons) on the c2 data (10M GMM points x 100k
    star polygons), no max_distance, all ties.  A step = the expanding-radius
    search (rounds of grown point boxes -> box query over the packed polygon
    tree -> fp64 pair distances -> nearest selection) -> lexsort -> distances
    of the result; tree built once.  Algorithmic bytes per step (lower bound):
    per round 32 per open query box + 16 per candidate pair read twice and its
    distance written, + 16 V for the polygons, + 24 per output pair."""
    import numpy as np
    import torch

    from vibespatial_amd import join as J
    from vibespatial_amd import synthetic as S
    from vibespatial_amd.device import DevicePoints, DevicePolygons

    _log("nearest: generating c2 workload")
    w = S.c2(n_points=args.n_points, n_polys=args.n_polys)
    Q = DevicePoints.from_host(w.points)
    P = DevicePolygons.from_host(w.polygons)
    tree = J.build_tree(J.polygon_bounds(P))
    qb = J.point_boxes(Q)
    tb = w.polygons.bounds()
    x0 = min(tb[:, 0].min(), w.points.x.min())
    y0 = min(tb[:, 1].min(), w.points.y.min())
    x1 = max(tb[:, 2].max(), w.points.x.max())
    y1 = max(tb[:, 3].max(), w.points.y.max())
    cover = float(np.hypot(x1 - x0, y1 - y0)) * 1.01
    start = float(np.sqrt((tb[:, 2].max() - tb[:, 0].min()) * (tb[:, 3].max() - tb[:, 1].min()) / len(tb)))
    torch.cuda.synchronize()

    def step():
        return J.nearest_pairs(qb, Q, tree, P, start_radius=start, cover_radius=cover)

    for _ in range(max(args.warmup, 1)):
        l, r, d = step()
    k = int(l.numel())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        l, r, d = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    assert int(l.numel()) == k
    n = len(w.points)
    V = int(len(w.polygons.x))
    zero = int((d == 0).sum().item())
    alg = 32 * n + 16 * V + 40 * k + 24 * k
    out = {
        "metric": "sjoin_nearest query rows/sec, 10M points x 100k polygons, all ties, fp64 distances",
        "value": n / (ms * 1e-3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{n:,} GMM points x {len(w.polygons):,} star polygons (0-3 holes), nearest "
                               "polygon of every point, no max_distance",
                   "pairs": k, "pairs_at_distance_0": zero, "start_radius": start, "cover_radius": cover},
        "roofline": {"bound": "hbm", "kernel": "whole step", "achieved": alg / (ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "alg_bytes_per_step_lower_bound": alg},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def run_relate(args):
    """sindex.query(polygons, predicate=P) between two star-polygon columns:
    the 100k c2 polygons (0-3 holes) query a tree over 100k more star polygons
    of the same kind (another seed).  A step = box query over the tree (built
    once) -> exact polygon x polygon refine (vs_refine_pairs mode 1) ->
    compaction.  Algorithmic bytes per step (lower bound): 32 per query box,
    16 per candidate pair, and per candidate both rows' coordinates
    (16 (V_a + V_b)), + 16 per kept pair."""
    import numpy as np
    import torch

    from vibespatial_amd import api
    from vibespatial_amd import join as J
    from vibespatial_amd import synthetic as S
    from vibespatial_amd.device import device_polygons

    _log(f"relate: generating polygons ({args.predicate})")
    A = S.c2(n_points=1, n_polys=args.n_polys).polygons
    rg = np.random.default_rng(11)
    m = args.n_polys
    holes = rg.integers(0, 4, m)
    nh = int(holes.sum())
    B = S.star_polygons(rg, rg.uniform(0, 1e5, m), rg.uniform(0, 1e5, m), rg.uniform(50, 300, m),
                        np.clip(np.round(rg.lognormal(np.log(24.0), 1.0, m)), 4, 10_000).astype(np.int64),
                        holes=holes, hole_vertices=rg.integers(4, 17, nh), hole_frac=rg.uniform(0.1, 0.4, nh))
    DA, DB = device_polygons(A), device_polygons(B)
    tree = J.build_tree(J.polygon_bounds(DB))
    qb = J.polygon_bounds(DA)
    torch.cuda.synchronize()

    def step():
        l, r, _ = J.query_boxes(qb, tree)
        kl, _ = api._refine(1, args.predicate, l, r, None, A, B)
        return int(l.numel()), int(kl.numel())

    for _ in range(max(args.warmup, 1)):
        cands, pairs = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        c2, p2 = step()
        assert (c2, p2) == (cands, pairs)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    alg = 32 * len(A) + 16 * cands + 16 * int(len(A.x) / max(len(A), 1) + len(B.x) / max(len(B), 1)) * cands \
        + 16 * pairs
    out = {
        "metric": f"sindex.query(predicate={args.predicate!r}) candidate pairs/sec, 100k x 100k star polygons, "
                  "exact relations",
        "value": cands / (ms * 1e-3),
        "unit": "candidate pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(A):,} x {len(B):,} star polygons (0-3 holes, lognormal vertices), "
                               f"predicate {args.predicate}", "bbox_candidates": cands, "pairs": pairs},
        "roofline": {"bound": "hbm", "kernel": "whole step", "achieved": alg / (ms * 1e-3) / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": None, "alg_bytes_per_step_lower_bound": alg},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


def run_wkb(args):
    """GPU WKB ingest of the c2 data: 10M point records followed by 100k
    star-polygon records (ISO WKB, little endian), payload and offsets resident
    in HBM.  A step = vs_wkb_scan -> result read by the host -> output buffers
    -> vs_wkb_fill.  Algorithmic bytes per step (lower bound): the payload and
    offsets read once, and the decoded buffers written once (16 per point + 8
    record number; 16 per coordinate + 4 per ring / part / polygon offset + 8
    record number)."""
    import numpy as np
    import torch

    from oracle import wkb_ref
    from vibespatial_amd import _lib
    from vibespatial_amd import synthetic as S
    from vibespatial_amd.geometry import GeometryArray
    from vibespatial_amd.io import decode_wkb_device, encode_wkb

    _log("wkb: generating c2 workload and encoding it")
    w = S.c2(n_points=args.n_points, n_polys=args.n_polys)
    n, m = len(w.points), len(w.polygons)
    tags = np.concatenate([np.zeros(n, np.int8), np.ones(m, np.int8)])
    ga = GeometryArray(tags, w.points, np.arange(n), w.polygons, np.arange(n, n + m))
    payload, offsets, valid = encode_wkb(ga)
    dev = torch.device("cuda", 0)
    d_pay = torch.from_numpy(payload).to(dev)
    d_off = torch.from_numpy(offsets).to(dev)
    d_val = torch.from_numpy(valid).to(dev)
    col = decode_wkb_device(d_pay, d_off, d_val)
    ws = torch.empty(int(_lib.lib().vs_wkb_workspace_bytes(n + m)), dtype=torch.uint8, device=dev)
    # parity of the measured decode with the source buffers (bit-exact)
    P = w.polygons
    assert torch.equal(col.points.x.cpu().view(torch.int64), torch.from_numpy(w.points.x).view(torch.int64))
    assert torch.equal(col.points.y.cpu().view(torch.int64), torch.from_numpy(w.points.y).view(torch.int64))
    assert np.array_equal(col.polygons.x.cpu().numpy(), P.x) and np.array_equal(col.polygons.y.cpu().numpy(), P.y)
    assert np.array_equal(col.polygons.ring_offsets.cpu().numpy(), P.ring_offsets)
    assert np.array_equal(col.polygons.part_offsets.cpu().numpy(), P.part_offsets)
    assert np.array_equal(col.polygons.geom_offsets.cpu().numpy(), P.geom_offsets)
    del col
    for _ in range(max(args.warmup, 1)):
        decode_wkb_device(d_pay, d_off, d_val, workspace=ws)
    torch.cuda.synchronize()
    timing = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        col = decode_wkb_device(d_pay, d_off, d_val, workspace=ws, timing=timing)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    scan_ms = statistics.mean(a.elapsed_time(b) for (a, b), _ in timing)
    fill_ms = statistics.mean(c.elapsed_time(d) for _, (c, d) in timing)
    V = len(P.x)
    n_rings, n_parts = len(P.ring_offsets) - 1, len(P.part_offsets) - 1
    out_bytes = 24 * n + 16 * V + 4 * (n_rings + n_parts + m + 3) + 8 * m
    alg = len(payload) + 8 * (n + m + 1) + (n + m) + out_bytes
    device_ms = scan_ms + fill_ms

    cpu = None
    if args.cpu_sample > 0:
        # the oracle (plain Python over struct) on the first 3M point records
        # and the first 30k polygon records
        ks, kp = min(3_000_000, n), min(30_000, m)
        pt_pay = payload[: offsets[ks]]
        pg_pay = payload[offsets[n]: offsets[n + kp]]
        pg_off = offsets[n: n + kp + 1] - offsets[n]
        _log(f"wkb: cpu baseline on {ks} point + {kp} polygon records")
        tc0 = time.perf_counter()
        wkb_ref.decode_column(pt_pay, offsets[: ks + 1], valid[:ks])
        wkb_ref.decode_column(pg_pay, pg_off, valid[n: n + kp])
        tc = time.perf_counter() - tc0
        cpu = {"value": (len(pt_pay) + len(pg_pay)) / tc / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"first {ks:,} point records + first {kp:,} polygon records "
                         f"({(len(pt_pay) + len(pg_pay)) / 1e6:.1f} MB) decoded by the Python WKB restatement in "
                         f"{tc:.1f} s"}

    out = {
        "metric": "WKB ingest GB/s of payload decoded to device SoA buffers, 10M point + 100k polygon records",
        "value": len(payload) / (ms * 1e-3) / 1e9,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8->f64",
        "data": "synthetic",
        "config": {"workload": f"{n:,} point records + {m:,} star-polygon records (0-3 holes) as ISO WKB, "
                               f"{len(payload):,} bytes",
                   "records_per_s": (n + m) / (ms * 1e-3), "coordinates": int(V),
                   "scan_ms": scan_ms, "fill_ms": fill_ms},
        "roofline": {"bound": "hbm", "kernel": "wkb decode on the device (scan + prefix sums + fill)",
                     "achieved": alg / (device_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (device_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel_ms": device_ms, "alg_bytes_per_step_lower_bound": alg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
