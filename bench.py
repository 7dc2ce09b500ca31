"""Headline benchmark: Barnes-Hut force iterations per second at 1M vertices.

Workload (the reference's headline GPU configuration): Barabasi-Albert graph,
n = 1,000,000, m = 3 (~3M undirected edges), LinLog forces, gravity 0.05,
prevent-overlap with radius 0.1 * (1 + sqrt(degree)), Barnes-Hut theta 0.7,
one node per cell, positions uniform in [-1, 1]^2 (seed 0). A step is one
gvizForceEmbedderStep-equivalent iteration on the GPU (quadtree rebuild,
traversal, attraction, FA2 speed update, position update), replayed from a
captured hipGraph with positions resident in HBM.

    python bench.py [--gpus 1] [--steps 200] [--warmup 5]

Prints one JSON line. The roofline object covers the dominant kernel (the
Barnes-Hut traversal), timed with HIP events on the engine's stream; the
cpu_baseline object times the same step on the host cores (the reference
library compiled from its own sources when it was built into oracle/_ref,
else this repository's FP64 restatement of it) on a bounded sample.

The other measured paths, one JSON line each:

    python bench.py --workload allpairs_er65k    exact FR all-pairs step, ER n=65,536 avg deg 8
    python bench.py --workload tutte_delaunay1m  Tutte Jacobi sweeps, Delaunay n=1,000,000
    python bench.py --workload grip_rgg4m        GRIP schedule and finest-layer rounds, RGG n=4,000,000
                    [--grip-dim 3|4]             the same in three or four dimensions
    python bench.py --workload batch4096         4,096 graphs (50..20,000 vertices) laid out together
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from grapher_amd import graphs as G  # noqa: E402
from grapher_amd import native  # noqa: E402

METRIC = "BH force-iters/s at 1M vertices; time-to-stress-tol warm-start vs plain, 1 GPU"
FP32_PEAK_TFLOPS = 157.3
FLOP_PER_INTERACTION = 15  # FR / LinLog repulsion as written in the reference


PROFILE = os.path.join(ROOT, "profiles", "force_step_ba1m_linlog_mixed.md")


def traversal_traffic_bytes():
    """HBM bytes per traversal launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size)."""
    import re
    try:
        text = open(PROFILE).read()
    except OSError:
        return None
    m = re.search(r"Median per launch: ([0-9.]+) MB read \+ ([0-9.]+) MB written", text)
    if not m:
        return None
    return int((float(m.group(1)) + float(m.group(2))) * 1e6)


def valu_issue(profile):
    """Steady-state VALU issue rate and lane use of a kernel from the
    committed rocprofv3 --pmc table of `profile` (median over dispatches
    after the first)."""
    import re
    import statistics
    try:
        text = open(os.path.join(ROOT, "profiles", profile)).read()
    except OSError:
        return None
    sec = text.split("## Vector ALU issue", 1)
    if len(sec) < 2:
        return None
    rows = re.findall(r"^\| \d+ \|.*\| ([0-9.]+) \| ([0-9.]+) \|$", sec[1], re.M)
    if len(rows) < 2:
        return None
    rows = rows[1:]
    return {"issue_frac": round(statistics.median(float(a) for a, _ in rows), 3),
            "lane_use": round(statistics.median(float(b) for _, b in rows), 3),
            "source": f"profiles/{profile} (pmc: SQ_INSTS_VALU, SQ_THREAD_CYCLES_VALU, "
                      "SQ_ACTIVE_INST_VALU, GRBM_GUI_ACTIVE)"}


def workload(n):
    g_t0 = time.time()
    n, edges = G.barabasi_albert(n, 3, 0)
    off, adj = G.csr_from_edges(n, edges)
    pos = G.uniform_positions(n, 0)
    return n, edges, off, adj, pos, time.time() - g_t0


PARAMS = dict(model=native.MODEL_LINLOG, barnes_hut=True, prevent_overlap=True,
              nodes_per_cell=1, edge_length=10.0, theta=0.7, gravity=0.05,
              overlap_constant=100.0, radius_base=0.1, radius_per_degree=1.0,
              jitter=1.0)


def cpu_baseline(n, edges, pos, seconds_budget=20.0):
    """Times the CPU path on a bounded sample of the same workload: steps
    from `pos`, the layout the GPU's timed steps start from, on the host
    threads one GPU's share of the machine allows (16; the reference's
    default pool takes every online CPU, which here would oversubscribe the
    share)."""
    threads = min(16, len(os.sched_getaffinity(0)))
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    iters = 3
    if os.path.exists(harness):
        import struct
        with tempfile.TemporaryDirectory() as td:
            fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
            with open(fi, "wb") as f:
                f.write(struct.pack("<qq", n, edges.shape[0]))
                f.write(np.ascontiguousarray(edges, dtype="<i8").tobytes())
                f.write(struct.pack("<qqqqqq", 1, 1, 1, 1, iters, threads - 1))
                f.write(struct.pack("<ddddddd", 10.0, 0.7, 0.05, 100.0, 0.1, 1.0, 1.0))
                f.write(np.ascontiguousarray(pos, dtype="<f8").tobytes())
            t0 = time.time()
            subprocess.run([harness, "time_force", fi, fo], check=True,
                           timeout=seconds_budget * 10)
            wall = time.time() - t0
            times = np.fromfile(fo, dtype="<f8")
        per = float(np.mean(times))
        return {"value": 1.0 / per, "unit": "iters/s", "cores": threads,
                "kind": "reference",
                "sample": f"{iters} gvizForceEmbedderStep calls of the reference library "
                          f"(compiled from its sources) on the same 1M-vertex graph, from "
                          f"the layout the GPU's timed steps start from, pool of "
                          f"{threads} threads (one GPU's share of the host) "
                          f"(caller + {threads - 1} workers); mean {per:.3f} s/step; "
                          f"harness wall {wall:.1f} s"}
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    g = O.Graph(*G.csr_from_edges(n, edges))
    p = O.params(model=1, barnes_hut=True, prevent_overlap=True, npc=1, theta=0.7,
                 gravity=0.05, radius_base=0.1, radius_per_degree=1.0)
    pool = O.Pool(threads)
    run = O.ForceRun(g, p, pos, pool)
    times = []
    for _ in range(iters):
        t0 = time.time()
        run.step()
        times.append(time.time() - t0)
    per = float(np.mean(times))
    return {"value": 1.0 / per, "unit": "iters/s", "cores": threads, "kind": "port",
            "sample": f"{iters} steps of the FP64 C restatement on the same 1M-vertex "
                      f"graph, from the layout the GPU's timed steps start from, "
                      f"{threads} threads; mean {per:.3f} s/step"}


HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def run_harness(mode, payload, timeout):
    """Runs the reference library's timing harness (built from the
    reference's sources into oracle/_ref) and returns its output bytes."""
    with tempfile.TemporaryDirectory() as td:
        fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fi, "wb") as f:
            f.write(payload)
        subprocess.run([HARNESS, mode, fi, fo], check=True, timeout=timeout)
        return open(fo, "rb").read()


def graph_bytes(n, edges):
    import struct
    return struct.pack("<qq", n, edges.shape[0]) + \
        np.ascontiguousarray(edges, dtype="<i8").tobytes()


def emit(line):
    print(json.dumps(line), flush=True)


def base_line(metric, value, unit, args, ms, dtype, workload, extra_config,
              higher=True):
    return {"metric": metric, "value": round(value, 4), "unit": unit, "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
            "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": "synthetic (seeded generators in grapher_amd/graphs.py)",
            "config": dict({"workload": workload, "parallelism": "single GPU"},
                           **extra_config)}


def bench_allpairs(args):
    """Exact FR all-pairs step (Barnes-Hut off), ER n=65,536, avg degree 8."""
    import struct
    n, edges = G.erdos_renyi(65536, 8.0, 0)
    off, adj = G.csr_from_edges(n, edges)
    pos = G.uniform_positions(n, 0)
    eng = native.ForceEngine(off, adj, precision=args.precision, model=native.MODEL_FR,
                             barnes_hut=False)
    eng.begin(pos)
    steps = args.steps if args.steps_given else 50
    args.steps = steps
    if args.warmup:
        eng.run(args.warmup, stats=False)
    t0 = time.perf_counter()
    _, ran = eng.run(steps, stats=False)
    el = time.perf_counter() - t0
    rep_ms, inter, _ = eng.profile(3)
    launch_s = rep_ms / 1e3 / 3
    flops = FLOP_PER_INTERACTION * inter / 3
    achieved = flops / launch_s / 1e12
    line = base_line("exact all-pairs FR force-iters/s at N=65,536", steps / el,
                     "iters/s", args, 1e3 * el / steps,
                     {"mixed": "f32 (f64 positions)", "fp32": "f32", "fp64": "f64"}[args.precision],
                     "ER n=65,536 avg degree 8, FR, Barnes-Hut off (exact all-pairs), "
                     "uniform [-1,1]^2 init",
                     {"n_vertices": n, "n_edges": int(edges.shape[0]),
                      "precision": args.precision})
    line["roofline"] = {"bound": "valu", "achieved": round(achieved, 3),
                        "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(achieved / FP32_PEAK_TFLOPS, 5), "traffic": None,
                        "kernel": "k_allpairs",
                        "note": "FP32 vector ALU (no matrix op on this path); 15 FLOP per "
                                "pair evaluation, N(N-1) evaluations per launch",
                        "launch_ms": round(launch_s * 1e3, 4),
                        "interactions_per_launch": int(inter / 3),
                        "valu": valu_issue("allpairs_er65k.md")}
    cpu = None
    if not args.no_cpu_baseline and os.path.exists(HARNESS):
        threads = min(16, len(os.sched_getaffinity(0)))
        iters = 2
        payload = graph_bytes(n, edges) + struct.pack("<qqqqqq", 0, 0, 0, 1, iters, threads - 1) \
            + struct.pack("<ddddddd", 10.0, 1.0, 0.0, 100.0, 0.0, 0.0, 1.0) \
            + np.ascontiguousarray(pos, dtype="<f8").tobytes()
        times = np.frombuffer(run_harness("time_force", payload, 600), dtype="<f8")
        per = float(np.mean(times))
        cpu = {"value": 1.0 / per, "unit": "iters/s", "cores": threads, "kind": "reference",
               "sample": f"{iters} gvizForceEmbedderStep calls (Barnes-Hut off) of the reference "
                         f"library on the same graph and positions, pool of {threads} "
                         f"threads; mean {per:.3f} s/step"}
    line["cpu_baseline"] = cpu
    if cpu:
        line["speedup_vs_cpu"] = round(line["value"] / cpu["value"], 1)
    emit(line)


def tutte_setup(n, hull, radius=200.0):
    """gvizTutteFixConvexPolygon + gvizTutteEmbedderSeedInterior
    (src/embedders/gvizTutteEmbedder.c:207-247): hull on a circle, every other
    vertex at the hull's centroid."""
    k = np.arange(len(hull), dtype=np.float64)
    ang = 2.0 * np.pi * k / len(hull)
    ring = np.stack([radius * np.cos(ang), radius * np.sin(ang)], 1)
    cx = 0.0
    cy = 0.0
    for x, y in ring:  # the reference's summation order
        cx += x
        cy += y
    pos = np.empty((n, 2))
    pos[:, 0] = cx / len(hull)
    pos[:, 1] = cy / len(hull)
    pos[hull] = ring
    isb = np.zeros(n, np.uint8)
    isb[hull] = 1
    return pos, isb


def bench_tutte(args):
    """Tutte Jacobi sweeps, Delaunay triangulation of n=1,000,000 uniform
    points, convex hull pinned on a circle of radius 200, alpha = 1."""
    import struct
    n, edges, _, hull = G.delaunay(1_000_000, 0)
    off, adj = G.csr_from_edges(n, edges)
    pos, isb = tutte_setup(n, hull)
    prec = "fp32" if args.precision == "fp32" else "fp64"
    eng = native.TutteEngine(off, adj, precision=prec)
    eng.upload(pos, isb)
    steps = args.steps if args.steps_given else 1000
    args.steps = steps
    if args.warmup:
        eng.sweeps(args.warmup, record=False)
    eng.upload(pos, isb)
    t0 = time.perf_counter()
    _, ran = eng.sweeps(steps, 1.0, -1.0, record=False)
    el = time.perf_counter() - t0
    assert ran == steps
    sweep_ms, _ = eng.profile(20)
    launch_s = sweep_ms / 1e3 / 20
    nnz = int(adj.shape[0])
    ps = 16 if prec == "fp64" else 8
    bytes_per = 4 * (n + 1) + 4 * nnz + 2 * ps * n + n
    achieved = bytes_per / launch_s / 1e9
    line = base_line("Tutte Jacobi sweeps/s, Delaunay N=1,000,000", steps / el, "sweeps/s",
                     args, 1e3 * el / steps, "f64" if prec == "fp64" else "f32",
                     "Delaunay triangulation of 1,000,000 uniform points, convex hull "
                     "pinned on a circle (radius 200), interior seeded at the centroid, "
                     "Jacobi alpha = 1", {"n_vertices": n, "n_edges": int(edges.shape[0]),
                                           "hull": int(len(hull)), "precision": prec})
    line["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0,
                        "unit": "GB/s", "frac": round(achieved / 8000.0, 5),
                        "traffic": None, "kernel": "k_sweep",
                        "note": "algorithmic bytes per sweep = 4(N+1) offsets + 4 nnz "
                                "indices + N x (read + write position) + N boundary "
                                "bytes; neighbour gathers are not counted",
                        "launch_ms": round(launch_s * 1e3, 4),
                        "bytes_per_launch": int(bytes_per)}
    cpu = None
    if not args.no_cpu_baseline and os.path.exists(HARNESS):
        iters = 60
        payload = graph_bytes(n, edges) + struct.pack("<qq", len(hull), iters) \
            + np.ascontiguousarray(hull, dtype="<i8").tobytes() + struct.pack("<d", 200.0)
        times = np.frombuffer(run_harness("time_tutte", payload, 600), dtype="<f8")
        per = float(np.mean(times))
        cpu = {"value": 1.0 / per, "unit": "sweeps/s", "cores": 1, "kind": "reference",
               "sample": f"{iters} gvizTutteEmbedderStep(dt=1) calls of the reference "
                         f"library on the same graph and boundary (its sweep is serial); "
                         f"mean {per * 1e3:.2f} ms/sweep"}
    line["cpu_baseline"] = cpu
    if cpu:
        line["speedup_vs_cpu"] = round(line["value"] / cpu["value"], 1)
    emit(line)


def regular_simplex2(side):
    """Three vertices of an equilateral triangle centred at the origin, built
    like makeRegularSimplex (src/embedders/gvizGRIPEmbedder.c:335-420) for
    n = 2: unit vectors with pairwise dot product -1/2, scaled to `side`."""
    v = np.zeros((3, 2))
    v[0] = (1.0, 0.0)
    v[1, 0] = -0.5
    v[1, 1] = np.sqrt(1.0 - 0.25)
    v[2, 0] = -0.5
    v[2, 1] = (-0.5 - v[2, 0] * v[1, 0]) / v[1, 1]
    edge = np.sqrt(((v[0] - v[1]) ** 2).sum())
    return v * (side / edge)


def regular_simplex(dim, side):
    """dim + 1 unit vectors with pairwise dot product -1/dim (the
    construction of makeRegularSimplex, src/embedders/gvizGRIPEmbedder.c:
    335-420), scaled so that an edge is `side` long."""
    if dim == 2:
        return regular_simplex2(side)
    v = np.zeros((dim + 1, dim))
    c = -1.0 / dim
    v[0, 0] = 1.0
    for k in range(1, dim + 1):
        for j in range(k):
            v[k, j] = (c - v[k, :j] @ v[j, :j]) / v[j, j]
        if k < dim:
            v[k, k] = np.sqrt(max(1.0 - v[k, :k] @ v[k, :k], 0.0))
    return v * (side / np.sqrt(((v[0] - v[1]) ** 2).sum()))


def prefer_batch(n_active, n_edges, visible, targets):
    """src/algorithms/search/gvizKNearest.c:149-166."""
    if visible == 0 or visible >= targets or visible > 256 or n_active == 0:
        return False
    if 2.0 * n_edges / n_active > 32.0:
        return True
    return targets >= 8 * visible


def bench_grip(args):
    """GRIP on an RGG with n=4,000,000 (largest component), k = 10, the
    reference's Embed schedule (<= 30 rounds per layer, stop below 0.01);
    value = refinement rounds/s at the finest layer."""
    import struct
    k = 10
    dim = args.grip_dim
    n, edges, _ = G.random_geometric(4_000_000, 8.0, 3)
    act, inside = G.largest_component(n, edges)
    off, adj = G.csr_from_edges(n, inside)
    na = int(act.sum())
    ne = int(inside.shape[0])
    eng = native.GripEngine(off, adj, active=act, kcap=k, dim=dim)
    t_all = time.perf_counter()
    filt, borders = eng.mis_filtration(dim)
    t_begin = time.perf_counter() - t_all
    layers = len(borders)
    pos = np.zeros((n, dim))
    pos[filt[:dim + 1]] = regular_simplex(dim, 10.0 * 1000)
    eng.upload(pos)
    vis = np.zeros(n, np.uint8)
    vis[filt[:dim + 1]] = 1
    layer = layers - 1
    end = int(borders[layer])
    eng.clear_decorators(end)
    eng.update_knn(end, vis, k, prefer_batch(na, ne, dim + 1, end))
    rounds_total = 0
    while True:
        active = int(borders[layer])
        for _ in range(30):
            st = eng.refine_round(active, layer)
            rounds_total += 1
            if st.maxDisp < 10.0 * 1e-3:
                break
        if layer == 0:
            break
        layer -= 1
        begin, end = int(borders[layer + 1]), int(borders[layer])
        eng.place(begin, end, vis, k, prefer_batch(na, ne, begin, end - begin))
        vis[filt[begin:end]] = 1
        eng.clear_decorators(end)
        eng.update_knn(end, vis, k, prefer_batch(na, ne, end, end))
    t_embed = time.perf_counter() - t_all
    steps = args.steps if args.steps_given else 20
    args.steps = steps
    for _ in range(args.warmup):
        eng.refine_round(na, 0)
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.refine_round(na, 0)
    el = time.perf_counter() - t0
    f_ms, r_ms = eng.profile_rounds(na, 0, 5)
    launch_s = f_ms / 1e3 / 5
    deg_sum = int(adj.shape[0])
    # the fixed convention for a GRIP round's algorithmic bytes: per active
    # vertex k kNN slots of a 4-byte id and a 2-byte distance, 4 bytes per
    # adjacency entry, an FP32 position read (4 dim) and written (4 dim), the
    # displacement and previous displacement (8 dim), heat and previous
    # cosine (8); plus the CSR offsets
    bytes_per = na * (6 * k + 16 * dim + 8) + 4 * deg_sum + 4 * (n + 1)
    # what this engine's k_refine itself moves once per active vertex at
    # layer 0 (FP64 state): its filtration entry (4), four dim-vectors of
    # FP64 (position read, displacement read and written, previous
    # displacement written), CSR offsets (8), kNN count (4), heat and cosine
    # read and written (32), last force magnitude (8), the computed flag
    # read and written (2); 4 B per adjacency entry and per kNN id
    bytes_engine = na * (58 + 32 * dim) + 4 * deg_sum + 4 * k * na
    achieved = bytes_per / launch_s / 1e9
    metric = "GRIP refinement rounds/s at the finest layer, RGG N=4,000,000, k=10"
    if dim != 2:
        metric += f", dim {dim}"
    line = base_line(metric,
                     steps / el, "rounds/s", args, 1e3 * el / steps, "f64",
                     "random geometric graph, 4,000,000 uniform points in the unit square, "
                     f"radius for average degree 8, largest component; GRIP dim {dim}, "
                     "kNN capacity 10, K = 10 constant; the reference's Embed schedule",
                     {"n_vertices": n, "active_vertices": na, "n_edges": ne,
                      "layers": layers, "borders": [int(b) for b in borders]})
    line["embed_s"] = round(t_embed, 3)
    line["filtration_s"] = round(t_begin, 3)
    line["schedule_rounds"] = rounds_total
    line["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0,
                        "unit": "GB/s", "frac": round(achieved / 8000.0, 5),
                        "traffic": None,
                        "kernel": "k_refine (layer 0)" if dim == 2 else
                                  f"k_refine_nd<{dim}> (layer 0)",
                        "note": "algorithmic bytes per launch = N_active (6k + 16 dim + 8) "
                                "+ 4 per adjacency entry + 4 (N + 1): kNN slots of a 4-byte "
                                "id and 2-byte distance, FP32 position read and written, "
                                "displacement and previous displacement, heat and previous "
                                "cosine, CSR; neighbour position gathers not counted",
                        "launch_ms": round(launch_s * 1e3, 4),
                        "round_ms": round(r_ms / 5, 4),
                        "bytes_per_launch": int(bytes_per),
                        "engine_bytes_per_launch": int(bytes_engine),
                        "engine_frac": round(bytes_engine / launch_s / 1e9 / 8000.0, 5)}
    cpu = None
    if dim == 2 and not args.no_cpu_baseline and os.path.exists(HARNESS):
        threads = min(16, len(os.sched_getaffinity(0)))
        sn, se, _ = G.random_geometric(500_000, 8.0, 3)
        sact, _ = G.largest_component(sn, se)
        payload = graph_bytes(sn, se) + np.ascontiguousarray(sact, dtype="<i8").tobytes() \
            + struct.pack("<qqqq", k, 2, 64, threads - 1)
        b = run_harness("time_grip", payload, 900)
        off_b = 16
        last = None
        while True:
            lay = struct.unpack_from("<q", b, off_b)[0]
            off_b += 8
            if lay == -1:
                break
            cnt = struct.unpack_from("<q", b, off_b)[0]
            off_b += 16
            rs = struct.unpack_from("<2d", b, off_b)
            off_b += 16
            last = (lay, cnt, rs)
        per = float(np.mean(last[2])) * (na / last[1])
        cpu = {"value": 1.0 / per, "unit": "rounds/s", "cores": threads, "kind": "reference",
               "sample": f"the reference's GRIP schedule on an RGG of 500,000 points (same "
                         f"generator and parameters, {last[1]} active vertices), 2 rounds per "
                         f"layer with its pool of {threads} threads; its finest-layer round "
                         f"({np.mean(last[2]):.3f} s) scaled linearly to {na} active vertices"}
    line["cpu_baseline"] = cpu
    if cpu:
        line["speedup_vs_cpu"] = round(line["value"] / cpu["value"], 1)
    emit(line)


def warm_start_line(graphs=100, n=200_000):
    """The metric's second half: wall time to the plain method's final
    stress (within 1%) from the learned start versus from the random start,
    on unseen connected SBM graphs of ~200k vertices (FR, Barnes-Hut 0.7,
    mixed precision; the plain method's final stress is its stress after 500
    steps; stress checked every 2 steps on both sides). Summed and median
    over the graphs; the bound is the plain time over the network's own
    time, which no start this network computes can beat."""
    from grapher_amd import warmstart as WS
    if not os.path.exists(WS.WEIGHTS):
        return {"error": f"{WS.WEIGHTS} missing"}
    ws = WS.WarmStart.load()
    rng = np.random.default_rng(2024)
    specs = [(int(rng.integers(2, 33)), 7000 + i) for i in range(graphs)]
    rows = WS.evaluate_many(ws, specs, n=n)
    out = {"workload": f"{graphs} unseen connected SBM graphs, ~{n} vertices, 2-32 blocks, "
                       "p_in = 8/size, p_out = 0.2/n; FR Barnes-Hut 0.7, mixed precision",
           "target": "stress within 1% of the plain method's stress after 500 steps"}
    z = np.load(WS.WEIGHTS, allow_pickle=False)
    if "graphs" in z.files and "n_min" in z.files:
        out["network"] = (f"trained on the GPU on FP64 FR layouts of {int(z['graphs'])} SBM "
                          f"graphs ({int(z['n_min'])}-{int(z['n_max'])} vertices, "
                          f"{int(z['layout_iters'])} steps each)")
    out.update(WS.summarize(rows))
    return out


def bench_batch(args):
    """4,096 graphs, FR Barnes-Hut (theta 1.0, the reference's default) in
    FP32, laid out together: every graph runs until its maxDisp < 1.0 (a
    tenth of the edge length) or 800 steps. value = vertex-steps per second over the batch; per-graph latency
    = time from the start of the run to the check that saw it stop."""
    import struct
    t0 = time.time()
    graphs, kinds = G.batch_mix()
    sizes = np.array([len(o) - 1 for o, _ in graphs])
    print(f"# batch of {len(graphs)} graphs, {sizes.sum()} vertices built in "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    starts = [G.uniform_positions(int(n), 1000 + i) for i, n in enumerate(sizes)]
    start = np.concatenate(starts)
    cfg = dict(model=native.MODEL_FR, precision="fp32", theta=1.0, edge_length=10.0,
               gravity=0.0)
    b = native.BatchEngine(graphs, **cfg)
    b.begin(start)
    order, spl, tw = b.autotune(4)
    b.begin(start)
    max_steps, eps = 800, 1.0
    t1 = time.perf_counter()
    launched, its, lat, _ = b.run(max_steps, epsilon=eps)
    el = time.perf_counter() - t1
    vsteps = float((its * sizes).sum())
    args.steps = int(launched)
    b.begin(start)
    trav_ms, inter, step_ms = b.profile(3)
    launch_s = trav_ms / 1e3 / 3
    achieved = FLOP_PER_INTERACTION * inter / 3 / launch_s / 1e12
    line = base_line("batched layouts: vertex-steps/s over 4,096 graphs", vsteps / el,
                     "vertex-steps/s", args, 1e3 * el / max(launched, 1), "f32",
                     "4,096 graphs: random trees, 2-D grids, Erdos-Renyi (average degree "
                     "4), vertex counts log-uniform in [50, 20,000]; FR Barnes-Hut theta 1.0 "
                     "FP32; each graph until maxDisp < 1.0 (a tenth of an edge) or 800 steps",
                     {"graphs": len(graphs), "vertices": int(sizes.sum()),
                      "edges": int(sum(len(a) for _, a in graphs) // 2),
                      "schedule": {"wave_order": order, "steps_per_launch": spl,
                                   "traversal_waves_per_workgroup": tw}})
    line["latency_ms"] = {"p50": round(float(np.percentile(lat, 50)), 2),
                          "p99": round(float(np.percentile(lat, 99)), 2),
                          "max": round(float(lat.max()), 2)}
    done = its < max_steps
    if done.any():
        line["latency_ms_converged"] = {
            "graphs": int(done.sum()),
            "p50": round(float(np.percentile(lat[done], 50)), 2),
            "p99": round(float(np.percentile(lat[done], 99)), 2)}
    line["graphs_per_s"] = round(len(graphs) / el, 1)
    line["steps_per_graph"] = {"p50": int(np.percentile(its, 50)),
                               "max": int(its.max()),
                               "stopped": int((its < max_steps).sum())}
    line["roofline"] = {"bound": "valu", "achieved": round(achieved, 3),
                        "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(achieved / FP32_PEAK_TFLOPS, 5), "traffic": None,
                        "kernel": "k_traverse (batch)",
                        "note": "FP32 vector ALU; 15 FLOP per repulsion evaluation over "
                                "all graphs of one step",
                        "launch_ms": round(launch_s * 1e3, 4),
                        "interactions_per_launch": int(inter / 3),
                        "step_ms_uncaptured": round(step_ms / 3, 4)}
    cpu = None
    if not args.no_cpu_baseline and os.path.exists(HARNESS):
        threads = min(16, len(os.sched_getaffinity(0)))
        order_idx = np.argsort(sizes)
        pick = order_idx[np.linspace(0, len(sizes) - 1, 24).astype(int)]
        tot_t, tot_vs = 0.0, 0.0
        for gi in pick.tolist():
            o, a = graphs[gi]
            n = len(o) - 1
            src = np.repeat(np.arange(n), np.diff(o))
            e = np.stack([src, a], 1)
            e = e[e[:, 0] < e[:, 1]]
            steps = int(min(its[gi], 10))
            payload = graph_bytes(n, e) + struct.pack("<qqqqqq", 0, 1, 0, 1, steps,
                                                      threads - 1) \
                + struct.pack("<ddddddd", 10.0, 1.0, 0.0, 100.0, 0.0, 0.0, 1.0) \
                + np.ascontiguousarray(starts[gi], dtype="<f8").tobytes()
            t = np.frombuffer(run_harness("time_force", payload, 600), dtype="<f8")
            tot_t += float(t.sum())
            tot_vs += steps * n
        cpu = {"value": tot_vs / tot_t, "unit": "vertex-steps/s", "cores": threads,
               "kind": "reference",
               "sample": f"24 graphs of the batch spread over its size range, up to 10 "
                         f"gvizForceEmbedderStep calls each (the reference lays out one "
                         f"graph at a time), pool of {threads} threads; "
                         f"{tot_vs:.0f} vertex-steps in {tot_t:.1f} s"}
    line["cpu_baseline"] = cpu
    if cpu:
        line["speedup_vs_cpu"] = round(line["value"] / cpu["value"], 1)
    emit(line)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="force_ba1m",
                    choices=["force_ba1m", "allpairs_er65k", "tutte_delaunay1m", "grip_rgg4m",
                             "batch4096"])
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--precision", default="mixed", choices=["fp32", "mixed", "fp64"])
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--grip-dim", type=int, default=2, choices=[2, 3, 4],
                    help="layout dimension of the grip_rgg4m workload")
    ap.add_argument("--no-warm-start", action="store_true")
    ap.add_argument("--edges-file", default=None,
                    help="force_ba1m settings on the graph of a .edges file (Network "
                         "Repository format, undirected) instead of the BA graph")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device-to-device copy (the known-size "
                         "kernel that scales FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print("copy calibration: 3 copies of 268435456 bytes")
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not "
              "supported (the multi-GPU design is described in DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    if native.device_count() <= 0:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    args.steps_given = args.steps is not None
    if args.workload != "force_ba1m":
        {"allpairs_er65k": bench_allpairs, "tutte_delaunay1m": bench_tutte,
         "grip_rgg4m": bench_grip, "batch4096": bench_batch}[args.workload](args)
        return
    if args.steps is None:
        args.steps = 200

    if args.edges_file:
        from grapher_amd import edges_file
        g_t0 = time.time()
        off, adj = edges_file.load_csr(args.edges_file)
        n = len(off) - 1
        # edge list for the CPU baseline's graph: each undirected edge once
        # (self loops left out)
        src = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
        keep = src < adj
        edges = np.stack([src[keep], adj[keep]], 1)
        pos = G.uniform_positions(n, 0)
        gen_s = time.time() - g_t0
        args.n = n
    else:
        n, edges, off, adj, pos, gen_s = workload(args.n)
    print(f"# workload built in {gen_s:.1f}s: n={n} edges={edges.shape[0]}",
          file=sys.stderr, flush=True)
    eng = native.ForceEngine(off, adj, precision=args.precision, **PARAMS)
    eng.begin(pos)
    if args.warmup:
        eng.run(args.warmup, stats=False)
    # the layout the timed steps start from (downloaded outside the timed
    # region): the CPU baseline steps from the same state
    pos_timed = eng.positions() if not args.no_cpu_baseline else pos
    t0 = time.perf_counter()
    stats, ran = eng.run(args.steps, stats=True)
    t1 = time.perf_counter()
    assert ran == args.steps
    elapsed = t1 - t0
    ips = args.steps / elapsed
    print(f"# timed {args.steps} steps in {elapsed:.3f}s; last maxDisp "
          f"{stats['maxDisp'][-1]:.4g}", file=sys.stderr, flush=True)

    # dominant kernel: HIP events around the traversal on the engine's stream
    trav_ms, inter, step_ms = eng.profile(args.profile_steps)
    ps = max(args.profile_steps, 1)
    trav_avg_s = trav_ms / 1e3 / ps
    flops = FLOP_PER_INTERACTION * inter / ps
    achieved = flops / trav_avg_s / 1e12 if trav_avg_s > 0 else 0.0
    roofline = {"bound": "valu", "achieved": round(achieved, 3),
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 5),
                "traffic": (traversal_traffic_bytes()
                            if args.n == 1_000_000 and not args.edges_file else None),
                "traffic_source": "profiles/force_step_ba1m_linlog_mixed.md (pmc)",
                "kernel": "k_traverse (Barnes-Hut repulsion)",
                "note": "compute-bound on the FP32 vector ALU (no matrix op on this "
                        "path); peak = MI355X FP32 vector; achieved = 15 FLOP x "
                        "repulsion evaluations per launch / mean launch time",
                "launch_ms": round(trav_avg_s * 1e3, 4),
                "interactions_per_launch": int(inter / ps),
                "step_ms_unCaptured": round(step_ms / ps, 4),
                "valu": valu_issue("force_step_ba1m_linlog_mixed.md")}

    warm = None
    if not args.no_warm_start:
        warm = warm_start_line()

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(n, edges, pos_timed)
        except Exception as e:  # reported, never silently replaced
            cpu = {"value": None, "error": str(e)}

    line = {
        "metric": METRIC, "value": round(ips, 3), "unit": "iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"mixed": "f32 (f64 positions)", "fp32": "f32", "fp64": "f64"}[args.precision],
        "data": ("synthetic (seeded Barabasi-Albert graph, seeded uniform positions)"
                 if not args.edges_file else
                 f"graph from {os.path.basename(args.edges_file)}, seeded uniform positions"),
        "config": {"workload": ("BA n=1,000,000 m=3" if not args.edges_file else
                                os.path.basename(args.edges_file)) +
                               " LinLog gravity 0.05 prevent-overlap "
                               "(r=0.1(1+sqrt(deg))) Barnes-Hut theta 0.7, uniform "
                               "[-1,1]^2 init", "n_vertices": n,
                   "n_edges": int(edges.shape[0]), "precision": args.precision,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "warm_start": warm,
    }
    if cpu and cpu.get("value"):
        line["speedup_vs_cpu"] = round(ips / cpu["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
