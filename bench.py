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
    """Times the CPU path on a bounded sample of the same workload: the first
    steps from the same initial positions, on up to 16 host threads."""
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
                          f"(compiled from its sources) on the same 1M-vertex graph and "
                          f"initial positions, pool of {threads} threads "
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
                      f"graph and initial positions, {threads} threads; mean {per:.3f} s/step"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--precision", default="mixed", choices=["fp32", "mixed", "fp64"])
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
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

    n, edges, off, adj, pos, gen_s = workload(args.n)
    print(f"# workload built in {gen_s:.1f}s: n={n} edges={edges.shape[0]}",
          file=sys.stderr, flush=True)
    eng = native.ForceEngine(off, adj, precision=args.precision, **PARAMS)
    eng.begin(pos)
    if args.warmup:
        eng.run(args.warmup, stats=False)
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
    roofline = {"bound": "mfma", "achieved": round(achieved, 3),
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 5),
                "traffic": traversal_traffic_bytes() if args.n == 1_000_000 else None,
                "traffic_source": "profiles/force_step_ba1m_linlog_mixed.md (pmc)",
                "kernel": "k_traverse (Barnes-Hut repulsion)",
                "note": "compute-bound on the FP32 vector ALU (no matrix op on this "
                        "path); peak = MI355X FP32 vector; achieved = 15 FLOP x "
                        "repulsion evaluations per launch / mean launch time",
                "launch_ms": round(trav_avg_s * 1e3, 4),
                "interactions_per_launch": int(inter / ps),
                "step_ms_unCaptured": round(step_ms / ps, 4)}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(n, edges, pos)
        except Exception as e:  # reported, never silently replaced
            cpu = {"value": None, "error": str(e)}

    line = {
        "metric": METRIC, "value": round(ips, 3), "unit": "iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"mixed": "f32 (f64 positions)", "fp32": "f32", "fp64": "f64"}[args.precision],
        "data": "synthetic (seeded Barabasi-Albert graph, seeded uniform positions)",
        "config": {"workload": "BA n=1,000,000 m=3 LinLog gravity 0.05 prevent-overlap "
                               "(r=0.1(1+sqrt(deg))) Barnes-Hut theta 0.7, uniform "
                               "[-1,1]^2 init", "n_vertices": n,
                   "n_edges": int(edges.shape[0]), "precision": args.precision,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu and cpu.get("value"):
        line["speedup_vs_cpu"] = round(ips / cpu["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
