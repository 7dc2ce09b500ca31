"""Writes the golden fixtures in tests/golden/*.npz by running the reference
library (oracle/_ref, compiled from the reference's own sources by
`make -C oracle ref`) on small seeded inputs.

Each fixture holds the inputs and the reference's outputs; nothing of the
reference's code. Run from the repository root:

    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))

import refio  # noqa: E402
from grapher_amd import graphs as G  # noqa: E402


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print(f"{name}: {os.path.getsize(path) / 1024:.1f} KiB")


def quadtree_cases():
    pts = G.uniform_positions(300, 3)
    r = refio.quadtree(pts, np.ones(300), 1)
    save("quadtree_uniform300", points=pts, masses=np.ones(300), npc=1, **r)

    rng = np.random.default_rng(7)
    centers = rng.uniform(-5, 5, size=(4, 2))
    pts = np.concatenate([c + rng.normal(scale=0.05, size=(100, 2)) for c in centers])
    pts[10] = pts[11]          # exact duplicates force overflow leaves
    pts[200] = pts[201] = pts[202]
    ms = 1.0 + rng.integers(0, 6, size=400).astype(np.float64)
    r = refio.quadtree(pts, ms, 3)
    save("quadtree_clustered_npc3", points=pts, masses=ms, npc=3, **r)

    pts = G.uniform_positions(200, 11) * 1000.0
    pts[5] = pts[6]
    pts[50] = pts[51]
    r = refio.quadtree(pts, np.ones(200), 1)
    save("quadtree_dup_npc1", points=pts, masses=np.ones(200), npc=1, **r)


FORCE_DEFAULTS = dict(npc=1, edge_length=10.0, theta=1.0, overlap_constant=100.0,
                      radius_base=0.0, radius_per_degree=0.0, jitter=1.0,
                      gravity=0.0, prevent_overlap=0)


def force_case(name, n, edges, pos, iters, dump_every, **kw):
    args = dict(FORCE_DEFAULTS)
    args.update(kw)
    stats, dumps = refio.force(n, edges, pos, iters=iters, dump_every=dump_every,
                               **args)
    out = dict(n=n, edges=edges, pos0=pos, iters=iters, stats=stats,
               dump_iters=np.array(sorted(dumps), dtype=np.int64))
    for k, v in args.items():
        out["param_" + k] = v
    for it, d in dumps.items():
        for k, v in d.items():
            out[f"it{it}_{k}"] = v
    save(name, **out)


def force_cases():
    n, e = G.grid(12, 12)
    force_case("force_fr_bh_grid12", n, e, G.uniform_positions(n, 0), 8, 1,
               model=0, barnes_hut=1, theta=0.5)
    n, e = G.barabasi_albert(300, 3, 5)
    force_case("force_linlog_bh_ba300", n, e, G.uniform_positions(n, 1), 8, 1,
               model=1, barnes_hut=1, theta=0.7, gravity=0.05,
               prevent_overlap=1, radius_base=0.1, radius_per_degree=1.0)
    n, e = G.erdos_renyi(200, 8, 9)
    force_case("force_fr_exact_er200", n, e, G.uniform_positions(n, 2), 5, 1,
               model=0, barnes_hut=0)
    n, e = G.grid(15, 15)
    force_case("force_fr_bh_npc2_grid15", n, e, G.uniform_positions(n, 4), 5, 1,
               model=0, barnes_hut=1, theta=0.8, npc=2)
    # The CPU-runnable headline case: 100x100 grid, FR, BH theta 0.5, 500 its.
    n, e = G.grid(100, 100)
    force_case("force_fr_bh_grid100", n, e, G.uniform_positions(n, 0), 500, 499,
               model=0, barnes_hut=1, theta=0.5)


def tutte_cases():
    n, e = G.grid(6, 6)
    ring = [0, 1, 2, 3, 4, 5, 11, 17, 23, 29, 35, 34, 33, 32, 31, 30, 24, 18, 12, 6]
    for gs in (0, 1):
        p0, deltas, p1 = refio.tutte(n, e, ring, 40, 200.0, 1.0, bool(gs))
        save(f"tutte_grid6_{'gs' if gs else 'jacobi'}", n=n, edges=e,
             boundary=np.array(ring), p0=p0, deltas=deltas, p1=p1, iters=40,
             radius=200.0, dt=1.0)
    n, e, pts, hull = G.delaunay(300, 3)
    p0, deltas, p1 = refio.tutte(n, e, hull, 200, 200.0, 1.0, False)
    save("tutte_delaunay300_jacobi", n=n, edges=e, boundary=hull, p0=p0,
         deltas=deltas, p1=p1, iters=200, radius=200.0, dt=1.0)
    # Relaxation below 1 (dt = 1/60: alpha = 5/60) exercises the blend.
    p0, deltas, p1 = refio.tutte(n, e, hull, 30, 200.0, 1.0 / 60.0, False)
    save("tutte_delaunay300_dt60", n=n, edges=e, boundary=hull, p0=p0,
         deltas=deltas, p1=p1, iters=30, radius=200.0, dt=1.0 / 60.0)
    k4 = np.array([[0, 1], [0, 2], [0, 3], [1, 2], [1, 3], [2, 3]])
    p0, deltas, p1 = refio.tutte(4, k4, [0, 1, 2], 5, 200.0, 1.0, False)
    save("tutte_k4", n=4, edges=k4, boundary=np.array([0, 1, 2]), p0=p0,
         deltas=deltas, p1=p1, iters=5, radius=200.0, dt=1.0)


def knn_cases():
    n, e, _ = G.random_geometric(300, 8.0, 1)
    rng = np.random.default_rng(0)
    vis = (rng.random(n) < 0.1).astype(np.int64)
    srcs = np.arange(n)
    per, rc, batch = refio.knn(n, e, vis, 6, srcs)
    save("knn_rgg300", n=n, edges=e, visible=vis, k=6, sources=srcs,
         per_count=np.array([c for c, _ in per]),
         per_list=np.stack([lst for _, lst in per]), batch_rc=rc,
         batch_targets=np.array([b[0] for b in batch]),
         batch_count=np.array([b[1] for b in batch]),
         batch_list=np.stack([b[2] for b in batch]))


def grip_case(name, n, e, kcap, rounds, kmax=None, dim=2):
    kmax = kcap if kmax is None else kmax
    borders, filt, rr = refio.grip(n, e, kcap, kmax, rounds, dim)
    out = dict(n=n, edges=e, kcap=kcap, kmax=kmax, borders=borders,
               filtration=filt, rounds=len(rr), dim=dim)
    for i, r in enumerate(rr):
        out[f"r{i}_layer"] = r["layer"]
        out[f"r{i}_pos0"] = r["pos0"]
        out[f"r{i}_dec0"] = r["dec0"]
        out[f"r{i}_knn"] = r["knn"]
        out[f"r{i}_pos1"] = r["pos1"]
        out[f"r{i}_dec1"] = r["dec1"]
        out[f"r{i}_stats"] = np.array(r["stats"])
    save(name, **out)


def grip_cases():
    n, e, _ = G.random_geometric(300, 8.0, 2)
    grip_case("grip_rgg300", n, e, 10, 2)
    n, e = G.grid(10, 10)
    grip_case("grip_grid10", n, e, 8, 2)
    n, e = G.barabasi_albert(200, 2, 1)
    grip_case("grip_ba200", n, e, 10, 1)


def grip_chain_cases():
    """Several rounds per layer with K below the kNN capacity: the rounds
    chain through the device state and every NextStage is included."""
    n, e = G.grid(16, 16)
    grip_case("grip_grid16_k4", n, e, 16, 4, kmax=4)


def grip_nd_cases():
    """Refinement rounds, placement, decorator clears and kNN updates in
    three and four dimensions (the 3-D rotation solve and the six 4-D
    coordinate planes), and gvizGRIPEmbedderEmbed end to end."""
    n, e, _ = G.random_geometric(300, 8.0, 2)
    grip_case("grip_rgg300_d3", n, e, 10, 2, dim=3)
    grip_case("grip_rgg300_d4", n, e, 10, 2, dim=4)
    n, e = G.grid(16, 16)
    grip_case("grip_grid16_k4_d3", n, e, 16, 4, kmax=4, dim=3)
    grip_case("grip_grid16_k4_d4", n, e, 16, 4, kmax=4, dim=4)
    n, e, _ = G.random_geometric(400, 8.0, 5)
    for dim in (3, 4):
        r = refio.api_grip(n, e, 12, 12, 12, 0, dim=dim)
        save(f"api_grip_rgg400_constant_d{dim}", n=n, edges=e, kcap=12, placement_k=12,
             refinement_k=12, policy=0, dim=dim, **r)


def layout_quality(n, e, hidden, pos, edge_length):
    """Stress and edge energy of the layout of the subgraph induced by every
    vertex but `hidden` (compact ids in ascending parent id)."""
    import pyoracle as O
    keep = np.ones(n, bool)
    keep[hidden] = False
    ids = np.flatnonzero(keep)
    index = -np.ones(n, np.int64)
    index[ids] = np.arange(ids.size)
    ee = e[keep[e[:, 0]] & keep[e[:, 1]]]
    off, adj = G.csr_from_edges(ids.size, index[ee])
    g = O.Graph(off, adj)
    pool = O.Pool(4)
    return O.stress(g, pos[ids], edge_length, None, pool), O.edge_energy(g, pos[ids])


SPREAD_RUNS = 8
SPREAD_NOISE = 1e-12


def api_force_case(name, n, e, hidden, seed, run_iters, epsilon, **kw):
    """The reference's run, plus SPREAD_RUNS runs of the reference from the
    same start with every coordinate scaled by 1 +- 1e-12: the spread of
    their final stress / edge energy / run length is how far rounding-size
    differences carry in this chaotic FA2 run, the yardstick for any other
    summation order (the engine's included)."""
    r = refio.api_force(n, e, hidden, seed=seed, run_iters=run_iters,
                        epsilon=epsilon, **kw)
    params = {"param_" + k: v for k, v in kw.items()}
    st, en, rc = [], [], []
    for k in range(SPREAD_RUNS):
        q = refio.api_force(n, e, hidden, seed=seed, run_iters=run_iters,
                            epsilon=epsilon, perturb=SPREAD_NOISE, perturb_seed=k + 1, **kw)
        a, b = layout_quality(n, e, hidden, q["pos_step"], kw["edge_length"])
        st.append(a)
        en.append(b)
        rc.append(q["run_rc"])
    save(name, n=n, edges=e, hidden=hidden, seed=seed, run_iters=run_iters,
         epsilon=epsilon, **params, **r, spread_stress=np.array(st),
         spread_energy=np.array(en), spread_run_rc=np.array(rc, np.int64),
         spread_noise=SPREAD_NOISE)


def api_cases():
    """Embedder API runs end to end (Begin's own random placement included)."""
    n, e = G.barabasi_albert(400, 3, 21)
    api_force_case("api_force_linlog_ba400", n, e,
                   np.array([3, 17, 250, 399], dtype=np.int64), 777, 40, 1e-3,
                   model=1, barnes_hut=1, prevent_overlap=1, npc=1,
                   edge_length=10.0, theta=0.7, gravity=0.05,
                   overlap_constant=100.0, radius_base=0.1,
                   radius_per_degree=1.0, jitter=1.0)
    n, e = G.grid(20, 20)
    api_force_case("api_force_fr_grid20", n, e, np.zeros(0, np.int64), 4242, 300,
                   0.5, model=0, barnes_hut=1, prevent_overlap=0, npc=1,
                   edge_length=10.0, theta=1.0, gravity=0.0,
                   overlap_constant=100.0, radius_base=0.0,
                   radius_per_degree=0.0, jitter=1.0)


def api_grip_case(name, n, e, kcap, placement_k, refinement_k, policy):
    """gvizGRIPEmbedderEmbed, plus SPREAD_RUNS runs of the reference's Embed
    schedule from Begin's placement scaled by 1 +- 1e-12: the stress / edge
    energy spread of their final layouts measures how far rounding-size
    differences carry through the heat feedback (the yardstick for any other
    reduction order, the engine's included)."""
    r = refio.api_grip(n, e, kcap, placement_k, refinement_k, policy)
    st, en = [], []
    for k in range(SPREAD_RUNS):
        q = refio.api_grip_spread(n, e, kcap, placement_k, refinement_k, policy,
                                  SPREAD_NOISE, k + 1)
        a, b = layout_quality(n, e, np.zeros(0, np.int64), q, 10.0)
        st.append(a)
        en.append(b)
    save(name, n=n, edges=e, kcap=kcap, placement_k=placement_k,
         refinement_k=refinement_k, policy=policy, **r, spread_stress=np.array(st),
         spread_energy=np.array(en), spread_noise=SPREAD_NOISE)


def api_grip_cases():
    n, e, _ = G.random_geometric(400, 8.0, 5)
    api_grip_case("api_grip_rgg400_constant", n, e, 12, 12, 12, 0)
    n, e = G.grid(16, 16)
    api_grip_case("api_grip_grid16_decay", n, e, 16, 16, 8, 1)


def embedding_io_texts(saved):
    """Files for gvizEmbeddedGraphLoadEmbedding: the reference's own saved
    text, then edits that exercise the parser (scientific notation, tabs and
    blank lines, a long name line, a wrong vertex count, a wrong dimension, a
    truncated list)."""
    lines = saved.split(b"\n")
    body = lines[2:]
    sci = b"\n".join(lines[:2] + [b" ".join(f"{float(t):.6e}".encode() for t in ln.split())
                                   for ln in body if ln.strip()]) + b"\n"
    tabs = lines[0] + b"\n\n" + lines[1] + b"\n" + b"\t\n".join(
        ln.replace(b" ", b"\t") for ln in body)
    long_name = b"n" * 300 + b"\n" + b"\n".join(lines[1:])
    n, dim = (int(t) for t in lines[1].split())
    wrong_n = lines[0] + b"\n" + f"{n + 1} {dim}".encode() + b"\n" + b"\n".join(body)
    wrong_dim = lines[0] + b"\n" + f"{n} {dim + 1}".encode() + b"\n" + b"\n".join(body)
    truncated = b"\n".join(lines[:2 + n // 2]) + b"\n1.5 "
    return [("saved", saved), ("scientific", sci), ("tabs", tabs), ("long_name", long_name),
            ("wrong_count", wrong_n), ("wrong_dim", wrong_dim), ("truncated", truncated)]


def embedding_io_cases():
    """gvizEmbeddedGraphSaveEmbedding / LoadEmbedding of the reference
    (src/embedders/gvizEmbeddedGraph.c:428-485) on positions that stress the
    "%f" text: rounding at the sixth decimal, tiny and large magnitudes,
    negative zero."""
    import tempfile
    rng = np.random.default_rng(17)
    for dim in (2, 3):
        n = 40
        e = np.stack([np.arange(n - 1), np.arange(1, n)], 1)
        pos = rng.uniform(-1e3, 1e3, (n, dim))
        pos[0] = 0.0000005
        pos[1] = -0.0000015
        pos[2] = 1e-12
        pos[3] = -0.0
        pos[4] = 123456789.123456789
        pos[5] = 2.5e-7
        with tempfile.TemporaryDirectory() as td:
            save_path = os.path.join(td, "saved.txt")
            rc, saved, _ = refio.embedding_io(n, e, dim, pos, "layout", save_path, [])
            texts = embedding_io_texts(saved)
            paths = []
            for name, text in texts:
                p = os.path.join(td, name + ".txt")
                with open(p, "wb") as f:
                    f.write(text)
                paths.append(p)
            paths.append(os.path.join(td, "missing.txt"))
            rc2, saved2, loads = refio.embedding_io(n, e, dim, pos, "layout", save_path,
                                                    paths)
        assert rc == rc2 == 0 and saved == saved2
        arrays = dict(n=n, edges=e, dim=dim, pos=pos, save_rc=rc,
                      saved=np.frombuffer(saved, np.uint8),
                      names=np.array([t[0] for t in texts] + ["missing"]))
        for k, (name, text) in enumerate(texts):
            arrays[f"t{k}"] = np.frombuffer(text, np.uint8)
        for k, (lrc, lpos) in enumerate(loads):
            arrays[f"l{k}_rc"] = lrc
            arrays[f"l{k}_pos"] = lpos
        save(f"embedding_io_d{dim}", **arrays)


def edges_text_cases():
    """(name, text, directed, skip_header) for the .edges loader: the
    reference's own test files (tests/data/*.edges) and files that exercise
    each parsing rule."""
    rng = np.random.default_rng(5)
    big = "".join(f"{int(a)} {int(b)}\n" for a, b in rng.integers(0, 5000, size=(2000, 2)))
    return [
        ("sample", "% Small undirected test graph (1-indexed vertex IDs)\n%     4\n"
                   "%    /\n%   3\n%  / \\\n% 1---2\n1 2\n1 3\n2 3\n3 4\n", 0, 0),
        ("malformed", "% malformed edges file: second data line is non-numeric\n"
                      "1 2\nfoo bar\n2 3\n", 0, 0),
        ("directed", "% Small directed test graph (1-indexed vertex IDs)\n1 2\n2 3\n", 1, 0),
        ("gapped", "% Vertex IDs with gaps, to test id compaction\n10 20\n20 30\n", 0, 0),
        ("rules", "% header next\n5 7\n\n  % indented comment\n1 2 0.5\n\t3\t1\r\n"
                  "   4 2 weight\n3 3\n1 2\n+9 4\n1000000 6abc\n\r\n12 1000000", 0, 1),
        ("rules_directed", "1 2\n2 1\n2 2\n7 1\n1 2\n", 1, 0),
        ("zero_based", "0 1\n1 2\n2 0\n", 0, 0),
        ("header_not_numeric", "% c\nnodes edges\n1 2\n", 0, 1),
        ("negative", "1 2\n-3 4\n", 0, 0),
        ("comments_only", "% nothing\n\n%\n", 0, 0),
        ("single_number", "1 2\n5\n", 0, 0),
        ("random2000", big, 0, 0),
    ]


def edges_cases():
    import tempfile
    out = {}
    names = []
    with tempfile.TemporaryDirectory() as td:
        for i, (name, text, directed, skip) in enumerate(edges_text_cases()):
            path = os.path.join(td, name + ".edges")
            with open(path, "w", newline="") as f:
                f.write(text)
            rc, off, adj = refio.edges_file(path, directed, skip)
            names.append(name)
            out[f"c{i}_text"] = np.frombuffer(text.encode(), np.uint8)
            out[f"c{i}_opts"] = np.array([directed, skip], np.int64)
            out[f"c{i}_rc"] = np.array(rc, np.int64)
            out[f"c{i}_offsets"] = off if off is not None else np.zeros(0, np.int64)
            out[f"c{i}_adj"] = adj if adj is not None else np.zeros(0, np.int64)
    save("edges_files", names=np.array(names), **out)


def spring_tutte_cases():
    """Spring-driven Tutte (src/embedders/gvizSpringTutteEmbedder.c): steps
    and a Run to convergence, default and underdamped settings."""
    n, e = G.grid(6, 6)
    ring = [0, 1, 2, 3, 4, 5, 11, 17, 23, 29, 35, 34, 33, 32, 31, 30, 24, 18, 12, 6]
    cases = [("spring_tutte_grid6", n, e, ring, 12, 400, 1.0 / 60.0, 30.0, 6.0, 1e-5)]
    n2, e2, _, hull = G.delaunay(300, 3)
    cases.append(("spring_tutte_delaunay300", n2, e2, hull, 8, 3000, 1.0 / 60.0, 30.0, 6.0,
                  1e-6))
    cases.append(("spring_tutte_delaunay300_under", n2, e2, hull, 20, 120, 0.05, 80.0, 2.0,
                  1e-5))
    for name, nn, ee, bd, steps, run_to, dt, k, c, eps in cases:
        r = refio.spring_tutte(nn, ee, bd, steps, run_to, 200.0, dt, k, c, eps)
        save(name, n=nn, edges=ee, boundary=np.asarray(bd), steps=steps, run_to=run_to,
             radius=200.0, dt=dt, stiffness=k, damping=c, epsilon=eps,
             **{key: (v if v is not None else np.zeros(0)) for key, v in r.items()})


def tutte_begin_cases():
    """Begin through the planarity suite on 3-connected planar graphs (a
    Delaunay triangulation, a wheel-like grid), a 2-connected graph that is
    not 3-connected (a plain grid), a graph with a cut vertex (two grids
    sharing a corner) and a non-planar graph, then Run to convergence, for
    the Tutte and the spring-driven Tutte embedder."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(11)
    pts = rng.random((400, 2))
    e = set()
    for t in Delaunay(pts).simplices:
        for a in range(3):
            u, v = int(t[a]), int(t[(a + 1) % 3])
            e.add((min(u, v), max(u, v)))
    e = np.array(sorted(e), np.int64)
    n2, e2 = G.grid(10, 10)
    # close the grid into a 3-connected graph: the diagonals of every cell
    d = [(10 * r + c, 10 * (r + 1) + c + 1) for r in range(9) for c in range(9)]
    e2 = np.concatenate([np.asarray(e2, np.int64), np.array(d, np.int64)])
    k33 = np.array([(i, 3 + j) for i in range(3) for j in range(3)], np.int64)
    n3, e3 = G.grid(8, 8)
    e3 = np.asarray(e3, np.int64)
    n6, e6 = G.grid(6, 6)
    e6 = np.asarray(e6, np.int64)
    # the second grid's corner 0 is the first grid's corner n6 - 1
    e7 = np.concatenate([e6, e6 + (n6 - 1)])
    for name, n, ee in (("delaunay400", 400, e), ("trigrid10", n2, e2), ("k33", 6, k33),
                        ("grid8", n3, e3), ("twogrids6", 2 * n6 - 1, e7)):
        for spring in (0, 1):
            r = refio.tutte_begin(n, ee, spring, 20000, 1e-9)
            save(f"tutte_begin_{name}_{'spring' if spring else 'jacobi'}", n=n, edges=ee,
                 spring=spring, iters=20000, epsilon=1e-9,
                 **{k: np.asarray(v) for k, v in r.items()})


CASES = dict(quadtree=quadtree_cases, force=force_cases, tutte=tutte_cases,
             knn=knn_cases, grip=grip_cases, grip_chain=grip_chain_cases,
             grip_nd=grip_nd_cases, tutte_begin=tutte_begin_cases,
             api=api_cases, api_grip=api_grip_cases, edges=edges_cases,
             embedding_io=embedding_io_cases, spring_tutte=spring_tutte_cases)

if __name__ == "__main__":
    if not refio.harness_available():
        sys.exit("oracle/_ref/ref_harness is not built: make -C oracle ref")
    for name in (sys.argv[1:] or CASES):
        CASES[name]()
