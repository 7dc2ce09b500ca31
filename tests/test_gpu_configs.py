"""Every benchmark workload at its full size on the device, against the FP64
oracle on the same seeded inputs (the generators and starts bench.py uses).

  BA n = 1,000,000, m = 3, LinLog, gravity 0.05, prevent-overlap
      r = 0.1 (1 + sqrt deg), Barnes-Hut theta 0.7: FP64 forces within 1e-11
      normwise and 1e-9 per vertex of sum_j |f_ij| without overlap
      prevention, 1e-7 with it (near-contact terms, see the test), interaction
      sets equal, one whole step; MIXED within 1e-5 normwise and per vertex.
  The same configuration at n = 100,000 for 200 steps: MIXED final stress
      and edge-length energy within 1 % of the oracle's FP64 run.
  ER n = 65,536, mean degree 8, exact all-pairs FR: FP64 1e-11 / 1e-9;
      MIXED 1e-5 normwise and per vertex; FP32 (FP32 positions) 1e-5
      normwise and 1e-3 per vertex.
  Tutte, Delaunay n = 1,000,000, hull pinned on a radius-200 circle: 1,000
      FP64 Jacobi sweeps bit-exact against the oracle, plus the properties
      any Jacobi run has (boundary fixed, interior inside the polygon, max
      displacement non-increasing).
  GRIP, largest component of an RGG n = 4,000,000 (mean degree 8): the MIS
      filtration bit-exact, with the reference's invariants (nesting, a
      permutation of the subgraph, spacing >= 2^(i-1)+1 and maximality for
      every layer but the last, final layer dim + 1) checked for every
      vertex; 1,000 kNN lists per search path; two finest-layer refinement
      rounds within 1e-12.
  Batch, 4,096 graphs (trees, grids, ER; 50-20,000 vertices): 32 graphs
      across the size range bit-identical to their own single-graph runs in
      FP32, and the FP64 batch's first step against the oracle.

The per-vertex bound is SURVEY's FP32 parity definition: |dF_i| <= tol *
sum_j |f_ij|, where the sum runs over every term of vertex i's force
(oracle oc_force_term_sums).
"""
import os
import time

import numpy as np
import pytest

import pyoracle as O
from grapher_amd import graphs as G

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def log(msg):
    print(f"[configs] {msg}", flush=True)


@pytest.fixture(scope="module")
def pool():
    p = O.Pool(THREADS)
    yield p
    p.close()


def _force_errors(got, ref_disp, term_sum):
    d = got - ref_disp
    norm = float(np.linalg.norm(d) / np.linalg.norm(ref_disp))
    per = float((np.linalg.norm(d, axis=1) / term_sum).max())
    return norm, per


# ------------------------------------------------------------ BA 1M LinLog

LINLOG = dict(model=1, barnes_hut=True, prevent_overlap=True, nodes_per_cell=1,
              edge_length=10.0, theta=0.7, gravity=0.05, overlap_constant=100.0,
              radius_base=0.1, radius_per_degree=1.0)


def _oparams(cfg):
    return O.params(model=cfg["model"], barnes_hut=cfg["barnes_hut"],
                    prevent_overlap=cfg.get("prevent_overlap", False),
                    npc=cfg.get("nodes_per_cell", 1), edge_length=cfg["edge_length"],
                    theta=cfg.get("theta", 1.0), gravity=cfg.get("gravity", 0.0),
                    overlap_constant=cfg.get("overlap_constant", 100.0),
                    radius_base=cfg.get("radius_base", 0.0),
                    radius_per_degree=cfg.get("radius_per_degree", 0.0))


@pytest.fixture(scope="module")
def ba1m():
    t = time.time()
    n, e = G.barabasi_albert(1_000_000, 3, 0)
    off, adj = G.csr_from_edges(n, e)
    log(f"BA n={n} E={len(e)} built in {time.time() - t:.1f}s")
    return n, off, adj, G.uniform_positions(n, 0)


def test_ba1m_linlog_fp64_and_mixed_match_oracle(gpu, ba1m, pool):
    """FP64 and MIXED forces at 1M vertices, at the start and after 20 FP64
    steps. The FP64 bar with prevent-overlap on is 1e-7, not 1e-11: a
    repulsion term m_v m_u / (d - r_v - r_u) is ill-conditioned where the gap
    nears 0 (a pseudo-body's radius is 0 but the target's is not), and among
    ~1.5e8 interactions some gaps are ~1e-9. The device computes each cell's
    centre of mass as an exactly rounded quotient of double-double sums; the
    reference keeps a running FP64 mean, a few ulps away. Moving every cell's
    centre of mass of the oracle's own tree by one ulp changes its forces by
    2.6e-9 normwise here. With prevent-overlap off (every term well
    conditioned) the same 1M-vertex evaluation holds 1e-11 / 1e-9."""
    n, off, adj, pos0 = ba1m
    g = O.Graph(off, adj)
    p = _oparams(LINLOG)
    f64 = gpu.ForceEngine(off, adj, precision="fp64", **LINLOG)
    mixed = gpu.ForceEngine(off, adj, precision="mixed", **LINLOG)
    free = dict(LINLOG, prevent_overlap=False)
    f64_free = gpu.ForceEngine(off, adj, precision="fp64", **free)
    p_free = _oparams(free)
    # at the start, and after 20 FP64 steps (the layout has spread out)
    f64.begin(pos0)
    for label in ("start", "step 20"):
        pos = f64.positions()
        t = time.time()
        ref, ts, inter = O.force_term_sums(g, pos, p, pool, interactions=True)
        log(f"BA1M {label}: oracle evaluation {time.time() - t:.1f}s, "
            f"{inter.sum()} interactions")
        got = f64.evaluate(interactions=True)
        norm, per = _force_errors(got["disp"], ref, ts)
        log(f"BA1M {label} fp64: normwise {norm:.3g}, per vertex {per:.3g}")
        assert np.array_equal(got["interactions"], inter)
        assert norm <= 1e-7 and per <= 1e-7
        mixed.set_positions(pos)
        got = mixed.evaluate(interactions=True)
        norm, per = _force_errors(got["disp"], ref, ts)
        log(f"BA1M {label} mixed: normwise {norm:.3g}, per vertex {per:.3g}, "
            f"{int((got['interactions'] != inter).sum())} interaction counts differ")
        assert np.array_equal(got["interactions"], inter)
        assert norm <= 1e-5 and per <= 1e-5
        ref_f, ts_f, inter_f = O.force_term_sums(g, pos, p_free, pool, interactions=True)
        f64_free.set_positions(pos)
        got = f64_free.evaluate(interactions=True)
        norm, per = _force_errors(got["disp"], ref_f, ts_f)
        log(f"BA1M {label} fp64 without overlap prevention: normwise {norm:.3g}, "
            f"per vertex {per:.3g}")
        assert np.array_equal(got["interactions"], inter_f)
        assert norm <= 1e-11 and per <= 1e-9
        if label == "start":
            # one whole step (tree, forces, FA2 speed, position update)
            run = O.ForceRun(g, p, pos, pool)
            ost = run.step()
            st = f64.step()
            for k in ("maxDisp", "speed", "meanAtt", "meanRep", "meanGrav",
                      "speedEfficiency"):
                a, b = getattr(st, k), getattr(ost, k)
                log(f"BA1M step: {k} {a!r} vs {b!r}")
                assert abs(a - b) <= 1e-7 * max(abs(b), 1e-300), k
            got_pos = f64.positions()
            assert np.abs(got_pos - run.pos).max() <= 1e-7 * np.abs(run.pos).max()
            run.close()
            f64.run(19, stats=False)


def test_linlog_100k_200_steps_mixed_final_stress_and_energy(gpu, pool):
    """The headline configuration at 100,000 vertices: 200 MIXED steps on the
    device against 200 FP64 oracle steps from the same start. Trajectories
    part after a few dozen steps (the FA2 speed update amplifies round-off),
    so the end states are compared through layout stress (256 seeded BFS
    pivots) and edge-length energy, within 1 %."""
    n, e = G.barabasi_albert(100_000, 3, 0)
    off, adj = G.csr_from_edges(n, e)
    g = O.Graph(off, adj)
    pos0 = G.uniform_positions(n, 0)
    eng = gpu.ForceEngine(off, adj, precision="mixed", **LINLOG)
    eng.begin(pos0)
    eng.run(200, stats=False)
    got = eng.positions()
    t = time.time()
    run = O.ForceRun(g, _oparams(LINLOG), pos0, pool)
    for _ in range(200):
        run.step()
    log(f"LinLog 100k: 200 oracle steps in {time.time() - t:.1f}s")
    piv = np.random.default_rng(0).choice(n, 256, replace=False)
    s_ref = O.stress(g, run.pos, 10.0, piv, pool)
    s_got = O.stress(g, got, 10.0, piv, pool)
    e_ref, e_got = O.edge_energy(g, run.pos), O.edge_energy(g, got)
    log(f"LinLog 100k: stress {s_got:.6f} vs {s_ref:.6f}, energy {e_got:.6f} vs {e_ref:.6f}")
    assert abs(s_got - s_ref) <= 0.01 * s_ref
    assert abs(e_got - e_ref) <= 0.01 * e_ref


# ------------------------------------------------------ ER 65,536 all-pairs

def test_er65k_allpairs_fr_matches_oracle(gpu, pool):
    n, e = G.erdos_renyi(65536, 8.0, 0)
    off, adj = G.csr_from_edges(n, e)
    g = O.Graph(off, adj)
    pos = G.uniform_positions(n, 0)
    cfg = dict(model=0, barnes_hut=False, edge_length=10.0)
    t = time.time()
    ref, ts = O.force_term_sums(g, pos, _oparams(cfg), pool)
    log(f"ER65k: oracle all-pairs evaluation {time.time() - t:.1f}s")
    # FP32 positions cannot resolve the closest pairs: 65,536 uniform points
    # in [-1, 1]^2 have pairs ~2e-5 apart, where an FP32 coordinate's
    # rounding (6e-8) is ~3e-3 of the difference and that pair's L^2/d term
    # is a sizeable share of sum_j |f_ij|. MIXED differences FP64 positions.
    for prec, tol_n, tol_v in (("fp64", 1e-11, 1e-9), ("mixed", 1e-5, 1e-5),
                               ("fp32", 1e-5, 1e-3)):
        eng = gpu.ForceEngine(off, adj, precision=prec, **cfg)
        eng.set_positions(pos)
        got = eng.evaluate(interactions=True)
        norm, per = _force_errors(got["disp"], ref, ts)
        log(f"ER65k {prec}: normwise {norm:.3g}, per vertex {per:.3g}")
        assert np.all(got["interactions"] == n - 1)
        assert norm <= tol_n and per <= tol_v
        eng.close()


# ------------------------------------------------------ Tutte Delaunay 1M

def _tutte_setup(n, hull, radius=200.0):
    """gvizTutteFixConvexPolygon + SeedInterior (src/embedders/
    gvizTutteEmbedder.c:207-247 of the reference): hull on the circle through
    the oracle's polygon, interior at the polygon's centroid summed in order."""
    ring = O.tutte_polygon(len(hull), radius)
    cx = cy = 0.0
    for x, y in ring:
        cx += x
        cy += y
    pos = np.empty((n, 2))
    pos[:, 0] = cx / len(hull)
    pos[:, 1] = cy / len(hull)
    pos[hull] = ring
    isb = np.zeros(n, np.uint8)
    isb[hull] = 1
    return pos, isb, ring


def test_tutte_delaunay1m_1000_sweeps_bit_exact(gpu):
    t = time.time()
    n, e, _, hull = G.delaunay(1_000_000, 0)
    off, adj = G.csr_from_edges(n, e)
    g = O.Graph(off, adj)
    pos, isb, ring = _tutte_setup(n, hull)
    log(f"Delaunay n={n} E={len(e)} hull={len(hull)} built in {time.time() - t:.1f}s")
    eng = gpu.TutteEngine(off, adj, "fp64")
    eng.upload(pos, isb)
    ref = pos.copy()
    scratch = np.zeros_like(ref)
    sweeps = 1000
    t = time.time()
    ref_d = np.array([O.tutte_step(g, isb, ref, scratch) for _ in range(sweeps)])
    log(f"Tutte 1M: {sweeps} oracle sweeps in {time.time() - t:.1f}s")
    got_d, r = eng.sweeps(sweeps)
    got = eng.positions()
    assert r == sweeps
    assert np.array_equal(got_d, ref_d)
    assert np.array_equal(got, ref)
    # properties of any Jacobi run with alpha = 1
    assert np.array_equal(got[hull], ring)
    assert np.all(np.diff(got_d) <= 1e-12 * got_d[:-1])
    a = ring
    b = np.roll(ring, -1, axis=0)
    inner = got[isb == 0]
    cross = ((b[:, 0] - a[:, 0])[None, :] * (inner[:, 1:2] - a[:, 1][None, :])
             - (b[:, 1] - a[:, 1])[None, :] * (inner[:, 0:1] - a[:, 0][None, :]))
    assert cross.min() >= -1e-9  # every interior vertex inside the convex polygon


# ------------------------------------------------------------ GRIP RGG 4M

def _multi_source_bfs(off, adj, sources):
    """Breadth-first search from every source at once: depth to the nearest
    source and that source's id. The closest pair of sources is then
    min over edges (x, y) with different labels of depth x + depth y + 1."""
    n = len(off) - 1
    depth = np.full(n, -1, np.int64)
    label = np.full(n, -1, np.int64)
    frontier = np.asarray(sources, dtype=np.int64)
    depth[frontier] = 0
    label[frontier] = frontier
    d = 0
    while frontier.size:
        starts = off[frontier]
        cnt = off[frontier + 1] - starts
        tot = int(cnt.sum())
        if tot == 0:
            break
        par = np.repeat(frontier, cnt)
        first = np.repeat(starts - np.concatenate(([0], np.cumsum(cnt)[:-1])), cnt)
        nb = adj[first + np.arange(tot)]
        keep = depth[nb] < 0
        nb, par = nb[keep], par[keep]
        nb, idx = np.unique(nb, return_index=True)
        d += 1
        depth[nb] = d
        label[nb] = label[par[idx]]
        frontier = nb
    return depth, label


@pytest.fixture(scope="module")
def rgg4m():
    t = time.time()
    n, e, _ = G.random_geometric(4_000_000, 8.0, 3)
    act, e = G.largest_component(n, e)
    off, adj = G.csr_from_edges(n, e)
    log(f"RGG n={n}, component {int(act.sum())}, E={len(e)} built in {time.time() - t:.1f}s")
    return n, off, adj, act


def test_grip_rgg4m_filtration_invariants_knn_and_rounds(gpu, rgg4m):
    n, off, adj, act = rgg4m
    g = O.Graph(off, adj)
    eng = gpu.GripEngine(off, adj, active=act, kcap=10)
    filt, borders = eng.mis_filtration(2)
    t = time.time()
    of, ob = O.mis_filtration(g, act)
    log(f"GRIP 4M: oracle filtration {time.time() - t:.1f}s, layers {ob.tolist()}")
    assert np.array_equal(borders, ob)
    assert np.array_equal(filt, of)

    # the reference's invariants (tests/embedders/gvizGRIPTests.c:54-227)
    L = len(borders)
    assert borders[0] == act.sum() and borders[-1] == 3
    assert np.all(np.diff(borders) < 0)
    assert np.array_equal(np.sort(filt), np.flatnonzero(act))
    src = np.repeat(np.arange(n), np.diff(off))
    t = time.time()
    for i in range(1, L - 1):
        min_dist = (1 << (i - 1)) + 1
        layer = filt[: borders[i]]
        removed = filt[borders[i]: borders[i - 1]]
        depth, label = _multi_source_bfs(off, adj, layer)
        cut = label[src] != label[adj]
        closest = int((depth[src] + depth[adj] + 1)[cut].min())
        assert closest >= min_dist, (i, closest, min_dist)  # spacing
        assert depth[removed].max() < min_dist, i  # maximality
    log(f"GRIP 4M: spacing and maximality of layers 1..{L - 2} in {time.time() - t:.1f}s")

    rng = np.random.default_rng(5)
    # per-vertex searches as placement runs them: layer 3 visible, 1,000
    # vertices of layer 2 that it hides
    vis = np.zeros(n, np.uint8)
    vis[filt[: borders[3]]] = 1
    srcs = rng.choice(filt[borders[3]: borders[2]], 1000, replace=False)
    v, d, c = eng.knn(srcs, 10, vis)
    for i, s in enumerate(srcs):
        o = O.knn_bfs(g, int(s), 10, vis)
        assert c[i] == len(o) and np.array_equal(v[i, :c[i]], o[:, 0]) \
            and np.array_equal(d[i, :c[i]], o[:, 1]), int(s)
    # the batched search (one BFS per visible seed): the second-coarsest
    # layer's vertices as seeds, 300 targets
    sv = np.zeros(n, np.uint8)
    sv[filt[: borders[L - 2]]] = 1
    targets = rng.choice(filt[borders[L - 2]: borders[4]], 300, replace=False)
    ov, cnt = O.knn_batch(g, sv, 6, targets)
    v, d, c = eng.knn_batch(targets, 6, sv)
    assert np.array_equal(c, cnt)
    for i in range(len(targets)):
        assert np.array_equal(v[i, :c[i]], ov[i, :c[i], 0])
        assert np.array_equal(d[i, :c[i]], ov[i, :c[i], 1])
    log("GRIP 4M: 1,000 per-vertex and 300 batched kNN lists equal")

    # the finest layer: every subgraph vertex visible, kNN lists on the device
    na = int(borders[0])
    eng.clear_decorators(na)
    pos = rng.uniform(-2000.0, 2000.0, (n, 2))
    eng.upload(positions=pos)
    eng.update_knn(na, act, 10)
    pos_d, dec, kv, kd = eng.download()
    assert np.array_equal(pos_d, pos)
    for s in rng.choice(filt[:na], 1000, replace=False):
        o = O.knn_bfs(g, int(s), 10, act)
        c = int(dec["knnCount"][s])
        assert c == len(o) and np.array_equal(kv[s, :c], o[:, 0]) \
            and np.array_equal(kd[s, :c], o[:, 1]), int(s)
    knn = np.stack([kv, kd], axis=2)
    odec = np.zeros(n, dtype=O.GRIP_VERTEX_DTYPE)
    for k in ("dispX", "dispY", "oldDispX", "oldDispY", "heat", "oldCos",
              "lastForceMag", "knnCount"):
        odec[k] = dec[k]
    odec["dispCalculated"] = dec["dispCalculated"].astype(np.uint8)
    ref_pos = np.ascontiguousarray(pos.copy())
    ids = filt[:na]
    for rnd in range(2):
        t = time.time()
        ost = O.grip_refine_round(g, filt, na, 0, knn, 10, odec, ref_pos)
        log(f"GRIP 4M: oracle round {rnd} in {time.time() - t:.1f}s")
        st = eng.refine_round(na, 0)
        got_pos, got_dec, _, _ = eng.download()
        scale = np.abs(ref_pos).max()
        assert np.abs(got_pos - ref_pos).max() <= 1e-12 * scale, rnd
        for k in ("heat", "oldCos", "lastForceMag"):
            np.testing.assert_allclose(got_dec[k][ids], odec[k][ids], rtol=1e-12,
                                       atol=1e-300, err_msg=k)
        np.testing.assert_allclose([st.maxDisp, st.meanDisp, st.meanForce],
                                   [ost.maxDisp, ost.meanDisp, ost.meanForce], rtol=1e-11)


# ----------------------------------------------------------- batch 4,096

def test_batch4096_sampled_graphs_equal_single_runs_and_oracle(gpu, pool):
    t = time.time()
    graphs, _ = G.batch_mix()
    sizes = np.array([len(o) - 1 for o, _ in graphs])
    starts = [G.uniform_positions(int(m), 1000 + i) for i, m in enumerate(sizes)]
    start = np.concatenate(starts)
    log(f"batch: {len(graphs)} graphs, {sizes.sum()} vertices built in {time.time() - t:.1f}s")
    cfg = dict(model=0, theta=1.0, edge_length=10.0, gravity=0.0)
    pick = np.argsort(sizes, kind="stable")[np.linspace(0, len(sizes) - 1, 32).astype(int)]
    steps = 10
    b = gpu.BatchEngine(graphs, precision="fp32", **cfg)
    b.begin(start)
    launched, its, _, _ = b.run(steps)
    assert launched == steps and np.all(its == steps)
    got = b.positions()
    b.close()
    for gi in pick.tolist():
        o, a = graphs[gi]
        e = gpu.ForceEngine(o, a, precision="fp32", **cfg)
        e.begin(starts[gi])
        e.run(steps, stats=False)
        assert np.array_equal(got[b.start[gi]: b.start[gi + 1]], e.positions()), gi
        e.close()
    log(f"batch: 32 graphs ({sizes[pick].min()}-{sizes[pick].max()} vertices) "
        f"bit-identical to their single runs over {steps} FP32 steps")
    b = gpu.BatchEngine(graphs, precision="fp64", **cfg)
    b.begin(start)
    b.run(1)
    got = b.positions()
    p = O.params(model=0, theta=1.0, edge_length=10.0)
    worst = 0.0
    for gi in pick.tolist():
        o, a = graphs[gi]
        run = O.ForceRun(O.Graph(o, a), p, starts[gi], pool)
        run.step()
        mine = got[b.start[gi]: b.start[gi + 1]]
        err = np.abs(mine - run.pos).max() / np.abs(run.pos).max()
        worst = max(worst, err)
        assert err <= 1e-9, gi
        run.close()
    log(f"batch: FP64 first step of 32 graphs within {worst:.3g} of the oracle")
