"""HIP force engine against the reference's fixtures and the FP64 oracle.

Tolerances, per precision:
  fp64  forces normwise <= 1e-11 relative, per vertex <= 1e-9 of |att|+|rep|;
        repulsion interaction sets equal (same per-vertex counts)
  mixed/fp32 forces normwise <= 1e-5 relative (the 1e-5 FP32 bar)
Quadtree membership (leaf depth and quadrant path) is bit-exact everywhere.
"""
import numpy as np
import pytest

import pyoracle as O
from conftest import golden
from grapher_amd import graphs as G

pytestmark = pytest.mark.gpu

FORCE_FIXTURES = ["force_fr_bh_grid12", "force_linlog_bh_ba300",
                  "force_fr_exact_er200", "force_fr_bh_npc2_grid15"]


def _engine(native, f, precision):
    off, adj = G.csr_from_edges(int(f["n"]), f["edges"])
    return native.ForceEngine(
        off, adj, model=int(f["param_model"]),
        barnes_hut=bool(f["param_barnes_hut"]),
        prevent_overlap=bool(f["param_prevent_overlap"]), precision=precision,
        nodes_per_cell=int(f["param_npc"]), edge_length=float(f["param_edge_length"]),
        theta=float(f["param_theta"]), gravity=float(f["param_gravity"]),
        overlap_constant=float(f["param_overlap_constant"]),
        radius_base=float(f["param_radius_base"]),
        radius_per_degree=float(f["param_radius_per_degree"]),
        jitter=float(f["param_jitter"]))


def _check_forces(got, ref_disp, ref_att, ref_rep, tol_norm, tol_vertex):
    d = got["disp"] - ref_disp
    assert np.linalg.norm(d) <= tol_norm * np.linalg.norm(ref_disp)
    scale = ref_att + ref_rep + 1e-300
    per = np.linalg.norm(d, axis=1) / scale
    assert per.max() <= tol_vertex, (per.max(), int(per.argmax()))


@pytest.mark.parametrize("name", FORCE_FIXTURES)
def test_fp64_forces_match_reference_each_iteration(gpu, name):
    f = golden(name)
    eng = _engine(gpu, f, "fp64")
    off, adj = G.csr_from_edges(int(f["n"]), f["edges"])
    g = O.Graph(off, adj)
    p = O.params(model=int(f["param_model"]), barnes_hut=bool(f["param_barnes_hut"]),
                 prevent_overlap=bool(f["param_prevent_overlap"]),
                 npc=int(f["param_npc"]), theta=float(f["param_theta"]),
                 gravity=float(f["param_gravity"]),
                 radius_base=float(f["param_radius_base"]),
                 radius_per_degree=float(f["param_radius_per_degree"]))
    pos = f["pos0"]
    for it in f["dump_iters"]:
        eng.set_positions(pos)
        got = eng.evaluate(interactions=True)
        _check_forces(got, f[f"it{it}_disp"], f[f"it{it}_attMag"],
                      f[f"it{it}_repMag"], 1e-11, 1e-9)
        np.testing.assert_allclose(got["attMag"], f[f"it{it}_attMag"], rtol=1e-12,
                                   atol=1e-12)
        # identical Barnes-Hut interaction sets
        orc = O.force_compute(g, pos, p)
        assert np.array_equal(got["interactions"], orc["interactions"])
        pos = f[f"it{it}_pos_after"]


@pytest.mark.parametrize("name", FORCE_FIXTURES)
def test_fp64_trajectory_matches_reference(gpu, name):
    f = golden(name)
    eng = _engine(gpu, f, "fp64")
    eng.begin(f["pos0"])
    stats, r = eng.run(int(f["iters"]))
    assert r == int(f["iters"])
    ref = f["stats"]
    for k, col in (("maxDisp", 0), ("speed", 1), ("meanAtt", 2), ("meanRep", 3),
                   ("meanGrav", 4), ("speedEfficiency", 5)):
        np.testing.assert_allclose(stats[k], ref[:, col], rtol=1e-9, atol=1e-12,
                                   err_msg=k)
    last = int(f["dump_iters"][-1])
    np.testing.assert_allclose(eng.positions(), f[f"it{last}_pos_after"],
                               rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
@pytest.mark.parametrize("name", FORCE_FIXTURES)
def test_fp32_forces_within_1e5(gpu, name, precision):
    """FP32 force math within 1e-5 normwise and per vertex (against the sum
    of the magnitudes of the vertex's force terms). With prevent-overlap on, the
    repulsion magnitude m/gap is singular where two circles touch; the mixed
    mode recomputes near-contact gaps from FP64 positions and radii and keeps
    1e-5, while FP32 positions cannot resolve such gaps (bound 1e-3 there)."""
    f = golden(name)
    eng = _engine(gpu, f, precision)
    tol = 1e-5
    if precision == "fp32" and bool(f["param_prevent_overlap"]):
        tol = 1e-3
    off, adj = G.csr_from_edges(int(f["n"]), f["edges"])
    g = O.Graph(off, adj)
    p = O.params(model=int(f["param_model"]), barnes_hut=bool(f["param_barnes_hut"]),
                 prevent_overlap=bool(f["param_prevent_overlap"]),
                 npc=int(f["param_npc"]), theta=float(f["param_theta"]),
                 gravity=float(f["param_gravity"]),
                 radius_base=float(f["param_radius_base"]),
                 radius_per_degree=float(f["param_radius_per_degree"]))
    pos = f["pos0"]
    for it in f["dump_iters"]:
        eng.set_positions(pos)
        got = eng.evaluate()
        d = got["disp"] - f[f"it{it}_disp"]
        assert np.linalg.norm(d) <= tol * np.linalg.norm(f[f"it{it}_disp"])
        # per vertex: |dF_i| <= tol * sum_j |f_ij| over the terms of i's force
        ref, ts = O.force_term_sums(g, pos, p)
        assert np.array_equal(ref, f[f"it{it}_disp"])
        per = np.linalg.norm(d, axis=1) / ts
        assert per.max() <= tol, (it, per.max(), int(per.argmax()))
        pos = f[f"it{it}_pos_after"]


@pytest.mark.parametrize("name", ["quadtree_uniform300", "quadtree_dup_npc1",
                                  "quadtree_clustered_npc3"])
def test_quadtree_membership_bit_exact(gpu, name):
    f = golden(name)
    pts = f["points"]
    n = pts.shape[0]
    eng = gpu.ForceEngine(np.zeros(n + 1, np.int64), np.zeros(0, np.int64),
                          precision="fp64", nodes_per_cell=int(f["npc"]))
    eng.set_positions(pts)
    depth, kh, kl, root = eng.quadtree_leaves()
    assert np.array_equal(depth, f["leaf_depth"])
    assert np.array_equal(kh, f["leaf_keyHi"])
    assert np.array_equal(kl, f["leaf_keyLo"])
    if np.all(f["masses"] == 1.0):
        # branching cells: mass exact, centre of mass to FP64 round-off
        cells = eng.quadtree_cells()
        ref = f["nodes"]
        branching = {}
        for r in ref:
            if r["leaf"]:
                continue
            branching[(int(r["depth"]), int(r["keyHi"]), int(r["keyLo"]))] = r
        got = {(int(d), int(h), int(l)): (m, x, y) for d, h, l, m, x, y in zip(
            cells["depth"], cells["keyHi"], cells["keyLo"], cells["mass"],
            cells["comX"], cells["comY"])}
        for key, (m, x, y) in got.items():
            r = branching[key]
            assert m == r["mass"]
            assert abs(x - r["comX"]) <= 1e-12 * (1 + abs(r["comX"]))
            assert abs(y - r["comY"]) <= 1e-12 * (1 + abs(r["comY"]))


def test_grid100_500_iterations_stress_within_1pct(gpu):
    """The CPU-runnable headline case: 100x100 grid, FR, BH theta 0.5, 500
    iterations from the same seeded positions; final stress and edge energy
    within 1% of the reference's final layout, in every precision."""
    f = golden("force_fr_bh_grid100")
    off, adj = G.csr_from_edges(int(f["n"]), f["edges"])
    g = O.Graph(off, adj)
    pool = O.Pool(0)
    ref_pos = f["it499_pos_after"]
    rng = np.random.default_rng(0)
    src = rng.choice(int(f["n"]), 256, replace=False)
    s_ref = O.stress(g, ref_pos, 10.0, src, pool)
    e_ref = O.edge_energy(g, ref_pos)
    for prec in ("fp64", "mixed", "fp32"):
        eng = _engine(gpu, f, prec)
        eng.begin(f["pos0"])
        _, r = eng.run(500)
        assert r == 500
        pos = eng.positions()
        assert np.all(np.isfinite(pos))
        s = O.stress(g, pos, 10.0, src, pool)
        e = O.edge_energy(g, pos)
        assert abs(s - s_ref) <= 0.01 * s_ref, (prec, s, s_ref)
        assert abs(e - e_ref) <= 0.01 * e_ref + 1e-3, (prec, e, e_ref)


def test_run_stops_at_epsilon_and_step_matches_run(gpu):
    f = golden("force_fr_bh_grid12")
    a = _engine(gpu, f, "fp64")
    b = _engine(gpu, f, "fp64")
    a.begin(f["pos0"])
    b.begin(f["pos0"])
    stats, r = a.run(8)
    steps = [b.step().maxDisp for _ in range(8)]
    np.testing.assert_array_equal(stats["maxDisp"], np.array(steps))
    np.testing.assert_array_equal(a.positions(), b.positions())
    # epsilon above the third step's maxDisp stops right there
    c = _engine(gpu, f, "fp64")
    c.begin(f["pos0"])
    eps = stats["maxDisp"][2] * 1.0000001
    if stats["maxDisp"][0] >= eps and stats["maxDisp"][1] >= eps:
        _, r2 = c.run(8, eps)
        assert r2 == 3


def test_large_ba_linlog_mixed_matches_fp64_engine(gpu):
    """1M-vertex Barabasi-Albert LinLog + gravity + overlap, theta 0.7: the
    mixed-precision forces against the FP64 engine on the same positions
    (normwise 1e-5), and every quadtree leaf holds its own vertex (uniform
    positions have no duplicates, so no leaf at the depth limit holds two)."""
    n, e = G.barabasi_albert(1_000_000, 3, 0)
    off, adj = G.csr_from_edges(n, e)
    pos = G.uniform_positions(n, 0)
    kw = dict(model=1, theta=0.7, gravity=0.05, prevent_overlap=True,
              radius_base=0.1, radius_per_degree=1.0)
    hi = gpu.ForceEngine(off, adj, precision="fp64", **kw)
    lo = gpu.ForceEngine(off, adj, precision="mixed", **kw)
    hi.set_positions(pos)
    lo.set_positions(pos)
    a = hi.evaluate()["disp"]
    b = lo.evaluate()["disp"]
    assert np.all(np.isfinite(b))
    assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(a)
    depth, kh, kl, _ = lo.quadtree_leaves()
    leaves = np.unique(np.stack([depth.astype(np.uint64), kh, kl], 1), axis=0)
    assert len(leaves) == n  # one vertex per leaf: no two share a leaf


@pytest.mark.parametrize("model,overlap,precision", [
    (m, o, p) for m in (0, 1) for o, p in ((False, "mixed"), (False, "fp32"), (True, "mixed"))])
def test_exact_all_pairs_against_oracle(gpu, model, overlap, precision):
    """Exact all-pairs repulsion (Barnes-Hut off) with FP32 force math on 1,500
    vertices (six 256-target tiles, sources split over block rows), 11
    vertices stacked on one point and one more coincident pair: the
    displacement within 1e-5 normwise of the FP64 oracle, for both models,
    with and without overlap prevention. (FP32 positions with overlap
    prevention cannot resolve near-contact gaps among 1,500 overlapping
    circles; that mode keeps its 1e-3 bound on the reference's fixtures.)"""
    n, e = G.erdos_renyi(1500, 6.0, 3)
    off, adj = G.csr_from_edges(n, e)
    pos = G.uniform_positions(n, 5)
    pos[10:20] = pos[0]
    pos[300] = pos[1000]
    kw = dict(model=model, barnes_hut=False, prevent_overlap=overlap,
              gravity=0.05 if model == 1 else 0.0,
              radius_base=0.1 if overlap else 0.0,
              radius_per_degree=1.0 if overlap else 0.0)
    eng = gpu.ForceEngine(off, adj, precision=precision, **kw)
    eng.set_positions(pos)
    got = eng.evaluate()["disp"]
    ref = O.force_compute(O.Graph(off, adj), pos, O.params(**kw))["disp"]
    assert np.all(np.isfinite(got))
    tol = 1e-5
    assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref)
    # the stacked vertices are pushed apart along their hashed directions
    for v in (0, 10, 19, 300, 1000):
        assert np.linalg.norm(got[v] - ref[v]) <= tol * np.linalg.norm(ref[v]), v


@pytest.mark.parametrize("precision", ["fp64", "mixed", "fp32"])
def test_checkpoint_resume_is_bit_identical(gpu, precision):
    """gvizhip_force_save_state / load_state: a run resumed in a second engine
    from the state after 20 steps matches the uninterrupted run bit for bit
    over the next 30 steps (positions, speed, step count); a state of another
    vertex count is refused."""
    n, e = G.barabasi_albert(3000, 3, 4)
    off, adj = G.csr_from_edges(n, e)
    kw = dict(model=1, theta=0.7, gravity=0.05, prevent_overlap=True, radius_base=0.01,
              radius_per_degree=0.5, precision=precision)
    a = gpu.ForceEngine(off, adj, **kw)
    a.begin(G.uniform_positions(n, 1))
    a.run(20)
    state = a.save_state()
    stats_a, _ = a.run(30)
    b = gpu.ForceEngine(off, adj, **kw)
    b.begin(G.uniform_positions(n, 2))
    b.load_state(state)
    stats_b, _ = b.run(30)
    np.testing.assert_array_equal(a.positions(), b.positions())
    np.testing.assert_array_equal(stats_a["maxDisp"], stats_b["maxDisp"])
    assert a.speed() == b.speed()
    n2, e2 = G.barabasi_albert(2000, 3, 4)
    o2, a2 = G.csr_from_edges(n2, e2)
    c = gpu.ForceEngine(o2, a2, **kw)
    c.begin(G.uniform_positions(n2, 1))
    with pytest.raises(RuntimeError):
        c.load_state(state)


@pytest.mark.parametrize("npc", [1, 3])
def test_clustered_points_sort_and_tree_match_oracle(gpu, npc):
    """Points that share their first 16 quadrant levels reach the sort's run
    repair: 3,000 inside one 1e-7 square (heap-sort path), 20 inside a 1e-6
    square (insertion path), 50 exact duplicates and 10 near duplicates,
    among 5,000 uniform points. Every point's leaf (depth and quadrant path)
    equals the oracle's quadtree, and the FP64 repulsion interaction counts
    equal the oracle's: leaves that hold several points at the depth limit
    are accepted as one body wherever the reference accepts the chain of
    single-child cells above them."""
    rng = np.random.default_rng(31)
    parts = [rng.uniform(-1, 1, (5000, 2)),
             np.array([0.3, 0.3]) + rng.uniform(0, 1e-7, (3000, 2)),
             np.array([-0.61, 0.47]) + rng.uniform(0, 1e-6, (20, 2)),
             np.tile([[0.125, -0.5]], (50, 1)),
             np.array([0.7, 0.7]) + rng.uniform(0, 1e-13, (10, 2))]
    for with_tight in (True, False):
        pts = np.concatenate(parts if with_tight else parts[:1] + parts[2:])
        pts = pts[rng.permutation(len(pts))]
        n = len(pts)
        o = O.quadtree(pts, np.ones(n), npc)
        eng = gpu.ForceEngine(np.zeros(n + 1, np.int64), np.zeros(0, np.int64),
                              precision="fp64", nodes_per_cell=npc)
        eng.set_positions(pts)
        depth, kh, kl, _ = eng.quadtree_leaves()
        np.testing.assert_array_equal(depth, o["leaf_depth"])
        np.testing.assert_array_equal(kh, o["leaf_keyHi"])
        np.testing.assert_array_equal(kl, o["leaf_keyLo"])
        if True:
            got = eng.evaluate(interactions=True)
            g = O.Graph(np.zeros(n + 1, np.int64), np.zeros(0, np.int64))
            ref = O.force_compute(g, pts, O.params(npc=npc))
            np.testing.assert_array_equal(got["interactions"], ref["interactions"])


@pytest.mark.parametrize("npc", [1, 4])
def test_runs_across_workgroups_sort_and_tree_match_oracle(gpu, npc):
    """Runs of points that share their first 16 quadrant levels are put in
    full key order by rank, each run by the workgroup of 256 sorted positions
    that holds its first entry, runs longer than 64 by one thread. 400
    clusters of 2-70 points (each inside a 1e-6 square, a few exact
    duplicates among them) among 3,000 uniform points put run boundaries
    across most workgroup boundaries and take both paths: every point's leaf
    and the FP64 interaction counts equal the oracle's."""
    rng = np.random.default_rng(47)
    parts = [rng.uniform(-1, 1, (3000, 2))]
    sizes = rng.integers(2, 71, 400)
    for k, size in enumerate(sizes):
        c = rng.uniform(-0.9, 0.9, 2)
        if k % 10 == 0:
            parts.append(np.tile(c, (int(size), 1)))
        else:
            parts.append(c + rng.uniform(0, 1e-6, (int(size), 2)))
    pts = np.concatenate(parts)
    pts = pts[rng.permutation(len(pts))]
    n = len(pts)
    assert (sizes > 64).sum() >= 10 and n > 40 * 256
    o = O.quadtree(pts, np.ones(n), npc)
    eng = gpu.ForceEngine(np.zeros(n + 1, np.int64), np.zeros(0, np.int64),
                          precision="fp64", nodes_per_cell=npc)
    eng.set_positions(pts)
    depth, kh, kl, _ = eng.quadtree_leaves()
    np.testing.assert_array_equal(depth, o["leaf_depth"])
    np.testing.assert_array_equal(kh, o["leaf_keyHi"])
    np.testing.assert_array_equal(kl, o["leaf_keyLo"])
    got = eng.evaluate(interactions=True)
    g = O.Graph(np.zeros(n + 1, np.int64), np.zeros(0, np.int64))
    ref = O.force_compute(g, pts, O.params(npc=npc))
    np.testing.assert_array_equal(got["interactions"], ref["interactions"])


@pytest.mark.parametrize("barnes_hut", [True, False])
@pytest.mark.parametrize("model", [0, 1])
def test_tiny_and_degenerate_graphs_match_oracle(gpu, model, barnes_hut):
    """FP64 forces of graphs the tree and the traversal treat specially
    against the oracle: one vertex (gravity only), two vertices, two
    coincident vertices (the coincident direction rule), three vertices of
    which one is isolated, and a graph with isolated vertices at one point;
    with and without overlap prevention and gravity."""
    cases = [
        (1, np.zeros((0, 2), np.int64), np.array([[0.3, -0.2]])),
        (2, np.array([[0, 1]]), np.array([[0.0, 0.0], [1.5, -0.5]])),
        (2, np.array([[0, 1]]), np.array([[0.25, 0.25], [0.25, 0.25]])),
        (3, np.array([[0, 1]]), np.array([[0.0, 1.0], [1.0, 0.0], [-2.0, 0.5]])),
        (6, np.array([[0, 1], [1, 2]]),
         np.array([[0.0, 0.0], [1.0, 1.0], [2.0, 0.0], [5.0, 5.0], [5.0, 5.0], [5.0, 5.0]])),
    ]
    for n, e, pos in cases:
        off, adj = G.csr_from_edges(n, e)
        for overlap in (False, True):
            kw = dict(model=model, barnes_hut=barnes_hut, prevent_overlap=overlap,
                      gravity=0.05, radius_base=0.1 if overlap else 0.0,
                      radius_per_degree=0.5 if overlap else 0.0)
            eng = gpu.ForceEngine(off, adj, precision="fp64", **kw)
            eng.set_positions(pos)
            got = eng.evaluate()["disp"]
            ref = O.force_compute(O.Graph(off, adj), pos, O.params(**kw))["disp"]
            assert np.all(np.isfinite(got)), (n, overlap)
            assert np.linalg.norm(got - ref) <= 1e-11 * max(np.linalg.norm(ref), 1e-300), \
                (n, overlap, got, ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed", "fp32"])
@pytest.mark.parametrize("overlap", [True, False])
def test_counting_interactions_leaves_forces_unchanged(gpu, precision, overlap):
    """The traversal counts interactions per target only when asked (a
    separate pass over each flush's masks and the undecided cells' ballots),
    and a timed launch runs without counting: the forces of an evaluation
    with counts equal bit for bit those of one without, and a second
    evaluation repeats forces and counts."""
    n, e = G.barabasi_albert(20000, 3, 5)
    off, adj = G.csr_from_edges(n, e)
    pos = G.uniform_positions(n, 6) * 50.0
    eng = gpu.ForceEngine(off, adj, precision=precision, model=1, theta=0.7,
                          gravity=0.05, prevent_overlap=overlap, radius_base=0.1,
                          radius_per_degree=1.0)
    eng.set_positions(pos)
    plain = eng.evaluate()
    counted = eng.evaluate(interactions=True)
    again = eng.evaluate(interactions=True)
    for k in ("disp", "structForce", "repMag"):
        assert np.array_equal(plain[k], counted[k]), k
        assert np.array_equal(counted[k], again[k]), k
    assert np.array_equal(counted["interactions"], again["interactions"])
    assert counted["interactions"].min() > 0
