"""The FP64 oracle against the fixtures the reference library wrote
(tests/golden/make_golden.py). Every comparison is bit-exact: the oracle
restates the reference's arithmetic in the reference's order."""
import numpy as np
import pytest

import pyoracle as O
from conftest import golden
from grapher_amd import graphs as G


def _graph(f):
    off, adj = G.csr_from_edges(int(f["n"]), f["edges"])
    return O.Graph(off, adj)


@pytest.mark.parametrize("name", ["quadtree_uniform300", "quadtree_clustered_npc3",
                                  "quadtree_dup_npc1"])
def test_quadtree_membership_and_cells(name):
    f = golden(name)
    o = O.quadtree(f["points"], f["masses"], int(f["npc"]))
    assert np.array_equal(o["leaf_depth"], f["leaf_depth"])
    assert np.array_equal(o["leaf_keyHi"], f["leaf_keyHi"])
    assert np.array_equal(o["leaf_keyLo"], f["leaf_keyLo"])
    # node set in DFS order: depth, key, mass, COM and half size all equal
    nodes = o["nodes"]
    ref = f["nodes"]
    assert len(nodes) == len(ref)
    # the oracle stores nodes in allocation order; index them by (depth, key)
    mine = {(int(d), int(h), int(l)): (m, cx, cy, hs) for d, h, l, m, cx, cy, hs in
            zip(nodes["depth"], nodes["keyHi"], nodes["keyLo"], nodes["mass"],
                nodes["comX"], nodes["comY"], nodes["halfSize"])}
    for r in ref:
        key = (int(r["depth"]), int(r["keyHi"]), int(r["keyLo"]))
        m, cx, cy, hs = mine[key]
        assert m == r["mass"] and cx == r["comX"] and cy == r["comY"]
        assert hs == r["halfSize"]


def _params_from(f):
    return O.params(model=int(f["param_model"]), barnes_hut=bool(f["param_barnes_hut"]),
                    prevent_overlap=bool(f["param_prevent_overlap"]),
                    npc=int(f["param_npc"]), edge_length=float(f["param_edge_length"]),
                    theta=float(f["param_theta"]), gravity=float(f["param_gravity"]),
                    overlap_constant=float(f["param_overlap_constant"]),
                    radius_base=float(f["param_radius_base"]),
                    radius_per_degree=float(f["param_radius_per_degree"]),
                    jitter=float(f["param_jitter"]))


@pytest.mark.parametrize("name", ["force_fr_bh_grid12", "force_linlog_bh_ba300",
                                  "force_fr_exact_er200", "force_fr_bh_npc2_grid15"])
def test_force_step_bit_exact(name):
    f = golden(name)
    g = _graph(f)
    p = _params_from(f)
    run = O.ForceRun(g, p, f["pos0"])
    pool = O.Pool(4)
    for it in range(int(f["iters"])):
        before = run.pos.copy()
        forces = O.force_compute(g, before, p, pool)
        st = run.step()
        ref = f["stats"][it]
        assert (st.maxDisp, st.speed, st.meanAtt, st.meanRep, st.meanGrav,
                st.speedEfficiency) == tuple(ref)
        if it in f["dump_iters"]:
            for k in ("disp", "structForce", "attMag", "repMag"):
                assert np.array_equal(forces[k], f[f"it{it}_{k}"]), k
            assert np.array_equal(run.pos, f[f"it{it}_pos_after"])


def test_force_grid100_first_iteration_and_final_positions():
    f = golden("force_fr_bh_grid100")
    g = _graph(f)
    p = _params_from(f)
    forces = O.force_compute(g, f["pos0"], p, O.Pool(0))
    for k in ("disp", "structForce", "attMag", "repMag"):
        assert np.array_equal(forces[k], f[f"it0_{k}"]), k


@pytest.mark.parametrize("name,gs", [("tutte_grid6_jacobi", 0), ("tutte_grid6_gs", 1),
                                     ("tutte_delaunay300_jacobi", 0),
                                     ("tutte_delaunay300_dt60", 0), ("tutte_k4", 0)])
def test_tutte_bit_exact(name, gs):
    f = golden(name)
    g = _graph(f)
    n = int(f["n"])
    isb = np.zeros(n, np.uint8)
    isb[f["boundary"]] = 1
    # FixConvexPolygon positions and the interior seed are what p0 holds
    poly = O.tutte_polygon(len(f["boundary"]), float(f["radius"]))
    assert np.array_equal(poly, f["p0"][f["boundary"]])
    pos = f["p0"].copy()
    scratch = np.zeros_like(pos)
    deltas = [O.tutte_step(g, isb, pos, scratch, 5.0, float(f["dt"]), bool(gs))
              for _ in range(int(f["iters"]))]
    assert np.array_equal(np.array(deltas), f["deltas"])
    assert np.array_equal(pos, f["p1"])


def test_knn_both_paths_exact():
    f = golden("knn_rgg300")
    g = _graph(f)
    vis = f["visible"].astype(np.uint8)
    k = int(f["k"])
    for s in f["sources"]:
        o = O.knn_bfs(g, int(s), k, vis)
        c = f["per_count"][s]
        assert len(o) == c
        assert np.array_equal(o, f["per_list"][s][:c])
    out, counts = O.knn_batch(g, vis, k, f["batch_targets"])
    assert np.array_equal(counts, f["batch_count"])
    for i, c in enumerate(counts):
        assert np.array_equal(out[i][:c], f["batch_list"][i][:c])


@pytest.mark.parametrize("name", ["grip_rgg300", "grip_grid10", "grip_ba200"])
def test_grip_filtration_and_rounds_exact(name):
    f = golden(name)
    g = _graph(f)
    filt, borders = O.mis_filtration(g)
    assert np.array_equal(borders, f["borders"])
    assert np.array_equal(filt, f["filtration"])
    kcap = int(f["kcap"])
    n = int(f["n"])
    for i in range(int(f["rounds"])):
        layer = int(f[f"r{i}_layer"])
        d0 = f[f"r{i}_dec0"]
        dec = np.zeros(n, dtype=O.GRIP_VERTEX_DTYPE)
        dec["dispX"], dec["dispY"] = d0["disp"][:, 0], d0["disp"][:, 1]
        dec["oldDispX"], dec["oldDispY"] = d0["oldDisp"][:, 0], d0["oldDisp"][:, 1]
        for k in ("heat", "oldCos", "lastForceMag", "knnCount"):
            dec[k] = d0[k]
        dec["dispCalculated"] = d0["dispCalculated"]
        pos = f[f"r{i}_pos0"].copy()
        active = int(f["borders"][layer])
        # the kNN lists the reference used are the oracle's BFS over the
        # visible set (vertices of layers >= current)
        vis = np.zeros(n, np.uint8)
        vis[filt[:active]] = 1
        knn = f[f"r{i}_knn"]
        for v in filt[:active]:
            c = int(d0["knnCount"][v])
            o = O.knn_bfs(g, int(v), kcap, vis)
            assert np.array_equal(o[:c], knn[v][:c])
        st = O.grip_refine_round(g, filt, active, layer, knn, kcap, dec, pos)
        assert np.array_equal(pos, f[f"r{i}_pos1"])
        d1 = f[f"r{i}_dec1"]
        assert np.array_equal(dec["heat"], d1["heat"])
        assert (st.maxDisp, st.meanDisp, st.meanForce) == tuple(f[f"r{i}_stats"])


@pytest.mark.parametrize("name", ["api_force_linlog_ba400", "api_force_fr_grid20"])
def test_force_api_run_on_induced_subgraph_bit_exact(name):
    """gvizForceEmbedderRun after Begin(seed) on the subgraph induced by the
    non-hidden vertices: compact ids in ascending parent id, neighbours in
    the parent's order, stop once maxDisp < epsilon."""
    f = golden(name)
    n = int(f["n"])
    keep = np.ones(n, bool)
    keep[f["hidden"]] = False
    ids = np.flatnonzero(keep)
    index = -np.ones(n, np.int64)
    index[ids] = np.arange(ids.size)
    e = f["edges"]
    e = e[keep[e[:, 0]] & keep[e[:, 1]]]
    off, adj = G.csr_from_edges(ids.size, index[e])
    run = O.ForceRun(O.Graph(off, adj), _params_from(f), f["pos_begin"][ids])
    ref = np.stack([f["series_" + s] for s in ("maxDisp", "speed", "attractiveForce",
                                               "repulsiveForce", "gravityForce")], 1)
    steps = 0
    while steps < int(f["run_iters"]):
        st = run.step()
        assert (st.maxDisp, st.speed, st.meanAtt, st.meanRep,
                st.meanGrav) == tuple(ref[steps]), steps
        steps += 1
        if st.maxDisp < float(f["epsilon"]):
            break
    assert steps == int(f["run_rc"])
    assert np.array_equal(run.pos, f["pos_run"][ids])
    st = run.step()
    assert st.maxDisp == float(f["step_ret"])
    assert np.array_equal(run.pos, f["pos_step"][ids])


def test_grip_rounds_chain_bit_exact():
    """Rounds of one layer chained through the oracle's own per-vertex state
    (the reference run serially): every field after every round is equal."""
    f = golden("grip_grid16_k4")
    g = _graph(f)
    n, kcap = int(f["n"]), int(f["kcap"])
    filt = f["filtration"]
    prev, dec, pos = None, None, None
    for i in range(int(f["rounds"])):
        layer = int(f[f"r{i}_layer"])
        if layer != prev:
            d0 = f[f"r{i}_dec0"]
            dec = np.zeros(n, dtype=O.GRIP_VERTEX_DTYPE)
            dec["dispX"], dec["dispY"] = d0["disp"][:, 0], d0["disp"][:, 1]
            dec["oldDispX"], dec["oldDispY"] = d0["oldDisp"][:, 0], d0["oldDisp"][:, 1]
            for k in ("heat", "oldCos", "lastForceMag", "knnCount", "dispCalculated"):
                dec[k] = d0[k]
            pos = f[f"r{i}_pos0"].copy()
            prev = layer
        else:
            assert np.array_equal(pos, f[f"r{i}_pos0"])
        active = int(f["borders"][layer])
        st = O.grip_refine_round(g, filt, active, layer, f[f"r{i}_knn"], kcap, dec, pos)
        d1 = f[f"r{i}_dec1"]
        assert np.array_equal(pos, f[f"r{i}_pos1"]), i
        ids = filt[:active]
        for k, ref in (("dispX", d1["disp"][:, 0]), ("dispY", d1["disp"][:, 1]),
                       ("oldDispX", d1["oldDisp"][:, 0]), ("oldDispY", d1["oldDisp"][:, 1]),
                       ("heat", d1["heat"]), ("oldCos", d1["oldCos"]),
                       ("lastForceMag", d1["lastForceMag"]),
                       ("dispCalculated", d1["dispCalculated"])):
            assert np.array_equal(dec[k][ids], ref[ids]), (i, k)
        assert (st.maxDisp, st.meanDisp, st.meanForce) == tuple(f[f"r{i}_stats"])


def _nd_state(f, i, when, n, dim):
    d0 = f[f"r{i}_dec{when}"]
    dec = np.zeros(n, dtype=O.GRIP_VERTEX_ND_DTYPE)
    dec["disp"][:, :dim] = d0["disp"]
    dec["oldDisp"][:, :dim] = d0["oldDisp"]
    for k in ("heat", "oldCos", "lastForceMag", "knnCount", "dispCalculated"):
        dec[k] = d0[k]
    return dec


@pytest.mark.parametrize("name", ["grip_rgg300_d3", "grip_rgg300_d4",
                                  "grip_grid16_k4_d3", "grip_grid16_k4_d4"])
def test_grip_rounds_nd_bit_exact(name):
    """Every round of the reference's 3-D and 4-D runs from its own state:
    positions, displacements, heat and statistics bit for bit."""
    f = golden(name)
    g = _graph(f)
    n, dim, kcap = int(f["n"]), int(f["dim"]), int(f["kcap"])
    filt = f["filtration"]
    for i in range(int(f["rounds"])):
        layer = int(f[f"r{i}_layer"])
        active = int(f["borders"][layer])
        dec = _nd_state(f, i, 0, n, dim)
        pos = f[f"r{i}_pos0"].copy()
        st = O.grip_refine_round_nd(g, filt, active, layer, f[f"r{i}_knn"], kcap, dim,
                                    dec, pos)
        assert np.array_equal(pos, f[f"r{i}_pos1"]), (i, layer)
        d1 = f[f"r{i}_dec1"]
        assert np.array_equal(dec["disp"][:, :dim], d1["disp"]), i
        assert np.array_equal(dec["oldDisp"][:, :dim], d1["oldDisp"]), i
        for k in ("heat", "oldCos", "lastForceMag"):
            assert np.array_equal(dec[k], d1[k]), (i, k)
        assert (st.maxDisp, st.meanDisp, st.meanForce) == tuple(f[f"r{i}_stats"])


@pytest.mark.parametrize("dim", [2, 3, 4])
def test_grip_round_nd_with_zero_coordinates_is_the_2d_round(dim):
    """A layout whose coordinates beyond the first two are zero stays in the
    plane and moves exactly as the 2-D round moves it: the 3-D inertia
    system decouples (its z row holds the 2-D moment) and of the six 4-D
    planes only (0, 1) carries a rotation. Checked against the reference's
    2-D fixtures."""
    f = golden("grip_rgg300")
    g = _graph(f)
    n, kcap = int(f["n"]), int(f["kcap"])
    filt = f["filtration"]
    for i in range(int(f["rounds"])):
        layer = int(f[f"r{i}_layer"])
        active = int(f["borders"][layer])
        d0 = f[f"r{i}_dec0"]
        dec = np.zeros(n, dtype=O.GRIP_VERTEX_ND_DTYPE)
        dec["disp"][:, :2] = d0["disp"]
        dec["oldDisp"][:, :2] = d0["oldDisp"]
        for k in ("heat", "oldCos", "lastForceMag", "knnCount", "dispCalculated"):
            dec[k] = d0[k]
        pos = np.zeros((n, dim))
        pos[:, :2] = f[f"r{i}_pos0"]
        st = O.grip_refine_round_nd(g, filt, active, layer, f[f"r{i}_knn"], kcap, dim,
                                    dec, pos)
        assert np.array_equal(pos[:, :2], f[f"r{i}_pos1"]), (dim, i)
        assert not np.any(pos[:, 2:]), (dim, i)
        assert np.array_equal(dec["heat"], f[f"r{i}_dec1"]["heat"])
        assert (st.maxDisp, st.meanDisp, st.meanForce) == tuple(f[f"r{i}_stats"])
