"""HIP GRIP pieces against the reference's fixtures and the oracle: MIS
filtration and kNN lists bit-exact (integer results), refinement rounds to
FP64 round-off (1e-12 relative; only the reduction order differs)."""
import numpy as np
import pytest

import pyoracle as O
from conftest import golden
from grapher_amd import graphs as G

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["block", "wave"])
def search_mode(request, monkeypatch):
    """Every test runs twice: searches of calls with few queries on one block
    each (the default), and every search on one wavefront."""
    if request.param == "wave":
        monkeypatch.setenv("GVIZHIP_GRIP_BLOCK_QUERIES", "0")
    else:
        monkeypatch.delenv("GVIZHIP_GRIP_BLOCK_QUERIES", raising=False)
    return request.param


def _dec_from(d0, n):
    from grapher_amd import native
    dec = np.zeros(n, dtype=native.GRIP_VERTEX_DTYPE)
    dec["dispX"], dec["dispY"] = d0["disp"][:, 0], d0["disp"][:, 1]
    dec["oldDispX"], dec["oldDispY"] = d0["oldDisp"][:, 0], d0["oldDisp"][:, 1]
    for k in ("heat", "oldCos", "lastForceMag", "knnCount", "dispCalculated"):
        dec[k] = d0[k]
    return dec


@pytest.mark.parametrize("name", ["grip_rgg300", "grip_grid10", "grip_ba200"])
def test_mis_filtration_exact(gpu, name):
    f = golden(name)
    off, adj = G.csr_from_edges(int(f["n"]), f["edges"])
    eng = gpu.GripEngine(off, adj, kcap=int(f["kcap"]))
    filt, borders = eng.mis_filtration(2)
    assert np.array_equal(borders, f["borders"])
    assert np.array_equal(filt, f["filtration"])


@pytest.mark.parametrize("make", [
    lambda: G.random_geometric(20000, 8.0, 3)[:2],
    lambda: G.grid(60, 70),
    lambda: G.barabasi_albert(5000, 3, 2),
    lambda: G.random_tree(3000, 4),
])
def test_mis_filtration_matches_oracle_larger(gpu, make):
    """On the largest connected component (a subgraph with more than dim+1
    components never coarsens, in the reference as here)."""
    n, e = make()
    active, e = G.largest_component(n, e)
    off, adj = G.csr_from_edges(n, e)
    of, ob = O.mis_filtration(O.Graph(off, adj), active)
    eng = gpu.GripEngine(off, adj, active=active, kcap=10)
    filt, borders = eng.mis_filtration(2)
    assert np.array_equal(borders, ob)
    assert np.array_equal(filt, of)


def test_knn_per_vertex_and_batch_exact(gpu):
    f = golden("knn_rgg300")
    off, adj = G.csr_from_edges(int(f["n"]), f["edges"])
    eng = gpu.GripEngine(off, adj, kcap=16)
    k = int(f["k"])
    vis = f["visible"].astype(np.uint8)
    v, d, c = eng.knn(f["sources"], k, vis)
    assert np.array_equal(c, f["per_count"])
    for i in range(len(c)):
        assert np.array_equal(v[i, :c[i]], f["per_list"][i, :c[i], 0])
        assert np.array_equal(d[i, :c[i]], f["per_list"][i, :c[i], 1])
    v, d, c = eng.knn_batch(f["batch_targets"], k, vis)
    assert np.array_equal(c, f["batch_count"])
    for i in range(len(c)):
        assert np.array_equal(v[i, :c[i]], f["batch_list"][i, :c[i], 0])
        assert np.array_equal(d[i, :c[i]], f["batch_list"][i, :c[i], 1])


def test_knn_matches_oracle_on_larger_graph(gpu):
    n, e, _ = G.random_geometric(50000, 8.0, 7)
    off, adj = G.csr_from_edges(n, e)
    g = O.Graph(off, adj)
    rng = np.random.default_rng(1)
    vis = (rng.random(n) < 0.02).astype(np.uint8)
    src = rng.choice(n, 500, replace=False)
    eng = gpu.GripEngine(off, adj, kcap=10)
    v, d, c = eng.knn(src, 10, vis)
    for i, s in enumerate(src):
        o = O.knn_bfs(g, int(s), 10, vis)
        assert c[i] == len(o)
        assert np.array_equal(v[i, :c[i]], o[:, 0])
        assert np.array_equal(d[i, :c[i]], o[:, 1])
    targets = np.where(vis == 0)[0][:2000]
    sv = (rng.random(n) < 0.002).astype(np.uint8)
    sv[targets] = 0
    ov, cnt = O.knn_batch(g, sv, 6, targets)
    v, d, c = eng.knn_batch(targets, 6, sv)
    assert np.array_equal(c, cnt)
    for i in range(len(targets)):
        assert np.array_equal(v[i, :c[i]], ov[i, :c[i], 0])
        assert np.array_equal(d[i, :c[i]], ov[i, :c[i], 1])


@pytest.mark.parametrize("name", ["grip_rgg300", "grip_grid10", "grip_ba200"])
def test_refine_rounds_match_reference(gpu, name):
    f = golden(name)
    n = int(f["n"])
    kcap = int(f["kcap"])
    off, adj = G.csr_from_edges(n, f["edges"])
    eng = gpu.GripEngine(off, adj, kcap=kcap)
    eng.set_filtration(f["filtration"])
    for i in range(int(f["rounds"])):
        layer = int(f[f"r{i}_layer"])
        active = int(f["borders"][layer])
        knn = f[f"r{i}_knn"]
        eng.upload(f[f"r{i}_pos0"], _dec_from(f[f"r{i}_dec0"], n), knn[:, :, 0],
                   knn[:, :, 1])
        st = eng.refine_round(active, layer)
        pos, dec, _, _ = eng.download()
        ref = f[f"r{i}_pos1"]
        scale = np.abs(ref).max()
        assert np.abs(pos - ref).max() <= 1e-12 * scale, (i, layer)
        d1 = f[f"r{i}_dec1"]
        np.testing.assert_allclose(dec["heat"], d1["heat"], rtol=1e-12)
        np.testing.assert_allclose([st.maxDisp, st.meanDisp, st.meanForce],
                                   f[f"r{i}_stats"], rtol=1e-11)


def test_update_knn_matches_reference_lists(gpu):
    f = golden("grip_rgg300")
    n = int(f["n"])
    kcap = int(f["kcap"])
    off, adj = G.csr_from_edges(n, f["edges"])
    eng = gpu.GripEngine(off, adj, kcap=kcap)
    filt = f["filtration"]
    eng.set_filtration(filt)
    for i in range(int(f["rounds"])):
        layer = int(f[f"r{i}_layer"])
        active = int(f["borders"][layer])
        vis = np.zeros(n, np.uint8)
        vis[filt[:active]] = 1
        eng.update_knn(active, vis, kcap)
        _, dec, kv, kd = eng.download()
        d0 = f[f"r{i}_dec0"]
        ref = f[f"r{i}_knn"]
        for v in filt[:active]:
            c = int(d0["knnCount"][v])
            assert dec["knnCount"][v] == c
            assert np.array_equal(kv[v, :c], ref[v, :c, 0])
            assert np.array_equal(kd[v, :c], ref[v, :c, 1])


def _prefer_batch(n_vertices, n_edges, visible_count, target_count):
    """src/algorithms/search/gvizKNearest.c:149-166 (constants from
    include/algorithms/search/gvizKNearest.h:9-13)."""
    if visible_count == 0 or visible_count >= target_count or visible_count > 256:
        return False
    if 2.0 * n_edges / n_vertices > 32.0:
        return True
    return target_count >= 8 * visible_count


def _compare_state(i, when, pos, dec, kv, kd, ref_pos, ref_dec, ref_knn, active_ids):
    scale = np.abs(ref_pos).max()
    d = np.abs(pos - ref_pos).max()
    assert d <= 1e-12 * scale, f"round {i} {when}: positions differ by {d:.3g}"
    for k, a in (("dispX", ref_dec["disp"][:, 0]), ("dispY", ref_dec["disp"][:, 1]),
                 ("oldDispX", ref_dec["oldDisp"][:, 0]),
                 ("oldDispY", ref_dec["oldDisp"][:, 1]), ("heat", ref_dec["heat"]),
                 ("oldCos", ref_dec["oldCos"]), ("lastForceMag", ref_dec["lastForceMag"])):
        got = dec[k][active_ids]
        want = a[active_ids]
        tol = 1e-11 * max(np.abs(want).max(), 1e-300)
        bad = np.flatnonzero(np.abs(got - want) > tol)
        assert bad.size == 0, (f"round {i} {when}: {k} differs at vertex "
                               f"{active_ids[bad[0]]}: {got[bad[0]]!r} vs {want[bad[0]]!r}")
    for k in ("knnCount", "dispCalculated"):
        got, want = dec[k][active_ids], ref_dec[k][active_ids]
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, (f"round {i} {when}: {k} differs at vertex "
                               f"{active_ids[bad[0]]}: {got[bad[0]]} vs {want[bad[0]]}")
    if ref_knn is not None:
        for v in active_ids:
            c = int(ref_dec["knnCount"][v])
            assert np.array_equal(kv[v, :c], ref_knn[v, :c, 0]), (i, when, "knn", v)
            assert np.array_equal(kd[v, :c], ref_knn[v, :c, 1]), (i, when, "dist", v)


def test_grip_rounds_chain_through_device_state(gpu):
    """The whole schedule from the state after Begin, with nothing re-uploaded:
    every refinement round, NextStage placement, decorator clear and kNN
    update runs on the device state the previous phase left, and every field
    of every vertex is compared with the reference's after each phase."""
    f = golden("grip_grid16_k4")
    n = int(f["n"])
    kcap, k = int(f["kcap"]), int(f["kmax"])
    off, adj = G.csr_from_edges(n, f["edges"])
    ne = f["edges"].shape[0]
    eng = gpu.GripEngine(off, adj, kcap=kcap)
    filt, borders = f["filtration"], f["borders"]
    eng.set_filtration(filt)
    knn0 = f["r0_knn"]
    eng.upload(f["r0_pos0"], _dec_from(f["r0_dec0"], n), knn0[:, :, 0], knn0[:, :, 1])
    prev = int(f["r0_layer"])
    for i in range(int(f["rounds"])):
        layer = int(f[f"r{i}_layer"])
        active = int(borders[layer])
        ids = filt[:active]
        if layer != prev:
            begin, end = int(borders[layer + 1]), int(borders[layer])
            vis = np.zeros(n, np.uint8)
            vis[filt[:begin]] = 1
            eng.place(begin, end, vis, k, _prefer_batch(n, ne, begin, end - begin))
            vis[filt[:end]] = 1
            eng.clear_decorators(end)
            eng.update_knn(end, vis, k, _prefer_batch(n, ne, end, end))
            pos, dec, kv, kd = eng.download()
            _compare_state(i, "after NextStage", pos, dec, kv, kd, f[f"r{i}_pos0"],
                           f[f"r{i}_dec0"], f[f"r{i}_knn"], ids)
            prev = layer
        st = eng.refine_round(active, layer)
        pos, dec, kv, kd = eng.download()
        _compare_state(i, "after the round", pos, dec, kv, kd, f[f"r{i}_pos1"],
                       f[f"r{i}_dec1"], None, ids)
        np.testing.assert_allclose([st.maxDisp, st.meanDisp, st.meanForce],
                                   f[f"r{i}_stats"], rtol=1e-11)


def _dec_nd_from(d0, n, dim):
    from grapher_amd import native
    dec = np.zeros(n, dtype=native.GRIP_VERTEX_ND_DTYPE)
    dec["disp"][:, :dim] = d0["disp"]
    dec["oldDisp"][:, :dim] = d0["oldDisp"]
    for k in ("heat", "oldCos", "lastForceMag", "knnCount", "dispCalculated"):
        dec[k] = d0[k]
    return dec


def _compare_state_nd(i, when, dim, pos, dec, kv, kd, ref_pos, ref_dec, ref_knn, ids):
    scale = np.abs(ref_pos).max()
    d = np.abs(pos - ref_pos).max()
    assert d <= 1e-12 * scale, f"round {i} {when}: positions differ by {d:.3g}"
    for k in ("disp", "oldDisp"):
        got, want = dec[k][ids, :dim], ref_dec[k][ids]
        tol = 1e-11 * max(np.abs(want).max(), 1e-300)
        assert np.abs(got - want).max() <= tol, (i, when, k)
        assert not np.any(dec[k][ids, dim:]), (i, when, k, "coordinates beyond dim")
    for k in ("heat", "oldCos", "lastForceMag"):
        got, want = dec[k][ids], ref_dec[k][ids]
        tol = 1e-11 * max(np.abs(want).max(), 1e-300)
        assert np.abs(got - want).max() <= tol, (i, when, k)
    for k in ("knnCount", "dispCalculated"):
        assert np.array_equal(dec[k][ids], ref_dec[k][ids]), (i, when, k)
    if ref_knn is not None:
        for v in ids:
            c = int(ref_dec["knnCount"][v])
            assert np.array_equal(kv[v, :c], ref_knn[v, :c, 0]), (i, when, "knn", v)
            assert np.array_equal(kd[v, :c], ref_knn[v, :c, 1]), (i, when, "dist", v)


@pytest.mark.parametrize("name", ["grip_rgg300_d3", "grip_rgg300_d4"])
def test_refine_rounds_match_reference_nd(gpu, name):
    """Refinement rounds in three and four dimensions from the reference's
    state: the 3-D round removes one rotation solved from the inertia tensor
    (solve3x3), the 4-D round the six coordinate-plane rotations in turn."""
    f = golden(name)
    n, dim = int(f["n"]), int(f["dim"])
    kcap = int(f["kcap"])
    off, adj = G.csr_from_edges(n, f["edges"])
    eng = gpu.GripEngine(off, adj, kcap=kcap, dim=dim)
    eng.set_filtration(f["filtration"])
    print(f"grip {name}: {int(f['rounds'])} rounds in {dim}-D", flush=True)
    for i in range(int(f["rounds"])):
        layer = int(f[f"r{i}_layer"])
        active = int(f["borders"][layer])
        knn = f[f"r{i}_knn"]
        eng.upload(f[f"r{i}_pos0"], _dec_nd_from(f[f"r{i}_dec0"], n, dim), knn[:, :, 0],
                   knn[:, :, 1])
        st = eng.refine_round(active, layer)
        pos, dec, _, _ = eng.download()
        _compare_state_nd(i, "after the round", dim, pos, dec, None, None,
                          f[f"r{i}_pos1"], f[f"r{i}_dec1"], None, f["filtration"][:active])
        np.testing.assert_allclose([st.maxDisp, st.meanDisp, st.meanForce],
                                   f[f"r{i}_stats"], rtol=1e-11)


@pytest.mark.parametrize("name", ["grip_grid16_k4_d3", "grip_grid16_k4_d4"])
def test_grip_rounds_chain_through_device_state_nd(gpu, name):
    """The whole 3-D / 4-D schedule on the device state, compared with the
    reference after every NextStage (placement, decorator clear, kNN
    update) and every round."""
    f = golden(name)
    n, dim = int(f["n"]), int(f["dim"])
    kcap, k = int(f["kcap"]), int(f["kmax"])
    off, adj = G.csr_from_edges(n, f["edges"])
    ne = f["edges"].shape[0]
    eng = gpu.GripEngine(off, adj, kcap=kcap, dim=dim)
    filt, borders = f["filtration"], f["borders"]
    eng.set_filtration(filt)
    knn0 = f["r0_knn"]
    eng.upload(f["r0_pos0"], _dec_nd_from(f["r0_dec0"], n, dim), knn0[:, :, 0],
               knn0[:, :, 1])
    prev = int(f["r0_layer"])
    print(f"grip chain {name}: {int(f['rounds'])} rounds in {dim}-D", flush=True)
    for i in range(int(f["rounds"])):
        layer = int(f[f"r{i}_layer"])
        active = int(borders[layer])
        ids = filt[:active]
        if layer != prev:
            begin, end = int(borders[layer + 1]), int(borders[layer])
            vis = np.zeros(n, np.uint8)
            vis[filt[:begin]] = 1
            eng.place(begin, end, vis, k, _prefer_batch(n, ne, begin, end - begin))
            vis[filt[:end]] = 1
            eng.clear_decorators(end)
            eng.update_knn(end, vis, k, _prefer_batch(n, ne, end, end))
            pos, dec, kv, kd = eng.download()
            _compare_state_nd(i, "after NextStage", dim, pos, dec, kv, kd, f[f"r{i}_pos0"],
                              f[f"r{i}_dec0"], f[f"r{i}_knn"], ids)
            prev = layer
        st = eng.refine_round(active, layer)
        pos, dec, kv, kd = eng.download()
        _compare_state_nd(i, "after the round", dim, pos, dec, kv, kd, f[f"r{i}_pos1"],
                          f[f"r{i}_dec1"], None, ids)
        np.testing.assert_allclose([st.maxDisp, st.meanDisp, st.meanForce],
                                   f[f"r{i}_stats"], rtol=1e-11)


@pytest.mark.parametrize("dim", [3, 4])
@pytest.mark.parametrize("layer", [0, 1])
def test_refine_round_nd_matches_oracle_on_larger_graph(gpu, dim, layer):
    """One 3-D / 4-D round on a 20,000-vertex graph from a random state
    (positions, displacements, heat, cosines, half the vertices already
    computed, full kNN lists with graph distances 1..5), device against the
    oracle's serial round: layer 0 runs the FR-like edge and kNN forces,
    layer 1 the Kamada-Kawai kNN springs."""
    _round_nd_against_oracle(gpu, dim, layer, flat=())


@pytest.mark.parametrize("layer", [0, 1])
def test_refine_round_4d_with_a_flat_middle_plane_matches_oracle(gpu, layer):
    """The 4-D round with coordinates 1 and 2 the same for every vertex: the
    (1, 2) plane, fourth of the six in the reference's order, has no inertia
    and is skipped (its rotate flag 0), while the planes before and after it
    still rotate; every later pass recomputes the earlier removals per
    vertex, so the skip must reach the last pass exactly as the oracle's
    stored displacements do."""
    _round_nd_against_oracle(gpu, 4, layer, flat=((1, 3.0), (2, -7.0)))


def _round_nd_against_oracle(gpu, dim, layer, flat):
    n, kcap = 20000, 8
    _, e, _ = G.random_geometric(n, 8.0, 13)
    off, adj = G.csr_from_edges(n, e)
    rng = np.random.default_rng(100 * dim + layer)
    filt = rng.permutation(n).astype(np.int64)
    active = n - 1000 * layer  # layer 1: a prefix of the filtration
    pos = rng.uniform(-100.0, 100.0, (n, dim))
    for c, value in flat:
        pos[:, c] = value
    kv = rng.integers(0, n, (n, kcap)).astype(np.int64)
    kd = rng.integers(1, 6, (n, kcap)).astype(np.int64)
    dec = np.zeros(n, dtype=O.GRIP_VERTEX_ND_DTYPE)
    dec["disp"][:, :dim] = rng.normal(0.0, 2.0, (n, dim))
    dec["oldDisp"][:, :dim] = rng.normal(0.0, 2.0, (n, dim))
    dec["heat"] = rng.uniform(0.1, 5.0, n)
    dec["oldCos"] = rng.uniform(-1.0, 1.0, n)
    dec["knnCount"] = kcap
    dec["dispCalculated"] = rng.integers(0, 2, n)
    eng = gpu.GripEngine(off, adj, kcap=kcap, dim=dim)
    eng.set_filtration(filt)
    eng.upload(pos, dec, kv, kd)
    st = eng.refine_round(active, layer)
    got_pos, got_dec, _, _ = eng.download()
    knn = np.stack([kv, kd], axis=2)
    ref_dec = dec.copy()
    ref_pos = np.ascontiguousarray(pos.copy())
    ref = O.grip_refine_round_nd(O.Graph(off, adj), filt, active, layer, knn, kcap, dim,
                                 ref_dec, ref_pos)
    ids = filt[:active]
    scale = np.abs(ref_pos).max()
    assert np.abs(got_pos - ref_pos).max() <= 1e-12 * scale
    for k in ("heat", "oldCos", "lastForceMag"):
        np.testing.assert_allclose(got_dec[k][ids], ref_dec[k][ids], rtol=1e-12, err_msg=k)
    d = np.abs(got_dec["disp"][ids] - ref_dec["disp"][ids]).max()
    assert d <= 1e-11 * np.abs(ref_dec["disp"][ids]).max()
    np.testing.assert_allclose([st.maxDisp, st.meanDisp, st.meanForce],
                               [ref.maxDisp, ref.meanDisp, ref.meanForce], rtol=1e-11)


def _random_round_state(n, kcap, active, filt, rng, targets_in_layer):
    pos = rng.uniform(-100.0, 100.0, (n, 2))
    pool = filt[:active] if targets_in_layer else np.arange(n)
    kv = rng.choice(pool, (n, kcap)).astype(np.int64)
    kd = rng.integers(1, 6, (n, kcap)).astype(np.int64)
    from grapher_amd import native
    dec = np.zeros(n, dtype=native.GRIP_VERTEX_DTYPE)
    dec["dispX"], dec["dispY"] = rng.normal(0.0, 2.0, n), rng.normal(0.0, 2.0, n)
    dec["oldDispX"], dec["oldDispY"] = rng.normal(0.0, 2.0, n), rng.normal(0.0, 2.0, n)
    dec["heat"] = rng.uniform(0.1, 5.0, n)
    dec["oldCos"] = rng.uniform(-1.0, 1.0, n)
    dec["knnCount"] = kcap
    dec["dispCalculated"] = rng.integers(0, 2, n)
    return pos, dec, kv, kd


@pytest.mark.parametrize("targets_in_layer", [True, False])
@pytest.mark.parametrize("layer", [0, 1])
def test_refine_round_2d_compact_matches_oracle_on_larger_graph(gpu, layer,
                                                                targets_in_layer):
    """Layers of >= 32,768 active vertices run their rounds on a compact copy
    in the Morton order of the positions (k_compact_*). One round on a
    40,000-vertex graph from a random state against the oracle's serial
    round; kNN targets inside the layer (the compact path) or anywhere in
    the graph (a target outside the layer: the id-ordered path)."""
    n, kcap = 40000, 8
    _, e, _ = G.random_geometric(n, 8.0, 17)
    off, adj = G.csr_from_edges(n, e)
    rng = np.random.default_rng(7 + layer + 2 * targets_in_layer)
    filt = rng.permutation(n).astype(np.int64)
    active = n - 4000 * layer
    pos, dec, kv, kd = _random_round_state(n, kcap, active, filt, rng, targets_in_layer)
    eng = gpu.GripEngine(off, adj, kcap=kcap)
    eng.set_filtration(filt)
    eng.upload(pos, dec, kv, kd)
    st = eng.refine_round(active, layer)
    got_pos, got_dec, _, _ = eng.download()
    odec = np.zeros(n, dtype=O.GRIP_VERTEX_DTYPE)
    for k in ("dispX", "dispY", "oldDispX", "oldDispY", "heat", "oldCos",
              "lastForceMag", "knnCount"):
        odec[k] = dec[k]
    odec["dispCalculated"] = dec["dispCalculated"].astype(np.uint8)
    ref_pos = np.ascontiguousarray(pos.copy())
    ref = O.grip_refine_round(O.Graph(off, adj), filt, active, layer,
                              np.stack([kv, kd], axis=2), kcap, odec, ref_pos)
    ids = filt[:active]
    assert np.abs(got_pos - ref_pos).max() <= 1e-12 * np.abs(ref_pos).max()
    for k in ("heat", "oldCos", "lastForceMag"):
        np.testing.assert_allclose(got_dec[k][ids], odec[k][ids], rtol=1e-12, err_msg=k)
    for k in ("dispX", "dispY", "oldDispX", "oldDispY"):
        d = np.abs(got_dec[k][ids] - odec[k][ids]).max()
        assert d <= 1e-11 * np.abs(odec[k][ids]).max(), k
    assert np.array_equal(got_dec["dispCalculated"][ids], odec["dispCalculated"][ids])
    np.testing.assert_allclose([st.maxDisp, st.meanDisp, st.meanForce],
                               [ref.maxDisp, ref.meanDisp, ref.meanForce], rtol=1e-11)


def test_compact_rounds_unchanged_by_reads_in_between(gpu):
    """Downloads between compact rounds write the compact state back without
    dropping it: four rounds with a download after each equal four rounds
    without, bit for bit; an upload in between starts a fresh compact copy
    from the uploaded state."""
    n, kcap = 40000, 8
    _, e, _ = G.random_geometric(n, 8.0, 19)
    off, adj = G.csr_from_edges(n, e)
    rng = np.random.default_rng(3)
    filt = rng.permutation(n).astype(np.int64)
    pos, dec, kv, kd = _random_round_state(n, kcap, n, filt, rng, True)
    runs = []
    for reads in (False, True):
        eng = gpu.GripEngine(off, adj, kcap=kcap)
        eng.set_filtration(filt)
        eng.upload(pos, dec, kv, kd)
        for _ in range(4):
            eng.refine_round(n, 0)
            if reads:
                eng.download()
        runs.append(eng.download())
        if reads:
            p2, d2, kv2, kd2 = runs[-1]
            eng.upload(p2, d2, kv2, kd2)
            eng.refine_round(n, 0)
            after_upload = eng.download()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    eng = gpu.GripEngine(off, adj, kcap=kcap)
    eng.set_filtration(filt)
    eng.upload(*runs[0])
    eng.refine_round(n, 0)
    fresh = eng.download()
    assert np.array_equal(after_upload[0], fresh[0])
    assert np.array_equal(after_upload[1], fresh[1])
