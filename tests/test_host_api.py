"""The C11 host layer (libgviz.so, include/gviz/*.h): the reference's embedder
API on the gfx950 engine.

CPU tests: the library exports every declared function; host-side unit
checks (graph/CSR, actions, stats, draw mask, checkpoints, configuration
defaults and setters, K policies) in tests/host/api_driver.c; without a GPU
the embedders fail loudly instead of computing anything on the CPU.

GPU tests replay the reference's own API runs (fixtures written by the
reference library: Init, Configure*, Begin(seed), Run, Step, GRIP Embed,
Tutte polygon + seeding + sweeps) through this library and compare the
outputs: Begin's random placement, the GRIP filtration and schedule and the
Tutte sweeps bit for bit; force and GRIP trajectories step for step while
they stay within reach of the reference's (rounding differences grow over
a run) and by final layout quality after that.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden
import refio
from grapher_amd import graphs as G
from grapher_amd import native

DRIVER = os.path.join(ROOT, "tests", "host", "api_driver")
HOST_LIB = os.path.join(ROOT, "grapher_amd", "libgviz.so")


def _declared():
    names = set()
    hdir = os.path.join(ROOT, "include", "gviz")
    for h in os.listdir(hdir):
        text = open(os.path.join(hdir, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"typedef[^;]*;", "", text, flags=re.S)
        text = re.sub(r"^#define(?:[^\n]*\\\n)*[^\n]*$", "", text, flags=re.M)
        found = set(re.findall(r"\b(gviz[A-Z][A-Za-z0-9]*)\s*\(", text))
        # header-only helpers (static inline) are not library symbols
        found -= set(re.findall(r"static\s+inline\s+[\w\s\*]*?\b(gviz[A-Z]\w*)\s*\(", text))
        names.update(found)
    return sorted(names)


def test_host_library_exports_every_declared_function():
    lib = ctypes.CDLL(HOST_LIB)
    declared = _declared()
    assert len(declared) > 60
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing


def test_host_unit_checks():
    r = subprocess.run([DRIVER, "unit"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK (0 failed checks)" in r.stdout


def test_embedders_fail_loudly_without_gpu():
    if native.device_count() > 0:
        pytest.skip("a GPU is visible")
    f = golden("api_force_fr_grid20")
    with pytest.raises(subprocess.CalledProcessError):
        _api_force(f)


def _api_force(f, env=None):
    old = {}
    for k, v in (env or {}).items():
        old[k] = os.environ.get(k)
        os.environ[k] = v
    try:
        return refio.api_force(
            int(f["n"]), f["edges"], f["hidden"], model=int(f["param_model"]),
            barnes_hut=int(f["param_barnes_hut"]),
            prevent_overlap=int(f["param_prevent_overlap"]), npc=int(f["param_npc"]),
            seed=int(f["seed"]), run_iters=int(f["run_iters"]),
            edge_length=float(f["param_edge_length"]), theta=float(f["param_theta"]),
            gravity=float(f["param_gravity"]),
            overlap_constant=float(f["param_overlap_constant"]),
            radius_base=float(f["param_radius_base"]),
            radius_per_degree=float(f["param_radius_per_degree"]),
            jitter=float(f["param_jitter"]), epsilon=float(f["epsilon"]), exe=DRIVER)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


SERIES = ("maxDisp", "speed", "attractiveForce", "repulsiveForce", "gravityForce")


def _active(f):
    mask = np.ones(int(f["n"]), bool)
    mask[f["hidden"]] = False
    return mask


def _first_divergence(a, b, rtol):
    """Index of the first sample whose relative difference exceeds rtol."""
    n = min(len(a), len(b))
    rel = np.abs(a[:n] - b[:n]) / np.maximum(np.abs(b[:n]), 1e-300)
    bad = np.flatnonzero(rel > rtol)
    return int(bad[0]) if bad.size else n, rel


def _compact(f):
    """The active subgraph over compact ids (ascending parent id)."""
    import pyoracle as O
    from grapher_amd import graphs as G
    n = int(f["n"])
    keep = _active(f)
    ids = np.flatnonzero(keep)
    index = -np.ones(n, np.int64)
    index[ids] = np.arange(ids.size)
    e = f["edges"]
    e = e[keep[e[:, 0]] & keep[e[:, 1]]]
    off, adj = G.csr_from_edges(ids.size, index[e])
    return O.Graph(off, adj), ids


def _layout_quality(g, pos, edge_length=10.0):
    import pyoracle as O
    pool = O.Pool(4)
    return O.stress(g, pos, edge_length, None, pool), O.edge_energy(g, pos)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["api_force_linlog_ba400", "api_force_fr_grid20"])
def test_force_api_fp64_matches_reference(gpu, name):
    """FP64 engine behind the reference API. Begin's placement is bit-exact;
    each step's forces agree with the reference to ~1e-11 (the Barnes-Hut
    sums run in another order), so the first steps agree to 1e-8 and the
    chaotic FA2 dynamics then amplify the difference: the run is checked for
    its length and the final layout's stress and edge energy against the
    band the reference itself spans when its start is scaled by 1 +- 1e-12
    (eight such reference runs are stored in the fixture), widened by 1 %
    (3 % or 2 steps for the run length)."""
    f = golden(name)
    print(f"host api force {name}", flush=True)
    r = _api_force(f, {"GVIZ_TEST_PRECISION": "fp64", "GVIZ_TEST_RESIDENT": "0"})
    assert r["box_extent"] == float(f["box_extent"])
    assert r["begin_rc"] == 0
    assert np.array_equal(r["pos_begin"], f["pos_begin"])
    for s in SERIES:
        first, rel = _first_divergence(r["series_" + s], f["series_" + s], 1e-9)
        print(f"  {s}: first step beyond 1e-9 relative: {first}; "
              f"max rel over the run {rel.max():.3g}", flush=True)
        np.testing.assert_allclose(r["series_" + s][:8], f["series_" + s][:8],
                                   rtol=1e-8, atol=1e-12, err_msg=s)
    ref_rc = int(f["run_rc"])
    rcs = np.append(f["spread_run_rc"], ref_rc)
    slack = max(2, 0.03 * ref_rc)
    print(f"  run length {r['run_rc']} vs {ref_rc} (perturbed reference runs "
          f"{rcs.min()}..{rcs.max()})")
    assert rcs.min() - slack <= r["run_rc"] <= rcs.max() + slack
    act = _active(f)
    assert np.all(r["pos_run"][~act] == 0.0)
    assert r["iteration"] == r["run_rc"] + 1
    g, ids = _compact(f)
    s_ref, e_ref = _layout_quality(g, f["pos_step"][ids])
    s_gpu, e_gpu = _layout_quality(g, r["pos_step"][ids])
    ss = np.append(f["spread_stress"], s_ref)
    es = np.append(f["spread_energy"], e_ref)
    print(f"  final stress {s_gpu:.6g} vs {s_ref:.6g} (perturbed reference runs "
          f"{ss.min():.6g}..{ss.max():.6g}); energy {e_gpu:.6g} vs {e_ref:.6g} "
          f"({es.min():.6g}..{es.max():.6g})")
    assert ss.min() - 0.01 * abs(s_ref) <= s_gpu <= ss.max() + 0.01 * abs(s_ref)
    assert es.min() - 0.01 * abs(e_ref) <= e_gpu <= es.max() + 0.01 * abs(e_ref)


@pytest.mark.gpu
def test_force_api_resident_positions_identical(gpu):
    """Keeping positions on the device between steps changes nothing."""
    f = golden("api_force_linlog_ba400")
    a = _api_force(f, {"GVIZ_TEST_PRECISION": "fp64", "GVIZ_TEST_RESIDENT": "0"})
    b = _api_force(f, {"GVIZ_TEST_PRECISION": "fp64", "GVIZ_TEST_RESIDENT": "1"})
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_force_api_reduced_precision_tracks_reference(gpu, precision):
    """FP32 force math: the first steps agree closely with the reference's
    FP64 run; the later trajectory is only required to stay finite and in
    range (force layouts amplify rounding differences)."""
    f = golden("api_force_linlog_ba400")
    r = _api_force(f, {"GVIZ_TEST_PRECISION": precision})
    assert np.array_equal(r["pos_begin"], f["pos_begin"])
    assert r["run_rc"] == int(f["run_rc"])
    for s in SERIES:
        np.testing.assert_allclose(r["series_" + s][:3], f["series_" + s][:3],
                                   rtol=2e-4, err_msg=s)
    act = _active(f)
    assert np.all(np.isfinite(r["pos_step"]))
    ext = np.abs(f["pos_step"][act]).max()
    assert np.abs(r["pos_step"][act]).max() <= 2.0 * ext


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["api_grip_rgg400_constant", "api_grip_grid16_decay"])
def test_grip_embed_matches_reference(gpu, name):
    """gvizGRIPEmbedderEmbed end to end. The filtration, the draw mask and
    the layer schedule are exact. Every round agrees with the reference's to
    1e-12 from the same state (test_gpu_grip), but the per-round means and
    the rotation fix-up are reductions in another order, and the heat
    feedback (x1.45 per round while a vertex keeps its direction) amplifies
    those last-bit differences by about 2.5x per round: the series are
    compared to 1e-6 over the first 60 rounds and the final (unconverged)
    layouts by stress and edge energy: within 1 % of the reference, or
    inside the reference's own spread under 1e-12 start perturbations
    widened by 1 %."""
    f = golden(name)
    print(f"host api grip {name}", flush=True)
    r = refio.api_grip(int(f["n"]), f["edges"], int(f["kcap"]), int(f["placement_k"]),
                       int(f["refinement_k"]), int(f["policy"]), exe=DRIVER)
    assert r["embed_rc"] == 0
    assert np.array_equal(r["borders"], f["borders"])
    assert np.array_equal(r["visible"], f["visible"])
    for s in ("heat", "meanDisp", "maxDisp", "meanForce"):
        first, rel = _first_divergence(r["series_" + s], f["series_" + s], 1e-9)
        print(f"  {s}: {len(r['series_' + s])} rounds (reference "
              f"{len(f['series_' + s])}); first beyond 1e-9 relative: {first}",
              flush=True)
    ref_rounds = len(f["series_maxDisp"])
    assert abs(len(r["series_maxDisp"]) - ref_rounds) <= max(2, 0.05 * ref_rounds)
    head = min(60, ref_rounds, len(r["series_maxDisp"]))
    for s in ("heat", "meanDisp", "maxDisp", "meanForce"):
        np.testing.assert_allclose(r["series_" + s][:head], f["series_" + s][:head],
                                   rtol=1e-6, atol=1e-12, err_msg=s)
    from grapher_amd import graphs as G
    import pyoracle as O
    g = O.Graph(*G.csr_from_edges(int(f["n"]), f["edges"]))
    s_ref, e_ref = _layout_quality(g, f["pos"])
    s_gpu, e_gpu = _layout_quality(g, r["pos"])
    print(f"  final stress {s_gpu:.6g} vs {s_ref:.6g} (perturbed reference runs "
          f"{f['spread_stress'].min():.6g}..{f['spread_stress'].max():.6g}); energy "
          f"{e_gpu:.6g} vs {e_ref:.6g} ({f['spread_energy'].min():.6g}.."
          f"{f['spread_energy'].max():.6g})")
    # within 1 % of the reference's run, or inside the spread of the
    # reference's own runs from Begin's placement scaled by 1 +- 1e-12,
    # widened by 1 %
    for got, ref, spread in ((s_gpu, s_ref, f["spread_stress"]),
                             (e_gpu, e_ref, f["spread_energy"])):
        lo, hi = min(spread.min(), ref), max(spread.max(), ref)
        assert abs(got - ref) <= 0.01 * abs(ref) or lo * 0.99 <= got <= hi * 1.01


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tutte_grid6_jacobi", "tutte_delaunay300_jacobi",
                                  "tutte_delaunay300_dt60", "tutte_k4", "tutte_grid6_gs"])
def test_tutte_api_bit_exact(gpu, name):
    f = golden(name)
    print(f"host api tutte {name}", flush=True)
    p0, deltas, p1 = refio.tutte(int(f["n"]), f["edges"], f["boundary"], int(f["iters"]),
                                 float(f["radius"]), float(f["dt"]), name.endswith("_gs"),
                                 exe=DRIVER)
    assert np.array_equal(p0, f["p0"])
    assert np.array_equal(deltas, f["deltas"])
    assert np.array_equal(p1, f["p1"])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_force_state_checkpoint_resumes_exactly(gpu, tmp_path, monkeypatch, precision):
    """gvizForceEmbedderSaveState / LoadState: a second embedder (begun from
    another seed) that loads the state saved after 15 Run steps continues
    bit for bit with the first over the next 25 steps."""
    from grapher_amd import graphs as G
    n, e = G.barabasi_albert(800, 2, 3)
    monkeypatch.setenv("GVIZ_TEST_PRECISION", precision)
    (pa, ia, ga), (pb, ib, gb) = refio.force_checkpoint(
        n, e, 1, 7, 15, 25, str(tmp_path / "state.bin"), DRIVER)
    np.testing.assert_array_equal(pa, pb)
    assert ia == ib == 40 and ga == gb


@pytest.mark.gpu
def test_grip_state_checkpoint_resumes_exactly(gpu, tmp_path):
    """gvizGRIPEmbedderSaveState / LoadState: saved after 3 NextStage calls and
    4 rounds, the continuation (4 rounds, then NextStage + 4 rounds up to 3
    times) is bit-identical in the saved embedder and in a second one that
    Begins on the same graph and loads the state."""
    from grapher_amd import graphs as G
    n, e = G.grid(50, 50)
    (la, pa, va), (lb, pb, vb) = refio.grip_checkpoint(
        n, e, 10, 10, 10, 0, 3, 4, str(tmp_path / "grip.bin"), DRIVER)
    assert la == lb
    np.testing.assert_array_equal(va, vb)
    np.testing.assert_array_equal(pa, pb)
    assert np.all(np.isfinite(pa[va == 1]))


def _stress_nd(n, edges, pos, edge_length=10.0):
    """Kamada-Kawai stress of a layout in any dimension: the sum over pairs
    in one component of (|p_u - p_v| - L d_uv)^2 / (L d_uv)^2 with d the
    BFS distance."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import shortest_path
    a = sp.coo_matrix((np.ones(len(edges)), (edges[:, 0], edges[:, 1])), shape=(n, n))
    dist = shortest_path(a, directed=False, unweighted=True)
    iu = np.triu_indices(n, 1)
    d = dist[iu]
    keep = np.isfinite(d)  # pairs in one component
    ideal = edge_length * d[keep]
    real = np.linalg.norm(pos[iu[0][keep]] - pos[iu[1][keep]], axis=1)
    return float(np.sum((real - ideal) ** 2 / ideal ** 2))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["api_grip_rgg400_constant_d3", "api_grip_rgg400_constant_d4"])
def test_grip_embed_matches_reference_nd(gpu, name):
    """gvizGRIPEmbedderEmbed in three and four dimensions end to end, by the
    bars of the 2-D run: exact filtration, draw mask and schedule, the
    per-round series to 1e-6 over the first 60 rounds, and the final
    layout's stress within 1 %."""
    f = golden(name)
    n, dim = int(f["n"]), int(f["dim"])
    print(f"host api grip {name}", flush=True)
    r = refio.api_grip(n, f["edges"], int(f["kcap"]), int(f["placement_k"]),
                       int(f["refinement_k"]), int(f["policy"]), exe=DRIVER, dim=dim)
    assert r["embed_rc"] == 0
    assert r["pos"].shape == (n, dim)
    assert np.array_equal(r["borders"], f["borders"])
    assert np.array_equal(r["visible"], f["visible"])
    ref_rounds = len(f["series_maxDisp"])
    assert abs(len(r["series_maxDisp"]) - ref_rounds) <= max(2, 0.05 * ref_rounds)
    head = min(60, ref_rounds, len(r["series_maxDisp"]))
    for s in ("heat", "meanDisp", "maxDisp", "meanForce"):
        np.testing.assert_allclose(r["series_" + s][:head], f["series_" + s][:head],
                                   rtol=1e-6, atol=1e-12, err_msg=s)
    s_ref = _stress_nd(n, f["edges"], f["pos"])
    s_gpu = _stress_nd(n, f["edges"], r["pos"])
    print(f"  final stress {s_gpu:.6g} vs {s_ref:.6g}")
    assert abs(s_gpu - s_ref) <= 0.01 * abs(s_ref)


def _same_cycle_shift(a, b):
    """(shift, reversed) with a == roll(b or reversed b), or None."""
    a, b = list(a), list(b)
    if len(a) != len(b) or sorted(a) != sorted(b):
        return None
    for rev in (False, True):
        c = list(reversed(b)) if rev else b
        k = c.index(a[0])
        if c[k:] + c[:k] == a:
            return k, rev
    return None


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["delaunay400_jacobi", "delaunay400_spring",
                                  "trigrid10_jacobi", "trigrid10_spring",
                                  "grid8_jacobi", "grid8_spring",
                                  "k33_jacobi", "k33_spring"])
def test_tutte_begin_matches_reference_up_to_polygon_symmetry(gpu, name):
    """Tutte / spring-Tutte Begin (planarity test, largest face, convex
    polygon, seeded interior) and Run to convergence. The reference embeds
    with the Boyer-Myrvold suite, this engine with the left-right test; on
    the 3-connected graphs the largest face is the same cycle, which may
    start at another vertex or run the other way round, so the converged
    layout equals the reference's under that symmetry of the polygon (an
    orthogonal map fitted on the boundary). The plain 8 x 8 grid is
    2-connected but not 3-connected (a corner's two neighbours separate it);
    both embeddings still make its border the largest face. A non-planar
    graph returns -2 from both."""
    f = golden("tutte_begin_" + name)
    n = int(f["n"])
    print(f"tutte begin {name}", flush=True)
    r = refio.tutte_begin(n, f["edges"], int(f["spring"]), int(f["iters"]),
                          float(f["epsilon"]), exe=DRIVER)
    assert r["begin_rc"] == int(f["begin_rc"])
    if r["begin_rc"] != 0:
        return
    sym = _same_cycle_shift(r["boundary"], f["boundary"])
    assert sym is not None, (r["boundary"], f["boundary"])
    ref_run = int(f["run"])
    assert abs(r["run"] - ref_run) <= max(2, 0.01 * ref_run), (r["run"], ref_run)
    b = np.asarray(f["boundary"])
    # orthogonal Q with pos_ref[b] @ Q ~ pos[b] (both polygons are centred)
    u, _, vt = np.linalg.svd(f["pos"][b].T @ r["pos"][b])
    q = u @ vt
    assert np.allclose(q @ q.T, np.eye(2), atol=1e-12)
    err = np.abs(f["pos"] @ q - r["pos"]).max()
    print(f"  boundary {len(b)} (shift {sym[0]}, reversed {sym[1]}), run {r['run']} vs "
          f"{ref_run}, max deviation {err:.3g}", flush=True)
    assert err <= 1e-6 * 200.0


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["twogrids6_jacobi", "twogrids6_spring"])
def test_tutte_begin_with_a_cut_vertex_documented_difference(gpu, name):
    """Two 6 x 6 grids sharing one corner (a cut vertex). A graph with a cut
    vertex has several planar embeddings that differ in which face of one
    block holds the other. The reference's Boyer-Myrvold embedding keeps
    both grids in one outer face, whose walk of 40 darts visits the cut
    vertex twice; the left-right embedding built here places the second grid
    inside the square face at the first grid's corner, and its largest face
    has 24 darts (the square's two edges at the corner, the second grid's
    border, the cut vertex twice). That difference is asserted as it
    stands, together with what both layouts are: the boundary is a closed
    walk of the graph, Run converges, every boundary vertex sits on the
    radius-200 polygon and every other vertex at the mean of its
    neighbours, the Tutte fixed point."""
    f = golden("tutte_begin_" + name)
    n = int(f["n"])
    edges = np.asarray(f["edges"])
    print(f"tutte begin {name}", flush=True)
    r = refio.tutte_begin(n, edges, int(f["spring"]), int(f["iters"]), float(f["epsilon"]),
                          exe=DRIVER)
    assert int(f["begin_rc"]) == 0 and r["begin_rc"] == 0
    ref_b, b = [int(x) for x in f["boundary"]], [int(x) for x in r["boundary"]]
    cut = 35
    assert len(ref_b) == 40 and ref_b.count(cut) == 2
    assert len(b) == 24 and b.count(cut) == 2
    es = {(int(u), int(v)) for u, v in edges} | {(int(v), int(u)) for u, v in edges}
    assert all((b[i], b[(i + 1) % len(b)]) in es for i in range(len(b)))
    assert r["run"] < int(f["iters"])
    pos = r["pos"]
    on = np.zeros(n, bool)
    on[b] = True
    assert np.allclose(np.linalg.norm(pos[on], axis=1), 200.0, rtol=1e-12)
    nb = [[] for _ in range(n)]
    for u, v in edges:
        nb[int(u)].append(int(v))
        nb[int(v)].append(int(u))
    inner = [v for v in range(n) if not on[v]]
    dev = max(np.abs(pos[v] - pos[nb[v]].mean(axis=0)).max() for v in inner)
    print(f"  boundary {len(b)} darts (reference {len(ref_b)}), run {r['run']} vs "
          f"{int(f['run'])}, fixed-point deviation {dev:.3g}", flush=True)
    assert dev <= 1e-6 * 200.0


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["grip_rgg300", "grip_grid10", "grip_ba200", "grip_grid16_k4",
                                  "grip_rgg300_d3", "grip_rgg300_d4"])
def test_grip_white_box_stages_build_reference_filtration(gpu, name):
    """makeFirstMISPartition, iterMISFiltration and migrateOneToFinalLayer
    composed as the reference's createMISFiltration composes them, and
    createMISFiltration itself, both give the reference's filtration and
    borders (src/embedders/gvizGRIPInternal.h:22-40)."""
    f = golden(name)
    dim = int(f["dim"]) if "dim" in f.files else 2
    print(f"grip stages {name}", flush=True)
    for layers, borders, filt in refio.grip_stages(int(f["n"]), f["edges"], dim, exe=DRIVER):
        assert layers == len(f["borders"])
        assert np.array_equal(borders, f["borders"])
        assert np.array_equal(filt, f["filtration"])


def _face_walk(lists, u, v):
    """The reference's face walk (src/embedders/gvizPlanarEmbedder.c:220-254):
    (u, v) -> (v, the entry of v's list before u), until (u, v) returns."""
    out, su, sv = [], u, v
    while True:
        out.append(int(v))
        nb = [int(x) for x in lists[v]]
        w = nb[nb.index(u) - 1]
        u, v = v, w
        if (u, v) == (su, sv):
            return out


def _outer_face_case(kind):
    if kind == "grid":
        n, e = G.grid(6, 6)
        c = [2 * 6 + 2, 2 * 6 + 3, 3 * 6 + 3, 3 * 6 + 2]  # an interior cell
        return n, e, c, [(c[0], c[1]), (c[1], c[2]), (c[2], c[3]), (c[3], c[0])]
    n, e, pts, hull = G.delaunay(300, 9)
    from scipy.spatial import Delaunay
    tri = Delaunay(pts)
    onhull = set(int(h) for h in hull)
    t = next(s for s in tri.simplices if not (set(int(x) for x in s) & onhull))
    t = [int(x) for x in t]
    return n, e, t, [(t[0], t[1]), (t[1], t[2]), (t[2], t[0])]


@pytest.mark.gpu
@pytest.mark.parametrize("spring", [False, True])
@pytest.mark.parametrize("kind", ["grid", "delaunay"])
def test_tutte_fix_outer_face(gpu, kind, spring):
    """Begin writes a planar rotation system into the adjacency lists
    (Euler's formula over its faces, the same edges as before) and sets
    planarEmbedded; FixOuterFace pins the face walked from the highlight's
    first edge (src/embedders/gvizTutteEmbedder.c:105-143, 347-373) on the
    radius-200 polygon, and Run relaxes the rest inside it."""
    n, e, hv, he = _outer_face_case(kind)
    print(f"tutte outer face {kind} spring={spring}", flush=True)
    r = refio.tutte_outer_face(n, e, spring=spring, iters=20000, hl_vertices=hv, hl_edges=he,
                               exe=DRIVER)
    assert r["begin_rc"] == 0 and r["planar"] == 1
    lists = r["lists"]
    got = sorted((u, int(v)) for u in range(n) for v in lists[u])
    want = sorted([(int(a), int(b)) for a, b in e] + [(int(b), int(a)) for a, b in e])
    assert got == want
    seen, faces = set(), 0
    for u in range(n):
        for v in lists[u]:
            if (u, int(v)) in seen:
                continue
            faces += 1
            a, b = u, int(v)
            while (a, b) not in seen:
                seen.add((a, b))
                nb = [int(x) for x in lists[b]]
                a, b = b, nb[nb.index(a) - 1]
    assert n - len(e) + faces == 2
    # the highlight shows the entry (a -> b) of each listed edge (a, b); its
    # first edge is the smallest vertex's first shown entry
    u0 = min(hv)
    v0 = next(int(v) for v in lists[u0] if (u0, int(v)) in set(map(tuple, he)))
    walk = _face_walk(lists, u0, v0)
    assert r["fix_rc"] == 0
    assert list(r["boundary"]) == walk
    assert u0 in walk and v0 in walk and 3 <= len(walk) <= max(len(hv), 8)
    pos = r["positions"]
    k = np.arange(len(walk))
    ring = 200.0 * np.stack([np.cos(2 * np.pi * k / len(walk)),
                             np.sin(2 * np.pi * k / len(walk))], axis=1)
    assert np.array_equal(pos[walk], ring)
    inner = np.setdiff1d(np.arange(n), walk)
    assert np.all(np.hypot(pos[inner, 0], pos[inner, 1]) < 200.0)
    if not spring:
        assert r["iterations"] < 20000
        off, adj = G.csr_from_edges(n, e)
        bary = np.array([pos[adj[off[u]:off[u + 1]]].mean(axis=0) for u in inner])
        assert np.abs(bary - pos[inner]).max() < 1e-3
