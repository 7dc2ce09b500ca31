"""Pins for the oracle functions that no reference fixture covers.

* Pairwise force terms against the reference's own known-answer tests
  (tests/core/gvizVecTests.c:38-117 of the reference: KK, GRIP-FR repulsion
  and attraction, FR attraction and repulsion, the radius gap, overlap
  saturation, and the coincident-pair magnitude k^2 * C_ov), plus closed forms
  for the terms those tests leave out (LinLog attraction through math.log,
  LinLog repulsion, gravity).
* The coincident-pair direction. The reference hashes buffer addresses
  (include/core/gvizVec.h:385-394), which no other program can reproduce; its
  test fixes only that the force is finite with magnitude k^2 * C_ov
  (gvizVecTests.c:107-117). Here: that magnitude, a unit direction, and
  distinct pairs pushed in distinct directions (the property the reference's
  comment states for its hash).
* Layout stress and edge-length energy (oracle/gviz_oracle.c oc_stress,
  oc_edge_energy), which judge every end-to-end run: closed forms on small
  graphs, the minimum over the scale s found numerically with hop distances
  from scipy's own shortest-path routine, and invariance under rotation,
  translation and scale.
"""
import math

import numpy as np
import pytest
import scipy.optimize
import scipy.sparse
import scipy.sparse.csgraph

import pyoracle as O
from grapher_amd import graphs as G

KAT_TOL = 1e-12  # the reference's KAT tolerance (gvizVecTests.c)


def _near(a, b, tol=KAT_TOL):
    assert np.all(np.abs(np.asarray(a) - np.asarray(b)) <= tol), (a, b)


# ----------------------------------------------------------- reference KATs

def test_kat_kk():  # gvizVecTests.c:38-48
    _near(O.kk((0, 0), (3, 4), 5.0), (0.0, 0.0))
    _near(O.kk((0, 0), (3, 4), 10.0), (-1.5, -2.0))


def test_kat_grip_fr_rep():  # gvizVecTests.c:50-56
    _near(O.grip_fr_rep((0, 0), (2, 0), 1.0), (8.0, 0.0))


def test_kat_grip_fr_att():  # gvizVecTests.c:58-64
    _near(O.grip_fr_att((2, 0), (0, 0), 1.0, 0.05), (0.025, 0.0))


def test_kat_fr_att():  # gvizVecTests.c:66-72
    _near(O.fr_att((0, 0), (3, 4), 5.0), (3.0, 4.0))


def test_kat_fr_rep():  # gvizVecTests.c:74-80
    _near(O.fr_rep((0, 0), (2, 0), 1.0, 0.0, 100.0), (-0.5, 0.0))


def test_kat_fr_rep_radius_gap():  # gvizVecTests.c:82-91
    _near(O.fr_rep((0, 0), (2, 0), 1.0, 1.0, 100.0), (-1.0, 0.0))


def test_kat_fr_rep_overlap_saturates():  # gvizVecTests.c:93-102
    _near(O.fr_rep((0, 0), (1, 0), 1.0, 2.0, 10.0), (-10.0, 0.0))


def test_kat_fr_rep_coincident_magnitude():  # gvizVecTests.c:104-117
    k, cov = 2.0, 10.0
    acc = O.fr_rep((0, 0), (0, 0), k, 1.0, cov)
    assert np.all(np.isfinite(acc))
    assert abs(math.hypot(*acc) - k * k * cov) <= 1e-9


# ------------------------------------------- closed forms beyond the KATs

def test_linlog_att_closed_form():
    # magnitude log(1 + d) along (u - v): (0,0) -> (3,4), d = 5
    _near(O.linlog_att((0, 0), (3, 4)), (math.log(6.0) * 0.6, math.log(6.0) * 0.8))
    # below 1e-9 apart no force (gvizVec.h:493)
    _near(O.linlog_att((1, 1), (1 + 5e-10, 1)), (0.0, 0.0), 0.0)


def test_linlog_rep_closed_forms():
    # masses 2 and 3 at distance 4 with radius sum 1: 6 / 3 = 2 away from u
    _near(O.linlog_rep((0, 0), (0, 4), 2.0, 3.0, 1.0, 100.0), (0.0, -2.0))
    # overlap: 6 * C_ov
    _near(O.linlog_rep((0, 0), (0, 1), 2.0, 3.0, 2.0, 7.0), (0.0, -42.0))
    acc = O.linlog_rep((5, 5), (5, 5), 2.0, 3.0, 0.0, 7.0)
    assert abs(math.hypot(*acc) - 42.0) <= 1e-9


def test_gravity_closed_form():
    _near(O.gravity((3, 4), 2.0), (-1.2, -1.6))
    _near(O.gravity((1e-10, 0), 2.0), (0.0, 0.0), 0.0)


def test_fr_rep_inverse_distance_law():
    # k^2 / d for d in a range of scales, direction (u - v) / d, pushing v away
    rng = np.random.default_rng(3)
    for _ in range(200):
        v = rng.uniform(-100, 100, 2)
        d = 10.0 ** rng.uniform(-6, 4)
        ang = rng.uniform(0, 2 * math.pi)
        u = v + d * np.array([math.cos(ang), math.sin(ang)])
        k = 10.0 ** rng.uniform(-1, 2)
        got = O.fr_rep(v, u, k, 0.0, 100.0)
        dist = float(np.hypot(*(u - v)))
        want = -(k * k / dist) * (u - v) / dist
        assert np.allclose(got, want, rtol=1e-12, atol=0.0)


# ------------------------------------------------ coincident-pair direction

def test_coincident_direction_is_unit_and_spreads():
    angles = []
    for a in range(40):
        for b in range(40):
            if a == b:
                continue
            ux, uy = O.coincident_dir(a, b)
            assert abs(math.hypot(ux, uy) - 1.0) <= 1e-15
            angles.append(math.atan2(uy, ux))
    angles = np.array(angles)
    # distinct pairs get distinct directions, spread over the whole circle
    assert len(np.unique(np.round(angles, 9))) >= 0.99 * len(angles)
    hist, _ = np.histogram(angles, bins=8, range=(-math.pi, math.pi))
    assert hist.min() > 0.5 * len(angles) / 8


def test_coincident_force_magnitude_in_both_models():
    for self_id, other_id in [(0, 1), (7, 3), (123456, 98765)]:
        fr = O.fr_rep((2, 2), (2, 2), 3.0, 0.5, 11.0, self_id, other_id)
        assert abs(math.hypot(*fr) - 9.0 * 11.0) <= 1e-9
        ll = O.linlog_rep((2, 2), (2, 2), 4.0, 5.0, 0.0, 11.0, self_id, other_id)
        assert abs(math.hypot(*ll) - 20.0 * 11.0) <= 1e-9


# ----------------------------------------------------- stress and energy

def _graph(n, edges):
    off, adj = G.csr_from_edges(n, np.asarray(edges, dtype=np.int64).reshape(-1, 2))
    return O.Graph(off, adj)


def _hops(n, edges):
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    A = scipy.sparse.coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n))
    return scipy.sparse.csgraph.shortest_path(A, directed=False, unweighted=True)


def _stress_by_minimisation(pos, hops, L, sources):
    """min over s of sum over ordered reachable pairs (s D / (L d) - 1)^2,
    divided by the pair count, with s found by a numerical line search."""
    D, d = [], []
    for s in sources:
        for j in range(len(pos)):
            if j == s or not np.isfinite(hops[s, j]):
                continue
            D.append(np.hypot(*(pos[j] - pos[s])))
            d.append(L * hops[s, j])
    D, d = np.array(D), np.array(d)
    f = lambda sc: float(np.sum((sc * D / d - 1.0) ** 2))
    hi = 10.0 * float(np.max(d / np.maximum(D, 1e-300)))
    r = scipy.optimize.minimize_scalar(f, bounds=(0.0, hi), method="bounded",
                                       options={"xatol": 1e-14 * hi, "maxiter": 2000})
    return r.fun / len(D)


def test_stress_of_path_and_cycle_at_unit_spacing_is_zero():
    n = 12
    path = [(i, i + 1) for i in range(n - 1)]
    pos = np.stack([np.arange(n) * 3.5, np.zeros(n)], 1)
    assert abs(O.stress(_graph(n, path), pos, 10.0)) <= 1e-14
    assert abs(O.edge_energy(_graph(n, path), pos)) <= 1e-28
    # a cycle on a regular polygon: equal edges, energy 0 (stress is not 0:
    # chords are shorter than hop distances)
    cyc = path + [(n - 1, 0)]
    t = 2 * np.pi * np.arange(n) / n
    poly = np.stack([np.cos(t), np.sin(t)], 1) * 7.0
    assert abs(O.edge_energy(_graph(n, cyc), poly)) <= 1e-26


def test_stress_of_square_c4_closed_form():
    # from each vertex: two neighbours at D = a, d = L and the diagonal at
    # D = a sqrt 2, d = 2 L; with r = a / L, per source
    # A = r (2 + 2^-1/2), B = r^2 (2 + 1/2), C = 3 and
    # stress = 1 - A^2 / (B C) = (3 - 2 sqrt 2) / 7.5 for any side a and L.
    want = (3.0 - 2.0 * math.sqrt(2.0)) / 7.5
    g = _graph(4, [(0, 1), (1, 2), (2, 3), (3, 0)])
    for a, L in [(1.0, 1.0), (3.0, 10.0), (1e-3, 7.0)]:
        sq = np.array([[0, 0], [a, 0], [a, a], [0, a]], dtype=float)
        assert abs(O.stress(g, sq, L) - want) <= 1e-15
        # one source is enough by symmetry
        assert abs(O.stress(g, sq, L, sources=[2]) - want) <= 1e-15
        assert abs(O.edge_energy(g, sq)) <= 1e-30


def test_edge_energy_closed_forms():
    # star with edge lengths 1, 2, 3: mean 2, residuals -1/2, 0, 1/2 -> 1/6
    star = np.array([[0, 0], [1, 0], [0, 2], [-3, 0]], dtype=float)
    g = _graph(4, [(0, 1), (0, 2), (0, 3)])
    assert abs(O.edge_energy(g, star) - 1.0 / 6.0) <= 1e-15
    # path with lengths 1 and 3: mean 2 -> (1/4 + 1/4) / 2
    g = _graph(3, [(0, 1), (1, 2)])
    assert abs(O.edge_energy(g, np.array([[0, 0], [1, 0], [4, 0]], float)) - 0.25) <= 1e-15


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_stress_equals_numerical_minimum_over_scale(seed):
    rng = np.random.default_rng(seed)
    n = 30
    _, edges = G.erdos_renyi(n, 3.0, seed=seed)
    pos = rng.uniform(-5, 5, (n, 2))
    hops = _hops(n, edges)
    g = _graph(n, edges)
    L = 10.0
    want = _stress_by_minimisation(pos, hops, L, range(n))
    assert abs(O.stress(g, pos, L) - want) <= 1e-9
    src = [0, 5, 17]
    want_s = _stress_by_minimisation(pos, hops, L, src)
    assert abs(O.stress(g, pos, L, sources=src) - want_s) <= 1e-9


@pytest.mark.parametrize("seed", [0, 1])
def test_stress_and_energy_are_similarity_invariant(seed):
    n, edges = G.grid(8, 9)
    g = _graph(n, edges)
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, (n, 2))
    s0, e0 = O.stress(g, pos, 10.0), O.edge_energy(g, pos)
    assert s0 > 0.01 and e0 > 0.01
    for _ in range(5):
        ang = rng.uniform(0, 2 * math.pi)
        R = np.array([[math.cos(ang), -math.sin(ang)], [math.sin(ang), math.cos(ang)]])
        if rng.uniform() < 0.5:
            R = R @ np.diag([1.0, -1.0])  # a reflection too
        q = (pos @ R.T) * 10.0 ** rng.uniform(-3, 3) + rng.uniform(-1e3, 1e3, 2)
        assert abs(O.stress(g, q, 10.0) - s0) <= 1e-9 * s0
        assert abs(O.edge_energy(g, q) - e0) <= 1e-9 * e0
    # the edge length L only rescales s*, never the stress
    assert abs(O.stress(g, pos, 0.37) - s0) <= 1e-12


def test_stress_detects_a_distorted_layout():
    # a correct grid layout has lower stress than the same grid with two
    # vertices swapped: a mistake in the pair set or hop distances that made
    # stress insensitive to the layout fails here
    n, edges = G.grid(10, 10)
    g = _graph(n, edges)
    ij = np.array([(i, j) for i in range(10) for j in range(10)], dtype=float)
    good = ij.copy()
    bad = ij.copy()
    bad[[0, 99]] = bad[[99, 0]]
    assert O.stress(g, good, 10.0) < 0.5 * O.stress(g, bad, 10.0)
    assert O.edge_energy(g, good) <= 1e-28 < O.edge_energy(g, bad)


# --------------------------------------------- per-vertex term magnitude sum

def _pair_terms_bruteforce(g, pos, p):
    """sum_j |f_ij| from the pinned pair functions, all pairs (exact mode)."""
    n = g.n
    deg = g.degree
    mass = O.vertex_mass(p.model, deg)
    rad = (p.radiusBase * (1.0 + p.radiusPerDegree * np.sqrt(deg))
           if p.preventOverlap else np.zeros(n))
    out = np.zeros(n)
    for i in range(n):
        for e in range(g.offsets[i], g.offsets[i + 1]):
            u = g.adj[e]
            t = (O.linlog_att(pos[i], pos[u]) if p.model == O.MODEL_LINLOG
                 else O.fr_att(pos[i], pos[u], p.edgeLength))
            out[i] += math.hypot(*t)
        for j in range(n):
            if j == i:
                continue
            if p.model == O.MODEL_LINLOG:
                t = O.linlog_rep(pos[i], pos[j], mass[i], mass[j], rad[i] + rad[j],
                                 p.overlapConstant, i, j)
            else:
                t = O.fr_rep(pos[i], pos[j], p.edgeLength, rad[i] + rad[j],
                             p.overlapConstant, i, j) * float(int(mass[j]))
            out[i] += math.hypot(*t)
        out[i] += math.hypot(*O.gravity(pos[i], p.gravityK * (deg[i] + 1)))
    return out


@pytest.mark.parametrize("model", [O.MODEL_FR, O.MODEL_LINLOG])
def test_force_term_sums(model):
    n, edges = G.barabasi_albert(60, 2, 5)
    g = _graph(n, edges)
    pos = np.random.default_rng(4).uniform(-3, 3, (n, 2))
    pos[7] = pos[8]  # one coincident pair
    p = O.params(model=model, barnes_hut=False, prevent_overlap=True,
                 radius_base=0.05, radius_per_degree=0.5, gravity=0.1)
    disp, ts = O.force_term_sums(g, pos, p)
    ref = O.force_compute(g, pos, p)
    assert np.array_equal(disp, ref["disp"])  # the evaluation itself is unchanged
    np.testing.assert_allclose(ts, _pair_terms_bruteforce(g, pos, p), rtol=1e-12)
    assert np.all(ts >= np.linalg.norm(disp, axis=1) * (1 - 1e-12))
    # Barnes-Hut: same disp as force_compute, and a sum over fewer terms
    p.barnesHut = 1
    p.theta = 0.8
    disp_bh, ts_bh, inter = O.force_term_sums(g, pos, p, interactions=True)
    ref_bh = O.force_compute(g, pos, p)
    assert np.array_equal(disp_bh, ref_bh["disp"])
    assert np.array_equal(inter, ref_bh["interactions"])
    assert np.all(ts_bh >= np.linalg.norm(disp_bh, axis=1) * (1 - 1e-12))
