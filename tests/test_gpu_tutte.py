"""HIP Tutte Jacobi sweep against the reference's fixtures (FP64: bit-exact)
and the oracle; FP32 within 1e-5 relative of the layout span."""
import numpy as np
import pytest

import pyoracle as O
from conftest import golden
from grapher_amd import graphs as G

pytestmark = pytest.mark.gpu

JACOBI = ["tutte_grid6_jacobi", "tutte_delaunay300_jacobi", "tutte_delaunay300_dt60",
          "tutte_k4"]


def _setup(native, f, precision):
    n = int(f["n"])
    off, adj = G.csr_from_edges(n, f["edges"])
    t = native.TutteEngine(off, adj, precision)
    isb = np.zeros(n, np.uint8)
    isb[f["boundary"]] = 1
    t.upload(f["p0"], isb)
    return t


@pytest.mark.parametrize("name", JACOBI)
def test_fp64_sweeps_bit_exact(gpu, name):
    f = golden(name)
    t = _setup(gpu, f, "fp64")
    alpha = min(5.0 * float(f["dt"]), 1.0)
    deltas, r = t.sweeps(int(f["iters"]), alpha)
    assert r == int(f["iters"])
    assert np.array_equal(deltas, f["deltas"])
    assert np.array_equal(t.positions(), f["p1"])


@pytest.mark.parametrize("name", JACOBI)
def test_fp32_sweeps_close(gpu, name):
    f = golden(name)
    t = _setup(gpu, f, "fp32")
    alpha = min(5.0 * float(f["dt"]), 1.0)
    _, r = t.sweeps(int(f["iters"]), alpha)
    span = np.ptp(f["p1"])
    assert np.abs(t.positions() - f["p1"]).max() <= 1e-5 * span


def test_sweeps_in_pieces_equal_one_call(gpu):
    f = golden("tutte_delaunay300_jacobi")
    a = _setup(gpu, f, "fp64")
    b = _setup(gpu, f, "fp64")
    da, _ = a.sweeps(200)
    parts = []
    for k in (1, 2, 3, 57, 137):
        d, _ = b.sweeps(k)
        parts.append(d)
    assert np.array_equal(np.concatenate(parts), da)
    assert np.array_equal(a.positions(), b.positions())


def test_converges_to_laplacian_solution_and_stops(gpu):
    """Fixed point = solution of the Laplacian system with the hull pinned;
    Run-style stop once maxDelta < epsilon."""
    n, e, pts, hull = G.delaunay(2000, 5)
    off, adj = G.csr_from_edges(n, e)
    t = gpu.TutteEngine(off, adj, "fp64")
    isb = np.zeros(n, np.uint8)
    isb[hull] = 1
    p0 = np.zeros((n, 2))
    p0[hull] = O.tutte_polygon(len(hull), 200.0)
    p0[~isb.astype(bool)] = p0[hull].mean(axis=0)
    t.upload(p0, isb)
    deltas, r = t.sweeps(100000, 1.0, 1e-9)
    assert r < 100000 and deltas[-1] < 1e-9 and np.all(deltas[:-1] >= 1e-9)
    # dense solve of the harmonic system
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    deg = np.diff(off)
    rows = np.repeat(np.arange(n), deg)
    A = sp.csr_matrix((np.ones(len(adj)), (rows, adj)), shape=(n, n))
    L = sp.diags(deg.astype(float)) - A
    inner = np.where(isb == 0)[0]
    Lii = L[inner][:, inner].tocsc()
    Lib = L[inner][:, hull]
    x = spla.spsolve(Lii, -Lib @ p0[hull])
    got = t.positions()
    assert np.abs(got[inner] - x).max() <= 1e-5 * 200.0
    assert np.array_equal(got[hull], p0[hull])


def test_gs_sweeps_bit_exact_against_reference(gpu):
    """Gauss-Seidel sweeps (useGaussSeidel) against the reference's run."""
    f = golden("tutte_grid6_gs")
    t = _setup(gpu, f, "fp64")
    alpha = min(5.0 * float(f["dt"]), 1.0)
    deltas, r = t.gs_sweeps(int(f["iters"]), alpha)
    assert r == int(f["iters"])
    assert np.array_equal(deltas, f["deltas"])
    assert np.array_equal(t.positions(), f["p1"])


def _gs_case(kind):
    if kind == "delaunay":
        n, e, _, hull = G.delaunay(20000, 5)
        boundary = hull
    else:  # row-major grid: dependency chains run the length of the grid
        n, e = G.grid(70, 90)
        boundary = np.array([j for j in range(90)] + [i * 90 + 89 for i in range(1, 70)] +
                            [69 * 90 + j for j in range(88, -1, -1)] +
                            [i * 90 for i in range(68, 0, -1)], dtype=np.int64)
    off, adj = G.csr_from_edges(n, e)
    isb = np.zeros(n, np.uint8)
    isb[boundary] = 1
    ring = O.tutte_polygon(len(boundary), 200.0)
    pos = np.zeros((n, 2))
    pos[:] = ring.mean(0)
    pos[boundary] = ring
    pos[isb == 0] += np.random.default_rng(1).normal(0, 5.0, (int((isb == 0).sum()), 2))
    return off, adj, isb, pos


@pytest.mark.parametrize("alpha", [1.0, 0.35])
@pytest.mark.parametrize("kind", ["delaunay", "grid"])
def test_gs_sweeps_bit_exact_against_oracle(gpu, kind, alpha):
    off, adj, isb, pos = _gs_case(kind)
    g = O.Graph(off, adj)
    t = gpu.TutteEngine(off, adj, "fp64")
    t.upload(pos, isb)
    ref = pos.copy()
    scratch = np.zeros_like(ref)
    want = np.array([O.tutte_step(g, isb, ref, scratch, rate=alpha, dt=1.0,
                                  gauss_seidel=True) for _ in range(25)])
    got, r = t.gs_sweeps(25, alpha)
    assert r == 25
    assert np.array_equal(got, want)
    assert np.array_equal(t.positions(), ref)
    # a Jacobi sweep after the Gauss-Seidel ones continues from their result
    t.sweeps(1, alpha)
    O.tutte_step(g, isb, ref, scratch, rate=alpha, dt=1.0)
    assert np.array_equal(t.positions(), ref)


def test_gs_sweeps_stop_at_epsilon_like_the_reference_run(gpu):
    off, adj, isb, pos = _gs_case("delaunay")
    g = O.Graph(off, adj)
    ref = pos.copy()
    scratch = np.zeros_like(ref)
    eps = 0.5
    count = 0
    while count < 400:
        count += 1
        if O.tutte_step(g, isb, ref, scratch, gauss_seidel=True) < eps:
            break
    t = gpu.TutteEngine(off, adj, "fp64")
    t.upload(pos, isb)
    _, r = t.gs_sweeps(400, 1.0, eps)
    assert r == count < 400
    assert np.array_equal(t.positions(), ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_runs_of_sweeps_equal_single_sweeps(gpu, precision):
    """A run of Jacobi sweeps finishes each sweep inside the next one (one
    launch per sweep, the last finished after the run); single sweeps keep
    their own finish. On a 5,000-point Delaunay graph: runs from both buffer
    parities, of even and odd lengths, and a run that stops on epsilon give
    bit for bit the positions, the maxDelta records and the sweep counts of
    the same sweeps taken one at a time."""
    n, e, _, hull = G.delaunay(5000, 4)
    off, adj = G.csr_from_edges(n, e)
    isb = np.zeros(n, np.uint8)
    isb[hull] = 1
    ang = 2.0 * np.pi * np.arange(len(hull)) / len(hull)
    p0 = np.zeros((n, 2))
    p0[hull] = 200.0 * np.stack([np.cos(ang), np.sin(ang)], 1)
    p0[~isb.astype(bool)] = p0[hull].mean(0)
    a = gpu.TutteEngine(off, adj, precision)
    b = gpu.TutteEngine(off, adj, precision)
    a.upload(p0, isb)
    b.upload(p0, isb)
    for length in (1, 6, 7, 40, 33):
        da, ra = a.sweeps(length)
        db = []
        for _ in range(length):
            d, r = b.sweeps(1)
            assert r == 1
            db.extend(d)
        assert ra == length
        assert np.array_equal(da, np.asarray(db)), length
        assert np.array_equal(a.positions(), b.positions()), length
    # a run stopping on epsilon: the same sweep count, records and layout
    eps = float(np.asarray(db)[-1]) * 0.5
    da, ra = a.sweeps(10000, 1.0, eps)
    count, db = 0, []
    while count < 10000:
        d, r = b.sweeps(1, 1.0, eps)
        count += r
        db.extend(d)
        if d[-1] < eps:
            break
    assert 0 < ra < 10000 and ra == count
    assert np.array_equal(da, np.asarray(db))
    assert np.array_equal(a.positions(), b.positions())
