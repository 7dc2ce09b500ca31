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
