"""The reference's pairwise known-answer tests (tests/core/gvizVecTests.c:
66-117 of the reference) through the device engine: two vertices joined by
one edge, one FP64 force evaluation, expected forces written in closed form.

For vertex 0 at v and vertex 1 at u the engine's disp is attraction +
repulsion (+ gravity): FR attraction (d/L)(u - v), FR repulsion -L^2/gap along
(u - v)/d, saturating at L^2 C_ov when gap <= 0 (gvizForceEmbedder.c:76-126,
gvizVec.h:320-427 of the reference). The same cases run with Barnes-Hut on
and off: with two points every cell is a leaf, so both are exact.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EDGE = (np.array([0, 1, 2], dtype=np.int64), np.array([1, 0], dtype=np.int64))


def _eval(native, pos, bh, **kw):
    eng = native.ForceEngine(*EDGE, precision="fp64", barnes_hut=bh, **kw)
    eng.set_positions(np.asarray(pos, dtype=np.float64))
    out = eng.evaluate(interactions=True)
    eng.close()
    return out


def _near(a, b, tol=1e-12):
    a, b = np.asarray(a, float), np.asarray(b, float)
    assert np.all(np.abs(a - b) <= tol * np.maximum(1.0, np.abs(b))), (a, b)


@pytest.mark.parametrize("bh", [True, False])
def test_fr_attraction_and_repulsion(gpu, bh):
    # KATs :66-80 combined: att 2 (2, 0) = (4, 0), rep -(1/2)(1, 0)
    o = _eval(gpu, [[0, 0], [2, 0]], bh, edge_length=1.0)
    _near(o["disp"][0], (3.5, 0.0))
    _near(o["disp"][1], (-3.5, 0.0))
    _near(o["attMag"], (4.0, 4.0))
    _near(o["repMag"], (0.5, 0.5))
    assert list(o["interactions"]) == [1, 1]


@pytest.mark.parametrize("bh", [True, False])
def test_fr_radius_gap(gpu, bh):
    # KAT :82-91: radius sum 1 (0.5 each) turns d = 2 into gap 1 -> rep 1
    o = _eval(gpu, [[0, 0], [2, 0]], bh, edge_length=1.0, prevent_overlap=True,
              radius_base=0.5, radius_per_degree=0.0)
    _near(o["repMag"], (1.0, 1.0))
    _near(o["disp"][0], (4.0 - 1.0, 0.0))


@pytest.mark.parametrize("bh", [True, False])
def test_fr_overlap_saturates(gpu, bh):
    # KAT :93-102: d = 1, radius sum 2, C_ov = 10 -> rep 10
    o = _eval(gpu, [[0, 0], [1, 0]], bh, edge_length=1.0, prevent_overlap=True,
              radius_base=1.0, radius_per_degree=0.0, overlap_constant=10.0)
    _near(o["repMag"], (10.0, 10.0))
    _near(o["disp"][0], (1.0 - 10.0, 0.0))


@pytest.mark.parametrize("overlap", [True, False])
@pytest.mark.parametrize("bh", [True, False])
def test_coincident_pair_magnitude(gpu, bh, overlap):
    # KAT :104-117: coincident points, k = 2, C_ov = 10 -> |rep| = 40, finite;
    # attraction is skipped below 1e-9 apart
    kw = dict(radius_base=0.5, radius_per_degree=0.0) if overlap else {}
    o = _eval(gpu, [[3, -1], [3, -1]], bh, edge_length=2.0, prevent_overlap=overlap,
              overlap_constant=10.0, **kw)
    assert np.all(np.isfinite(o["disp"]))
    _near(o["attMag"], (0.0, 0.0), 0.0)
    _near(np.linalg.norm(o["disp"], axis=1), (40.0, 40.0), 1e-9)
    _near(o["repMag"], (40.0, 40.0), 1e-9)


@pytest.mark.parametrize("bh", [True, False])
def test_linlog_pair(gpu, bh):
    # masses 1 + deg = 2; d = 4, radius sum 1: rep 4 / 3; att log(5) along +y
    o = _eval(gpu, [[0, 0], [0, 4]], bh, model=gpu.MODEL_LINLOG, edge_length=1.0,
              prevent_overlap=True, radius_base=0.5, radius_per_degree=0.0)
    _near(o["disp"][0], (0.0, math.log(5.0) - 4.0 / 3.0))
    _near(o["disp"][1], (0.0, -(math.log(5.0) - 4.0 / 3.0)))
    _near(o["attMag"], (math.log(5.0),) * 2)
    _near(o["repMag"], (4.0 / 3.0,) * 2)


@pytest.mark.parametrize("bh", [True, False])
def test_gravity_and_struct_force(gpu, bh):
    # v = (3, 4), u = (-3, -4), d = 10, L = 1: att 10 (-6, -8), rep -(1/10)(-0.6, -0.8),
    # gravity gravityK (deg + 1) = 1 toward the origin; structForce leaves gravity out
    o = _eval(gpu, [[3, 4], [-3, -4]], bh, edge_length=1.0, gravity=0.5)
    att = np.array([-60.0, -80.0])
    rep = np.array([0.06, 0.08])
    grav = -np.array([3.0, 4.0]) / 5.0
    _near(o["structForce"][0], att + rep)
    _near(o["disp"][0], att + rep + grav)
    _near(o["disp"][1], -(att + rep + grav))
