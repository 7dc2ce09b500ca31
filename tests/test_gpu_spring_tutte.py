"""Spring-driven Tutte embedder (include/gviz/gvizSpringTutteEmbedder.h on
gvizhip_tutte_spring_steps) against the reference
(src/embedders/gvizSpringTutteEmbedder.c).

The fixtures were written by the reference library: FixConvexPolygon,
SeedInterior, Configure, a few Step(dt) calls and one Run, with every step's
max delta, the positions and velocities after the steps and after Run, and
the "springTutte.maxDelta" series. The host layer over the GPU reproduces
them bit for bit in FP64 (the kernel keeps the reference's operation order,
no FMA contraction). At full size the spring dynamics must settle at the
Jacobi sweep's fixed point (both converge to the barycentric embedding).
"""
import os

import numpy as np
import pytest

from conftest import ROOT, golden
import refio
from grapher_amd import graphs as G

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "tests", "host", "api_driver")


@pytest.mark.parametrize("name", ["spring_tutte_grid6", "spring_tutte_delaunay300",
                                  "spring_tutte_delaunay300_under"])
def test_spring_tutte_bit_exact(gpu, name):
    f = golden(name)
    r = refio.spring_tutte(int(f["n"]), f["edges"], f["boundary"], int(f["steps"]),
                           int(f["run_to"]), float(f["radius"]), float(f["dt"]),
                           float(f["stiffness"]), float(f["damping"]), float(f["epsilon"]),
                           exe=DRIVER)
    for key in ("p0", "step", "p1", "v1", "p2", "v2", "series"):
        np.testing.assert_array_equal(r[key], f[key], err_msg=key)
    assert r["run"] == int(f["run"])


def test_spring_settles_at_the_jacobi_fixed_point(gpu):
    """Delaunay, 5,000 points: a lightly damped spring run (dt 0.05, stiffness
    30, damping 0.5: the slow modes settle in ~8,000 steps, where damping
    that is critical for the fast modes would need ~10^5) and the Jacobi
    sweeps from the same seeded start end at the same barycentric embedding
    (within 1e-6 of the polygon radius); the velocities vanish."""
    n, e, _, hull = G.delaunay(5000, 11)
    off, adj = G.csr_from_edges(n, e)
    radius = 200.0
    b = np.zeros(n, np.uint8)
    b[hull] = 1
    pos = np.zeros((n, 2))
    ang = 2 * np.pi * np.arange(len(hull)) / len(hull)
    pos[hull, 0] = radius * np.cos(ang)
    pos[hull, 1] = radius * np.sin(ang)
    pos[b == 0] = pos[hull].mean(0)
    jac = gpu.TutteEngine(off, adj)
    jac.upload(pos, b)
    _, ran = jac.sweeps(60000, 1.0, 1e-11, record=False)
    spr = gpu.TutteEngine(off, adj)
    spr.upload(pos, b)
    spr.set_velocity(None)
    _, ran2 = spr.spring_steps(60000, 0.05, 30.0, 0.5, 1e-11, record=False)
    assert 0 < ran < 60000 and 0 < ran2 < 60000
    pj, ps = jac.positions(), spr.positions()
    assert np.abs(pj - ps).max() <= 1e-6 * radius
    np.testing.assert_array_equal(ps[hull], pos[hull])
    assert np.abs(spr.velocity()).max() <= 1e-6
