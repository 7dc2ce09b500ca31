"""The C11 host layer (libgviz.so, include/gviz/*.h): the reference's embedder
API on the gfx950 engine.

CPU tests: the library exports every declared function; host-side unit
checks (graph/CSR, actions, stats, draw mask, checkpoints, configuration
defaults and setters, K policies) in tests/host/api_driver.c; without a GPU
the embedders fail loudly instead of computing anything on the CPU.

GPU tests replay the reference's own API runs (fixtures written by the
reference library: Init, Configure*, Begin(seed), Run, Step, GRIP Embed,
Tutte polygon + seeding + sweeps) through this library and compare every
output: Begin's random placement and the Tutte sweeps bit for bit, the FP64
force and GRIP trajectories to 1e-9 relative.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden
import refio
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
        names.update(re.findall(r"\b(gviz[A-Z][A-Za-z0-9]*)\s*\(", text))
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


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["api_force_linlog_ba400", "api_force_fr_grid20"])
@pytest.mark.parametrize("resident", ["0", "1"])
def test_force_api_fp64_matches_reference(gpu, name, resident):
    f = golden(name)
    print(f"host api force {name} resident={resident}", flush=True)
    r = _api_force(f, {"GVIZ_TEST_PRECISION": "fp64", "GVIZ_TEST_RESIDENT": resident})
    assert r["box_extent"] == float(f["box_extent"])
    assert r["begin_rc"] == 0
    # Begin's placement is host rand_r, identical to the reference's
    assert np.array_equal(r["pos_begin"], f["pos_begin"])
    assert r["run_rc"] == int(f["run_rc"])
    for s in SERIES:
        np.testing.assert_allclose(r["series_" + s], f["series_" + s], rtol=1e-9,
                                   atol=1e-12, err_msg=s)
    act = _active(f)
    scale = np.abs(f["pos_run"][act]).max()
    assert np.abs(r["pos_run"] - f["pos_run"])[act].max() <= 1e-9 * scale
    assert np.all(r["pos_run"][~act] == 0.0)
    assert np.abs(r["pos_step"] - f["pos_step"])[act].max() <= 1e-9 * scale
    np.testing.assert_allclose(r["step_ret"], f["step_ret"], rtol=1e-9)
    np.testing.assert_allclose([r["speed"], r["efficiency"]],
                               [f["speed"], f["efficiency"]], rtol=1e-9)
    assert r["iteration"] == int(f["iteration"])


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
    f = golden(name)
    print(f"host api grip {name}", flush=True)
    r = refio.api_grip(int(f["n"]), f["edges"], int(f["kcap"]), int(f["placement_k"]),
                       int(f["refinement_k"]), int(f["policy"]), exe=DRIVER)
    assert r["embed_rc"] == 0
    assert np.array_equal(r["borders"], f["borders"])
    assert np.array_equal(r["visible"], f["visible"])
    for s in ("heat", "meanDisp", "maxDisp", "meanForce"):
        assert r["series_" + s].shape == f["series_" + s].shape, s
        np.testing.assert_allclose(r["series_" + s], f["series_" + s], rtol=1e-9,
                                   atol=1e-12, err_msg=s)
    scale = np.abs(f["pos"]).max()
    assert np.abs(r["pos"] - f["pos"]).max() <= 1e-9 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["tutte_grid6_jacobi", "tutte_delaunay300_jacobi",
                                  "tutte_delaunay300_dt60", "tutte_k4"])
def test_tutte_api_bit_exact(gpu, name):
    f = golden(name)
    print(f"host api tutte {name}", flush=True)
    p0, deltas, p1 = refio.tutte(int(f["n"]), f["edges"], f["boundary"], int(f["iters"]),
                                 float(f["radius"]), float(f["dt"]), False, exe=DRIVER)
    assert np.array_equal(p0, f["p0"])
    assert np.array_equal(deltas, f["deltas"])
    assert np.array_equal(p1, f["p1"])
