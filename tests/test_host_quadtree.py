"""gvizQuadtree* through the host layer (libgviz.so): the reference's host
tree as a view of keys built on the GPU.

- Standalone trees equal the reference library's trees bit for bit (node
  shape, leaf contents in order, mass, running-mean centre of mass, half
  size) on the fixtures it wrote (uniform points, coincident points with
  overflow leaves, clustered points with nodesPerCell 3), and agree with the
  oracle's restatement of src/ds/gvizQuadtree.c.
- The reference's own quadtree unit cases (tests/ds/gvizQuadtreeTests.c)
  pass, restated in tests/host/api_driver.c (quadtree_cases).
- gvizForceEmbedderState.quadtree is the tree of the last step's starting
  positions (src/embedders/gvizForceEmbedder.c:430-452) with the state's
  masses, and stays empty with Barnes-Hut off.
"""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O
import refio
from conftest import ROOT, golden
from grapher_amd import graphs as G

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "tests", "host", "api_driver")
TREES = ["quadtree_uniform300", "quadtree_dup_npc1", "quadtree_clustered_npc3"]


def _same_tree(got, want):
    for k in ("nodes", "leaf_depth", "leaf_keyHi", "leaf_keyLo"):
        assert np.array_equal(got[k], want[k]), k


def _matches_oracle(tree, points, masses, npc):
    """Every node of the dump has the oracle's node at the same path, with
    equal mass, centre of mass, half size and point count; the counts of
    nodes agree."""
    o = O.quadtree(points, masses, npc)["nodes"]
    by_key = {(int(r["depth"]), int(r["keyHi"]), int(r["keyLo"])): r for r in o}
    nodes = tree["nodes"]
    assert len(nodes) == len(o)
    for r in nodes:
        ref = by_key[(int(r["depth"]), int(r["keyHi"]), int(r["keyLo"]))]
        assert r["mass"] == ref["mass"]
        assert r["comX"] == ref["comX"] and r["comY"] == ref["comY"]
        assert r["halfSize"] == ref["halfSize"]
        if r["leaf"]:
            assert r["count"] == ref["ptCount"]


@pytest.mark.parametrize("name", TREES)
def test_standalone_tree_equals_reference(gpu, name):
    f = golden(name)
    print(f"quadtree host view {name}", flush=True)
    got = refio.quadtree(f["points"], f["masses"], int(f["npc"]), exe=DRIVER)
    _same_tree(got, f)
    _matches_oracle(got, f["points"], f["masses"], int(f["npc"]))


def test_reference_quadtree_unit_cases(gpu, tmp_path):
    src = tmp_path / "in.bin"
    src.write_bytes(b"")
    out = tmp_path / "out.bin"
    r = subprocess.run([DRIVER, "quadtree_cases", str(src), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.parametrize("npc", [1, 2, 5])
def test_standalone_tree_weighted_and_larger(gpu, npc):
    """Uneven masses (the running mean's rounding shows) and 20k points."""
    rng = np.random.default_rng(npc)
    pts = rng.normal(0.0, 50.0, (20000, 2))
    pts[:40] = pts[0]  # a cluster of coincident points: an overflow leaf
    ms = rng.uniform(0.5, 7.0, 20000)
    got = refio.quadtree(pts, ms, npc, exe=DRIVER)
    _matches_oracle(got, pts, ms, npc)
    o = O.quadtree(pts, ms, npc)
    assert np.array_equal(got["leaf_depth"], o["leaf_depth"])
    assert np.array_equal(got["leaf_keyHi"], o["leaf_keyHi"])
    assert np.array_equal(got["leaf_keyLo"], o["leaf_keyLo"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("steps", [0, 1, 3])
def test_force_state_tree_is_last_steps_tree(gpu, monkeypatch, precision, steps):
    n, e = G.barabasi_albert(400, 3, seed=2)
    pos = G.uniform_positions(n, seed=5, extent=30.0)
    monkeypatch.setenv("GVIZ_TEST_PRECISION", precision)
    print(f"force state quadtree {precision} steps={steps}", flush=True)
    before, has_root, ready, pts, ms, tree = refio.force_quadtree(
        n, e, pos, model=1, barnes_hut=1, npc=1, steps=steps, theta=0.7, exe=DRIVER)
    assert has_root and ready == 1
    assert np.array_equal(pts, before)
    off, _ = G.csr_from_edges(n, e)
    assert np.array_equal(ms, 1.0 + np.diff(off))  # LinLog masses
    want = refio.quadtree(before, ms, 1, exe=DRIVER)
    _same_tree(tree, want)
    _matches_oracle(tree, before, ms, 1)


def test_force_state_tree_empty_without_barnes_hut(gpu):
    n, e = G.grid(8, 8)
    pos = G.uniform_positions(n, seed=1, extent=10.0)
    _, has_root, ready, _, _, tree = refio.force_quadtree(
        n, e, pos, model=0, barnes_hut=0, npc=1, steps=2, theta=0.7, exe=DRIVER)
    assert not has_root and ready == 0
    assert len(tree["nodes"]) == 0
