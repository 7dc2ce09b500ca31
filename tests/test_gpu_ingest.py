"""Graphs read from .edges files go through the engines: the file is parsed
by this library's loader (gvizGraphLoadEdgesFileCSR, checked against the
reference's loader in tests/test_graph_loader.py), and the CSR it gives is
laid out on the device and compared with the FP64 oracle on the same graph
and positions: force evaluations (FP64 1e-11 normwise with equal interaction
sets, MIXED 1e-5 normwise and per vertex), eight FP64 steps (the first three to 1e-9), and the GRIP
MIS filtration of the largest component (bit-exact)."""
import numpy as np
import pytest

import pyoracle as O
from conftest import golden
from grapher_amd import edges_file
from grapher_amd import graphs as G

pytestmark = pytest.mark.gpu


def _loaded(tmp_path, name):
    z = golden("edges_files")
    i = [str(s) for s in z["names"]].index(name)
    p = tmp_path / f"{name}.edges"
    p.write_bytes(z[f"c{i}_text"].tobytes())
    directed, skip = (int(v) for v in z[f"c{i}_opts"])
    off, adj = edges_file.load_csr(str(p), directed=bool(directed), skip_header=bool(skip))
    np.testing.assert_array_equal(off, z[f"c{i}_offsets"])
    np.testing.assert_array_equal(adj, z[f"c{i}_adj"])
    return off, adj


@pytest.mark.parametrize("name,model", [("random2000", 0), ("random2000", 1), ("sample", 0)])
def test_loaded_graph_force_engine_matches_oracle(gpu, tmp_path, name, model):
    off, adj = _loaded(tmp_path, name)
    n = len(off) - 1
    g = O.Graph(off, adj)
    kw = dict(model=model, theta=0.7, gravity=0.05, prevent_overlap=True,
              radius_base=0.1, radius_per_degree=1.0)
    p = O.params(model=model, theta=0.7, gravity=0.05, prevent_overlap=True,
                 radius_base=0.1, radius_per_degree=1.0)
    pos = G.uniform_positions(n, 3) * 50.0
    ref, ts, inter = O.force_term_sums(g, pos, p, interactions=True)
    f64 = gpu.ForceEngine(off, adj, precision="fp64", **kw)
    f64.set_positions(pos)
    got = f64.evaluate(interactions=True)
    assert np.array_equal(got["interactions"], inter)
    assert np.linalg.norm(got["disp"] - ref) <= 1e-11 * np.linalg.norm(ref)
    mixed = gpu.ForceEngine(off, adj, precision="mixed", **kw)
    mixed.set_positions(pos)
    d = mixed.evaluate()["disp"] - ref
    assert np.linalg.norm(d) <= 1e-5 * np.linalg.norm(ref)
    assert (np.linalg.norm(d, axis=1) / ts).max() <= 1e-5
    # eight steps: the FA2 speed update amplifies last-bit differences in the
    # sums (near-contact terms start it early here), so the first three
    # steps hold 1e-9 and all eight 1e-5
    f64.begin(pos)
    stats, ran = f64.run(8)
    run = O.ForceRun(g, p, pos)
    want = [run.step().maxDisp for _ in range(8)]
    np.testing.assert_allclose(stats["maxDisp"][:3], want[:3], rtol=1e-9)
    np.testing.assert_allclose(stats["maxDisp"], want, rtol=1e-5)
    scale = np.abs(run.pos).max()
    assert np.abs(f64.positions() - run.pos).max() <= 1e-5 * scale


def test_loaded_graph_grip_filtration_matches_oracle(gpu, tmp_path):
    off, adj = _loaded(tmp_path, "random2000")
    n = len(off) - 1
    src = np.repeat(np.arange(n), np.diff(off))
    e = np.stack([src, adj], 1)
    e = e[e[:, 0] < e[:, 1]]
    act, inside = G.largest_component(n, e)
    o2, a2 = G.csr_from_edges(n, inside)
    # edges of the component keep the loader's adjacency order
    keep = act[src].astype(bool) & act[adj].astype(bool)
    o_keep = np.concatenate([[0], np.cumsum(np.bincount(src[keep], minlength=n))])
    np.testing.assert_array_equal(o2, o_keep)
    eng = gpu.GripEngine(o_keep, adj[keep], active=act, kcap=10)
    filt, borders = eng.mis_filtration(2)
    of, ob = O.mis_filtration(O.Graph(o_keep, adj[keep]), act)
    assert np.array_equal(borders, ob) and np.array_equal(filt, of)
