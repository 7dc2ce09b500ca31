"""The warm-start network on the GPU against its CPU restatement
(oracle/gnn_oracle.py, itself pinned by finite differences), the device
stress meter against the oracle's stress, and training on force-embedder
layouts.

Tolerances: weights and activations are BF16 on both sides (the oracle
rounds at the same points), products accumulate in FP32 on the GPU and in
FP64 in the oracle, so a rounding boundary can fall differently; outputs
agree to 1e-2 normwise; gradients, rounded to BF16 after every layer of
the backward pass on both sides, agree to 5e-2 normwise per parameter block
and point the same way (cosine >= 0.998)."""
import numpy as np
import pytest

import gnn_oracle as N
import pyoracle as O
from grapher_amd import graphs as G
from grapher_amd import warmstart as W

pytestmark = pytest.mark.gpu


def _case(seed=0, n=600, blocks=6):
    n, e, _ = W.sbm(n, blocks, seed)
    off, adj = G.csr_from_edges(n, e)
    x = W.features(off, seed + 11)
    return n, off, adj, x


def _theta(seed):
    t = N.init_params(seed)
    t[N.PARAMS - N.O:N.PARAMS - N.O + 2] = [0.05, -0.03]
    return t.astype(np.float32)


def _norm_err(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def test_forward_matches_oracle(gpu):
    n, off, adj, x = _case()
    th = _theta(1)
    eng = gpu.GnnEngine(n, len(adj))
    eng.set_params(th)
    eng.set_graph(off, adj, x)
    y_gpu = eng.forward()
    y_ref = N.forward(th.astype(np.float64), off, adj, x.astype(np.float64), rb=True)[0][:, :2]
    assert _norm_err(y_gpu, y_ref) <= 1e-2, _norm_err(y_gpu, y_ref)


def test_device_features_match_oracle(gpu):
    """Features built on the device (random-walk diffusion in FP32) against
    the oracle's FP64 construction, seen through the network's output."""
    n, off, adj, _ = _case(4, 800, 5)
    start = G.uniform_positions(n, 9)
    th = _theta(6)
    eng = gpu.GnnEngine(n, len(adj))
    eng.set_params(th)
    eng.set_graph_diffused(off, adj, start)
    y_gpu = eng.forward()
    X = N.diffused_features(off, adj, start)
    y_ref = N.forward(th.astype(np.float64), off, adj, X, rb=True)[0][:, :2]
    assert _norm_err(y_gpu, y_ref) <= 2e-2, _norm_err(y_gpu, y_ref)


def test_seeded_start_and_proposal_match_oracle(gpu):
    """Features from the device-drawn start (SplitMix64, restated in the
    oracle) and the proposal: the network's output centred and scaled to
    the requested RMS radius on the device."""
    n, off, adj, _ = _case(5, 700, 4)
    th = _theta(8)
    eng = gpu.GnnEngine(n, len(adj))
    eng.set_params(th)
    eng.set_graph_diffused(off, adj, np.zeros((n, 2)))
    eng.set_start_seeded(77)
    y_gpu = eng.forward()
    X = N.diffused_features(off, adj, N.uniform_start(n, 77))
    y_ref = N.forward(th.astype(np.float64), off, adj, X, rb=True)[0][:, :2]
    assert _norm_err(y_gpu, y_ref) <= 2e-2, _norm_err(y_gpu, y_ref)
    p = eng.propose(77, 123.0)
    c = y_gpu - y_gpu.mean(0)
    want = c * (123.0 / np.sqrt((c * c).sum(1).mean()))
    assert _norm_err(p, want) <= 1e-6, _norm_err(p, want)
    assert abs(np.sqrt((p * p).sum(1).mean()) - 123.0) <= 1e-9 * 123.0
    assert np.abs(p.mean(0)).max() <= 1e-9 * 123.0


def test_gradients_match_oracle(gpu):
    n, off, adj, x = _case(2, 900, 9)
    th = _theta(3)
    target = np.random.default_rng(5).normal(size=(n, 2)) * 40.0
    eng = gpu.GnnEngine(n, len(adj))
    eng.set_params(th)
    eng.set_graph(off, adj, x)
    loss, g = eng.gradients(target)
    loss_ref, g_ref = N.loss_and_grads(th.astype(np.float64), off, adj,
                                       x.astype(np.float64), target, rb=True)
    assert abs(loss - loss_ref) <= 1e-3 * max(abs(loss_ref), 1e-3)
    o = 0
    for li, (k, m) in enumerate(N.SHAPES):
        for name, size in (("W", k * m), ("b", m)):
            a, b = g[o:o + size], g_ref[o:o + size]
            o += size
            if li == 3 and name == "b":
                # the head bias gradient is the column sum of dL/dP, which is
                # exactly zero (P is centred); both sides hold rounding noise
                w4 = np.linalg.norm(g_ref[o - size - k * m:o - size])
                assert np.abs(a).max() <= 1e-2 * w4 and np.abs(b).max() <= 1e-2 * w4
                continue
            if np.linalg.norm(b) == 0:
                assert np.abs(a).max() == 0.0
                continue
            cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))
            assert _norm_err(a, b) <= 5e-2 and cos >= 0.998, (k, m, name,
                                                             _norm_err(a, b), cos)


def test_training_on_force_layouts_reduces_loss(gpu):
    data = []
    for i in range(6):
        n, e, _ = W.sbm(400 + 150 * i, 3 + i, 100 + i)
        off, adj = G.csr_from_edges(n, e)
        data.append((off, adj, W.force_layout(off, adj, seed=i, iters=150)))
    ws, hist = W.train(data, epochs=25, lr=3e-3, seed=0, log=None)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.8 * hist[0], hist
    p = ws.positions(*data[0][:2], seed=3)
    assert np.all(np.isfinite(p)) and p.shape == (len(data[0][0]) - 1, 2)


def test_stress_meter_matches_oracle(gpu):
    n, e, _ = W.sbm(3000, 5, 7)
    off, adj = G.csr_from_edges(n, e)
    pos = G.uniform_positions(n, 3) * 17.0
    piv = np.random.default_rng(0).choice(n, 32, replace=False)
    meter = gpu.StressMeter(off, adj, piv)
    ref = O.stress(O.Graph(off, adj), pos, 10.0, piv, O.Pool(4))
    assert abs(meter(pos) - ref) <= 1e-9 * max(abs(ref), 1e-12)
    # invariant under scale, rotation and translation
    th = 1.1
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    assert abs(meter(3.0 * pos @ R + 5.0) - meter(pos)) <= 1e-9


def test_trained_warm_start_gains_over_random_start(gpu):
    """The committed network (trained on the GPU on FP64 FR layouts of SBM
    graphs) against the method without it, on unseen connected SBM graphs:
    after the same number of refinement steps the warm start's layout has
    the lower stress."""
    from grapher_amd import native
    ws = W.WarmStart.load()
    wins = 0
    for seed, blocks in ((501, 7), (502, 19)):
        m, e = W.sbm_connected(50000, blocks, seed)
        off, adj = G.csr_from_edges(m, e)
        meter = native.StressMeter(off, adj,
                                   np.random.default_rng(seed).choice(m, 64, replace=False))
        _, _, plain = W.time_to_stress(off, adj, G.uniform_positions(m, seed), meter,
                                       None, 20, chunk=10)
        _, _, warm = W.time_to_stress(off, adj, ws.positions(off, adj, seed=seed), meter,
                                      None, 20, chunk=10)
        print(f"SBM {m} vertices, {blocks} blocks: stress after 10/20 steps "
              f"plain {plain[1][1]:.4f}/{plain[2][1]:.4f} warm {warm[1][1]:.4f}/{warm[2][1]:.4f}")
        wins += warm[2][1] < plain[2][1]
    assert wins == 2


def _connected_sbm(n, blocks, seed):
    m, e = W.sbm_connected(n, blocks, seed)
    off, adj = G.csr_from_edges(m, e)
    return m, off, adj


@pytest.mark.parametrize("shape", ["grid", "sbm"])
def test_pivot_mds_matches_oracle(gpu, shape):
    """Device BFS from 64 pivots, double centring, Gram matrix and Jacobi
    eigenpairs against the oracle's FP64 restatement (coordinates are
    returned in FP32)."""
    if shape == "grid":
        n, e = G.grid(30, 45)
        off, adj = G.csr_from_edges(n, e)
    else:
        n, off, adj = _connected_sbm(3000, 3, 21)
    piv = W.pick_pivots(n, 5)
    eng = gpu.GnnEngine(n, len(adj))
    eng.set_graph_diffused(off, adj, G.uniform_positions(n, 1))
    P = eng.set_pivots(piv)
    R = N.pivot_mds(off, adj, piv)
    print(f"{shape}: max |device - oracle| {np.abs(P - R).max():.2e}")
    assert np.abs(P - R).max() <= 1e-4


def test_forward_and_gradients_with_skip_match_oracle(gpu):
    n, off, adj = _connected_sbm(1500, 4, 8)
    start = G.uniform_positions(n, 2)
    piv = W.pick_pivots(n, 3)
    th = _theta(7)
    eng = gpu.GnnEngine(n, len(adj))
    eng.set_params(th)
    eng.set_graph_diffused(off, adj, start)
    P = eng.set_pivots(piv)
    X = N.diffused_features(off, adj, start)
    X[:, 12:14] = P
    y_gpu = eng.forward()
    y_ref = N.forward(th.astype(np.float64), off, adj, X, rb=True, skip=P)[0][:, :2]
    assert _norm_err(y_gpu, y_ref) <= 2e-2, _norm_err(y_gpu, y_ref)
    target = W.force_layout(off, adj, seed=1, iters=60)
    loss, g = eng.gradients(target)
    loss_ref, g_ref = N.loss_and_grads(th.astype(np.float64), off, adj, X, target,
                                       rb=True, skip=P)
    assert abs(loss - loss_ref) <= 2e-2 * max(abs(loss_ref), 1e-3), (loss, loss_ref)
    cos = float(g @ g_ref / (np.linalg.norm(g) * np.linalg.norm(g_ref)))
    assert cos >= 0.99, cos
    # removing the pivots removes the skip
    eng.set_pivots([])
    y0 = eng.forward()
    assert _norm_err(y_gpu - y0, P) <= 1e-2
