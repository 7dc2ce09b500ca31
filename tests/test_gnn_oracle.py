"""The warm-start network's CPU restatement (oracle/gnn_oracle.py) against
finite differences and closed forms: the backward pass it gives is the
derivative of its forward pass, and its Procrustes loss is invariant under
the motions it is meant to ignore."""
import numpy as np

import gnn_oracle as N
from grapher_amd import graphs as G


def _small(seed=0, n=40):
    n, e = G.erdos_renyi(n, 4.0, seed)
    off, adj = G.csr_from_edges(n, e)
    rng = np.random.default_rng(seed)
    X = np.zeros((n, N.F))
    X[:, :2] = rng.uniform(-1, 1, (n, 2))
    X[:, 2] = np.log1p(np.diff(off)) / 4
    X[:, 3] = 1.0
    T = rng.normal(size=(n, 2))
    return off, adj, X, T


def test_finite_differences_match_backward():
    off, adj, X, T = _small()
    theta = N.init_params(3)
    theta[N.PARAMS - N.O:N.PARAMS - N.O + 2] = [0.1, -0.2]
    loss, grad = N.loss_and_grads(theta, off, adj, X, T)
    rng = np.random.default_rng(1)
    # every head/bias parameter that matters plus a random sample of weights
    live = np.flatnonzero(np.abs(grad) > 0)
    idx = rng.choice(live, size=60, replace=False)
    h = 1e-6
    for i in idx:
        tp = theta.copy()
        tm = theta.copy()
        tp[i] += h
        tm[i] -= h
        lp = N.loss_and_grads(tp, off, adj, X, T)[0]
        lm = N.loss_and_grads(tm, off, adj, X, T)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, grad[i])


def test_padded_head_columns_get_no_gradient():
    off, adj, X, T = _small(2)
    theta = N.init_params(5)
    _, grad = N.loss_and_grads(theta, off, adj, X, T)
    W4 = grad[N.PARAMS - N.O - N.H * N.O:N.PARAMS - N.O].reshape(N.H, N.O)
    assert np.all(W4[:, 2:] == 0.0) and np.all(grad[N.PARAMS - N.O + 2:] == 0.0)


def test_procrustes_loss_ignores_rigid_motion_and_scale():
    rng = np.random.default_rng(4)
    T = rng.normal(size=(50, 2))
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    P = 3.5 * T @ R + np.array([4.0, -2.0])
    assert N.procrustes(P, T)[0] < 1e-12
    P2 = P * np.array([1.0, -1.0])  # a reflection
    assert N.procrustes(P2, T)[0] < 1e-12
    assert 0.5 < N.procrustes(rng.normal(size=(50, 2)), T)[0] <= 1.0


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3])
    r = N.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8
