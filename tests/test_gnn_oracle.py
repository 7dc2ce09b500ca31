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


def _classical_mds(off, adj):
    """Textbook classical MDS of all-pairs hop distances: B = -1/2 J D^2 J,
    axes sqrt(lambda) u for the top two eigenpairs, centred, RMS radius 1."""
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg
    n = len(off) - 1
    A = sp.csr_matrix((np.ones(len(adj)), adj, off), shape=(n, n))
    D = cg.shortest_path(A, unweighted=True)
    J = np.eye(n) - 1.0 / n
    B = -0.5 * J @ (D * D) @ J
    w, U = np.linalg.eigh(B)
    X = U[:, [-1, -2]] * np.sqrt(w[[-1, -2]])
    X -= X.mean(0)
    return X / np.sqrt((X * X).sum(1).mean())


def test_pivot_mds_with_every_vertex_is_classical_mds():
    """With all vertices as pivots C is the double-centred distance matrix
    B, C^T C = B^2, and C v / mu^(1/4) = sqrt(lambda) u: classical MDS,
    axis by axis up to sign."""
    n, e = G.erdos_renyi(30, 4.0, 7)
    act, inside = G.largest_component(n, e)
    ids = np.flatnonzero(act)
    idx = -np.ones(n, np.int64)
    idx[ids] = np.arange(ids.size)
    off, adj = G.csr_from_edges(ids.size, idx[inside])
    P = N.pivot_mds(off, adj, np.arange(ids.size))
    M = _classical_mds(off, adj)
    for a in range(2):
        s = np.sign(P[:, a] @ M[:, a])
        assert np.abs(P[:, a] - s * M[:, a]).max() <= 1e-9


def test_pivot_mds_of_a_cycle_is_a_regular_polygon():
    """Hop distances of a 12-cycle embed classically as a regular 12-gon:
    every radius is 1 (RMS radius 1) and neighbours sit 30 degrees apart."""
    n = 12
    e = np.array([[i, (i + 1) % n] for i in range(n)])
    off, adj = G.csr_from_edges(n, e)
    P = N.pivot_mds(off, adj, np.arange(n))
    r = np.sqrt((P * P).sum(1))
    assert np.abs(r - 1.0).max() <= 1e-9
    ang = np.unwrap(np.arctan2(P[:, 1], P[:, 0]))
    step = np.abs(np.diff(ang))
    assert np.abs(step - 2 * np.pi / n).max() <= 1e-9


def test_skip_term_enters_output_and_gradients():
    """The skip is added to the head's outputs; finite differences still
    match the backward pass with it."""
    off, adj, X, T = _small(6)
    skip = np.random.default_rng(8).normal(size=(len(off) - 1, 2))
    theta = N.init_params(4)
    y0 = N.forward(theta, off, adj, X)[0]
    y1 = N.forward(theta, off, adj, X, skip=skip)[0]
    assert np.abs(y1[:, :2] - y0[:, :2] - skip).max() <= 1e-12
    assert np.array_equal(y1[:, 2:], y0[:, 2:])
    loss, grad = N.loss_and_grads(theta, off, adj, X, T, skip=skip)
    rng = np.random.default_rng(2)
    live = np.flatnonzero(np.abs(grad) > 0)
    h = 1e-6
    for i in rng.choice(live, size=30, replace=False):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (N.loss_and_grads(tp, off, adj, X, T, skip=skip)[0] -
              N.loss_and_grads(tm, off, adj, X, T, skip=skip)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, grad[i])


def _dense_lazy_walk(off, adj):
    n = len(off) - 1
    P = np.zeros((n, n))
    for v in range(n):
        d = off[v + 1] - off[v]
        P[v, v] = 0.5 if d else 1.0
        for u in adj[off[v]:off[v + 1]]:
            P[v, u] += 0.5 / d
    return P


def test_cheb_filter_of_a_cycle_eigenvector_is_the_chebyshev_value():
    """On the cycle C_n, cos(2 pi j v / n) is an eigenvector of the lazy walk
    with eigenvalue (1 + cos(2 pi j / n)) / 2; the filter multiplies it by
    T_d(mu), mu = (2 lambda - c) / c, T_d the Chebyshev polynomial."""
    n = 60
    e = np.array([(i, (i + 1) % n) for i in range(n)])
    off, adj = G.csr_from_edges(n, e)
    v = np.arange(n)
    for j in (1, 3, 11):
        x = np.stack([np.cos(2 * np.pi * j * v / n), np.sin(2 * np.pi * j * v / n)], 1)
        lam = 0.5 * (1 + np.cos(2 * np.pi * j / n))
        mu = (2 * lam - N.CHEB_C) / N.CHEB_C
        got = N.cheb_filters(off, adj, x)
        for d, y in zip(N.CHEB_MARKS, got):
            t = np.polynomial.chebyshev.chebval(mu, [0] * d + [1])
            assert np.allclose(y, t * x, rtol=1e-9, atol=1e-9 * max(1.0, abs(t)))


def test_cheb_filter_equals_the_matrix_polynomial():
    """On an irregular graph with an isolated vertex: T_d(mu(P)) from the
    eigendecomposition-free matrix polynomial (numpy's Chebyshev series of
    the dense P) equals the recurrence."""
    n, e = G.erdos_renyi(30, 3.0, 7)
    off, adj = G.csr_from_edges(n + 1, e)  # vertex n is isolated
    P = _dense_lazy_walk(off, adj)
    M = (2.0 * P - N.CHEB_C * np.eye(n + 1)) / N.CHEB_C
    x = np.random.default_rng(2).normal(size=(n + 1, 2))
    for d, y in zip(N.CHEB_MARKS, N.cheb_filters(off, adj, x)):
        T0, T1 = np.eye(n + 1), M
        for _ in range(d - 1):
            T0, T1 = T1, 2 * M @ T1 - T0
        assert np.allclose(y, T1 @ x, rtol=1e-10, atol=1e-10 * np.abs(T1 @ x).max())


def test_uniform_start_is_splitmix64():
    """SplitMix64's published first outputs from state 0 (Vigna's reference
    C, which adds the gamma before mixing): 0xE220A8397B1DCDAF,
    0x6E789E6AA1B965F4, 0x06C45D188009454F; coordinates are their bit fields
    40-63 and 16-39, and every value lies in [-1, 1)."""
    first = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    S = N.uniform_start(3, 0)
    for v, z in enumerate(first):
        assert S[v, 0] == (z >> 40) / 2 ** 24 * 2 - 1
        assert S[v, 1] == ((z >> 16) & 0xFFFFFF) / 2 ** 24 * 2 - 1
    U = N.uniform_start(4000, 12345)
    assert U.min() >= -1.0 and U.max() < 1.0
    assert abs(U.mean()) < 0.03 and abs(U.var() - 1 / 3) < 0.02
