"""NumPy restatement of the warm-start network's forward and backward passes.

TEST INFRASTRUCTURE ONLY: the checker for grapher_amd/csrc/gnn.hip (tests,
bench cpu_baseline). Nothing in the product imports it.

The model has no counterpart in the reference (it proposes the start that
gvizForceEmbedderBegin, src/embedders/gvizForceEmbedder.c:412-443, would
otherwise draw at random):

    x (n x 16) -> h1 = relu([x  | A x ] W1 + b1)
               -> h2 = relu([h1 | A h1] W2 + b2)
               -> h3 = relu([h2 | A h2] W3 + b3)
               -> y  = h3 W4 + b4,   P = y[:, :2]

with A the row-normalised adjacency (mean over neighbours; 0 for isolated
vertices); with pivots set, P = y[:, :2] + the pivot-MDS coordinates. Loss: L = |s Pc R - Tc|^2 / |Tc|^2 with Pc, Tc centred and (R, s)
the orthogonal Procrustes fit (rotation or reflection) of Pc to Tc.

`bf16=True` rounds weights, inputs, activations and back-propagated signals
to bfloat16 at the points where the GPU kernels store them, so the GPU's
gradients can be compared closely; `bf16=False` is exact FP64, the form the
finite-difference checks pin.
"""
from __future__ import annotations

import numpy as np

F, H, O = 16, 64, 16
SHAPES = [(2 * F, H), (2 * H, H), (2 * H, H), (H, O)]
PARAMS = sum(k * n + n for k, n in SHAPES)


def unpack(theta):
    """Split the flat parameter vector into [(W, b)] per layer."""
    out, o = [], 0
    for k, n in SHAPES:
        W = theta[o:o + k * n].reshape(k, n)
        o += k * n
        b = theta[o:o + n]
        o += n
        out.append((W, b))
    return out


def bf16(x):
    """Round to bfloat16 (nearest even) and back, as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def mean_agg(off, adj, X):
    n = len(off) - 1
    deg = np.diff(off)
    S = np.zeros_like(X)
    np.add.at(S, np.repeat(np.arange(n), deg), X[adj])
    inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    return S * inv[:, None], inv


def mean_agg_T(off, adj, D, inv):
    """Transpose of mean_agg for a symmetric graph: out[u] = sum over
    neighbours v of D[v] / deg(v)."""
    n = len(off) - 1
    deg = np.diff(off)
    S = np.zeros_like(D)
    np.add.at(S, np.repeat(np.arange(n), deg), (D * inv[:, None])[adj])
    return S


CHEB_C = 0.9
CHEB_MARKS = (2, 4, 8, 16)


def lazy_walk(off, adj, x):
    """P x for the lazy random walk P = (I + D^-1 A) / 2; isolated vertices
    keep x."""
    deg = np.diff(off)
    m, _ = mean_agg(off, adj, x)
    return np.where(deg[:, None] > 0, 0.5 * (x + m), x)


def cheb_filters(off, adj, x, degrees=CHEB_MARKS, c=CHEB_C):
    """T_d(mu(P)) x for each d in `degrees`, mu(P) = (2 P - c I) / c, by the
    three-term recurrence y_1 = mu y_0, y_{k+1} = 2 mu y_k - y_{k-1}."""
    mu = lambda y: (2.0 / c) * lazy_walk(off, adj, y) - y
    y_prev, y = None, np.asarray(x, dtype=np.float64)
    out = {0: y}
    for k in range(1, max(degrees) + 1):
        y_next = mu(y) if k == 1 else 2.0 * mu(y) - y_prev
        y_prev, y = y, y_next
        out[k] = y
    return [out[d] for d in degrees]


def uniform_start(n, seed):
    """The device-drawn start (gvizhip_gnn_set_start_seeded): SplitMix64
    (Steele, Lea & Flood 2014) of seed + (v + 1) * 0x9E3779B97F4A7C15, bits
    40-63 and 16-39 as 2 b / 2^24 - 1."""
    M = (1 << 64) - 1
    out = np.zeros((n, 2))
    for v in range(n):
        z = (seed + (v + 1) * 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        out[v, 0] = (z >> 40) / 16777216.0 * 2.0 - 1.0
        out[v, 1] = ((z >> 16) & 0xFFFFFF) / 16777216.0 * 2.0 - 1.0
    return out


def diffused_features(off, adj, start):
    """The device-built inputs (gvizhip_gnn_set_graph_diffused): the start and
    its Chebyshev low-pass filters of degree 2, 4, 8, 16 (cheb_filters), each
    pair centred and scaled to unit variance per coordinate;
    log(1 + degree) / 4; a constant one."""
    n = len(off) - 1
    deg = np.diff(off)
    x = np.asarray(start, dtype=np.float64).reshape(n, 2)
    X = np.zeros((n, F))
    for f, y in enumerate([x] + cheb_filters(off, adj, x)):
        c = y - y.mean(0)
        sd = np.sqrt((c * c).mean(0))
        X[:, 2 * f:2 * f + 2] = c / np.where(sd > 1e-15, sd, np.inf)
    X[:, 10] = np.log1p(deg) / 4.0
    X[:, 11] = 1.0
    return X


def pivot_mds(off, adj, pivots):
    """Pivot MDS (Brandes & Pich) as gvizhip_gnn_set_pivots computes it: BFS
    hop distances D to the pivots (unreached = farthest + 1), C = -1/2 D^2
    double-centred (row means, column means, mean of the column means), the
    top two eigenpairs (v, mu) of C^T C with the largest-magnitude component
    of v positive, axes C v / mu^(1/4), centred and scaled to RMS radius 1."""
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg
    n = len(off) - 1
    A = sp.csr_matrix((np.ones(len(adj)), np.asarray(adj), np.asarray(off)), shape=(n, n))
    D = cg.shortest_path(A, unweighted=True, indices=np.asarray(pivots)).T
    fin = np.isfinite(D)
    D[~fin] = (D[fin].max() if fin.any() else 0.0) + 1.0
    D2 = D * D
    colm = D2.mean(0)
    C = -0.5 * (D2 - D2.mean(1)[:, None] - colm[None, :] + colm.mean())
    w, V = np.linalg.eigh(C.T @ C)
    X = np.zeros((n, 2))
    for a, c in enumerate(np.argsort(w)[::-1][:2]):
        if not w[c] > 0:
            continue
        v = V[:, c]
        v = v if v[np.argmax(np.abs(v))] > 0 else -v
        X[:, a] = C @ (v / w[c] ** 0.25)
    X -= X.mean(0)
    rms = np.sqrt((X * X).sum(1).mean())
    return X / rms if rms > 0 else X


def pivot_features(off, adj, start, pivots):
    """gvizhip_gnn_set_graph_diffused then gvizhip_gnn_set_pivots: the
    diffused features with the pivot-MDS coordinates in columns 12-13, and
    those coordinates (the skip term)."""
    X = diffused_features(off, adj, start)
    P = pivot_mds(off, adj, pivots)
    X[:, 12:14] = P
    return X, P


def forward(theta, off, adj, X, rb=False, skip=None):
    """Head output y (n x 16); `skip` (n x 2, the pivot-MDS coordinates) is
    added to outputs 0-1 after the head, as the device does."""
    r = bf16 if rb else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(np.asarray(theta, dtype=np.float64))
    h = r(X)
    cache = []
    for li, (W, b) in enumerate(layers[:3]):
        m, inv = mean_agg(off, adj, h)
        m = r(m)
        a = np.concatenate([h, m], 1)
        z = a @ r(W) + b
        hn = r(np.maximum(z, 0.0))
        cache.append((a, hn, inv))
        h = hn
    W4, b4 = layers[3]
    y = h @ r(W4) + b4
    if skip is not None:
        y[:, :2] += skip
    return y, cache


def procrustes(P, T):
    """(loss, R, s, Pc, Tc) of the orthogonal Procrustes fit of P to T."""
    Pc = P - P.mean(0)
    Tc = T - T.mean(0)
    M = Pc.T @ Tc
    U, S, Vt = np.linalg.svd(M)
    R = U @ Vt
    pp = (Pc ** 2).sum()
    tt = (Tc ** 2).sum()
    tr = S.sum()
    s = tr / pp
    loss = 1.0 - tr * tr / (pp * tt)
    return loss, R, s, Pc, Tc


def loss_and_grads(theta, off, adj, X, T, rb=False, skip=None):
    """Loss and the gradient of every parameter (flat, like theta)."""
    r = bf16 if rb else (lambda a: np.asarray(a, dtype=np.float64))
    theta = np.asarray(theta, dtype=np.float64)
    layers = unpack(theta)
    y, cache = forward(theta, off, adj, X, rb, skip)
    P = y[:, :2]
    loss, R, s, Pc, Tc = procrustes(P, T)
    tt = (Tc ** 2).sum()
    dP = 2.0 * s * (s * Pc @ R - Tc) @ R.T / tt
    dY = np.zeros_like(y)
    dY[:, :2] = dP
    dY = r(dY)
    grads = []
    h3 = cache[2][1]
    W4 = r(layers[3][0])
    grads.append((h3.T @ dY, dY.sum(0)))
    dh = dY @ W4.T
    dz = r(np.where(h3 > 0, dh, 0.0))
    for li in (2, 1, 0):
        a, hn, inv = cache[li]
        W = r(layers[li][0])
        grads.append((a.T @ dz, dz.sum(0)))
        if li == 0:
            break
        da = dz @ W.T
        k = a.shape[1] // 2
        dh = da[:, :k] + mean_agg_T(off, adj, da[:, k:], inv)
        hprev = cache[li - 1][1]
        dz = r(np.where(hprev > 0, dh, 0.0))
    grads = grads[::-1]  # layer order 1..4
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
    return loss, flat


def init_params(seed=0, head=0.5):
    """He-scaled normal weights, zero biases; head columns 2..15 zero and
    the head scaled by `head` (0: the output starts as the skip term)."""
    rng = np.random.default_rng(seed)
    parts = []
    for li, (k, n) in enumerate(SHAPES):
        W = rng.normal(0.0, np.sqrt(2.0 / k), size=(k, n))
        if li == 3:
            W[:, 2:] = 0.0
            W *= head
        parts.append(W.ravel())
        parts.append(np.zeros(n))
    return np.concatenate(parts)
