"""Seeded synthetic graphs and initial positions for the layout engine.

Every generator returns ``(n, edges)`` where ``edges`` is an ``(m, 2)`` int64
array in insertion order: the order in which the reference would call
``gvizGraphAddEdge``. Insertion order is semantics for the engine (attraction
summation order, BFS discovery order, MIS tail order), so ``csr_from_edges``
keeps it: each vertex's adjacency lists its neighbours in the order its edges
were added, exactly like the reference's per-vertex ``gvizArray``
(src/ds/gvizGraph.c:163-172).
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """The ``count`` outputs of splitmix64 after ``offset`` steps from ``seed``."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Doubles in [0, 1) from the top 53 bits of splitmix64."""
    return (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (
        1.0 / 9007199254740992.0)


def uniform_positions(n: int, seed: int = 0, extent: float = 1.0) -> np.ndarray:
    """(n, 2) positions uniform in [-extent, extent]^2: x then y per vertex."""
    u = uniform01(seed, 2 * n)
    return (extent * (2.0 * u - 1.0)).reshape(n, 2)


def grid(rows: int, cols: int):
    """``build_rect_mesh`` (src/utils/graphs.c:283-312): vertex i*cols+j, for
    each (i, j) the right edge then the down edge."""
    edges = []
    for i in range(rows):
        for j in range(cols):
            idx = i * cols + j
            if j + 1 < cols:
                edges.append((idx, idx + 1))
            if i + 1 < rows:
                edges.append((idx, idx + cols))
    return rows * cols, np.asarray(edges, dtype=np.int64).reshape(-1, 2)


def barabasi_albert(n: int, m: int, seed: int = 0):
    """Preferential attachment: a clique on m+1 seed vertices, then each new
    vertex v links to m distinct earlier vertices drawn proportionally to
    degree (edges (v, t) in draw order). No multi-edges, no self loops."""
    rng = np.random.Generator(np.random.PCG64(seed))
    edges = []
    ends = []
    for a in range(m + 1):
        for b in range(a + 1, m + 1):
            edges.append((a, b))
            ends += [a, b]
    ends_arr = np.empty(2 * m * n + 2 * (m + 1) ** 2, dtype=np.int64)
    ends_arr[: len(ends)] = ends
    cnt = len(ends)
    draws = rng.random(size=(n, 8 * m))
    for v in range(m + 1, n):
        chosen = []
        row = draws[v]
        k = 0
        while len(chosen) < m:
            if k < row.shape[0]:
                r = row[k]
            else:
                r = rng.random()
            k += 1
            t = int(ends_arr[int(r * cnt)])
            if t not in chosen:
                chosen.append(t)
        for t in chosen:
            edges.append((v, t))
            ends_arr[cnt] = v
            ends_arr[cnt + 1] = t
            cnt += 2
    return n, np.asarray(edges, dtype=np.int64)


def erdos_renyi(n: int, avg_degree: float, seed: int = 0):
    """G(n, M) with M = round(n*avg_degree/2) distinct undirected edges, in
    draw order."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = int(round(n * avg_degree / 2.0))
    seen = set()
    edges = []
    while len(edges) < m:
        need = m - len(edges)
        a = rng.integers(0, n, size=2 * need + 16)
        b = rng.integers(0, n, size=2 * need + 16)
        for u, v in zip(a.tolist(), b.tolist()):
            if u == v:
                continue
            key = (u, v) if u < v else (v, u)
            if key in seen:
                continue
            seen.add(key)
            edges.append((u, v))
            if len(edges) == m:
                break
    return n, np.asarray(edges, dtype=np.int64).reshape(-1, 2)


def erdos_renyi_fast(n: int, avg_degree: float, seed: int = 0):
    """G(n, M) with M = round(n * avg_degree / 2) distinct undirected edges
    drawn with numpy (rejection of loops and repeats); edges sorted (u < v).
    A vectorised companion of erdos_renyi for large batches; the two draw
    different graphs from the same seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = int(round(n * avg_degree / 2.0))
    if n < 2 or m == 0:
        return n, np.zeros((0, 2), dtype=np.int64)
    m = min(m, n * (n - 1) // 2)
    keys = np.zeros(0, dtype=np.int64)
    while keys.size < m:
        need = m - keys.size
        u = rng.integers(0, n, size=2 * need + 16)
        v = rng.integers(0, n, size=2 * need + 16)
        ok = u != v
        a = np.minimum(u[ok], v[ok])
        b = np.maximum(u[ok], v[ok])
        keys = np.unique(np.concatenate([keys, a * n + b]))
    keys = keys[:m] if keys.size == m else np.sort(
        rng.choice(keys, size=m, replace=False))
    return n, np.stack([keys // n, keys % n], axis=1)


def random_tree(n: int, seed: int = 0):
    """Vertex v > 0 attaches to a uniform earlier vertex."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if n <= 1:
        return n, np.zeros((0, 2), dtype=np.int64)
    parents = (rng.random(n - 1) * np.arange(1, n)).astype(np.int64)
    return n, np.stack([np.arange(1, n, dtype=np.int64), parents], axis=1)


def random_geometric(n: int, avg_degree: float = 8.0, seed: int = 0):
    """Points uniform in the unit square (ids in generation order), edges
    between points closer than r = sqrt(avg_degree / (pi n)), listed by
    ascending (u, v) with u < v. Returns (n, edges, points)."""
    from scipy.spatial import cKDTree

    rng = np.random.Generator(np.random.PCG64(seed))
    pts = rng.random((n, 2))
    r = np.sqrt(avg_degree / (np.pi * n))
    pairs = cKDTree(pts).query_pairs(r, output_type="ndarray").astype(np.int64)
    pairs.sort(axis=1)
    order = np.lexsort((pairs[:, 1], pairs[:, 0]))
    return n, pairs[order], pts


def delaunay(n: int, seed: int = 0):
    """Delaunay triangulation of n uniform points in the unit square.
    Edges are unique (u < v) in ascending order; the convex hull is returned
    counter-clockwise. Returns (n, edges, points, hull)."""
    from scipy.spatial import Delaunay

    rng = np.random.Generator(np.random.PCG64(seed))
    pts = rng.random((n, 2))
    tri = Delaunay(pts)
    s = tri.simplices.astype(np.int64)
    e = np.concatenate([s[:, [0, 1]], s[:, [1, 2]], s[:, [0, 2]]])
    e.sort(axis=1)
    key = e[:, 0] * n + e[:, 1]
    key = np.unique(key)
    edges = np.stack([key // n, key % n], axis=1)
    hull = _ccw_hull(tri, pts)
    return n, edges, pts, hull


def _ccw_hull(tri, pts):
    hv = tri.convex_hull  # boundary edges
    nxt = {}
    for a, b in hv.tolist():
        nxt.setdefault(a, []).append(b)
        nxt.setdefault(b, []).append(a)
    start = int(hv[0, 0])
    order = [start]
    prev, cur = None, start
    while True:
        cand = [x for x in nxt[cur] if x != prev]
        nx = cand[0]
        if nx == start:
            break
        order.append(nx)
        prev, cur = cur, nx
        if len(order) > len(nxt):
            break
    o = np.asarray(order, dtype=np.int64)
    p = pts[o]
    area = np.sum(p[:, 0] * np.roll(p[:, 1], -1) - np.roll(p[:, 0], -1) * p[:, 1])
    if area < 0:
        o = o[::-1].copy()
    return o


def csr_from_edges(n: int, edges: np.ndarray):
    """Undirected CSR (offsets[n+1], adj[2m]) keeping insertion order."""
    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    m = edges.shape[0]
    src = np.empty(2 * m, dtype=np.int64)
    dst = np.empty(2 * m, dtype=np.int64)
    src[0::2] = edges[:, 0]
    dst[0::2] = edges[:, 1]
    src[1::2] = edges[:, 1]
    dst[1::2] = edges[:, 0]
    order = np.argsort(src, kind="stable")
    adj = dst[order]
    counts = np.bincount(src, minlength=n).astype(np.int64)
    offsets = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    return offsets, np.ascontiguousarray(adj)


def largest_component(n: int, edges: np.ndarray):
    """The vertex-induced subgraph on the largest connected component (what
    the reference's GRIP benches lay out, tests/embedders/gvizGRIPPlaceBench.c:
    41-95): returns (active mask uint8[n], edges inside the component, in the
    original insertion order)."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    A = sp.coo_matrix((np.ones(len(edges)), (edges[:, 0], edges[:, 1])), shape=(n, n))
    ncomp, labels = connected_components(A, directed=False)
    big = np.bincount(labels).argmax()
    active = (labels == big).astype(np.uint8)
    keep = active[edges[:, 0]].astype(bool) & active[edges[:, 1]].astype(bool)
    return active, edges[keep]


def batch_mix(count=4096, seed=0, n_min=50, n_max=20000):
    """The batched-layout workload: random trees, 2-D grids and Erdos-Renyi
    graphs (average degree 4), vertex counts log-uniform in [n_min, n_max],
    seeded. Returns ([(offsets, adj)], kinds) with kind 0 tree, 1 grid, 2 ER."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.exp(rng.uniform(np.log(n_min), np.log(n_max), count)).astype(np.int64)
    kinds = rng.integers(0, 3, count)
    out = []
    for i, (size, kind) in enumerate(zip(sizes.tolist(), kinds.tolist())):
        if kind == 0:
            n, e = random_tree(size, seed * 10007 + i)
        elif kind == 1:
            w = max(1, int(round(np.sqrt(size))))
            n, e = grid(w, max(1, int(round(size / w))))
        else:
            n, e = erdos_renyi_fast(size, 4.0, seed * 10007 + i)
        out.append(csr_from_edges(n, e))
    return out, kinds
