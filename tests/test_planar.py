"""Planarity test and largest face behind the Tutte embedders' Begin
(grapher_amd/host/gviz_planar.c; reference src/embedders/gvizTutteEmbedder.c:
145-173). The reference decides planarity with the Boyer-Myrvold suite; the
decision is checked here against networkx's independent left-right
implementation on random graphs of both kinds, and the face against what
geometry fixes: a Delaunay triangulation's outer face is its convex hull, a
grid's is its border cycle, a wheel's is its rim."""
import ctypes as C
import os

import numpy as np
import pytest

from grapher_amd import graphs as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "grapher_amd", "libgviz.so")


@pytest.fixture(scope="module")
def planar():
    if not os.path.exists(LIB):
        pytest.skip("libgviz.so is not built")
    lib = C.CDLL(LIB)
    f = lib.gviz_planar_largest_face
    f.restype = C.c_int
    f.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                  C.POINTER(C.c_size_t)]

    def run(n, edges):
        off, adj = G.csr_from_edges(n, np.asarray(edges, np.int64).reshape(-1, 2))
        off = np.ascontiguousarray(off, np.int64)
        adj = np.ascontiguousarray(adj, np.int64)
        out = np.zeros(max(2 * len(adj), 1), np.uint64)
        cnt = C.c_size_t(0)
        rc = f(n, off.ctypes.data, adj.ctypes.data if len(adj) else None,
               out.ctypes.data, out.shape[0], C.byref(cnt))
        return rc, out[: cnt.value].astype(np.int64)
    return run


def _same_cycle(a, b):
    """a equals b up to the starting vertex and the direction."""
    a, b = list(a), list(b)
    if len(a) != len(b) or sorted(a) != sorted(b):
        return False
    k = b.index(a[0])
    r = b[k:] + b[:k]
    if r == a:
        return True
    rev = list(reversed(b))
    k = rev.index(a[0])
    return rev[k:] + rev[:k] == a


def _is_face_walk(n, edges, walk):
    es = {(int(u), int(v)) for u, v in edges} | {(int(v), int(u)) for u, v in edges}
    return all((int(walk[i]), int(walk[(i + 1) % len(walk)])) in es for i in range(len(walk)))


def test_kuratowski_graphs_are_not_planar(planar):
    k5 = [(i, j) for i in range(5) for j in range(i + 1, 5)]
    assert planar(5, k5)[0] == -2
    k33 = [(i, 3 + j) for i in range(3) for j in range(3)]
    assert planar(6, k33)[0] == -2
    # a subdivision of K3,3 (every edge split once) is sparse but not planar
    edges, nxt = [], 6
    for u, v in k33:
        edges += [(u, nxt), (nxt, v)]
        nxt += 1
    assert planar(nxt, edges)[0] == -2
    # the Petersen graph
    outer = [(i, (i + 1) % 5) for i in range(5)]
    inner = [(5 + i, 5 + (i + 2) % 5) for i in range(5)]
    spokes = [(i, 5 + i) for i in range(5)]
    assert planar(10, outer + inner + spokes)[0] == -2


def test_decision_matches_networkx_on_random_graphs(planar):
    nx = pytest.importorskip("networkx")
    rng = np.random.default_rng(7)
    seen = {0: 0, -2: 0}
    for trial in range(400):
        n = int(rng.integers(5, 40))
        m = int(rng.integers(n - 1, min(3 * n - 3, n * (n - 1) // 2)))
        pairs = set()
        while len(pairs) < m:
            u, v = rng.integers(0, n, 2)
            if u != v:
                pairs.add((int(min(u, v)), int(max(u, v))))
        edges = sorted(pairs)
        g = nx.Graph()
        g.add_nodes_from(range(n))
        g.add_edges_from(edges)
        want = 0 if nx.check_planarity(g)[0] else -2
        rc, walk = planar(n, edges)
        assert rc == want, (trial, n, edges)
        seen[rc] += 1
        if rc == 0:
            assert len(walk) >= 3 and _is_face_walk(n, edges, walk), trial
    assert seen[0] > 50 and seen[-2] > 50, seen


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_delaunay_outer_face_is_the_convex_hull(planar, seed):
    from scipy.spatial import ConvexHull, Delaunay
    rng = np.random.default_rng(seed)
    pts = rng.random((2000, 2))
    tri = Delaunay(pts)
    edges = set()
    for s in tri.simplices:
        for a in range(3):
            u, v = int(s[a]), int(s[(a + 1) % 3])
            edges.add((min(u, v), max(u, v)))
    edges = sorted(edges)
    rc, walk = planar(len(pts), edges)
    assert rc == 0
    hull = ConvexHull(pts).vertices
    assert _same_cycle(walk, hull)


def test_grid_wheel_and_shuffled_ids(planar):
    n, e = G.grid(12, 12)
    rc, walk = planar(n, e)
    assert rc == 0
    border = [i for i in range(12)] + [12 * r + 11 for r in range(1, 12)] + \
             [12 * 11 + c for c in range(10, -1, -1)] + [12 * r for r in range(10, 0, -1)]
    assert _same_cycle(walk, border)
    # wheel: hub 0 and a rim of 30; every inner face is a triangle
    rim = [(1 + i, 1 + (i + 1) % 30) for i in range(30)]
    rc, walk = planar(31, rim + [(0, 1 + i) for i in range(30)])
    assert rc == 0 and _same_cycle(walk, list(range(1, 31)))
    # the same grid under a random relabelling and edge order
    rng = np.random.default_rng(3)
    perm = rng.permutation(n)
    e2 = perm[np.asarray(e)]
    rng.shuffle(e2)
    rc, walk = planar(n, e2)
    assert rc == 0 and _same_cycle(walk, perm[border])


def test_paths_trees_and_tiny_graphs(planar):
    assert planar(0, [])[0] == -1           # no face at all
    assert planar(3, [(0, 1), (1, 2)])[0] == 0  # a path: one face of 4 darts
    rc, walk = planar(3, [(0, 1), (1, 2), (2, 0)])
    assert rc == 0 and sorted(walk) == [0, 1, 2]
    # a long path (deep DFS) stays on the explicit stacks
    n = 200000
    rc, walk = planar(n, [(i, i + 1) for i in range(n - 1)])
    assert rc == 0 and len(walk) == 2 * (n - 1)
