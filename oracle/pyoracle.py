"""ctypes view of liboracle.so, the FP64 CPU restatement of the reference's
layout hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the thing that
is measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

MODEL_FR = 0
MODEL_LINLOG = 1


class CSR(C.Structure):
    _fields_ = [("n", C.c_int64), ("offsets", C.c_void_p), ("adj", C.c_void_p)]


class ForceParams(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("barnesHut", C.c_int32),
        ("preventOverlap", C.c_int32), ("pad", C.c_int32),
        ("nodesPerCell", C.c_int64), ("edgeLength", C.c_double),
        ("theta", C.c_double), ("gravityK", C.c_double),
        ("overlapConstant", C.c_double), ("radiusBase", C.c_double),
        ("radiusPerDegree", C.c_double), ("jitterTolerance", C.c_double),
    ]


class StepStats(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "maxDisp", "speed", "meanAtt", "meanRep", "meanGrav",
        "speedEfficiency", "totalSwinging", "totalTraction")]


class QNode(C.Structure):
    _fields_ = [
        ("cx", C.c_double), ("cy", C.c_double), ("halfSize", C.c_double),
        ("comX", C.c_double), ("comY", C.c_double), ("mass", C.c_double),
        ("child", C.c_int32 * 4), ("depth", C.c_int32), ("pad", C.c_int32),
        ("ptStart", C.c_int64), ("ptCount", C.c_int64),
        ("keyHi", C.c_uint64), ("keyLo", C.c_uint64),
    ]


class GripVertex(C.Structure):
    _fields_ = [
        ("dispX", C.c_double), ("dispY", C.c_double),
        ("oldDispX", C.c_double), ("oldDispY", C.c_double),
        ("heat", C.c_double), ("oldCos", C.c_double),
        ("lastForceMag", C.c_double), ("knnCount", C.c_int64),
        ("dispCalculated", C.c_uint8),
    ]


class GripRoundStats(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("maxDisp", "meanDisp", "meanForce",
                                          "meanHeat")]


GRIP_VERTEX_DTYPE = np.dtype([
    ("dispX", "f8"), ("dispY", "f8"), ("oldDispX", "f8"), ("oldDispY", "f8"),
    ("heat", "f8"), ("oldCos", "f8"), ("lastForceMag", "f8"),
    ("knnCount", "i8"), ("dispCalculated", "u1"), ("_pad", "V7")])
assert GRIP_VERTEX_DTYPE.itemsize == C.sizeof(GripVertex)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.oc_pool_create.restype = C.c_void_p
        L.oc_pool_create.argtypes = [C.c_int]
        L.oc_pool_destroy.argtypes = [C.c_void_p]
        L.oc_quadtree_build.restype = C.c_void_p
        L.oc_quadtree_build.argtypes = [_f64p, _f64p, C.c_int64, C.c_int64]
        L.oc_quadtree_free.argtypes = [C.c_void_p]
        L.oc_quadtree_node_count.restype = C.c_int64
        L.oc_quadtree_node_count.argtypes = [C.c_void_p]
        L.oc_quadtree_nodes.restype = C.POINTER(QNode)
        L.oc_quadtree_nodes.argtypes = [C.c_void_p]
        L.oc_quadtree_leaf_points.restype = C.POINTER(C.c_int64)
        L.oc_quadtree_leaf_points.argtypes = [C.c_void_p]
        L.oc_quadtree_root.restype = C.c_int32
        L.oc_quadtree_root.argtypes = [C.c_void_p]
        L.oc_quadtree_root_params.argtypes = [
            C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
            C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        L.oc_quadtree_point_leaves.argtypes = [C.c_void_p, _i32p, _u64p, _u64p]
        L.oc_path_key.argtypes = [C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_int32,
                                  C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.oc_force_compute.argtypes = [
            C.POINTER(CSR), _f64p, _f64p, _i64p, C.POINTER(ForceParams),
            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
            C.c_void_p, C.c_void_p]
        L.oc_force_term_sums.argtypes = [
            C.POINTER(CSR), _f64p, _f64p, _i64p, C.POINTER(ForceParams),
            C.c_void_p, C.c_void_p, C.c_void_p, _f64p]
        L.oc_force_state_create.restype = C.c_void_p
        L.oc_force_state_create.argtypes = [C.POINTER(CSR),
                                            C.POINTER(ForceParams)]
        L.oc_force_state_free.argtypes = [C.c_void_p]
        L.oc_force_state_reset.argtypes = [C.c_void_p]
        L.oc_force_step.restype = C.c_double
        L.oc_force_step.argtypes = [C.c_void_p, C.POINTER(CSR),
                                    C.POINTER(ForceParams), _f64p, C.c_void_p,
                                    C.POINTER(StepStats)]
        L.oc_fa2_update.argtypes = [C.c_int64, C.c_double, C.c_double,
                                    C.c_double, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]
        L.oc_tutte_step.restype = C.c_double
        L.oc_tutte_step.argtypes = [C.POINTER(CSR), _u8p, _f64p, _f64p,
                                    C.c_double, C.c_double, C.c_int]
        L.oc_tutte_polygon.argtypes = [C.c_int64, C.c_double, _f64p]
        L.oc_stress.restype = C.c_double
        L.oc_stress.argtypes = [C.POINTER(CSR), _f64p, C.c_double, C.c_void_p,
                                C.c_int64, C.c_void_p]
        L.oc_edge_energy.restype = C.c_double
        L.oc_edge_energy.argtypes = [C.POINTER(CSR), _f64p]
        L.oc_knn_bfs.restype = C.c_int64
        L.oc_knn_bfs.argtypes = [C.POINTER(CSR), C.c_int64, C.c_int64,
                                 C.c_void_p, _i64p]
        L.oc_knn_batch.restype = C.c_int
        L.oc_knn_batch.argtypes = [C.POINTER(CSR), _u8p, C.c_int64, _i64p,
                                   C.c_int64, _i64p, _i64p]
        L.oc_knn_prefer_batch.restype = C.c_int
        L.oc_knn_prefer_batch.argtypes = [C.POINTER(CSR), C.c_int64, C.c_int64,
                                          C.c_int64, C.c_int64]
        L.oc_mis_filtration.restype = C.c_int64
        L.oc_mis_filtration.argtypes = [C.POINTER(CSR), C.c_void_p, C.c_int64,
                                        _i64p, _i64p]
        L.oc_mis_first.restype = C.c_int64
        L.oc_mis_first.argtypes = [C.POINTER(CSR), C.c_void_p, _i64p, _u8p]
        L.oc_grip_refine_round.argtypes = [
            C.POINTER(CSR), _i64p, C.c_int64, C.c_int64, _i64p, C.c_int64,
            C.c_void_p, _f64p, C.POINTER(GripRoundStats)]
        L.oc_grip_refine_round_nd.argtypes = [
            C.POINTER(CSR), _i64p, C.c_int64, C.c_int64, _i64p, C.c_int64, C.c_int64,
            C.c_void_p, _f64p, C.POINTER(GripRoundStats)]
        L.oc_regular_simplex.argtypes = [C.c_int64, C.c_double, _f64p]
        _d = C.c_double
        _dp = C.POINTER(C.c_double)
        L.oc_fr_att.argtypes = [_f64p, _f64p, _d, _f64p]
        L.oc_linlog_att.argtypes = [_f64p, _f64p, _f64p]
        L.oc_fr_rep.argtypes = [_f64p, _f64p, _d, _d, _d, C.c_int64, C.c_int64,
                                _f64p]
        L.oc_linlog_rep.argtypes = [_f64p, _f64p, _d, _d, _d, _d, C.c_int64,
                                    C.c_int64, _f64p]
        L.oc_gravity.argtypes = [_f64p, _d, _f64p]
        L.oc_kk.argtypes = [_f64p, _f64p, _d, _f64p]
        L.oc_grip_fr_rep.argtypes = [_f64p, _f64p, _d, _f64p]
        L.oc_grip_fr_att.argtypes = [_f64p, _f64p, _d, _d, _f64p]
        L.oc_coincident_dir.argtypes = [C.c_int64, C.c_int64, _dp, _dp]
        _lib = L
    return _lib


def _v2(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(2)


def _acc(acc):
    return np.zeros(2) if acc is None else np.ascontiguousarray(acc, dtype=np.float64).copy()


# Pairwise force terms (include/core/gvizVec.h, n = 2). Each adds into a copy
# of `acc` (zero by default) and returns it, as the reference's gvizVecAcc*
# functions accumulate into their last argument.
def fr_att(v, u, k, acc=None):
    a = _acc(acc)
    lib().oc_fr_att(_v2(v), _v2(u), k, a)
    return a


def linlog_att(v, u, acc=None):
    a = _acc(acc)
    lib().oc_linlog_att(_v2(v), _v2(u), a)
    return a


def fr_rep(v, u, k, radius_sum, overlap_constant, self_id=0, other_id=1, acc=None):
    a = _acc(acc)
    lib().oc_fr_rep(_v2(v), _v2(u), k, radius_sum, overlap_constant, self_id,
                    other_id, a)
    return a


def linlog_rep(v, u, v_mass, other_mass, radius_sum, overlap_constant,
               self_id=0, other_id=1, acc=None):
    a = _acc(acc)
    lib().oc_linlog_rep(_v2(v), _v2(u), v_mass, other_mass, radius_sum,
                        overlap_constant, self_id, other_id, a)
    return a


def gravity(v, magnitude, acc=None):
    a = _acc(acc)
    lib().oc_gravity(_v2(v), magnitude, a)
    return a


def kk(v, u, length, acc=None):
    a = _acc(acc)
    lib().oc_kk(_v2(v), _v2(u), length, a)
    return a


def grip_fr_rep(v, u, edge_length, acc=None):
    a = _acc(acc)
    lib().oc_grip_fr_rep(_v2(v), _v2(u), edge_length, a)
    return a


def grip_fr_att(v, u, edge_length, fr_scale, acc=None):
    a = _acc(acc)
    lib().oc_grip_fr_att(_v2(v), _v2(u), edge_length, fr_scale, a)
    return a


def coincident_dir(self_id, other_id):
    ux, uy = C.c_double(), C.c_double()
    lib().oc_coincident_dir(self_id, other_id, C.byref(ux), C.byref(uy))
    return ux.value, uy.value


def _ptr(a):
    return None if a is None else a.ctypes.data


class Graph:
    """CSR holder keeping the numpy buffers alive for the C struct."""

    def __init__(self, offsets, adj):
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.adj = np.ascontiguousarray(adj, dtype=np.int64)
        self.n = self.offsets.shape[0] - 1
        self.c = CSR(self.n, self.offsets.ctypes.data, self.adj.ctypes.data)

    @property
    def degree(self):
        return np.diff(self.offsets)


class Pool:
    def __init__(self, threads=0):
        self.h = lib().oc_pool_create(int(threads))

    def close(self):
        if self.h:
            lib().oc_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def params(model=MODEL_FR, barnes_hut=True, prevent_overlap=False, npc=1,
           edge_length=10.0, theta=1.0, gravity=0.0, overlap_constant=100.0,
           radius_base=0.0, radius_per_degree=0.0, jitter=1.0):
    return ForceParams(model, int(barnes_hut), int(prevent_overlap), 0, npc,
                       edge_length, theta, gravity, overlap_constant,
                       radius_base, radius_per_degree, jitter)


def vertex_mass(model, degree):
    degree = np.asarray(degree, dtype=np.int64)
    return (1.0 + degree.astype(np.float64)) if model == MODEL_LINLOG else \
        np.ones(degree.shape, dtype=np.float64)


def force_compute(g: Graph, pos, p: ForceParams, pool: Pool | None = None):
    """One force evaluation. Returns dict of disp, structForce, attMag,
    repMag (n, 2)/(n,) and the per-vertex repulsion interaction count."""
    n = g.n
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(n, 2)
    mass = vertex_mass(p.model, g.degree)
    deg = np.ascontiguousarray(g.degree, dtype=np.int64)
    disp = np.zeros((n, 2))
    sf = np.zeros((n, 2))
    am = np.zeros(n)
    rm = np.zeros(n)
    inter = np.zeros(n, dtype=np.int64)
    lib().oc_force_compute(C.byref(g.c), pos.reshape(-1), mass, deg,
                           C.byref(p), None, pool.h if pool else None,
                           _ptr(disp), _ptr(sf), _ptr(am), _ptr(rm),
                           _ptr(inter))
    return {"disp": disp, "structForce": sf, "attMag": am, "repMag": rm,
            "interactions": inter}


def force_term_sums(g: Graph, pos, p: ForceParams, pool: Pool | None = None,
                    interactions=False):
    """(disp (n, 2), termSum (n,)[, interactions (n,)]): the evaluation of
    force_compute and, per vertex, the sum of the magnitudes of every term of
    its force."""
    n = g.n
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(n, 2)
    mass = vertex_mass(p.model, g.degree)
    deg = np.ascontiguousarray(g.degree, dtype=np.int64)
    disp = np.zeros((n, 2))
    ts = np.zeros(n)
    inter = np.zeros(n, dtype=np.int64) if interactions else None
    lib().oc_force_term_sums(C.byref(g.c), pos.reshape(-1), mass, deg, C.byref(p),
                             pool.h if pool else None, _ptr(disp), _ptr(inter), ts)
    return (disp, ts, inter) if interactions else (disp, ts)


class ForceRun:
    """Stateful FP64 embedder: the oracle of gvizForceEmbedderStep."""

    def __init__(self, g: Graph, p: ForceParams, pos, pool: Pool | None = None):
        self.g = g
        self.p = p
        self.pool = pool
        self.pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 2).copy()
        self.h = lib().oc_force_state_create(C.byref(g.c), C.byref(p))

    def step(self):
        st = StepStats()
        lib().oc_force_step(self.h, C.byref(self.g.c), C.byref(self.p),
                            self.pos.reshape(-1), self.pool.h if self.pool else None,
                            C.byref(st))
        return st

    def close(self):
        if self.h:
            lib().oc_force_state_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def fa2_update(n, jitter, sw, tr, gs, eff):
    g = C.c_double(gs)
    e = C.c_double(eff)
    lib().oc_fa2_update(n, jitter, sw, tr, C.byref(g), C.byref(e))
    return g.value, e.value


def quadtree(points, masses, npc=1):
    """Oracle quadtree. Returns (nodes structured array, per-point leaf
    depth/keys, root params)."""
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1)
    ms = np.ascontiguousarray(masses, dtype=np.float64)
    n = ms.shape[0]
    L = lib()
    h = L.oc_quadtree_build(pts, ms, n, npc)
    try:
        cnt = L.oc_quadtree_node_count(h)
        nodes_p = L.oc_quadtree_nodes(h)
        nodes = np.ctypeslib.as_array(
            C.cast(nodes_p, C.POINTER(C.c_uint8)),
            shape=(cnt * C.sizeof(QNode),)).copy() if cnt else np.zeros(0, np.uint8)
        dt = np.dtype([("cx", "f8"), ("cy", "f8"), ("halfSize", "f8"),
                       ("comX", "f8"), ("comY", "f8"), ("mass", "f8"),
                       ("child", "i4", 4), ("depth", "i4"), ("pad", "i4"),
                       ("ptStart", "i8"), ("ptCount", "i8"), ("keyHi", "u8"),
                       ("keyLo", "u8")])
        nodes = nodes.view(dt)
        ld = np.zeros(n, dtype=np.int32)
        lh = np.zeros(n, dtype=np.uint64)
        ll = np.zeros(n, dtype=np.uint64)
        if n:
            L.oc_quadtree_point_leaves(h, ld, lh, ll)
        cx, cy, hs, md = C.c_double(), C.c_double(), C.c_double(), C.c_int32()
        if n:
            L.oc_quadtree_root_params(h, C.byref(cx), C.byref(cy), C.byref(hs),
                                      C.byref(md))
        lp = L.oc_quadtree_leaf_points(h)
        leafpts = np.ctypeslib.as_array(lp, shape=(n,)).copy() if n else np.zeros(0, np.int64)
        return {"nodes": nodes, "leaf_depth": ld, "leaf_keyHi": lh,
                "leaf_keyLo": ll, "root": (cx.value, cy.value, hs.value, md.value),
                "leaf_points": leafpts}
    finally:
        L.oc_quadtree_free(h)


def path_key(px, py, cx, cy, half, max_depth):
    kh, kl = C.c_uint64(), C.c_uint64()
    lib().oc_path_key(px, py, cx, cy, half, max_depth, C.byref(kh), C.byref(kl))
    return kh.value, kl.value


def tutte_step(g: Graph, is_boundary, pos, scratch, rate=5.0, dt=1.0,
               gauss_seidel=False):
    return lib().oc_tutte_step(C.byref(g.c), is_boundary, pos.reshape(-1),
                               scratch.reshape(-1), rate, dt, int(gauss_seidel))


def tutte_polygon(count, radius):
    out = np.zeros(2 * count)
    lib().oc_tutte_polygon(count, radius, out)
    return out.reshape(count, 2)


def stress(g: Graph, pos, edge_length=10.0, sources=None, pool: Pool | None = None):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    if sources is None:
        return lib().oc_stress(C.byref(g.c), pos, edge_length, None, 0,
                               pool.h if pool else None)
    s = np.ascontiguousarray(sources, dtype=np.int64)
    return lib().oc_stress(C.byref(g.c), pos, edge_length, s.ctypes.data,
                           s.shape[0], pool.h if pool else None)


def edge_energy(g: Graph, pos):
    return lib().oc_edge_energy(C.byref(g.c),
                                np.ascontiguousarray(pos, dtype=np.float64).reshape(-1))


def knn_bfs(g: Graph, source, k, filt=None):
    out = np.full(2 * max(k, 1), -1, dtype=np.int64)
    f = None if filt is None else np.ascontiguousarray(filt, dtype=np.uint8)
    c = lib().oc_knn_bfs(C.byref(g.c), source, k, _ptr(f), out)
    return out.reshape(-1, 2)[:c].copy()


def knn_batch(g: Graph, visible, k, targets):
    visible = np.ascontiguousarray(visible, dtype=np.uint8)
    targets = np.ascontiguousarray(targets, dtype=np.int64)
    out = np.full(2 * k * max(len(targets), 1), -1, dtype=np.int64)
    counts = np.zeros(max(len(targets), 1), dtype=np.int64)
    lib().oc_knn_batch(C.byref(g.c), visible, k, targets, len(targets), out,
                       counts)
    return out.reshape(-1, k, 2)[: len(targets)], counts[: len(targets)]


def mis_filtration(g: Graph, active=None, dim=2):
    filt = np.zeros(g.n, dtype=np.int64)
    borders = np.zeros(256, dtype=np.int64)
    a = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
    layers = lib().oc_mis_filtration(C.byref(g.c), _ptr(a), dim, filt, borders)
    if layers < 0:
        raise ValueError("MIS filtration does not terminate: more than dim+1 components")
    return filt[: borders[0]].copy(), borders[:layers].copy()


def mis_first(g: Graph, active=None):
    filt = np.zeros(g.n, dtype=np.int64)
    inmis = np.zeros(g.n, dtype=np.uint8)
    a = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
    border = lib().oc_mis_first(C.byref(g.c), _ptr(a), filt, inmis)
    return border, filt, inmis


def grip_refine_round(g: Graph, filtration, active, layer, knn, kcap, dec, pos):
    """knn: (n, kcap, 2) int64 lists; dec: GRIP_VERTEX_DTYPE array (n,),
    updated in place with pos (n, 2)."""
    st = GripRoundStats()
    knn = np.ascontiguousarray(knn, dtype=np.int64)
    lib().oc_grip_refine_round(C.byref(g.c),
                               np.ascontiguousarray(filtration, dtype=np.int64),
                               active, layer, knn.reshape(-1), kcap,
                               dec.ctypes.data, pos.reshape(-1), C.byref(st))
    return st


# oc_grip_vertex_nd
GRIP_VERTEX_ND_DTYPE = np.dtype([
    ("disp", "f8", 4), ("oldDisp", "f8", 4), ("heat", "f8"), ("oldCos", "f8"),
    ("lastForceMag", "f8"), ("knnCount", "i8"), ("dispCalculated", "i8")])


def grip_refine_round_nd(g: Graph, filtration, active, layer, knn, kcap, dim, dec, pos):
    """The round in dim = 2, 3 or 4 coordinates: dec a GRIP_VERTEX_ND_DTYPE
    array (n,) and pos (n, dim), both updated in place."""
    assert dec.dtype == GRIP_VERTEX_ND_DTYPE and pos.shape == (g.n, dim)
    assert pos.flags.c_contiguous and pos.dtype == np.float64
    st = GripRoundStats()
    knn = np.ascontiguousarray(knn, dtype=np.int64)
    lib().oc_grip_refine_round_nd(C.byref(g.c),
                                  np.ascontiguousarray(filtration, dtype=np.int64),
                                  active, layer, knn.reshape(-1), kcap, dim,
                                  dec.ctypes.data, pos.reshape(-1), C.byref(st))
    return st


def regular_simplex(n, side):
    out = np.zeros((n + 1) * n)
    lib().oc_regular_simplex(n, side, out)
    return out.reshape(n + 1, n)
