"""ctypes binding of libgvizhip.so (include/gviz_hip.h).

This is the Python side of the drop-in boundary: tests and bench.py drive the
HIP engine through exactly the C entry points a C caller would use. Nothing
here computes anything itself, and there is no fallback: if the library is
not built or no GPU is visible, the constructors raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(HERE, "libgvizhip.so")

PRECISION_FP32 = 0
PRECISION_FP64 = 1
PRECISION_MIXED = 2
PRECISIONS = {"fp32": PRECISION_FP32, "fp64": PRECISION_FP64, "mixed": PRECISION_MIXED}

MODEL_FR = 0
MODEL_LINLOG = 1


class GvizHipError(RuntimeError):
    pass


class ForceConfig(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("barnesHut", C.c_int32),
        ("preventOverlap", C.c_int32), ("precision", C.c_int32),
        ("nodesPerCell", C.c_int64), ("edgeLength", C.c_double),
        ("theta", C.c_double), ("gravityK", C.c_double),
        ("overlapConstant", C.c_double), ("radiusBase", C.c_double),
        ("radiusPerDegree", C.c_double), ("jitterTolerance", C.c_double),
    ]


class ForceStats(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "maxDisp", "speed", "meanAtt", "meanRep", "meanGrav",
        "speedEfficiency", "totalSwinging", "totalTraction")]


STATS_DTYPE = np.dtype([(k, "f8") for k, _ in ForceStats._fields_])


class GripVertex(C.Structure):
    _fields_ = [
        ("dispX", C.c_double), ("dispY", C.c_double),
        ("oldDispX", C.c_double), ("oldDispY", C.c_double),
        ("heat", C.c_double), ("oldCos", C.c_double),
        ("lastForceMag", C.c_double), ("knnCount", C.c_int64),
        ("dispCalculated", C.c_int64),
    ]


GRIP_VERTEX_DTYPE = np.dtype([
    ("dispX", "f8"), ("dispY", "f8"), ("oldDispX", "f8"), ("oldDispY", "f8"),
    ("heat", "f8"), ("oldCos", "f8"), ("lastForceMag", "f8"),
    ("knnCount", "i8"), ("dispCalculated", "i8")])
# gvizhip_grip_vertex_nd: disp / oldDisp of up to four coordinates
GRIP_VERTEX_ND_DTYPE = np.dtype([
    ("disp", "f8", 4), ("oldDisp", "f8", 4), ("heat", "f8"), ("oldCos", "f8"),
    ("lastForceMag", "f8"), ("knnCount", "i8"), ("dispCalculated", "i8")])


class GripRoundStats(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("maxDisp", "meanDisp", "meanForce",
                                          "meanHeat")]


_lib = None
_vp = C.c_void_p
_i64 = C.c_int64


def _p(a):
    return None if a is None else a.ctypes.data


def lib():
    """Loads libgvizhip.so, raising if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(HIP_LIB):
        raise GvizHipError(f"{HIP_LIB} is not built (python -m grapher_amd.build)")
    L = C.CDLL(HIP_LIB)
    L.gvizhip_last_error.restype = C.c_char_p
    L.gvizhip_device_count.restype = C.c_int
    sig = {
        "gvizhip_force_create": (C.c_int, [C.POINTER(_vp), _i64, _vp, _vp, C.POINTER(ForceConfig)]),
        "gvizhip_force_configure": (C.c_int, [_vp, C.POINTER(ForceConfig)]),
        "gvizhip_force_begin": (C.c_int, [_vp, _vp]),
        "gvizhip_force_upload_positions": (C.c_int, [_vp, _vp]),
        "gvizhip_force_download_positions": (C.c_int, [_vp, _vp]),
        "gvizhip_force_step": (C.c_int, [_vp, C.POINTER(ForceStats)]),
        "gvizhip_force_run": (_i64, [_vp, _i64, C.c_double, _vp]),
        "gvizhip_force_evaluate": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "gvizhip_force_get_speed": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(_i64)]),
        "gvizhip_force_set_speed": (C.c_int, [_vp, C.c_double, C.c_double]),
        "gvizhip_force_state_bytes": (_i64, [_vp]),
        "gvizhip_force_save_state": (C.c_int, [_vp, _vp, _i64]),
        "gvizhip_force_load_state": (C.c_int, [_vp, _vp, _i64]),
        "gvizhip_force_quadtree_leaves": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "gvizhip_force_quadtree_cells": (_i64, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "gvizhip_force_quadtree_keys": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "gvizhip_force_profile": (C.c_int, [_vp, _i64, C.POINTER(C.c_double), C.POINTER(_i64), C.POINTER(C.c_double)]),
        "gvizhip_qtree_create": (C.c_int, [C.POINTER(_vp)]),
        "gvizhip_qtree_build": (C.c_int, [_vp, _i64, _vp, _vp, _vp, _vp, _vp]),
        "gvizhip_qtree_destroy": (None, [_vp]),
        "gvizhip_force_destroy": (None, [_vp]),
        "gvizhip_tutte_create": (C.c_int, [C.POINTER(_vp), _i64, _vp, _vp, C.c_int32]),
        "gvizhip_tutte_upload": (C.c_int, [_vp, _vp, _vp]),
        "gvizhip_tutte_download": (C.c_int, [_vp, _vp]),
        "gvizhip_tutte_sweeps": (_i64, [_vp, _i64, C.c_double, C.c_double, _vp]),
        "gvizhip_tutte_gs_sweeps": (_i64, [_vp, _i64, C.c_double, C.c_double, _vp]),
        "gvizhip_tutte_spring_steps": (_i64, [_vp, _i64, C.c_double, C.c_double, C.c_double,
                                              C.c_double, _vp]),
        "gvizhip_tutte_set_velocity": (C.c_int, [_vp, _vp]),
        "gvizhip_tutte_get_velocity": (C.c_int, [_vp, _vp]),
        "gvizhip_tutte_profile": (C.c_int, [_vp, _i64, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "gvizhip_tutte_destroy": (None, [_vp]),
        "gvizhip_tutte_sweeps_noalloc": (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp, _i64,
                                                   C.c_double, C.c_double, _vp, _vp, _vp]),
        "gvizhip_grip_create": (C.c_int, [C.POINTER(_vp), _i64, _vp, _vp, _vp, _i64]),
        "gvizhip_grip_mis_filtration": (_i64, [_vp, _i64, _vp, _vp]),
        "gvizhip_grip_set_filtration": (C.c_int, [_vp, _vp, _i64]),
        "gvizhip_grip_mis_first": (_i64, [_vp, _vp, _vp]),
        "gvizhip_grip_mis_layer": (_i64, [_vp, _i64, _vp, _i64, _vp]),
        "gvizhip_grip_migrate": (C.c_int, [_vp, _vp, _i64, _vp, _i64, _i64]),
        "gvizhip_grip_knn": (C.c_int, [_vp, _vp, _vp, _i64, _i64, _vp, _vp, _vp]),
        "gvizhip_grip_knn_batch": (C.c_int, [_vp, _vp, _vp, _i64, _i64, _vp, _vp, _vp]),
        "gvizhip_grip_upload": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "gvizhip_grip_download": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "gvizhip_grip_set_dimension": (C.c_int, [_vp, C.c_int64]),
        "gvizhip_grip_upload_nd": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "gvizhip_grip_download_nd": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "gvizhip_grip_clear_decorators": (C.c_int, [_vp, _i64]),
        "gvizhip_grip_update_knn": (C.c_int, [_vp, _i64, _vp, _i64, C.c_int32]),
        "gvizhip_grip_place": (C.c_int, [_vp, _i64, _i64, _vp, _i64, C.c_int32]),
        "gvizhip_grip_refine_round": (C.c_int, [_vp, _i64, _i64, C.POINTER(GripRoundStats)]),
        "gvizhip_grip_profile_rounds": (C.c_int, [_vp, _i64, _i64, _i64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "gvizhip_grip_destroy": (None, [_vp]),
        "gvizhip_batch_create": (C.c_int, [C.POINTER(_vp), _i64, _vp, _vp, _vp, C.POINTER(ForceConfig)]),
        "gvizhip_batch_begin": (C.c_int, [_vp, _vp]),
        "gvizhip_batch_schedule": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32]),
        "gvizhip_batch_autotune": (C.c_int, [_vp, _i64, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                             C.POINTER(C.c_int32)]),
        "gvizhip_batch_run": (_i64, [_vp, _i64, C.c_double, _vp, _vp, _vp]),
        "gvizhip_batch_profile": (C.c_int, [_vp, _i64, C.POINTER(C.c_double), C.POINTER(_i64), C.POINTER(C.c_double)]),
        "gvizhip_batch_download_positions": (C.c_int, [_vp, _vp]),
        "gvizhip_batch_destroy": (None, [_vp]),
        "gvizhip_stress_create": (C.c_int, [C.POINTER(_vp), _i64, _vp, _vp, _vp, _i64]),
        "gvizhip_stress_evaluate": (C.c_int, [_vp, _vp, C.POINTER(C.c_double)]),
        "gvizhip_stress_destroy": (None, [_vp]),
        "gvizhip_gnn_create": (C.c_int, [C.POINTER(_vp), _i64, _i64]),
        "gvizhip_gnn_param_count": (_i64, []),
        "gvizhip_gnn_set_params": (C.c_int, [_vp, _vp]),
        "gvizhip_gnn_get_params": (C.c_int, [_vp, _vp]),
        "gvizhip_gnn_set_graph": (C.c_int, [_vp, _i64, _vp, _vp, _vp]),
        "gvizhip_gnn_set_graph_diffused": (C.c_int, [_vp, _i64, _vp, _vp, _vp]),
        "gvizhip_gnn_set_pivots": (C.c_int, [_vp, _vp, _i64, _vp]),
        "gvizhip_gnn_set_start": (C.c_int, [_vp, _vp]),
        "gvizhip_gnn_set_start_seeded": (C.c_int, [_vp, C.c_uint64]),
        "gvizhip_gnn_propose": (C.c_int, [_vp, C.c_uint64, C.c_double, _vp]),
        "gvizhip_gnn_forward": (C.c_int, [_vp, _vp]),
        "gvizhip_gnn_gradients": (C.c_int, [_vp, _vp, C.POINTER(C.c_double), _vp]),
        "gvizhip_gnn_train_step": (C.c_int, [_vp, _vp, C.c_double, C.POINTER(C.c_double)]),
        "gvizhip_gnn_destroy": (None, [_vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


# The entry points include/gviz_hip.h declares; tests check every one is
# exported by the built library.
DECLARED = [
    "gvizhip_last_error", "gvizhip_device_count",
    "gvizhip_force_create", "gvizhip_force_configure", "gvizhip_force_begin",
    "gvizhip_force_upload_positions", "gvizhip_force_download_positions",
    "gvizhip_force_step", "gvizhip_force_run", "gvizhip_force_evaluate",
    "gvizhip_force_get_speed", "gvizhip_force_set_speed",
    "gvizhip_force_state_bytes", "gvizhip_force_save_state", "gvizhip_force_load_state",
    "gvizhip_force_quadtree_leaves", "gvizhip_force_quadtree_cells",
    "gvizhip_force_quadtree_keys",
    "gvizhip_qtree_create", "gvizhip_qtree_build", "gvizhip_qtree_destroy",
    "gvizhip_force_profile", "gvizhip_force_destroy",
    "gvizhip_tutte_create", "gvizhip_tutte_upload", "gvizhip_tutte_download",
    "gvizhip_tutte_sweeps", "gvizhip_tutte_profile", "gvizhip_tutte_destroy",
    "gvizhip_tutte_gs_sweeps",
    "gvizhip_tutte_spring_steps", "gvizhip_tutte_set_velocity", "gvizhip_tutte_get_velocity",
    "gvizhip_tutte_sweeps_noalloc",
    "gvizhip_grip_create", "gvizhip_grip_mis_filtration",
    "gvizhip_grip_set_filtration", "gvizhip_grip_knn", "gvizhip_grip_knn_batch",
    "gvizhip_grip_mis_first", "gvizhip_grip_mis_layer", "gvizhip_grip_migrate",
    "gvizhip_grip_upload", "gvizhip_grip_download", "gvizhip_grip_set_dimension",
    "gvizhip_grip_upload_nd", "gvizhip_grip_download_nd",
    "gvizhip_grip_clear_decorators", "gvizhip_grip_update_knn",
    "gvizhip_grip_place", "gvizhip_grip_refine_round", "gvizhip_grip_profile_rounds",
    "gvizhip_grip_destroy",
    "gvizhip_batch_create", "gvizhip_batch_begin", "gvizhip_batch_schedule",
    "gvizhip_batch_autotune", "gvizhip_batch_run", "gvizhip_batch_profile",
    "gvizhip_batch_download_positions",
    "gvizhip_batch_destroy",
    "gvizhip_stress_create", "gvizhip_stress_evaluate", "gvizhip_stress_destroy",
    "gvizhip_gnn_create", "gvizhip_gnn_param_count", "gvizhip_gnn_set_params",
    "gvizhip_gnn_get_params", "gvizhip_gnn_set_graph", "gvizhip_gnn_set_graph_diffused",
    "gvizhip_gnn_set_pivots", "gvizhip_gnn_set_start", "gvizhip_gnn_set_start_seeded",
    "gvizhip_gnn_propose", "gvizhip_gnn_forward",
    "gvizhip_gnn_gradients", "gvizhip_gnn_train_step", "gvizhip_gnn_destroy",
]


def _err(what):
    msg = lib().gvizhip_last_error()
    raise GvizHipError(f"{what}: {msg.decode() if msg else 'unknown error'}")


def device_count() -> int:
    return int(lib().gvizhip_device_count())


def _csr(offsets, adj):
    o = np.ascontiguousarray(offsets, dtype=np.int64)
    a = np.ascontiguousarray(adj, dtype=np.int64)
    if a.size == 0:
        a = np.zeros(1, dtype=np.int64)
    return o, a


class ForceEngine:
    """One device embedder: gvizhip_force_* on a CSR graph."""

    def __init__(self, offsets, adj, *, model=MODEL_FR, barnes_hut=True,
                 prevent_overlap=False, precision="fp32", nodes_per_cell=1,
                 edge_length=10.0, theta=1.0, gravity=0.0,
                 overlap_constant=100.0, radius_base=0.0,
                 radius_per_degree=0.0, jitter=1.0):
        self._o, self._a = _csr(offsets, adj)
        self.n = self._o.shape[0] - 1
        self.cfg = ForceConfig(model, int(barnes_hut), int(prevent_overlap),
                               PRECISIONS.get(precision, precision), nodes_per_cell,
                               edge_length, theta, gravity, overlap_constant,
                               radius_base, radius_per_degree, jitter)
        h = _vp()
        if lib().gvizhip_force_create(C.byref(h), self.n, _p(self._o), _p(self._a),
                                      C.byref(self.cfg)) != 0:
            _err("gvizhip_force_create")
        self.h = h

    def configure(self, **kw):
        names = {"model": "model", "barnes_hut": "barnesHut",
                 "prevent_overlap": "preventOverlap", "precision": "precision",
                 "nodes_per_cell": "nodesPerCell", "edge_length": "edgeLength",
                 "theta": "theta", "gravity": "gravityK",
                 "overlap_constant": "overlapConstant", "radius_base": "radiusBase",
                 "radius_per_degree": "radiusPerDegree", "jitter": "jitterTolerance"}
        for k, v in kw.items():
            if k == "precision":
                v = PRECISIONS.get(v, v)
            setattr(self.cfg, names[k], int(v) if isinstance(v, bool) else v)
        if lib().gvizhip_force_configure(self.h, C.byref(self.cfg)) != 0:
            _err("gvizhip_force_configure")

    def begin(self, positions):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(self.n, 2)
        if lib().gvizhip_force_begin(self.h, _p(pos)) != 0:
            _err("gvizhip_force_begin")

    def set_positions(self, positions):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(self.n, 2)
        if lib().gvizhip_force_upload_positions(self.h, _p(pos)) != 0:
            _err("gvizhip_force_upload_positions")

    def positions(self):
        out = np.zeros((self.n, 2))
        if lib().gvizhip_force_download_positions(self.h, _p(out)) != 0:
            _err("gvizhip_force_download_positions")
        return out

    def step(self):
        st = ForceStats()
        if lib().gvizhip_force_step(self.h, C.byref(st)) != 0:
            _err("gvizhip_force_step")
        return st

    def run(self, iters, epsilon=-1.0, stats=True):
        out = np.zeros(max(iters, 1), dtype=STATS_DTYPE) if stats else None
        r = lib().gvizhip_force_run(self.h, iters, epsilon, _p(out))
        if r < 0:
            _err("gvizhip_force_run")
        return (out[:r] if stats else None), int(r)

    def evaluate(self, interactions=False):
        n = self.n
        disp = np.zeros((n, 2))
        sf = np.zeros((n, 2))
        am = np.zeros(n)
        rm = np.zeros(n)
        it = np.zeros(n, dtype=np.int32) if interactions else None
        if lib().gvizhip_force_evaluate(self.h, _p(disp), _p(sf), _p(am), _p(rm),
                                        _p(it)) != 0:
            _err("gvizhip_force_evaluate")
        out = {"disp": disp, "structForce": sf, "attMag": am, "repMag": rm}
        if interactions:
            out["interactions"] = it
        return out

    def speed(self):
        g, e, it = C.c_double(), C.c_double(), _i64()
        if lib().gvizhip_force_get_speed(self.h, C.byref(g), C.byref(e), C.byref(it)) != 0:
            _err("gvizhip_force_get_speed")
        return g.value, e.value, it.value

    def set_speed(self, gs, eff):
        if lib().gvizhip_force_set_speed(self.h, gs, eff) != 0:
            _err("gvizhip_force_set_speed")

    def save_state(self):
        """Exact checkpoint (positions, previous structForce, speed state,
        step count) as bytes."""
        size = lib().gvizhip_force_state_bytes(self.h)
        if size < 0:
            _err("gvizhip_force_state_bytes")
        buf = C.create_string_buffer(int(size))
        if lib().gvizhip_force_save_state(self.h, buf, size) != 0:
            _err("gvizhip_force_save_state")
        return buf.raw

    def load_state(self, state):
        buf = C.create_string_buffer(bytes(state), len(state))
        if lib().gvizhip_force_load_state(self.h, buf, len(state)) != 0:
            _err("gvizhip_force_load_state")

    def quadtree_leaves(self):
        n = self.n
        d = np.zeros(n, dtype=np.int32)
        kh = np.zeros(n, dtype=np.uint64)
        kl = np.zeros(n, dtype=np.uint64)
        root = np.zeros(4)
        if lib().gvizhip_force_quadtree_leaves(self.h, _p(d), _p(kh), _p(kl), _p(root)) != 0:
            _err("gvizhip_force_quadtree_leaves")
        return d, kh, kl, root

    def quadtree_cells(self):
        cap = max(self.n, 1)
        d = np.zeros(cap, dtype=np.int32)
        kh = np.zeros(cap, dtype=np.uint64)
        kl = np.zeros(cap, dtype=np.uint64)
        m = np.zeros(cap)
        cx = np.zeros(cap)
        cy = np.zeros(cap)
        cnt = np.zeros(cap, dtype=np.int64)
        k = lib().gvizhip_force_quadtree_cells(self.h, cap, _p(d), _p(kh), _p(kl),
                                               _p(m), _p(cx), _p(cy), _p(cnt))
        if k < 0:
            _err("gvizhip_force_quadtree_cells")
        k = min(k, cap)
        return {"depth": d[:k], "keyHi": kh[:k], "keyLo": kl[:k], "mass": m[:k],
                "comX": cx[:k], "comY": cy[:k], "count": cnt[:k]}

    def profile(self, steps):
        """(traversal ms, repulsion evaluations, step ms) summed over steps."""
        t, i, s = C.c_double(), _i64(), C.c_double()
        if lib().gvizhip_force_profile(self.h, steps, C.byref(t), C.byref(i),
                                       C.byref(s)) != 0:
            _err("gvizhip_force_profile")
        return t.value, i.value, s.value

    def close(self):
        if getattr(self, "h", None):
            lib().gvizhip_force_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class TutteEngine:
    """Device Jacobi sweeps: gvizhip_tutte_*."""

    def __init__(self, offsets, adj, precision="fp64"):
        self._o, self._a = _csr(offsets, adj)
        self.n = self._o.shape[0] - 1
        h = _vp()
        if lib().gvizhip_tutte_create(C.byref(h), self.n, _p(self._o), _p(self._a),
                                      PRECISIONS.get(precision, precision)) != 0:
            _err("gvizhip_tutte_create")
        self.h = h

    def upload(self, positions, is_boundary):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(self.n, 2)
        b = np.ascontiguousarray(is_boundary, dtype=np.uint8)
        if lib().gvizhip_tutte_upload(self.h, _p(pos), _p(b)) != 0:
            _err("gvizhip_tutte_upload")

    def positions(self):
        out = np.zeros((self.n, 2))
        if lib().gvizhip_tutte_download(self.h, _p(out)) != 0:
            _err("gvizhip_tutte_download")
        return out

    def sweeps(self, iters, alpha=1.0, epsilon=-1.0, record=True):
        out = np.zeros(max(iters, 1)) if record else None
        r = lib().gvizhip_tutte_sweeps(self.h, iters, alpha, epsilon, _p(out))
        if r < 0:
            _err("gvizhip_tutte_sweeps")
        return (out[:r] if record else None), int(r)

    def gs_sweeps(self, iters, alpha=1.0, epsilon=-1.0, record=True):
        """Gauss-Seidel sweeps (gvizhip_tutte_gs_sweeps): (max deltas, sweeps)."""
        out = np.zeros(max(iters, 1)) if record else None
        r = lib().gvizhip_tutte_gs_sweeps(self.h, iters, alpha, epsilon, _p(out))
        if r < 0:
            _err("gvizhip_tutte_gs_sweeps")
        return (out[:r] if record else None), int(r)

    def spring_steps(self, iters, dt, stiffness=30.0, damping=6.0, epsilon=-1.0,
                     record=True):
        """Spring-driven steps (gvizhip_tutte_spring_steps): (max deltas, steps)."""
        out = np.zeros(max(iters, 1)) if record else None
        r = lib().gvizhip_tutte_spring_steps(self.h, iters, dt, stiffness, damping, epsilon,
                                             _p(out))
        if r < 0:
            _err("gvizhip_tutte_spring_steps")
        return (out[:r] if record else None), int(r)

    def set_velocity(self, velocity=None):
        v = None if velocity is None else np.ascontiguousarray(
            velocity, dtype=np.float64).reshape(self.n, 2)
        if lib().gvizhip_tutte_set_velocity(self.h, _p(v)) != 0:
            _err("gvizhip_tutte_set_velocity")

    def velocity(self):
        out = np.zeros((self.n, 2))
        if lib().gvizhip_tutte_get_velocity(self.h, _p(out)) != 0:
            _err("gvizhip_tutte_get_velocity")
        return out

    def profile(self, sweeps, alpha=1.0):
        """(summed sweep-kernel ms, total ms) over `sweeps` sweeps."""
        a, b = C.c_double(), C.c_double()
        if lib().gvizhip_tutte_profile(self.h, sweeps, alpha, C.byref(a), C.byref(b)) != 0:
            _err("gvizhip_tutte_profile")
        return a.value, b.value

    def close(self):
        if getattr(self, "h", None):
            lib().gvizhip_tutte_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class GripEngine:
    """Device GRIP state: gvizhip_grip_*."""

    def __init__(self, offsets, adj, active=None, kcap=10, dim=2):
        self._o, self._a = _csr(offsets, adj)
        self.n = self._o.shape[0] - 1
        self.kcap = kcap
        self.dim = dim
        act = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
        h = _vp()
        if lib().gvizhip_grip_create(C.byref(h), self.n, _p(self._o), _p(self._a),
                                     _p(act), kcap) != 0:
            _err("gvizhip_grip_create")
        self.h = h
        if dim != 2 and lib().gvizhip_grip_set_dimension(h, dim) != 0:
            _err("gvizhip_grip_set_dimension")

    def mis_filtration(self, dim=2):
        filt = np.zeros(self.n, dtype=np.int64)
        borders = np.zeros(128, dtype=np.int64)
        layers = lib().gvizhip_grip_mis_filtration(self.h, dim, _p(filt), _p(borders))
        if layers < 0:
            _err("gvizhip_grip_mis_filtration")
        return filt[: borders[0]].copy(), borders[:layers].copy()

    def set_filtration(self, filt):
        f = np.ascontiguousarray(filt, dtype=np.int64)
        if lib().gvizhip_grip_set_filtration(self.h, _p(f), f.shape[0]) != 0:
            _err("gvizhip_grip_set_filtration")

    def knn(self, sources, k, visible=None):
        s = np.ascontiguousarray(sources, dtype=np.int64)
        vis = None if visible is None else np.ascontiguousarray(visible, dtype=np.uint8)
        ov = np.zeros(max(len(s) * k, 1), dtype=np.int64)
        od = np.zeros(max(len(s) * k, 1), dtype=np.int64)
        cnt = np.zeros(max(len(s), 1), dtype=np.int64)
        if lib().gvizhip_grip_knn(self.h, _p(vis), _p(s), len(s), k, _p(ov), _p(od),
                                  _p(cnt)) != 0:
            _err("gvizhip_grip_knn")
        return ov[: len(s) * k].reshape(len(s), k), od[: len(s) * k].reshape(len(s), k), cnt[: len(s)]

    def knn_batch(self, targets, k, visible):
        t = np.ascontiguousarray(targets, dtype=np.int64)
        vis = np.ascontiguousarray(visible, dtype=np.uint8)
        ov = np.zeros(max(len(t) * k, 1), dtype=np.int64)
        od = np.zeros(max(len(t) * k, 1), dtype=np.int64)
        cnt = np.zeros(max(len(t), 1), dtype=np.int64)
        if lib().gvizhip_grip_knn_batch(self.h, _p(vis), _p(t), len(t), k, _p(ov),
                                        _p(od), _p(cnt)) != 0:
            _err("gvizhip_grip_knn_batch")
        return ov[: len(t) * k].reshape(len(t), k), od[: len(t) * k].reshape(len(t), k), cnt[: len(t)]

    def upload(self, positions=None, dec=None, knn_vertex=None, knn_dist=None):
        """Positions (n, dim); dec in GRIP_VERTEX_DTYPE for dim 2, in
        GRIP_VERTEX_ND_DTYPE for dim 3 and 4."""
        pos = None if positions is None else np.ascontiguousarray(positions, dtype=np.float64)
        kv = None if knn_vertex is None else np.ascontiguousarray(knn_vertex, dtype=np.int64)
        kd = None if knn_dist is None else np.ascontiguousarray(knn_dist, dtype=np.int64)
        if self.dim != 2:
            d = None if dec is None else np.ascontiguousarray(dec, dtype=GRIP_VERTEX_ND_DTYPE)
            if lib().gvizhip_grip_upload_nd(self.h, _p(pos), _p(d), _p(kv), _p(kd)) != 0:
                _err("gvizhip_grip_upload_nd")
            return
        d = None if dec is None else np.ascontiguousarray(dec, dtype=GRIP_VERTEX_DTYPE)
        if lib().gvizhip_grip_upload(self.h, _p(pos), _p(d), _p(kv), _p(kd)) != 0:
            _err("gvizhip_grip_upload")

    def download(self):
        pos = np.zeros((self.n, self.dim))
        kv = np.zeros((self.n, self.kcap), dtype=np.int64)
        kd = np.zeros((self.n, self.kcap), dtype=np.int64)
        if self.dim != 2:
            dec = np.zeros(self.n, dtype=GRIP_VERTEX_ND_DTYPE)
            if lib().gvizhip_grip_download_nd(self.h, _p(pos), _p(dec), _p(kv), _p(kd)) != 0:
                _err("gvizhip_grip_download_nd")
            return pos, dec, kv, kd
        dec = np.zeros(self.n, dtype=GRIP_VERTEX_DTYPE)
        if lib().gvizhip_grip_download(self.h, _p(pos), _p(dec), _p(kv), _p(kd)) != 0:
            _err("gvizhip_grip_download")
        return pos, dec, kv, kd

    def clear_decorators(self, active):
        if lib().gvizhip_grip_clear_decorators(self.h, active) != 0:
            _err("gvizhip_grip_clear_decorators")

    def update_knn(self, active, visible, k, batch=False):
        vis = np.ascontiguousarray(visible, dtype=np.uint8)
        if lib().gvizhip_grip_update_knn(self.h, active, _p(vis), k, int(batch)) != 0:
            _err("gvizhip_grip_update_knn")

    def place(self, begin, end, visible, k, batch=False):
        vis = np.ascontiguousarray(visible, dtype=np.uint8)
        if lib().gvizhip_grip_place(self.h, begin, end, _p(vis), k, int(batch)) != 0:
            _err("gvizhip_grip_place")

    def refine_round(self, active, layer):
        st = GripRoundStats()
        if lib().gvizhip_grip_refine_round(self.h, active, layer, C.byref(st)) != 0:
            _err("gvizhip_grip_refine_round")
        return st

    def profile_rounds(self, active, layer, rounds):
        """(summed spring-force kernel ms, summed round ms)."""
        a, b = C.c_double(), C.c_double()
        if lib().gvizhip_grip_profile_rounds(self.h, active, layer, rounds, C.byref(a),
                                             C.byref(b)) != 0:
            _err("gvizhip_grip_profile_rounds")
        return a.value, b.value

    def close(self):
        if getattr(self, "h", None):
            lib().gvizhip_grip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class GnnEngine:
    """The warm-start network on the device: gvizhip_gnn_*."""

    def __init__(self, max_vertices, max_entries):
        h = _vp()
        if lib().gvizhip_gnn_create(C.byref(h), int(max_vertices), int(max_entries)) != 0:
            _err("gvizhip_gnn_create")
        self.h = h
        self.count = int(lib().gvizhip_gnn_param_count())
        self.n = 0

    def set_params(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        assert t.shape == (self.count,)
        if lib().gvizhip_gnn_set_params(self.h, _p(t)) != 0:
            _err("gvizhip_gnn_set_params")

    def params(self):
        t = np.zeros(self.count, dtype=np.float32)
        if lib().gvizhip_gnn_get_params(self.h, _p(t)) != 0:
            _err("gvizhip_gnn_get_params")
        return t

    def set_graph(self, offsets, adj, features):
        o = np.ascontiguousarray(offsets, dtype=np.int64)
        a = np.ascontiguousarray(adj, dtype=np.int64)
        f = np.ascontiguousarray(features, dtype=np.float32)
        n = o.shape[0] - 1
        assert f.shape == (n, 16)
        if lib().gvizhip_gnn_set_graph(self.h, n, _p(o), _p(a) if a.size else None,
                                       _p(f)) != 0:
            _err("gvizhip_gnn_set_graph")
        self.n = n

    def set_graph_diffused(self, offsets, adj, start):
        o = np.ascontiguousarray(offsets, dtype=np.int64)
        a = np.ascontiguousarray(adj, dtype=np.int64)
        n = o.shape[0] - 1
        st = np.ascontiguousarray(start, dtype=np.float64).reshape(n, 2)
        if lib().gvizhip_gnn_set_graph_diffused(self.h, n, _p(o), _p(a) if a.size else None,
                                                _p(st)) != 0:
            _err("gvizhip_gnn_set_graph_diffused")
        self.n = n

    def set_start(self, start):
        """New diffused features of the current graph from `start` (n x 2)."""
        st = np.ascontiguousarray(start, dtype=np.float64).reshape(self.n, 2)
        if lib().gvizhip_gnn_set_start(self.h, _p(st)) != 0:
            _err("gvizhip_gnn_set_start")

    def set_start_seeded(self, seed):
        """New diffused features of the current graph from a start drawn on
        the device from `seed`."""
        if lib().gvizhip_gnn_set_start_seeded(self.h, int(seed) & (2**64 - 1)) != 0:
            _err("gvizhip_gnn_set_start_seeded")

    def propose(self, seed, radius):
        """Network output from a device-drawn start, centred and scaled to RMS
        radius `radius` on the device (n x 2)."""
        out = np.empty((self.n, 2))
        if lib().gvizhip_gnn_propose(self.h, int(seed) & (2**64 - 1), float(radius),
                                     _p(out)) != 0:
            _err("gvizhip_gnn_propose")
        return out

    def set_pivots(self, pivots):
        """Pivot-MDS features and skip term from BFS distances to `pivots`
        (at most 64 vertex ids; empty removes the skip). Returns the
        normalised coordinates (n x 2)."""
        p = np.ascontiguousarray(pivots, dtype=np.int64)
        out = np.zeros((self.n, 2))
        if lib().gvizhip_gnn_set_pivots(self.h, _p(p) if p.size else None, int(p.size),
                                        _p(out)) != 0:
            _err("gvizhip_gnn_set_pivots")
        return out

    def forward(self):
        out = np.zeros((self.n, 2))
        if lib().gvizhip_gnn_forward(self.h, _p(out)) != 0:
            _err("gvizhip_gnn_forward")
        return out

    def gradients(self, target):
        t = np.ascontiguousarray(target, dtype=np.float64)
        loss = C.c_double()
        g = np.zeros(self.count, dtype=np.float32)
        if lib().gvizhip_gnn_gradients(self.h, _p(t), C.byref(loss), _p(g)) != 0:
            _err("gvizhip_gnn_gradients")
        return loss.value, g

    def train_step(self, target, lr):
        t = np.ascontiguousarray(target, dtype=np.float64)
        loss = C.c_double()
        if lib().gvizhip_gnn_train_step(self.h, _p(t), float(lr), C.byref(loss)) != 0:
            _err("gvizhip_gnn_train_step")
        return loss.value

    def close(self):
        if getattr(self, "h", None):
            lib().gvizhip_gnn_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class StressMeter:
    """Pivot-based normalised stress on the device: gvizhip_stress_*."""

    def __init__(self, offsets, adj, pivots):
        o = np.ascontiguousarray(offsets, dtype=np.int64)
        a = np.ascontiguousarray(adj, dtype=np.int64)
        p = np.ascontiguousarray(pivots, dtype=np.int64)
        self.n = o.shape[0] - 1
        h = _vp()
        if lib().gvizhip_stress_create(C.byref(h), self.n, _p(o), _p(a) if a.size else None,
                                       _p(p), p.shape[0]) != 0:
            _err("gvizhip_stress_create")
        self.h = h

    def __call__(self, positions):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(self.n, 2)
        out = C.c_double()
        if lib().gvizhip_stress_evaluate(self.h, _p(pos), C.byref(out)) != 0:
            _err("gvizhip_stress_evaluate")
        return out.value

    def close(self):
        if getattr(self, "h", None):
            lib().gvizhip_stress_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class BatchEngine:
    """Many independent graphs laid out together: gvizhip_batch_*.

    graphs: list of (offsets, adj) CSR pairs over each graph's own ids; they
    are concatenated in order (graph g's vertex i becomes start[g] + i)."""

    def __init__(self, graphs, *, model=MODEL_FR, prevent_overlap=False,
                 precision="fp32", nodes_per_cell=1, edge_length=10.0, theta=1.0,
                 gravity=0.0, overlap_constant=100.0, radius_base=0.0,
                 radius_per_degree=0.0, jitter=1.0):
        sizes = [len(o) - 1 for o, _ in graphs]
        self.start = np.zeros(len(graphs) + 1, dtype=np.int64)
        self.start[1:] = np.cumsum(sizes)
        offs, adjs, base_e = [np.zeros(1, np.int64)], [], 0
        for (o, a), s in zip(graphs, self.start[:-1]):
            o = np.asarray(o, dtype=np.int64)
            offs.append(o[1:] + base_e)
            adjs.append(np.asarray(a, dtype=np.int64) + s)
            base_e += int(o[-1])
        self._o = np.ascontiguousarray(np.concatenate(offs))
        self._a = np.ascontiguousarray(np.concatenate(adjs) if adjs else np.zeros(0, np.int64))
        self.n = int(self.start[-1])
        self.G = len(graphs)
        self.cfg = ForceConfig(model, 1, int(prevent_overlap),
                               PRECISIONS.get(precision, precision), nodes_per_cell,
                               edge_length, theta, gravity, overlap_constant,
                               radius_base, radius_per_degree, jitter)
        h = _vp()
        if lib().gvizhip_batch_create(C.byref(h), self.G, _p(self.start), _p(self._o),
                                      _p(self._a) if self._a.size else None,
                                      C.byref(self.cfg)) != 0:
            _err("gvizhip_batch_create")
        self.h = h

    def begin(self, positions):
        p = np.ascontiguousarray(positions, dtype=np.float64).reshape(self.n, 2)
        if lib().gvizhip_batch_begin(self.h, _p(p)) != 0:
            _err("gvizhip_batch_begin")

    def schedule(self, wave_order=0, steps_per_launch=4, traversal_waves=1):
        if lib().gvizhip_batch_schedule(self.h, wave_order, steps_per_launch,
                                        traversal_waves) != 0:
            _err("gvizhip_batch_schedule")

    def autotune(self, trial_steps=8):
        """(wave order, steps per launch, traversal waves per workgroup) of the
        fastest schedule, which the engine keeps."""
        o, s, t = C.c_int32(), C.c_int32(), C.c_int32()
        if lib().gvizhip_batch_autotune(self.h, trial_steps, C.byref(o), C.byref(s),
                                        C.byref(t)) != 0:
            _err("gvizhip_batch_autotune")
        return o.value, s.value, t.value

    def run(self, iters, epsilon=-1.0):
        """(steps launched, per-graph steps, per-graph latency ms, last maxDisp)."""
        it = np.zeros(self.G, dtype=np.int64)
        lat = np.zeros(self.G)
        md = np.zeros(self.G)
        r = lib().gvizhip_batch_run(self.h, iters, epsilon, _p(it), _p(lat), _p(md))
        if r < 0:
            _err("gvizhip_batch_run")
        return int(r), it, lat, md

    def profile(self, steps):
        """(traversal ms, repulsion evaluations, step ms) summed over steps."""
        t, i, s = C.c_double(), _i64(), C.c_double()
        if lib().gvizhip_batch_profile(self.h, steps, C.byref(t), C.byref(i),
                                       C.byref(s)) != 0:
            _err("gvizhip_batch_profile")
        return t.value, i.value, s.value

    def positions(self):
        out = np.zeros((self.n, 2))
        if lib().gvizhip_batch_download_positions(self.h, _p(out)) != 0:
            _err("gvizhip_batch_download_positions")
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().gvizhip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
