"""The entry point that allocates nothing (gvizhip_tutte_sweeps_noalloc)
against the one it varies (gvizhip_tutte_sweeps): FP64 positions and every
sweep's max displacement bit-identical, on two calls in a row, the second on
a smaller graph and stopping early at epsilon. The test owns every device
buffer (hipMalloc through the HIP runtime); the entry point gets only
pointers to them."""
import ctypes as C
import glob
import os

import numpy as np
import pytest

from grapher_amd import graphs as G

pytestmark = pytest.mark.gpu


def _hip():
    for cand in ["/opt/rocm/lib/libamdhip64.so"] + sorted(glob.glob("/opt/rocm*/lib/libamdhip64.so*")):
        if os.path.exists(cand):
            h = C.CDLL(cand)
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            return h
    raise RuntimeError("libamdhip64.so not found")


H2D, D2H = 1, 2


class DeviceArrays:
    """Device copies of host arrays, freed together."""

    def __init__(self, hip):
        self.hip = hip
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        if out.nbytes:
            assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _case(seed, n):
    n, e, _, hull = G.delaunay(n, seed)
    off, adj = G.csr_from_edges(n, e)
    ring = np.stack([200 * np.cos(2 * np.pi * np.arange(len(hull)) / len(hull)),
                     200 * np.sin(2 * np.pi * np.arange(len(hull)) / len(hull))], 1)
    pos = np.zeros((n, 2))
    pos[:] = ring.mean(0)
    pos[hull] = ring
    isb = np.zeros(n, np.uint8)
    isb[hull] = 1
    return off, adj, pos, isb


def _noalloc(native, hip, off, adj, pos, isb, iters, eps):
    L = native.lib()
    d = DeviceArrays(hip)
    try:
        n = len(off) - 1
        d_off = d.put(off.astype(np.int32))
        d_adj = d.put(adj.astype(np.int32))
        d_b = d.put(isb)
        d_a = d.put(pos)
        d_bb = d.put(np.zeros_like(pos))
        md0 = np.full(max(iters, 1), -7.0)
        d_md = d.put(md0)
        d_cnt = d.put(np.array([-1], np.int64))
        rc = L.gvizhip_tutte_sweeps_noalloc(n, d_off, d_adj, d_b, d_a, d_bb, iters, 1.0,
                                            eps, d_md, d_cnt, None)
        assert rc == 0, native.lib().gvizhip_last_error()
        cnt = int(d.get(d_cnt, np.zeros(1, np.int64))[0])
        md = d.get(d_md, md0)
        out = d.get(d_a, pos)
        return out, md, cnt
    finally:
        d.free()


def test_noalloc_sweeps_equal_engine_sweeps(gpu):
    hip = _hip()
    for seed, n, iters, eps in ((1, 3000, 301, -1.0), (2, 1200, 500, 0.3)):
        off, adj, pos, isb = _case(seed, n)
        got, md, cnt = _noalloc(gpu, hip, off, adj, pos, isb, iters, eps)
        eng = gpu.TutteEngine(off, adj, "fp64")
        eng.upload(pos, isb)
        ref_d, ran = eng.sweeps(iters, 1.0, eps)
        ref = eng.positions()
        eng.close()
        print(f"seed {seed}: n={len(off) - 1}, sweeps {cnt} (engine {ran})", flush=True)
        assert cnt == ran
        if eps >= 0:
            assert ran < iters  # the stop test ended the call early
        assert np.array_equal(md[:cnt], ref_d)
        assert np.all(md[cnt:] == -7.0)  # entries past the last sweep untouched
        assert np.array_equal(got, ref)


def test_noalloc_rejects_invalid_arguments(gpu):
    L = gpu.lib()
    assert L.gvizhip_tutte_sweeps_noalloc(10, None, None, None, None, None, 5, 1.0, -1.0,
                                          None, None, None) == -1
    assert b"invalid" in L.gvizhip_last_error()
    assert L.gvizhip_tutte_sweeps_noalloc(1 << 31, C.c_void_p(8), C.c_void_p(8),
                                          C.c_void_p(8), C.c_void_p(8), C.c_void_p(8),
                                          0, 1.0, -1.0, None, None, None) == -1
    assert b"32-bit" in L.gvizhip_last_error()
