""".edges files (Network Repository format) straight into the CSR arrays the
engines take, through the host layer's gvizGraphLoadEdgesFileCSR
(include/gviz/gvizGraphLoader.h): the reference's parsing rules and id
compaction (src/utils/graphLoader.c:790-825), each row in insertion order.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(HERE, "libgviz.so")

_lib = None


class _Options(C.Structure):
    _fields_ = [("directed", C.c_int), ("skip_header", C.c_int)]


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is not built: python -m grapher_amd.build")
        lib = C.CDLL(HOST_LIB)
        p64 = C.POINTER(C.c_int64)
        lib.gvizGraphLoadEdgesFileCSR.argtypes = [
            C.c_char_p, C.POINTER(_Options), C.POINTER(C.c_int64),
            C.POINTER(p64), C.POINTER(p64)]
        lib.gvizGraphLoadEdgesFileCSR.restype = C.c_int
        _lib = lib
    return _lib


def load_csr(path, directed=False, skip_header=False):
    """(offsets[n+1], adj) as int64 arrays; RuntimeError when the file cannot
    be read or parsed."""
    lib = _load()
    opts = _Options(int(bool(directed)), int(bool(skip_header)))
    n = C.c_int64()
    off = C.POINTER(C.c_int64)()
    adj = C.POINTER(C.c_int64)()
    if lib.gvizGraphLoadEdgesFileCSR(os.fsencode(path), C.byref(opts), C.byref(n),
                                     C.byref(off), C.byref(adj)) != 0:
        raise RuntimeError(f"cannot load {path}")
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    try:
        offsets = np.ctypeslib.as_array(off, shape=(n.value + 1,)).copy()
        m = int(offsets[-1])
        a = np.ctypeslib.as_array(adj, shape=(m,)).copy() if m else np.zeros(0, np.int64)
    finally:
        libc.free(C.cast(off, C.c_void_p))
        libc.free(C.cast(adj, C.c_void_p))
    return offsets, a
