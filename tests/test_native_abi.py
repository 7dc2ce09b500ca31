"""The C-ABI library loads and exports every entry point include/gviz_hip.h
declares (no compute: runs without a GPU)."""
import ctypes
import os
import re

from grapher_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_in_header():
    text = open(os.path.join(ROOT, "include", "gviz_hip.h")).read()
    return sorted(set(re.findall(r"\b(gvizhip_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert sorted(native.DECLARED) == _declared_in_header()


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(native.HIP_LIB)
    missing = [s for s in _declared_in_header() if not hasattr(lib, s)]
    assert not missing, missing


def test_device_count_is_callable():
    assert native.device_count() >= 0
