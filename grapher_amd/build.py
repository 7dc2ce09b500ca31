"""Builds the native libraries in-tree:

  grapher_amd/libgvizhip.so  hand-written HIP kernels + C ABI (include/gviz_hip.h)
  grapher_amd/libgviz.so     C11 host layer with the reference's embedder API
                             (include/gviz/*.h), linked against libgvizhip.so

Run:  python -m grapher_amd.build   (or __graft_entry__.build()).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
HIP_LIB = os.path.join(HERE, "libgvizhip.so")
HOST_LIB = os.path.join(HERE, "libgviz.so")

ARCH = os.environ.get("GVIZ_OFFLOAD_ARCH", "gfx950")

# Translation units that must round like the reference (FP64 paths, the
# double-double prefix sums): no FMA contraction.
NO_CONTRACT = {"force_common.hip", "force_f64.hip", "tutte.hip", "grip.hip",
               "batch_api.hip"}


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def hip_sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def host_sources():
    if not os.path.isdir(HOST):
        return []
    return sorted(f for f in os.listdir(HOST) if f.endswith(".c"))


def build(verbose: bool = False) -> None:
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    base = [cc, "-c", "-fPIC", "-O3", "-std=c++17", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-I", os.path.join(ROOT, "include")]
    jobs = []
    objs = []
    for src in hip_sources():
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        flags = ["-ffp-contract=off"] if src in NO_CONTRACT else []
        cmd = base + flags + [os.path.join(CSRC, src), "-o", obj]
        jobs.append(cmd)
        objs.append(obj)
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=max(workers, 1)) as ex:
        for r in ex.map(_run, jobs):
            if verbose and r.stderr.strip():
                print(r.stderr, file=sys.stderr)
    tmp = HIP_LIB + ".tmp"
    _run([cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, HIP_LIB)

    hs = host_sources()
    if hs:
        cmd = ["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", "-fPIC", "-shared", "-Wall",
               "-Wno-unused-function", "-pthread",
               "-I", os.path.join(ROOT, "include"),
               "-o", HOST_LIB + ".tmp"] + [os.path.join(HOST, s) for s in hs] + [
                   "-L", HERE, "-lgvizhip", "-Wl,-rpath,$ORIGIN", "-lm"]
        _run(cmd)
        os.replace(HOST_LIB + ".tmp", HOST_LIB)
        build_test_driver()


TEST_DRIVER_SRC = os.path.join(ROOT, "tests", "host", "api_driver.c")
TEST_DRIVER = os.path.join(ROOT, "tests", "host", "api_driver")


def build_test_driver() -> None:
    """The tests' C driver of the host layer (tests/host/api_driver.c)."""
    if not os.path.exists(TEST_DRIVER_SRC):
        return
    _run(["gcc", "-O2", "-std=gnu11", "-Wall", "-I", os.path.join(ROOT, "include"),
          "-o", TEST_DRIVER + ".tmp", TEST_DRIVER_SRC, "-L", HERE, "-lgviz",
          "-lgvizhip", "-Wl,-rpath,$ORIGIN/../../grapher_amd", "-lm"])
    os.replace(TEST_DRIVER + ".tmp", TEST_DRIVER)


if __name__ == "__main__":
    build(verbose=True)
    print("built", HIP_LIB, HOST_LIB if host_sources() else "")
