"""gfx950 layout engine for gviz's hot path.

Native pieces (built in-tree by ``python -m grapher_amd.build``):
  libgvizhip.so  HIP kernels + C ABI (include/gviz_hip.h)
  libgviz.so     C11 host layer mirroring the reference's embedder API
Python modules:
  native  ctypes binding of the C ABI
  graphs  seeded graph generators and insertion-order CSR
"""
