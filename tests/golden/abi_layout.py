"""Struct layouts of the drop-in boundary: sizeof and offsetof of every field
of the types a caller compiled against the reference's headers touches
(state structs read directly by callers, and every type nested in them).

    python tests/golden/abi_layout.py --reference   # writes abi_layout.json
                                                    # from /root/reference/include
    python tests/golden/abi_layout.py               # prints this library's
                                                    # layouts (include/gviz)

The probe is a generated C file compiled with gcc; only names of types and
fields are listed here (they are the interface).
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
FIXTURE = os.path.join(HERE, "abi_layout.json")

# (type, [fields]) -- every field of each struct, in declaration order
STRUCTS = [
    ("gvizArray", ["arr", "elementSize", "count", "capacity"]),
    ("gvizDeque", ["arr", "begin", "count", "capacity", "elementSize"]),
    ("gvizVertex", ["data", "neighbors"]),
    ("gvizGraphLayout", ["vertexOffsets", "nvertices", "edgeCount"]),
    ("gvizGraph", ["vertices", "map", "directed", "layout"]),
    ("gvizEdgeSubset", ["bitset", "layout"]),
    ("gvizSubgraph", ["g", "vs", "es", "vertexCapacity"]),
    ("gvizBitArrayIterator", ["arr", "size", "idx"]),
    ("gvizSubgraphVertexIterator", ["sg", "it"]),
    ("gvizSubgraphNeighborIterator", ["sg", "u", "base", "it", "adj_idx", "nb", "mode"]),
    ("gvizEmbedding", ["dim", "vertexPositions"]),
    ("gvizActionPayload", ["worldX", "worldY", "deltaTime", "iarg", "darg"]),
    ("gvizAction", ["name", "handler", "userData"]),
    ("gvizActionRegistry", ["actions", "count", "capacity"]),
    ("gvizStatSeries", ["name", "kind", "samples", "count", "capacity", "revision"]),
    ("gvizStatRegistry", ["series", "count", "capacity"]),
    ("gvizDrawMask", ["visibleVertices", "edgePolicy", "revision"]),
    ("gvizEmbeddedGraph", ["subgraph", "embedding", "actions", "drawMask", "stats",
                           "planarEmbedded", "highlight"]),
    ("gvizForceModel", ["vertexMass", "attractive", "repulsive"]),
    ("gvizQuadtreeNode", ["cx", "cy", "halfSize", "comX", "comY", "mass", "pointIndices",
                          "pointCount", "pointCapacity", "children"]),
    ("gvizQuadtree", ["root", "nodesPerCell", "points", "masses", "pointCount",
                      "nodeBlocks", "pointBlocks", "nodesUsed", "overflowBuffers"]),
    ("gvizForceEmbedderState", [
        "graph", "vertices", "vertexCount", "model", "mass", "degree", "radiusBase",
        "radiusPerDegree", "overlapConstant", "preventOverlap", "meanDegreePlusOne",
        "disp", "positionsScratch", "attForceMag", "repForceMag", "structForce",
        "oldStructForce", "appliedDisp", "swinging", "traction", "jitterTolerance",
        "globalSpeed", "speedEfficiency", "edgeLength", "boxExtent", "gravityK",
        "iteration", "lastMaxDisplacement", "begun", "theta", "nodesPerCell",
        "barnesHutEnabled", "quadtree", "quadtreeReady", "pool"]),
    ("gvizFoundVertex", ["v", "dist"]),
    ("gvizKNearestScratch", ["stamp", "nvertices", "epoch", "queue"]),
    ("gvizKNNBatchTarget", ["vertex", "out", "count"]),
    ("gvizGRIPRoundStats", ["maxDisplacement", "meanDisplacement", "meanForce"]),
    ("gvizGRIPDecorators", ["knn", "disp", "oldDisp", "oldCos", "heat", "lastForceMag"]),
    ("gvizGRIPState", [
        "graph", "layerCount", "currLayer", "currRound", "misFiltration", "misBorder",
        "dec", "dispCalculated", "radiusBfsStamp", "radiusBfsEpoch", "radiusBfsQueue",
        "knnScratch", "knnScratchCount", "pool", "placementKMax", "refinementKMax",
        "knnCapacity", "kPolicy", "lastRoundStats", "statsRegisterMode",
        "statsEnabled"]),
    ("gvizTutteState", ["graph", "boundary", "boundaryCount", "isBoundary", "scratch",
                        "iteration", "lastMaxDelta", "converged", "useGaussSeidel",
                        "epsilon", "relaxationRate", "begun"]),
    ("gvizSpringTutteState", ["graph", "boundary", "boundaryCount", "isBoundary",
                              "scratchPos", "velocity", "iteration", "lastMaxDelta",
                              "converged", "epsilon", "stiffness", "damping", "begun"]),
]
# enums and scalar typedefs whose size is part of the layout
SCALARS = ["gvizForceModelKind", "gvizGRIPKPolicy", "gvizStatChartKind",
           "gvizDrawEdgePolicy", "GVIZ_BIT_UNIT", "gvizVertexSubset",
           "gvizSubgraphNeighborIterMode", "gvizQuadrant"]

REF_INCLUDES = ["ds/gvizArray.h", "ds/gvizDeque.h", "ds/gvizBitArray.h", "ds/gvizGraph.h",
                "ds/gvizSubgraph.h", "ds/gvizQuadtree.h", "embedders/gvizEmbeddedGraph.h",
                "embedders/gvizForceModel.h", "embedders/gvizForceEmbedder.h",
                "algorithms/search/gvizKNearest.h", "embedders/gvizGRIPEmbedder.h",
                "embedders/gvizTutteEmbedder.h", "embedders/gvizSpringTutteEmbedder.h"]
OWN_INCLUDES = ["gviz/gvizGraph.h", "gviz/gvizEmbeddedGraph.h", "gviz/gvizForceEmbedder.h",
                "gviz/gvizKNearest.h", "gviz/gvizGRIPEmbedder.h", "gviz/gvizTutteEmbedder.h",
                "gviz/gvizSpringTutteEmbedder.h"]


def probe_source(includes):
    lines = ["#include <stdint.h>", "#include <stdbool.h>", "#include <stddef.h>",
             "#include <stdio.h>"]
    lines += [f'#include "{h}"' for h in includes]
    lines.append("int main(void) {")
    lines.append('  printf("{\\n");')
    first = True
    for t, fields in STRUCTS:
        sep = "" if first else ","
        first = False
        lines.append(f'  printf("{sep}\\"{t}\\": {{\\"sizeof\\": %zu, \\"align\\": %zu", '
                     f"sizeof({t}), _Alignof({t}));")
        for f in fields:
            lines.append(f'  printf(", \\"{f}\\": [%zu, %zu]", offsetof({t}, {f}), '
                         f"sizeof((({t} *)0)->{f}));")
        lines.append('  printf("}\\n");')
    for t in SCALARS:
        lines.append(f'  printf(",\\"{t}\\": {{\\"sizeof\\": %zu}}\\n", sizeof({t}));')
    lines.append('  printf("}\\n");')
    lines.append("  return 0;")
    lines.append("}")
    return "\n".join(lines) + "\n"


def layouts(include_dirs, includes, extra=()):
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "probe.c")
        exe = os.path.join(td, "probe")
        with open(src, "w") as f:
            f.write(probe_source(includes))
        cmd = ["gcc", "-std=gnu11", "-w", *extra]
        for d in include_dirs:
            cmd += ["-I", d]
        subprocess.run(cmd + ["-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def reference_layouts(ref="/root/reference"):
    return layouts([os.path.join(ref, "include"), os.path.join(ref, "src")], REF_INCLUDES,
                   extra=("-include", "stdint.h"))


def own_layouts():
    return layouts([os.path.join(ROOT, "include")], OWN_INCLUDES)


if __name__ == "__main__":
    if "--reference" in sys.argv:
        data = reference_layouts()
        with open(FIXTURE, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        print(f"wrote {FIXTURE}: {len(data)} types")
    else:
        print(json.dumps(own_layouts(), indent=1, sort_keys=True))
