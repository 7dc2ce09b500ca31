"""Input/output framing for oracle/_ref/ref_harness (the reference library
driven on fixed inputs). Used by make_golden.py and, in a container that has
the reference tree, by tests that cross-check the oracle against it live.
"""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "..", "..", "oracle", "_ref", "ref_harness")


def harness_available() -> bool:
    return os.path.exists(HARNESS)


class _W:
    def __init__(self):
        self.parts = []

    def i(self, *vals):
        for v in vals:
            self.parts.append(struct.pack("<q", int(v)))

    def d(self, *vals):
        for v in vals:
            self.parts.append(struct.pack("<d", float(v)))

    def ia(self, arr):
        self.parts.append(np.ascontiguousarray(arr, dtype="<i8").tobytes())

    def da(self, arr):
        self.parts.append(np.ascontiguousarray(arr, dtype="<f8").tobytes())

    def graph(self, n, edges):
        edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        self.i(n, edges.shape[0])
        self.ia(edges.reshape(-1))

    def bytes(self):
        return b"".join(self.parts)


class _R:
    def __init__(self, buf):
        self.buf = buf
        self.off = 0

    def i(self):
        v = struct.unpack_from("<q", self.buf, self.off)[0]
        self.off += 8
        return v

    def d(self):
        v = struct.unpack_from("<d", self.buf, self.off)[0]
        self.off += 8
        return v

    def ia(self, n):
        a = np.frombuffer(self.buf, dtype="<i8", count=n, offset=self.off).copy()
        self.off += 8 * n
        return a

    def da(self, n):
        a = np.frombuffer(self.buf, dtype="<f8", count=n, offset=self.off).copy()
        self.off += 8 * n
        return a


def _run(mode, payload: bytes, exe=None) -> bytes:
    """Runs `exe mode in out` (default: the reference harness); the same
    framing drives tests/host/api_driver over this repository's library."""
    with tempfile.TemporaryDirectory() as td:
        fi = os.path.join(td, "in.bin")
        fo = os.path.join(td, "out.bin")
        with open(fi, "wb") as f:
            f.write(payload)
        subprocess.run([exe or HARNESS, mode, fi, fo], check=True)
        with open(fo, "rb") as f:
            return f.read()


def quadtree(points, masses, npc, exe=None):
    points = np.asarray(points, dtype=np.float64).reshape(-1, 2)
    n = points.shape[0]
    w = _W()
    w.i(n, npc)
    w.da(points.reshape(-1))
    w.da(masses)
    return _read_tree(_R(_run("quadtree", w.bytes(), exe)), n)


def force_quadtree(n, edges, pos, *, model, barnes_hut, npc, steps, theta, exe):
    """The embedder state's quadtree (host API driver only): positions before
    the last step, whether a root exists, quadtreeReady, the tree's point and
    mass buffers, and the tree."""
    w = _W()
    w.graph(n, edges)
    w.i(model, barnes_hut, npc, steps)
    w.d(theta)
    w.da(np.asarray(pos, dtype=np.float64).reshape(-1))
    r = _R(_run("force_quadtree", w.bytes(), exe))
    before = r.da(2 * n).reshape(n, 2)
    has_root, ready = r.i(), r.i()
    points = r.da(2 * n).reshape(n, 2)
    masses = r.da(n)
    tree = _read_tree(r, n)
    return before, bool(has_root), ready, points, masses, tree


def _read_tree(r, n):
    nodes = []
    # node records are written in DFS order before the per-point table; read
    # them until the table fits exactly at the end of the buffer.
    total = (len(r.buf) - r.off) // 8
    node_words = total - 3 * n - 1
    count = node_words // 9
    for _ in range(count):
        depth, kh, kl, leaf, cnt = r.i(), r.i(), r.i(), r.i(), r.i()
        mass, cx, cy, h = r.d(), r.d(), r.d(), r.d()
        nodes.append((depth, kh & 0xFFFFFFFFFFFFFFFF, kl & 0xFFFFFFFFFFFFFFFF,
                      leaf, cnt, mass, cx, cy, h))
    leaf = r.ia(3 * n).reshape(n, 3)
    node_count = r.i()
    assert node_count == count
    dt = np.dtype([("depth", "i8"), ("keyHi", "u8"), ("keyLo", "u8"),
                   ("leaf", "i8"), ("count", "i8"), ("mass", "f8"),
                   ("comX", "f8"), ("comY", "f8"), ("halfSize", "f8")])
    arr = np.array(nodes, dtype=dt)
    return {
        "nodes": arr,
        "leaf_depth": leaf[:, 0].copy(),
        "leaf_keyHi": leaf[:, 1].astype(np.uint64),
        "leaf_keyLo": leaf[:, 2].astype(np.uint64),
    }


def force(n, edges, pos, *, model, barnes_hut, prevent_overlap, npc, iters,
          dump_every, edge_length, theta, gravity, overlap_constant,
          radius_base, radius_per_degree, jitter):
    w = _W()
    w.graph(n, edges)
    w.i(model, barnes_hut, prevent_overlap, npc, iters, dump_every)
    w.d(edge_length, theta, gravity, overlap_constant, radius_base,
        radius_per_degree, jitter)
    w.da(np.asarray(pos, dtype=np.float64).reshape(-1))
    r = _R(_run("force", w.bytes()))
    stats = []
    dumps = {}
    for it in range(iters):
        stats.append([r.d() for _ in range(6)])
        if dump_every > 0 and (it % dump_every == 0 or it == iters - 1):
            dumps[it] = {
                "disp": r.da(2 * n).reshape(n, 2),
                "structForce": r.da(2 * n).reshape(n, 2),
                "attMag": r.da(n),
                "repMag": r.da(n),
                "pos_after": r.da(2 * n).reshape(n, 2),
            }
    return np.asarray(stats), dumps


def tutte(n, edges, boundary, iters, radius, dt, gauss_seidel=False, exe=None):
    w = _W()
    w.graph(n, edges)
    w.i(len(boundary), iters, 1 if gauss_seidel else 0)
    w.ia(np.asarray(boundary, dtype=np.int64))
    w.d(radius, dt)
    r = _R(_run("tutte", w.bytes(), exe))
    p0 = r.da(2 * n).reshape(n, 2)
    deltas = r.da(iters)
    p1 = r.da(2 * n).reshape(n, 2)
    return p0, deltas, p1


def knn(n, edges, visible, k, sources, exe=None):
    w = _W()
    w.graph(n, edges)
    w.i(k, len(sources))
    w.ia(np.asarray(visible, dtype=np.int64))
    w.ia(np.asarray(sources, dtype=np.int64))
    r = _R(_run("knn", w.bytes(), exe))
    per = []
    for _ in sources:
        c = r.i()
        lst = r.ia(2 * k).reshape(k, 2)
        per.append((c, lst))
    rc = r.i()
    nt = r.i()
    batch = []
    for _ in range(nt):
        v = r.i()
        c = r.i()
        lst = r.ia(2 * k).reshape(k, 2)
        batch.append((v, c, lst))
    return per, rc, batch


def grip(n, edges, kcap, kmax, rounds_per_layer, dim=2):
    w = _W()
    w.graph(n, edges)
    w.i(kcap, kmax, rounds_per_layer, dim)
    r = _R(_run("grip", w.bytes()))
    layers = r.i()
    borders = r.ia(layers)
    filtration = r.ia(int(borders[0]))

    def state():
        pos = r.da(dim * n).reshape(n, dim)
        dt = np.dtype([
            ("disp", "<f8", dim), ("oldDisp", "<f8", dim), ("heat", "<f8"),
            ("oldCos", "<f8"), ("lastForceMag", "<f8"), ("knnCount", "<i8"),
            ("dispCalculated", "<i8")])
        dec = np.frombuffer(r.buf, dtype=dt, count=n, offset=r.off).copy()
        r.off += n * dt.itemsize
        return pos, dec

    rounds = []
    while True:
        layer = r.i()
        if layer == -1:
            break
        pos0, dec0 = state()
        knnl = r.ia(n * kcap * 2).reshape(n, kcap, 2)
        pos1, dec1 = state()
        st = (r.d(), r.d(), r.d())
        rounds.append(dict(layer=layer, pos0=pos0, dec0=dec0, knn=knnl,
                           pos1=pos1, dec1=dec1, stats=st))
    return borders, filtration, rounds


def api_force(n, edges, hidden, *, model, barnes_hut, prevent_overlap, npc, seed,
              run_iters, edge_length, theta, gravity, overlap_constant,
              radius_base, radius_per_degree, jitter, epsilon, perturb=0.0,
              perturb_seed=0, exe=None):
    """gvizForceEmbedderInit/Configure*/Begin(seed)/Run/Step on the subgraph
    induced by every vertex but `hidden`; `perturb` > 0 multiplies each
    coordinate after Begin by 1 + perturb * u (u uniform in [-1, 1), seeded)."""
    w = _W()
    w.graph(n, edges)
    w.i(len(hidden))
    w.ia(hidden)
    w.i(model, barnes_hut, prevent_overlap, npc, seed, run_iters)
    w.d(edge_length, theta, gravity, overlap_constant, radius_base,
        radius_per_degree, jitter, epsilon, perturb)
    w.i(perturb_seed)
    r = _R(_run("api_force", w.bytes(), exe))
    out = dict(box_extent=r.d(), begin_rc=r.i(), pos_begin=r.da(2 * n).reshape(n, 2))
    out["run_rc"] = r.i()
    for name in ("maxDisp", "speed", "attractiveForce", "repulsiveForce",
                 "gravityForce"):
        c = r.i()
        out["series_" + name] = r.da(c)
    out["pos_run"] = r.da(2 * n).reshape(n, 2)
    out["step_ret"] = r.d()
    out["pos_step"] = r.da(2 * n).reshape(n, 2)
    out["speed"] = r.d()
    out["efficiency"] = r.d()
    out["iteration"] = r.i()
    return out


def api_grip(n, edges, kcap, placement_k, refinement_k, policy, exe=None, dim=2):
    """gvizGRIPEmbedderEmbed with ConfigureKnnCapacity / ConfigureK in `dim`
    dimensions."""
    w = _W()
    w.graph(n, edges)
    w.i(kcap, placement_k, refinement_k, policy, dim)
    r = _R(_run("api_grip", w.bytes(), exe))
    out = dict(embed_rc=r.i())
    layers = r.i()
    out["borders"] = r.ia(layers)
    out["pos"] = r.da(dim * n).reshape(n, dim)
    for name in ("heat", "meanDisp", "maxDisp", "meanForce"):
        c = r.i()
        out["series_" + name] = r.da(c)
    out["visible"] = r.ia(n)
    return out


def api_grip_spread(n, edges, kcap, placement_k, refinement_k, policy, perturb,
                    perturb_seed, exe=None, dim=2):
    """Final positions of the reference's Embed schedule run from Begin's
    placement with every coordinate scaled by 1 + perturb * u (u uniform in
    [-1, 1), seeded)."""
    w = _W()
    w.graph(n, edges)
    w.i(kcap, placement_k, refinement_k, policy, dim)
    w.d(perturb)
    w.i(perturb_seed)
    r = _R(_run("api_grip_spread", w.bytes(), exe))
    return r.da(dim * n).reshape(n, dim)


def embedding_io(n, edges, dim, positions, name, save_path, load_paths, exe=None):
    """gvizEmbeddedGraphSaveEmbedding of `positions` to save_path (its return
    and the bytes written), then for each of load_paths, from the same
    positions, gvizEmbeddedGraphLoadEmbedding's return and the positions
    after it."""
    w = _W()
    w.graph(n, edges)
    w.i(dim)
    w.da(np.ascontiguousarray(positions, dtype=np.float64).reshape(-1))
    for s in (name, save_path):
        raw = os.fsencode(s)
        w.i(len(raw))
        w.parts.append(raw)
    w.i(len(load_paths))
    for s in load_paths:
        raw = os.fsencode(s)
        w.i(len(raw))
        w.parts.append(raw)
    r = _R(_run("embedding_io", w.bytes(), exe))
    rc = r.i()
    length = r.i()
    text = r.buf[r.off:r.off + length]
    r.off += length
    loads = []
    for _ in load_paths:
        lrc = r.i()
        loads.append((lrc, r.da(n * dim).reshape(n, dim)))
    return rc, text, loads


def edges_file(path, directed=0, skip_header=0, exe=None, csr=False):
    """The .edges loader on `path`: (rc, offsets, adj) of the adjacency lists
    in order (offsets/adj None when rc != 0). With csr=True (this
    repository's driver) also the CSR loader's (rc, offsets, adj)."""
    w = _W()
    raw = os.fsencode(path)
    w.i(directed, skip_header, len(raw))
    r = _R(_run("edges_file", w.bytes() + raw, exe))

    def lists():
        rc = r.i()
        if rc != 0:
            return rc, None, None
        n = r.i()
        off = [0]
        adj = []
        for _ in range(n):
            c = r.i()
            adj.extend(r.ia(c).tolist())
            off.append(off[-1] + c)
        return rc, np.array(off, np.int64), np.array(adj, np.int64)

    first = lists()
    if not csr:
        return first
    rc = r.i()
    if rc != 0:
        return first, (rc, None, None)
    n = r.i()
    off = r.ia(n + 1)
    adj = r.ia(int(off[-1]))
    return first, (rc, off, adj)


def force_checkpoint(n, edges, model, seed, k1, k2, path, exe):
    """This repository's driver: embedder A runs k1 steps, saves its state to
    `path` and runs k2 more; embedder B loads the state and runs k2. Returns
    [(positions, steps, globalSpeed)] for A and B."""
    w = _W()
    w.graph(n, edges)
    raw = os.fsencode(path)
    w.i(model, seed, k1, k2, len(raw))
    r = _R(_run("force_checkpoint", w.bytes() + raw, exe))
    out = []
    for _ in range(2):
        pos = r.da(2 * n).reshape(n, 2)
        out.append((pos, r.i(), r.d()))
    return out


def spring_tutte(n, edges, boundary, steps, run_to, radius, dt, stiffness, damping,
                 epsilon, exe=None):
    """Spring-driven Tutte through the reference library (or `exe`): seeded
    positions, Step returns, positions/velocities after the steps, Run's
    return, positions/velocities after it, the maxDelta series."""
    w = _W()
    w.graph(n, edges)
    w.i(len(boundary), steps, run_to)
    w.ia(np.asarray(boundary, dtype=np.int64))
    w.d(radius, dt, stiffness, damping, epsilon)
    r = _R(_run("spring_tutte", w.bytes(), exe))
    out = {"p0": r.da(2 * n).reshape(n, 2), "step": r.da(steps),
           "p1": r.da(2 * n).reshape(n, 2), "v1": r.da(2 * n).reshape(n, 2),
           "run": r.i(), "p2": r.da(2 * n).reshape(n, 2), "v2": r.da(2 * n).reshape(n, 2)}
    c = r.i()
    out["series"] = r.da(c) if c >= 0 else None
    return out


def grip_checkpoint(n, edges, kcap, pk, rk, policy, stages, rounds, path, exe):
    """This repository's driver: a GRIP run saved mid-schedule and the same
    continuation in the saved embedder and in one that loaded the state.
    Returns [(current layer, positions, visibility)] for both."""
    w = _W()
    w.graph(n, edges)
    raw = os.fsencode(path)
    w.i(kcap, pk, rk, policy, stages, rounds, len(raw))
    r = _R(_run("grip_checkpoint", w.bytes() + raw, exe))
    out = []
    for _ in range(2):
        layer = r.i()
        pos = r.da(2 * n).reshape(n, 2)
        vis = r.ia(n)
        out.append((layer, pos, vis))
    return out


def tutte_begin(n, edges, spring, iters, epsilon, exe=None):
    """Tutte (spring=0) or spring-Tutte (spring=1) Begin then Run(iters):
    Begin's return, and when 0 the boundary cycle, Run's return and the
    positions (n, 2)."""
    w = _W()
    w.graph(n, edges)
    w.i(spring, iters)
    w.d(epsilon)
    r = _R(_run("tutte_begin", w.bytes(), exe))
    out = dict(begin_rc=r.i())
    if out["begin_rc"] == 0:
        c = r.i()
        out["boundary"] = r.ia(c)
        out["run"] = r.i()
        out["pos"] = r.da(2 * n).reshape(n, 2)
    return out


def grip_stages(n, edges, dim, exe):
    """The filtration built through the white-box stage functions and by
    createMISFiltration (host API driver only): two (layers, borders,
    filtration) triples."""
    w = _W()
    w.graph(n, edges)
    w.i(dim)
    r = _R(_run("grip_stages", w.bytes(), exe))
    out = []
    for _ in range(2):
        layers = r.i()
        nb = r.i()
        borders = r.ia(nb)
        filt = r.ia(int(borders[0]))
        out.append((layers, borders, filt))
    return out


def tutte_outer_face(n, edges, *, spring, iters, hl_vertices, hl_edges, exe):
    """Tutte Begin, the rotated adjacency lists, FixOuterFace on a highlight,
    Run (host API driver only)."""
    w = _W()
    w.graph(n, edges)
    hl_edges = np.asarray(hl_edges, dtype=np.int64).reshape(-1, 2)
    w.i(1 if spring else 0, iters, len(hl_vertices), len(hl_edges))
    w.ia(np.asarray(hl_vertices, dtype=np.int64))
    w.ia(hl_edges.reshape(-1))
    r = _R(_run("tutte_outer_face", w.bytes(), exe))
    begin_rc, planar = r.i(), r.i()
    lists = []
    for _ in range(n):
        c = r.i()
        lists.append(r.ia(c))
    fix_rc = r.i()
    bc = r.i()
    boundary = r.ia(bc)
    iters_run = r.i()
    pos = r.da(2 * n).reshape(n, 2)
    return {"begin_rc": begin_rc, "planar": planar, "lists": lists, "fix_rc": fix_rc,
            "boundary": boundary, "iterations": iters_run, "positions": pos}
