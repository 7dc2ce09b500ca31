"""The precision split (north-star form 1) against the FP64 oracle, quantity
by quantity, on the headline workload at 100,000 vertices (BA m = 3, LinLog,
gravity 0.05, Barnes-Hut theta 0.7, prevent-overlap r = 0.1 (1 + sqrt deg)),
with overlap prevention on and off, at the uniform [-1, 1]^2 start and
after 20 FP64 steps (the layout spread out).

Modes: MIXED (FP64 positions differenced, then FP32 force math and
accumulation) and FP32 (positions rounded to FP32 before any force
arithmetic, FP32 cell classification and cell data: the split in its
purest form). In both, quadtree keys, cell mass and centre-of-mass sums,
the FA2 swing/traction reduction and global speed, and the position update
are FP64.

Quantities, per vertex against sum_j |f_ij| over the vertex's force terms
(the per-vertex FP32 parity definition): the attraction and the repulsion
magnitudes and the whole force; the whole force normwise; and from the
start state one whole step's FA2 swing and traction sums, global speed,
speed efficiency, maxDisp and new positions (relative). Gravity is a closed
form of the vertex's own position and degree (one reciprocal square root
and one reciprocal, within 2 ulp in FP32); next to repulsion sums of 1e5 and
more it cannot be read back separately from the FP32 force output.

Measured (MI355X): MIXED within 1e-6 for every quantity, with and without
overlap prevention, at the start and after 20 steps (one step: swing and
traction 1e-8, speed and efficiency exact, positions 3e-7). FP32 keeps the
attraction (6e-8) and the FA2 scalars close, but its repulsion carries the
rounding of the positions themselves: a source u at distance d from v moves
by up to 2^-24 |x| per coordinate, so the term's relative error is about
2^-24 |x| / d; 100,000 points in [-1, 1]^2 have pairs 1e-5 apart, and with
overlap prevention the 1/(d - r_u - r_v) factor multiplies the error by
d / gap near contact. Per vertex: 0.113 at the start and 1.4e-3 after 20
steps with overlap prevention, 3.9e-4 and 5.8e-3 without; one step moves
positions by up to 0.058 (relative) of the FP64 step. That is why MIXED, the
default, differences FP64 positions before its FP32 force math.
"""
import time

import numpy as np
import pytest

import pyoracle as O
from grapher_amd import graphs as G

pytestmark = pytest.mark.gpu

CFG = dict(model=1, barnes_hut=True, nodes_per_cell=1, edge_length=10.0, theta=0.7,
           gravity=0.05, overlap_constant=100.0, radius_base=0.1, radius_per_degree=1.0)


def log(msg):
    print(f"[precision] {msg}", flush=True)


def _oparams(cfg):
    return O.params(model=cfg["model"], barnes_hut=True,
                    prevent_overlap=cfg["prevent_overlap"], npc=1,
                    edge_length=cfg["edge_length"], theta=cfg["theta"],
                    gravity=cfg["gravity"], overlap_constant=cfg["overlap_constant"],
                    radius_base=cfg["radius_base"],
                    radius_per_degree=cfg["radius_per_degree"])


@pytest.fixture(scope="module")
def ba100k():
    n, e = G.barabasi_albert(100_000, 3, 0)
    off, adj = G.csr_from_edges(n, e)
    return n, off, adj


@pytest.fixture(scope="module")
def pool():
    p = O.Pool(8)
    yield p
    p.close()


# MIXED meets 1e-5 everywhere. FP32's repulsion bars are the measured
# figures (in the comments) with room: what FP32 positions can resolve.
_MIXED = dict(attraction=1e-5, repulsion=1e-5, normwise=1e-5, per_vertex=1e-5, step=1e-5)
BARS = {("mixed", True): _MIXED, ("mixed", False): _MIXED,
        # start: repulsion 0.113, normwise 0.0436; after 20 steps 1.37e-3 /
        # 3.5e-5; one step: maxDisp and positions 0.058
        ("fp32", True): dict(attraction=1e-5, repulsion=0.5, normwise=0.2, per_vertex=0.5,
                             step=0.2),
        # start: repulsion 3.9e-4, normwise 2.4e-5; one step: positions
        # 8.3e-4. After 20 steps the figure depends on the layout the FP64 run
        # reaches (its FA2 branches make it sensitive to summation order): from
        # 2e-4 / 4e-6 up to 0.158 / 0.026 when two vertices end 2.4e-3 apart at
        # coordinates near 6e3, where FP32 positions carry 2.4e-4 of rounding
        ("fp32", False): dict(attraction=1e-5, repulsion=0.5, normwise=0.2,
                              per_vertex=0.5, step=5e-3)}


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("overlap", [True, False])
def test_precision_split_error_per_quantity(gpu, ba100k, pool, overlap):
    n, off, adj = ba100k
    cfg = dict(CFG, prevent_overlap=overlap)
    g = O.Graph(off, adj)
    p = _oparams(cfg)
    f64 = gpu.ForceEngine(off, adj, precision="fp64", **cfg)
    f64.begin(G.uniform_positions(n, 0))
    for label in ("start", "step 20"):
        if label == "step 20":
            f64.run(20, stats=False)
        pos = f64.positions()
        ref = O.force_compute(g, pos, p, pool)
        _, ts = O.force_term_sums(g, pos, p, pool)
        for prec in ("mixed", "fp32"):
            eng = gpu.ForceEngine(off, adj, precision=prec, **cfg)
            eng.set_positions(pos)
            got = eng.evaluate()
            # structForce = attraction + repulsion; disp adds gravity
            err = {
                "attraction": float((np.abs(got["attMag"] - ref["attMag"]) / ts).max()),
                "repulsion": float((np.abs(got["repMag"] - ref["repMag"]) / ts).max()),
                "normwise": _rel(got["disp"], ref["disp"]),
                "per_vertex": float((np.linalg.norm(got["disp"] - ref["disp"], axis=1) /
                                     ts).max()),
            }
            log(f"overlap={overlap} {label} {prec}: " +
                ", ".join(f"{k} {v:.3g}" for k, v in err.items()))
            for k, v in err.items():
                assert v <= BARS[prec, overlap][k], (k, v)
            if label == "start":
                run = O.ForceRun(g, p, pos, pool)
                ost = run.step()
                eng.begin(pos)
                st = eng.step()
                errs = {k: abs(getattr(st, k) - getattr(ost, k)) / max(abs(getattr(ost, k)),
                                                                     1e-300)
                        for k in ("totalSwinging", "totalTraction", "speed",
                                  "speedEfficiency", "maxDisp")}
                errs["positions"] = float(np.abs(eng.positions() - run.pos).max() /
                                          np.abs(run.pos).max())
                run.close()
                log(f"overlap={overlap} one step {prec}: " +
                    ", ".join(f"{k} {v:.3g}" for k, v in errs.items()))
                for k, v in errs.items():
                    assert v <= BARS[prec, overlap]["step"], (k, v)
            eng.close()
    f64.close()
