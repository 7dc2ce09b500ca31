"""The batch engine (many graphs in one set of launches) against the
single-graph engine: each graph's trajectory must be the one it follows
alone, bit for bit, whatever the schedule the autotuner picks."""
import numpy as np
import pytest

from grapher_amd import graphs as G

pytestmark = pytest.mark.gpu


def _mixed_graphs(seed=0):
    rng = np.random.default_rng(seed)
    out = []
    specs = [("tree", 1), ("tree", 2), ("grid", 7), ("er", 65), ("tree", 300),
             ("grid", 31), ("er", 1200), ("tree", 2500), ("er", 64), ("grid", 12),
             ("er", 4000), ("tree", 129)]
    for kind, size in specs:
        if kind == "tree":
            n, e = G.random_tree(size, int(rng.integers(1 << 30)))
        elif kind == "grid":
            n, e = G.grid(size, size)
        else:
            n, e = G.erdos_renyi(size, 4.0, int(rng.integers(1 << 30)))
        out.append(G.csr_from_edges(n, e))
    return out


PARAMS = dict(theta=0.7, edge_length=10.0, gravity=0.0)


@pytest.mark.parametrize("precision", ["fp32", "mixed", "fp64"])
def test_batch_equals_single_runs(gpu, precision):
    graphs = _mixed_graphs(1)
    starts = [G.uniform_positions(len(o) - 1, 10 + i) * 5.0 for i, (o, _) in enumerate(graphs)]
    b = gpu.BatchEngine(graphs, precision=precision, **PARAMS)
    b.begin(np.concatenate(starts))
    steps = 12
    launched, its, _, _ = b.run(steps)
    assert launched == steps and np.all(its == steps)
    got = b.positions()
    for gi, ((o, a), st) in enumerate(zip(graphs, starts)):
        e = gpu.ForceEngine(o, a, precision=precision, **PARAMS)
        e.begin(st)
        e.run(steps, stats=False)
        ref = e.positions()
        mine = got[b.start[gi]:b.start[gi + 1]]
        print(f"graph {gi}: n={len(o) - 1}, max |diff| {np.abs(mine - ref).max():.3g}")
        assert np.array_equal(mine, ref), gi


def test_batch_schedules_do_not_change_results(gpu):
    graphs = _mixed_graphs(2)
    start = np.concatenate([G.uniform_positions(len(o) - 1, i) for i, (o, _) in enumerate(graphs)])
    ref = None
    for order, spl, tw in ((0, 1, 1), (1, 4, 1), (1, 16, 2), (0, 4, 4)):
        b = gpu.BatchEngine(graphs, precision="fp32", **PARAMS)
        b.begin(start)
        b.schedule(order, spl, tw)
        b.run(20)
        p = b.positions()
        if ref is None:
            ref = p
        assert np.array_equal(p, ref), (order, spl, tw)
    b = gpu.BatchEngine(graphs, precision="fp32", **PARAMS)
    b.begin(start)
    choice = b.autotune(4)
    b.run(20)
    assert np.array_equal(b.positions(), ref), choice


def test_batch_stops_each_graph_like_its_own_run(gpu):
    graphs = _mixed_graphs(3)[2:8]
    starts = [G.uniform_positions(len(o) - 1, 40 + i) for i, (o, _) in enumerate(graphs)]
    eps = 0.5
    b = gpu.BatchEngine(graphs, precision="mixed", **PARAMS)
    b.begin(np.concatenate(starts))
    _, its, lat, _ = b.run(400, epsilon=eps)
    for gi, ((o, a), st) in enumerate(zip(graphs, starts)):
        e = gpu.ForceEngine(o, a, precision="mixed", **PARAMS)
        e.begin(st)
        _, ran = e.run(400, epsilon=eps, stats=False)
        assert its[gi] == ran, (gi, its[gi], ran)
    assert np.all(lat > 0)
