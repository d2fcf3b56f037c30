"""Hub tables (RunConfig.hub_tables, on by default): rows of more than 512
entries are evaluated from per-vertex community tables that the first full
evaluation of a level's rounds builds and the neighbours' moves keep current
(in the one-workgroup decoupling, the grid marks and the multi-kernel path).
With unit weights every table weight is an exact integer, so the engine stays
bit-exact with the oracle; with real weights the bar is |dQ| <= 1e-3."""
import numpy as np
import pytest

import hit_leiden_amd as hl
from hit_leiden_amd import gen
from parity_util import assert_classes_connected, assert_same_state, assert_supergraph_consistent, pair
from test_gpu_parity import random_batch, say

pytestmark = pytest.mark.gpu

THR = hl.RunMode.THROUGHPUT


def with_hubs(n, hub_deg, hubs, seed):
    w = gen.sbm(n, 20, 8.0, 0.2, seed=seed)
    rng = np.random.default_rng(seed + 1)
    hu = [np.full(hub_deg, h, np.uint32) for h in range(hubs)]
    hv = [rng.choice(np.arange(hubs, n, dtype=np.uint32), size=hub_deg, replace=False) for _ in hu]
    u = np.concatenate([w.u] + hu)
    v = np.concatenate([w.v] + hv)
    key = np.unique(np.minimum(u, v).astype(np.int64) * n + np.maximum(u, v))
    return (key // n).astype(np.uint32), (key % n).astype(np.uint32)


@pytest.mark.parametrize("hub_deg,n", [(5000, 30000), (20000, 40000), (300000, 320000)])
def test_tables_hubs(hub_deg, n):
    u, v = with_hubs(n, hub_deg, 3, 5)
    say(f"hub tables deg={hub_deg}")
    g, o = pair(THR, 2)
    g.load(hl.GraphInput(n, u, v))
    o.load(n, u, v)
    assert_same_state(g, o, 2, "hubs load")
    for step in range(3):
        bu, bv, ba = random_batch(n, u, v, None, 600, 70 + step)
        g.hit_leiden(hl.GraphInput(n, bu, bv, ba))
        o.apply(bu, bv, ba)
        assert_same_state(g, o, 2, f"hubs batch {step}")
    g.close()


def test_tables_rmat():
    """R-MAT scale 16: rows above 512 entries, four levels, deletions."""
    w = gen.rmat(16, seed=3)
    w.make_batches(3, 800, 100, mode=1, seed=4)
    P = 4
    say(f"hub tables rmat n={w.n} m={w.m}")
    g, o = pair(THR, P)
    g.load(hl.GraphInput(w.n, w.u, w.v))
    o.load(w.n, w.u, w.v)
    assert_same_state(g, o, P, "rmat load")
    for i, (u, v, a) in enumerate(w.batches):
        g.hit_leiden(hl.GraphInput(w.n, u, v, a))
        o.apply(u, v, a)
        assert_same_state(g, o, P, f"rmat batch {i}")
    assert_supergraph_consistent(g, P)
    g.close()


def test_tables_on_equals_off():
    w = gen.rmat(17, seed=8)
    w.make_batches(3, 2000, 200, mode=1, seed=9)
    rows = []
    for on in (True, False):
        e = hl.HitLeiden(hl.RunConfig(mode=THR, levels=4, hub_tables=on))
        e.load(hl.GraphInput(w.n, w.u, w.v))
        for (u, v, a) in w.batches:
            e.hit_leiden(hl.GraphInput(w.n, u, v, a))
        rows.append((e.hierarchy(), e.modularity(hl.PART_COMMUNITY)))
        e.close()
    say(f"tables on/off Q {rows[0][1]} {rows[1][1]}")
    assert np.array_equal(rows[0][0], rows[1][0])
    # same partitions; the modularity pass sums community totals with atomics,
    # so two evaluations of one partition may differ in the last bits
    assert abs(rows[0][1] - rows[1][1]) <= 1e-12


def test_tables_weighted():
    """Real weights: table weights are sums kept by additions and subtractions,
    so the engine is held to |dQ| <= 1e-3 against the oracle, connected refined
    classes and a consistent supergraph."""
    n = 20000
    u, v = with_hubs(n, 6000, 4, 11)
    rng = np.random.default_rng(12)
    w = rng.uniform(0.5, 1.5, len(u))
    g, o = pair(THR, 3)
    g.load(hl.GraphInput(n, u, v, w))
    o.load(n, u, v, w)
    for step in range(3):
        bu, bv, ba = random_batch(n, u, v, w, 500, 30 + step, weighted=True)
        g.hit_leiden(hl.GraphInput(n, bu, bv, ba))
        o.apply(bu, bv, ba)
    qg = g.modularity(hl.PART_COMMUNITY)
    qo = o.modularity(o.hierarchy()[-1])
    say(f"weighted tables Q gpu {qg} oracle {qo}")
    assert abs(qg - qo) <= 1e-3
    off, col, ww = g.supergraph(0)
    r = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    keep = r < col
    assert_classes_connected(n, r[keep], col[keep].astype(np.int64), g.partition(0, hl.MAP_S_CUR))
    assert_supergraph_consistent(g, 3, tol=1e-9)
    g.close()


def test_tables_in_the_batched_form():
    """Several graphs with hubs in one engine: the tables are per vertex, so each
    graph still gets exactly the update it gets alone."""
    from hit_leiden_amd import _first_seen
    graphs = []
    for i, (n, hub) in enumerate([(6000, 5000), (9000, 700), (4000, 0), (12000, 9000)]):
        if hub:
            u, v = with_hubs(n, hub, 2, 40 + i)
        else:
            w = gen.sbm(n, 8, 8.0, 0.2, seed=40 + i)
            u, v = w.u, w.v
        graphs.append(hl.GraphInput(n, u, v))
    batches = [random_batch(g.node_count, g.src, g.dst, None, 300, 60 + i) for i, g in enumerate(graphs)]
    cfg = hl.RunConfig(mode=THR, levels=3)
    e = hl.HitLeiden(cfg)
    e.load_graphs(graphs)
    e.apply_graphs({i: hl.GraphInput(g.node_count, *batches[i]) for i, g in enumerate(graphs)})
    qs = e.modularity_graphs()
    for i, g in enumerate(graphs):
        a = hl.HitLeiden(cfg)
        a.load(g)
        a.hit_leiden(hl.GraphInput(g.node_count, *batches[i]))
        for p, (x, y) in enumerate(zip(a.hierarchy(), e.graph_hierarchy(i))):
            np.testing.assert_array_equal(_first_seen(x), _first_seen(y), err_msg=f"graph {i} row {p}")
        assert abs(qs[i] - a.modularity(hl.PART_COMMUNITY)) < 1e-12
        a.close()
    e.close()
