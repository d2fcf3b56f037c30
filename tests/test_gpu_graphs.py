"""Several independent graphs in one engine (the batched form), on the GPU.

Each graph must get exactly the update it gets alone: its partition and
hierarchy (label values aside), its per-graph modularity, and no change at all
for graphs without records in a pass.  Alone means a one-graph engine on the
same inputs, which the parity tests pin to the oracle; a few graphs are also
compared with the oracle directly.  Unit weights throughout, so every sum is an
exact integer and the comparison is exact.
"""
import numpy as np
import pytest

import hit_leiden_amd as hl
from hit_leiden_amd import _first_seen, gen
from oracle import hl_oracle as O

pytestmark = pytest.mark.gpu

P = 3


def say(*a):
    print(*a, flush=True)


def sbm_graph(n, blocks, seed):
    w = gen.sbm(n, blocks, 8.0, 0.25, seed=seed)
    return hl.GraphInput(w.n, w.u, w.v)


def insert_batch(g, k, seed, dels=0):
    """k new edges (uniform endpoints) plus dels deletions of existing edges."""
    rng = np.random.default_rng(seed)
    have = set(zip(g.src.tolist(), g.dst.tolist()))
    bu, bv, ba = [], [], []
    if dels:
        for i in rng.choice(len(g), size=min(dels, len(g)), replace=False):
            bu.append(g.src[i])
            bv.append(g.dst[i])
            ba.append(-1.0)
    while len(bu) < k + dels:
        a, b = rng.integers(0, g.node_count, 2)
        if a == b or (min(a, b), max(a, b)) in have:
            continue
        have.add((min(a, b), max(a, b)))
        bu.append(a)
        bv.append(b)
        ba.append(1.0)
    return hl.GraphInput(g.node_count, np.array(bu, np.uint32), np.array(bv, np.uint32), np.array(ba))


def canon_rows(h):
    return [_first_seen(r) for r in h]


def make_graphs(count, seed):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(40), np.log(3000), count)).astype(int)
    return [sbm_graph(int(n), int(rng.integers(2, 9)), seed * 1000 + i) for i, n in enumerate(sizes)]


ALONE = dict(mode=hl.RunMode.THROUGHPUT, levels=P)


def alone(graph, batches):
    e = hl.HitLeiden(hl.RunConfig(**ALONE))
    e.load(graph)
    for b in batches:
        e.hit_leiden(b)
    rows = canon_rows(e.hierarchy())
    q = e.modularity(hl.PART_COMMUNITY)
    e.close()
    return rows, q


def test_each_graph_updates_as_if_alone():
    graphs = make_graphs(40, 5)
    say(f"graphs {len(graphs)} vertices {sum(g.node_count for g in graphs)}")
    u = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=P))
    u.load_graphs(graphs)
    assert u.graph_count == len(graphs)
    # two passes, each over a different subset of the graphs
    rng = np.random.default_rng(9)
    history = {i: [] for i in range(len(graphs))}
    for step in range(2):
        chosen = sorted(rng.choice(len(graphs), size=25, replace=False).tolist())
        deltas = {i: insert_batch(graphs[i], int(rng.integers(5, 60)), 100 * step + i, dels=3) for i in chosen}
        before = {i: u.graph_hierarchy(i) for i in range(len(graphs)) if i not in deltas}
        u.apply_graphs(deltas)
        for i, d in deltas.items():
            history[i].append(d)
        for i, rows in before.items():  # graphs without records are untouched
            for a, b in zip(rows, u.graph_hierarchy(i)):
                np.testing.assert_array_equal(a, b, err_msg=f"graph {i} changed without records")
    qs = u.modularity_graphs()
    for i, g in enumerate(graphs):
        rows, q = alone(g, history[i])
        got = u.graph_hierarchy(i)
        for p, (a, b) in enumerate(zip(rows, got)):
            np.testing.assert_array_equal(a, b, err_msg=f"graph {i} hierarchy row {p}")
        assert abs(qs[i] - q) < 1e-12, f"graph {i}: Q {qs[i]} vs {q}"
    u.close()


def test_graphs_match_the_oracle():
    graphs = make_graphs(6, 17)
    # small loading passes: the graphs load over several passes, as large sets do
    u = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=P, load_pass_records=3000))
    u.load_graphs(graphs)
    deltas = {i: insert_batch(g, 40, 7 + i, dels=4) for i, g in enumerate(graphs)}
    u.apply_graphs(deltas)
    qs = u.modularity_graphs()
    for i, g in enumerate(graphs):
        o = O.OracleState(levels=P, gamma=1.0, mode=O.PAR)
        o.load(g.node_count, g.src, g.dst, None)
        d = deltas[i]
        o.apply(d.src, d.dst, d.weights)
        want = canon_rows(o.hierarchy())
        got = u.graph_hierarchy(i)
        for p, (a, b) in enumerate(zip(want, got)):
            np.testing.assert_array_equal(a, b, err_msg=f"graph {i} row {p} vs the oracle")
        assert abs(qs[i] - o.modularity(o.hierarchy()[-1])) < 1e-9
    u.close()


def test_served_emission_before_the_pass_ends():
    """One pass over graphs whose top frontiers empty at different rounds: a
    graph with a one-edge request is handed out at the round its frontier
    empties, while a graph with a large request keeps the pass going; the
    early labels equal the final ones."""
    graphs = make_graphs(8, 41)
    u = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=P))
    u.load_graphs(graphs)
    big = max(range(len(graphs)), key=lambda i: graphs[i].node_count)
    reqs = [(big, insert_batch(graphs[big], 800, 5, dels=40), 0.0)]
    reqs += [(g, insert_batch(graphs[g], 1, 9 + g), 0.0) for g in range(len(graphs)) if g != big]
    got = {}
    done, pass_of, st = u.serve(reqs, max_pass_records=1 << 20, coalesce=True,
                                emit=lambda r, lab: got.__setitem__(r, lab))
    em = u.last_emitted
    say(f"one pass: {st.passes}; emitted/done ms: " +
        ", ".join(f"{1e3 * a:.2f}/{1e3 * b:.2f}" for a, b in zip(em, done)))
    assert st.passes == 1 and np.all(em <= done)
    assert np.sum(em < done) >= 1
    for i, (g, _, _) in enumerate(reqs):
        np.testing.assert_array_equal(_first_seen(got[i]), u.graph_hierarchy(g)[-1], err_msg=f"graph {g}")
    u.close()


def test_graph_set_errors():
    graphs = make_graphs(3, 23)
    u = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=2))
    u.load_graphs(graphs)
    off = u.graph_offsets
    with pytest.raises(hl.InvalidInput):  # a record joining two graphs
        u.hit_leiden(hl.GraphInput(int(off[-1]), np.array([0], np.uint32), np.array([off[1]], np.uint32)))
    with pytest.raises(hl.HitLeidenError):
        u.modularity(hl.PART_COMMUNITY)
    u.close()
    d = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.DETERMINISTIC, levels=2))
    with pytest.raises(hl.InvalidInput):
        d.load_graphs(graphs)
    d.close()


def served_requests(graphs, count, rate, seed):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    zipf = 1.0 / np.arange(1, len(graphs) + 1) ** 1.1
    pick = rng.choice(len(graphs), size=count, p=zipf / zipf.sum())
    return [(int(g), insert_batch(graphs[g], int(rng.integers(5, 60)), seed * 7 + i), float(t[i]))
            for i, g in enumerate(pick)]


@pytest.mark.parametrize("budget,coalesce", [(1 << 20, False), (0, False), (1 << 20, True)])
def test_served_results_match_alone(budget, coalesce):
    """Continuous batching: every request's communities equal the ones its graph
    gets alone (its earlier requests applied one by one, or, when coalesced, the
    requests of each pass applied as one batch); budget 0 is the plain
    one-request-per-pass path."""
    graphs = make_graphs(10, 31)
    u = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=P))
    u.load_graphs(graphs)
    reqs = served_requests(graphs, 60, 2000.0, 4)
    got = {}
    done, pass_of, st = u.serve(reqs, max_pass_records=budget, coalesce=coalesce,
                                emit=lambda r, lab: got.__setitem__(r, lab))
    em = u.last_emitted
    say(f"budget={budget} coalesce={coalesce} passes={st.passes} max_pass_requests={st.max_pass_requests} "
        f"handed out before the pass end: {int(np.sum(em < done))} of {len(reqs)}")
    assert len(got) == len(reqs) and np.all(done > 0)
    # the labels are handed out when they are final, never after the pass ends
    assert np.all(em > 0) and np.all(em <= done)
    assert st.max_pass_bytes > 0 and st.footprint_cut_passes == 0
    if budget == 0:
        assert st.passes == len(reqs)
    else:
        assert st.passes < len(reqs)  # arrivals faster than passes were batched
    for g, graph in enumerate(graphs):
        mine = sorted((t, i) for i, (gg, _, t) in enumerate(reqs) if gg == g)
        ps = [int(pass_of[i]) for _, i in mine]
        assert ps == sorted(ps)  # a graph's requests complete in arrival order
        if not coalesce:
            assert len(set(ps)) == len(ps)
        e = hl.HitLeiden(hl.RunConfig(**ALONE))
        e.load(graph)
        for p in sorted(set(ps)):
            group = [i for _, i in mine if int(pass_of[i]) == p]
            b = [reqs[i][1] for i in group]
            e.hit_leiden(hl.GraphInput(graph.node_count, np.concatenate([x.src for x in b]),
                                       np.concatenate([x.dst for x in b]), np.concatenate([x.weights for x in b])))
            want = _first_seen(e.hierarchy()[-1])
            for i in group:
                np.testing.assert_array_equal(_first_seen(got[i]), want, err_msg=f"request {i} of graph {g}")
        e.close()
    u.close()
