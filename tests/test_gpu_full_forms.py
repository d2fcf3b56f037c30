"""The batched and served configurations at full size on the GPU.

* batched: 4,096 SBM graphs in one engine, one pass over every graph's batch;
  32 sampled graphs each equal to a one-graph engine on the same inputs.
* served: 256 resident LFR graphs (101 M vertices) under a Poisson stream with
  continuous batching; the requests of 12 sampled graphs emit exactly what
  their graph gets alone."""
import numpy as np
import pytest

import hit_leiden_amd as hl
from hit_leiden_amd import _first_seen, gen
from full_config_util import P, say

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------- batched: 4,096 graphs

def canon_rows(h):
    return [_first_seen(r) for r in h]


@pytest.mark.timeout(900)
def test_batched_4096_graphs_sample_alone():
    ws = gen.batched_config(count=4096, steps=1)
    graphs = [hl.GraphInput(w.n, w.u, w.v) for w in ws]
    say(f"batched: {len(graphs)} graphs, {sum(g.node_count for g in graphs)} vertices")
    u = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=P))
    u.load_graphs(graphs)
    deltas = {i: hl.GraphInput(w.n, *w.batches[0]) for i, w in enumerate(ws)}
    u.apply_graphs(deltas)
    qs = u.modularity_graphs()
    rng = np.random.default_rng(41)
    for i in sorted(rng.choice(len(graphs), size=32, replace=False).tolist()):
        e = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=P))
        e.load(graphs[i])
        e.hit_leiden(deltas[i])
        want = canon_rows(e.hierarchy())
        for p, (a, b) in enumerate(zip(want, u.graph_hierarchy(i))):
            np.testing.assert_array_equal(a, b, err_msg=f"graph {i} row {p}")
        assert abs(e.modularity(hl.PART_COMMUNITY) - qs[i]) < 1e-12
        e.close()
    u.close()


# ---------------------------------------------------------------- served: 256 graphs

@pytest.mark.timeout(900)
def test_served_256_graphs_sample_alone():
    ws = gen.served_config(count=256)
    sizes = [w.n for w in ws]
    say(f"served: {len(ws)} graphs, {sum(sizes)} vertices")
    u = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=P))
    u.load_graphs([hl.GraphInput(w.n, w.u, w.v) for w in ws])
    rq = gen.served_requests(sizes, 2000.0, 400, seed=77)
    reqs = [(g, hl.GraphInput(sizes[g], a, b), t) for g, a, b, t in rq]
    got = {}
    done, pass_of, st = u.serve(reqs, max_pass_records=1 << 20, coalesce=True,
                                emit=lambda r, lab: got.__setitem__(r, lab))
    say(f"served: passes={st.passes} max_pass_requests={st.max_pass_requests}")
    u.close()
    assert len(got) == len(reqs) and st.passes < len(reqs)
    hit = sorted({g for g, _, _ in reqs})
    rng = np.random.default_rng(5)
    for gi in sorted(rng.choice(hit, size=min(12, len(hit)), replace=False).tolist()):
        w = ws[gi]
        mine = sorted((t, i) for i, (gg, _, t) in enumerate(reqs) if gg == gi)
        e = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=P))
        e.load(hl.GraphInput(w.n, w.u, w.v))
        for p in sorted({int(pass_of[i]) for _, i in mine}):
            group = [i for _, i in mine if int(pass_of[i]) == p]
            b = [reqs[i][1] for i in group]
            e.hit_leiden(hl.GraphInput(w.n, np.concatenate([x.src for x in b]), np.concatenate([x.dst for x in b])))
            want = _first_seen(e.hierarchy()[-1])
            for i in group:
                np.testing.assert_array_equal(_first_seen(got[i]), want, err_msg=f"request {i} of graph {gi}")
        e.close()
