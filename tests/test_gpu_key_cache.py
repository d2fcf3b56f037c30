"""The throughput movement with the neighbour-community cache on
(RunConfig.key_cache = 1), against the oracle on the same inputs: every proposal kernel then reads
the cache instead of the community array, and the moves keep it current, in the
one-workgroup decoupling and in the multi-kernel path.  Unit weights keep every
sum exact, so the comparison is bit-exact."""
import numpy as np
import pytest

import hit_leiden_amd as hl
from hit_leiden_amd import gen
from parity_util import assert_same_state, assert_supergraph_consistent, pair
from test_gpu_parity import random_batch, random_graph, say

pytestmark = pytest.mark.gpu

THR = hl.RunMode.THROUGHPUT


@pytest.mark.parametrize("levels", [2, 4])
def test_cache_random_graphs(levels):
    for seed in range(3):
        n = 300 + 100 * seed
        u, v, _ = random_graph(n, 4 * n, seed)
        say(f"cache random P={levels} seed={seed}")
        g, o = pair(THR, levels, key_cache=1)
        g.load(hl.GraphInput(n, u, v))
        o.load(n, u, v)
        assert_same_state(g, o, levels, f"load seed {seed}")
        for b in range(3):
            bu, bv, ba = random_batch(n, u, v, None, 60, 100 * seed + b)
            g.hit_leiden(hl.GraphInput(n, bu, bv, ba))
            o.apply(bu, bv, ba)
            assert_same_state(g, o, levels, f"seed {seed} batch {b}")
        assert_supergraph_consistent(g, levels)
        g.close()


def test_cache_small_config():
    w = gen.small_config()
    P = 4
    g, o = pair(THR, P, key_cache=1)
    g.load(hl.GraphInput(w.n, w.u, w.v))
    o.load(w.n, w.u, w.v)
    assert_same_state(g, o, P, "small load")
    for i, (u, v, a) in enumerate(w.batches):
        g.hit_leiden(hl.GraphInput(w.n, u, v, a))
        o.apply(u, v, a)
        assert_same_state(g, o, P, f"small batch {i}")
    g.close()


@pytest.mark.parametrize("hub_deg,n", [(300, 30000), (12000, 30000)])
def test_cache_hubs(hub_deg, n):
    """Hub rows: every move next to a hub updates an entry deep in its row."""
    w = gen.sbm(n, 20, 8.0, 0.2, seed=5)
    rng = np.random.default_rng(6)
    hubs = [np.full(hub_deg, h, np.uint32) for h in (0, 1, 2)]
    nbrs = [rng.choice(np.arange(3, n, dtype=np.uint32), size=hub_deg, replace=False) for _ in hubs]
    u = np.concatenate([w.u] + hubs)
    v = np.concatenate([w.v] + nbrs)
    key = np.unique(np.minimum(u, v).astype(np.int64) * n + np.maximum(u, v))
    u, v = (key // n).astype(np.uint32), (key % n).astype(np.uint32)
    say(f"cache hubs deg={hub_deg}")
    g, o = pair(THR, 2, key_cache=1)
    g.load(hl.GraphInput(n, u, v))
    o.load(n, u, v)
    assert_same_state(g, o, 2, "hubs load")
    for step in range(2):
        bu, bv, ba = random_batch(n, u, v, None, 500, 90 + step)
        g.hit_leiden(hl.GraphInput(n, bu, bv, ba))
        o.apply(bu, bv, ba)
        assert_same_state(g, o, 2, f"hubs batch {step}")
    g.close()


def test_cache_multi_kernel_decoupling():
    """A load from singletons proposes far more than 4,096 moves in its first
    rounds, which the multi-kernel decoupling takes; the batches after it use
    the one-workgroup path."""
    w = gen.sbm(60000, 120, 16.0, 0.3, seed=31)
    w.make_batches(3, 1000, 200, mode=1, seed=32)
    P = 3
    say("cache multi-kernel")
    g, o = pair(THR, P, key_cache=1)
    g.load(hl.GraphInput(w.n, w.u, w.v))
    o.load(w.n, w.u, w.v)
    assert_same_state(g, o, P, "load")
    for i, (u, v, a) in enumerate(w.batches):
        g.hit_leiden(hl.GraphInput(w.n, u, v, a))
        o.apply(u, v, a)
        assert_same_state(g, o, P, f"batch {i}")
    g.close()


def test_cache_on_equals_off():
    w = gen.sbm(40000, 60, 12.0, 0.25, seed=41)
    w.make_batches(4, 800, 100, mode=1, seed=42)
    rows = []
    for kc in (1, 0):
        e = hl.HitLeiden(hl.RunConfig(mode=THR, levels=4, key_cache=kc))
        e.load(hl.GraphInput(w.n, w.u, w.v))
        for (u, v, a) in w.batches:
            e.hit_leiden(hl.GraphInput(w.n, u, v, a))
        rows.append((e.hierarchy(), e.modularity(hl.PART_COMMUNITY)))
        e.close()
    assert np.array_equal(rows[0][0], rows[1][0])
    assert rows[0][1] == rows[1][1]


def test_key_cache_config_checked():
    with pytest.raises(hl.HitLeidenError):
        hl.HitLeiden(hl.RunConfig(mode=THR, levels=2, key_cache=2))
