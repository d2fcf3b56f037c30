"""Persistent rounds: a level's throughput movement rounds in one cooperative
kernel (grid barriers between the proposals, the decoupling and the marks)
give the states of the per-round launches, bit-exact with unit weights, and
the same counters (rounds, evaluations, moves).  Covered: rounds that stay in
the kernel to convergence, a first round with more candidates than the
decoupling workgroup takes (the kernel leaves it and the rest to the per-round
launches), rows longer than the kernel takes (the level skips it), round
caps that end the kernel early, grids of 2 to 64 workgroups, and a second
process, to which the runtime refuses cooperative launches while this one
holds them (its rounds take the per-round launches)."""
import numpy as np
import pytest

import hit_leiden_amd as hl
from hit_leiden_amd import gen
from parity_util import MAPS, SETS, assert_same_state, pair

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def engine(levels, persist, persist_blocks=0, **cfg):
    if persist and persist_blocks:
        persist = persist_blocks
    return hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=levels, persist_rounds=persist, **cfg))


def assert_engines_equal(a, b, levels, tag):
    for p in range(levels):
        assert a.level_size(p) == b.level_size(p), f"{tag} level {p} size"
        for which in MAPS:
            assert np.array_equal(a.partition(p, which), b.partition(p, which)), f"{tag} level {p} map {which}"
        for which in SETS:
            assert np.array_equal(a.frontier(p, which), b.frontier(p, which)), f"{tag} level {p} set {which}"
        for x, y in zip(a.supergraph(p), b.supergraph(p)):
            assert np.array_equal(x, y), f"{tag} level {p} CSR"


def assert_counters_equal(sa, sb, levels, tag):
    for key in ("rounds", "evaluated", "moved", "active", "affected", "refined"):
        assert list(getattr(sa, key)[:levels]) == list(getattr(sb, key)[:levels]), (tag, key)


def run_both(n, u, v, batches, levels, **cfg):
    """Loads and applies the batches with persistent rounds on and off;
    returns the rounds the kernel ran, summed over the calls."""
    on, off = engine(levels, True, **cfg), engine(levels, False, **cfg)
    ran = 0
    try:
        on.load(hl.GraphInput(n, u, v))
        off.load(hl.GraphInput(n, u, v))
        assert off.last_stats.persist_rounds == 0
        ran += on.last_stats.persist_rounds
        assert_counters_equal(on.last_stats, off.last_stats, levels, "load")
        assert_engines_equal(on, off, levels, "load")
        for i, (bu, bv, ba) in enumerate(batches):
            on.hit_leiden(hl.GraphInput(n, bu, bv, ba))
            off.hit_leiden(hl.GraphInput(n, bu, bv, ba))
            assert off.last_stats.persist_rounds == 0
            ran += on.last_stats.persist_rounds
            assert_counters_equal(on.last_stats, off.last_stats, levels, f"batch {i}")
            assert_engines_equal(on, off, levels, f"batch {i}")
    finally:
        on.close()
        off.close()
    return ran


def random_edges(n, m, rng):
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    key = np.unique(np.minimum(u[keep], v[keep]).astype(np.int64) * n + np.maximum(u[keep], v[keep]))
    return (key // n).astype(np.uint32), (key % n).astype(np.uint32)


def insert_batches(n, count, k, rng):
    out = []
    for _ in range(count):
        a, b = random_edges(n, k, rng)
        out.append((a, b, np.ones(len(a))))
    return out


def test_small_config_persist_matches_launches():
    """LFR-style n=9,877 + 5 x 1,000 new edges, 4 levels: the batches' rounds
    run in the kernel, states and counters equal the per-round launches'."""
    w = gen.small_config()
    say("persist: small config")
    ran = run_both(w.n, w.u, w.v, w.batches, 4)
    say(f"persist: small config, {ran} rounds in the kernel")
    assert ran > 0


def test_small_config_persist_oracle():
    """The same with the oracle (parallel mode) as the checker."""
    w = gen.small_config()
    P = 4
    g, o = pair(hl.RunMode.THROUGHPUT, P, persist_rounds=True)
    g.load(hl.GraphInput(w.n, w.u, w.v))
    o.load(w.n, w.u, w.v)
    assert_same_state(g, o, P, "persist load")
    for i, (u, v, a) in enumerate(w.batches):
        g.hit_leiden(hl.GraphInput(w.n, u, v, a))
        o.apply(u, v, a)
        assert g.last_stats.persist_rounds > 0
        assert_same_state(g, o, P, f"persist batch {i}")
    g.close()


def test_first_round_leaves_the_kernel():
    """A fresh load of 12,000 vertices: the first round's candidates exceed
    the decoupling workgroup, so the kernel stops before applying anything and
    the per-round launches take over; the batches after it stay in the kernel."""
    rng = np.random.default_rng(31)
    n = 12000
    u, v = random_edges(n, 36000, rng)
    say("persist: fresh load n=12000")
    ran = run_both(n, u, v, insert_batches(n, 4, 300, rng), 3)
    assert ran > 0


def test_long_rows_skip_the_kernel():
    """A hub of 400 neighbours: level 0 takes the per-round launches (rows
    above 256 entries), the levels above may take the kernel."""
    rng = np.random.default_rng(37)
    n = 5000
    u, v = random_edges(n, 12000, rng)
    hub = np.arange(1, 401, dtype=np.uint32)
    u = np.concatenate([u, np.zeros(400, np.uint32)])
    v = np.concatenate([v, hub])
    key = np.unique(np.minimum(u, v).astype(np.int64) * n + np.maximum(u, v))
    u, v = (key // n).astype(np.uint32), (key % n).astype(np.uint32)
    say("persist: hub of 400")
    run_both(n, u, v, insert_batches(n, 3, 200, rng), 3)


def test_long_lists_leave_the_kernel():
    """A fresh load of 30,000 vertices (a list longer than the kernel takes)
    and batches whose frontiers grow past it: those rounds take the per-round
    launches."""
    rng = np.random.default_rng(43)
    n = 30000
    u, v = random_edges(n, 90000, rng)
    say("persist: fresh load n=30000")
    run_both(n, u, v, insert_batches(n, 2, 6000, rng), 3)


@pytest.mark.parametrize("blocks", [2, 7, 64])
def test_grid_sizes(blocks):
    """Any grid the kernel is given gives the same states."""
    w = gen.small_config()
    say(f"persist: small config, {blocks} workgroups")
    assert run_both(w.n, w.u, w.v, w.batches[:3], 4, persist_blocks=blocks) > 0


@pytest.mark.parametrize("max_rounds", [1, 2, 3])
def test_round_caps(max_rounds):
    """max_rounds ends the kernel's loop where it ends the launches'."""
    rng = np.random.default_rng(41 + max_rounds)
    n = 4000
    u, v = random_edges(n, 10000, rng)
    say(f"persist: max_rounds={max_rounds}")
    run_both(n, u, v, insert_batches(n, 3, 150, rng), 3, max_rounds=max_rounds)


_CHILD = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import hit_leiden_amd as hl
from hit_leiden_amd import gen
w = gen.small_config()
e = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=4))
e.load(hl.GraphInput(w.n, w.u, w.v))
ran = 0
for (u, v, a) in w.batches[:2]:
    e.hit_leiden(hl.GraphInput(w.n, u, v, a))
    ran += e.last_stats.persist_rounds
print(json.dumps({"ran": int(ran), "f": [e.partition(p, hl.MAP_F).tolist() for p in range(4)]}))
"""


def test_second_process_takes_the_launches():
    """While this process holds the device's cooperative launches, the runtime
    refuses them to a second process: its rounds take the per-round launches
    and give the same hierarchy."""
    import json
    import os
    import subprocess
    import sys
    w = gen.small_config()
    e = engine(4, True)
    e.load(hl.GraphInput(w.n, w.u, w.v))
    ran = 0
    for (u, v, a) in w.batches[:2]:
        e.hit_leiden(hl.GraphInput(w.n, u, v, a))
        ran += e.last_stats.persist_rounds
    assert ran > 0
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    say("persist: a second process")
    r = subprocess.run([sys.executable, "-c", _CHILD, root], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    say(f"persist: the second process ran {out['ran']} rounds in the kernel")
    for p in range(4):
        assert np.array_equal(np.array(out["f"][p], np.uint32), e.partition(p, hl.MAP_F)), f"level {p}"
    e.close()
