"""The multi-kernel decoupling with few priority rounds (RunConfig.decide_rounds):
after them the undecided candidates are settled in priority order by one
thread, which must give the sequential greedy's result (throughput.rs:204-236).
A load from singletons proposes far more than 4,096 moves in its first rounds,
so it takes the multi-kernel path; unit weights keep every sum exact, so the
comparisons with the oracle are bit-exact."""
import numpy as np
import pytest

import hit_leiden_amd as hl
from hit_leiden_amd import gen
from parity_util import assert_same_state, assert_supergraph_consistent, pair
from test_gpu_parity import say

pytestmark = pytest.mark.gpu

THR = hl.RunMode.THROUGHPUT


@pytest.mark.parametrize("rounds", [1, 3])
def test_settle_matches_oracle(rounds):
    w = gen.sbm(60000, 120, 16.0, 0.3, seed=31)
    w.make_batches(3, 1000, 200, mode=1, seed=32)
    P = 3
    say(f"decide_rounds={rounds} sbm n={w.n}")
    g, o = pair(THR, P, decide_rounds=rounds)
    g.load(hl.GraphInput(w.n, w.u, w.v))
    o.load(w.n, w.u, w.v)
    assert_same_state(g, o, P, "load")
    for i, (u, v, a) in enumerate(w.batches):
        g.hit_leiden(hl.GraphInput(w.n, u, v, a))
        o.apply(u, v, a)
        assert_same_state(g, o, P, f"batch {i}")
    assert_supergraph_consistent(g, P)
    g.close()


def test_settle_rmat_matches_oracle():
    """R-MAT scale 16: hubs, whose moves chain through shared communities."""
    w = gen.rmat(16, seed=3)
    w.make_batches(2, 800, 100, mode=1, seed=4)
    P = 4
    say(f"decide_rounds=1 rmat n={w.n} m={w.m}")
    g, o = pair(THR, P, decide_rounds=1)
    g.load(hl.GraphInput(w.n, w.u, w.v))
    o.load(w.n, w.u, w.v)
    assert_same_state(g, o, P, "rmat load")
    for i, (u, v, a) in enumerate(w.batches):
        g.hit_leiden(hl.GraphInput(w.n, u, v, a))
        o.apply(u, v, a)
        assert_same_state(g, o, P, f"rmat batch {i}")
    g.close()


def test_settle_equals_rounds():
    """One priority round then the in-order settle, against the default rounds."""
    w = gen.rmat(17, seed=8)
    w.make_batches(3, 2000, 200, mode=1, seed=9)
    rows = []
    for rounds in (1, 0):
        say(f"decide_rounds={rounds} rmat17")
        e = hl.HitLeiden(hl.RunConfig(mode=THR, levels=4, decide_rounds=rounds))
        e.load(hl.GraphInput(w.n, w.u, w.v))
        for (u, v, a) in w.batches:
            e.hit_leiden(hl.GraphInput(w.n, u, v, a))
        rows.append(e.hierarchy())
        e.close()
    assert np.array_equal(rows[0], rows[1])
