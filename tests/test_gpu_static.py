"""The static path on the GPU, hl_run_static (run() / multilevel_leiden,
hit_leiden.rs:30-118, 167-323), against the oracle's restatement.

With unit or dyadic weights every sum the path forms (degrees, community
degrees, w(v, C), w(S, C)) is an exact f64 whatever its order, so the GPU and
the oracle take the same decisions and the comparison is exact: labels,
hierarchy rows and the movement call count.  With arbitrary real weights the
device's atomic sums may round differently, and the bar is the reference's
own throughput bar on the modularity (equivalence.rs:41, 0.001).  The oracle
is pinned by test_static_oracle (the reference's own unit tests, exact
simulations of both steps) and test_graphio (the shuffle's generator).
"""
import numpy as np
import pytest

import hit_leiden_amd as hl
from hit_leiden_amd import gen
from oracle import hl_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def eng():
    e = hl.HitLeiden(hl.RunConfig())
    yield e
    e.close()


def run_both(e, n, u, v, w=None, max_levels=10, seed=0, gamma=1.0):
    g = hl.GraphInput(n, u, v, w)
    got = e.static_leiden(g, max_levels=max_levels, seed=seed)
    want = O.static_run(n, u, v, w, gamma=gamma, max_levels=max_levels, seed=seed)
    return got, want


def assert_same(got, want, tag):
    np.testing.assert_array_equal(got[0], want[0], err_msg=f"{tag}: labels")
    assert got[1].shape == want[1].shape, (tag, got[1].shape, want[1].shape)
    np.testing.assert_array_equal(got[1], want[1], err_msg=f"{tag}: hierarchy")
    assert got[2] == want[2], (tag, got[2], want[2])


def test_two_cliques(eng):
    n, u, v = 6, [0, 1, 0, 3, 4, 3, 2], [1, 2, 2, 4, 5, 5, 3]
    got, want = run_both(eng, n, u, v)
    assert got[0].tolist() == [0, 0, 0, 1, 1, 1]
    assert_same(got, want, "two cliques")


@pytest.mark.parametrize("n", [0, 1, 4, 7, 49])
def test_isolated_vertices(eng, n):
    got, want = run_both(eng, n, [], [])
    assert len(got[0]) == n and len(np.unique(got[0])) == n
    assert_same(got, want, f"{n} isolated")


def graphs():
    w = gen.small_config()
    yield "small", w.n, w.u, w.v
    w = gen.sbm(3000, 12, 8.0, 0.2, seed=3)
    yield "sbm3000", w.n, w.u, w.v
    w = gen.lfr(4000, 10.0, 0.3, seed=4)
    yield "lfr4000", w.n, w.u, w.v
    w = gen.planted(2500, 25, 6.0, 0.35, seed=5)
    yield "planted2500", w.n, w.u, w.v


@pytest.mark.parametrize("seed", [0, 5, 2**40 + 7])
def test_matches_oracle_unit_weights(eng, seed):
    for tag, n, u, v in graphs():
        got, want = run_both(eng, n, u, v, seed=seed)
        say(f"{tag} seed {seed}: {len(np.unique(got[0]))} communities, {got[1].shape[0]} rows, {got[2]} calls")
        assert_same(got, want, f"{tag} seed {seed}")


@pytest.mark.parametrize("max_levels", [1, 2, 3])
def test_level_cap(eng, max_levels):
    w = gen.sbm(3000, 12, 8.0, 0.2, seed=3)
    got, want = run_both(eng, w.n, w.u, w.v, max_levels=max_levels, seed=9)
    assert got[1].shape[0] <= max_levels + 1
    assert_same(got, want, f"max_levels {max_levels}")


def test_dyadic_weights_and_duplicates(eng):
    rng = np.random.default_rng(21)
    w = gen.sbm(2000, 10, 8.0, 0.25, seed=21)
    wt = rng.choice([0.5, 1.0, 2.0, 4.0], size=len(w.u))
    # duplicate records and self-loops: summed into one row entry each
    u = np.concatenate([w.u, w.u[:50], np.arange(20, dtype=np.uint32)])
    v = np.concatenate([w.v, w.v[:50], np.arange(20, dtype=np.uint32)])
    x = np.concatenate([wt, wt[:50], np.full(20, 1.5)])
    got, want = run_both(eng, w.n, u, v, x, seed=3)
    assert_same(got, want, "dyadic")


def test_real_weights_quality(eng):
    rng = np.random.default_rng(22)
    w = gen.planted(3000, 30, 6.0, 0.3, seed=22)
    x = rng.uniform(0.1, 3.0, size=len(w.u))
    got, want = run_both(eng, w.n, w.u, w.v, x, seed=1)
    m = hl.HitLeiden(hl.RunConfig())
    m.load_graph(hl.GraphInput(w.n, w.u, w.v, x))
    qg, qo = m.modularity_labels(got[0]), m.modularity_labels(want[0])
    m.close()
    say(f"real weights: Q gpu {qg:.6f} oracle {qo:.6f}, labels equal {np.array_equal(got[0], want[0])}")
    assert abs(qg - qo) <= 1e-3


def test_resolution(eng):
    w = gen.sbm(2000, 10, 8.0, 0.25, seed=31)
    for gamma in (0.5, 2.0):
        e = hl.HitLeiden(hl.RunConfig(resolution=gamma))
        got, want = run_both(e, w.n, w.u, w.v, seed=4, gamma=gamma)
        e.close()
        assert_same(got, want, f"gamma {gamma}")


def test_replay_and_state_untouched(eng):
    w = gen.sbm(1500, 8, 8.0, 0.2, seed=41)
    g = hl.GraphInput(w.n, w.u, w.v)
    e = hl.HitLeiden(hl.RunConfig(levels=2))
    e.load(g)
    before = e.hierarchy()
    a = e.static_leiden(g, seed=11)
    b = e.static_leiden(g, seed=11)
    for x, y in zip(a[:2], b[:2]):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(before, e.hierarchy()):
        np.testing.assert_array_equal(x, y)
    e.close()


def test_rejects_bad_input(eng):
    with pytest.raises(hl.InvalidInput):
        eng.static_leiden(hl.GraphInput(3, [0], [3]))
    with pytest.raises(hl.InvalidInput):
        eng.static_leiden(hl.GraphInput(3, [0], [1], [float("nan")]))
    with pytest.raises(hl.InvalidInput):
        eng.static_leiden(hl.GraphInput(3, [0], [1]), max_levels=0)
