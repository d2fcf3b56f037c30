"""The large configuration at full size on the GPU: R-MAT scale 24 (16.8 M
vertices, 260 M edges), 4 levels, batches of 95,000 R-MAT inserts + 5,000
deletes.  The oracle needs minutes for this graph, so the full size is checked
through properties that do not depend on it (a consistent hierarchy, refined
communities connected in G^0, every level's weighted degrees equal to its
children's sums, exact contraction on the upper levels), and the scale-20
instance of the same parameters is compared with the oracle (bit-exact for
unit weights, modularity within 1e-3)."""
import numpy as np
import pytest

import hit_leiden_amd as hl
from hit_leiden_amd import gen
from full_config_util import P, THREADS, check_contraction, check_degrees_add_up, refined_connected, say
from oracle import hl_oracle as O
from parity_util import assert_hierarchy_consistent

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------- R-MAT scale 24

@pytest.fixture(scope="module")
def rmat24():
    w = gen.large_config(batches=3)
    say(f"R-MAT-24: n={w.n} m={w.m}")
    g = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=P))
    g.load(hl.GraphInput(w.n, w.u, w.v))
    for (u, v, a) in w.batches:
        g.hit_leiden(hl.GraphInput(w.n, u, v, a))
    say(f"R-MAT-24 after 3 batches: Q={g.modularity():.5f} levels={[g.level_size(p) for p in range(P)]}")
    n = w.n
    del w
    yield n, g
    g.close()


@pytest.mark.timeout(600)
def test_rmat24_hierarchy_consistent(rmat24):
    n, g = rmat24
    assert_hierarchy_consistent(g, P, "R-MAT-24")


@pytest.mark.timeout(600)
def test_rmat24_refined_connected(rmat24):
    n, g = rmat24
    refined_connected(g, n, P)


@pytest.mark.timeout(600)
def test_rmat24_aggregation(rmat24):
    n, g = rmat24
    check_degrees_add_up(g, P, rtol=0.0)  # unit weights: every sum is an exact integer
    for p in range(1, P - 1):
        check_contraction(g, p, tol=0.0)


@pytest.mark.timeout(900)
def test_rmat20_tracks_the_oracle():
    """The scale-20 instance of the same parameters, load + 2 batches."""
    w = gen.large_config(scale=20, batches=2)
    say(f"R-MAT-20: n={w.n} m={w.m}; oracle with {THREADS} threads")
    g = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=P))
    o = O.OracleState(levels=P, mode=O.PAR, threads=THREADS)
    g.load(hl.GraphInput(w.n, w.u, w.v))
    o.load(w.n, w.u, w.v)
    for i, (u, v, a) in enumerate(w.batches):
        g.hit_leiden(hl.GraphInput(w.n, u, v, a))
        o.apply(u, v, a)
        qg, qo = g.modularity(hl.PART_COMMUNITY), o.modularity(o.hierarchy()[-1])
        say(f"R-MAT-20 batch {i}: Q gpu {qg:.6f} oracle {qo:.6f}")
        assert abs(qg - qo) <= 1e-3
    # unit weights: the throughput path is bit-exact with the oracle
    assert np.array_equal(g.hierarchy(), o.hierarchy())
    g.close()


