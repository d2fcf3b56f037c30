"""The mid configuration at full size on the GPU: planted partition n = 10^6,
mean degree 32, w ~ U[0.5,1.5], 4 levels, batches of 9,000 inserts + 1,000
deletes, against the oracle's parallel mode on the same inputs: modularity
within 1e-3 after the load and after each of 3 batches (real weights: the
GPU's f64 atomics sum in another order), every refined community connected,
G^{p+1} = the contraction of G^p under s^p, a consistent hierarchy."""
import numpy as np
import pytest

import hit_leiden_amd as hl
from hit_leiden_amd import gen
from full_config_util import P, THREADS, check_contraction, refined_connected, say
from oracle import hl_oracle as O
from parity_util import assert_hierarchy_consistent

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------- mid

@pytest.fixture(scope="module")
def mid():
    w = gen.mid_config(batches=3)
    say(f"mid: n={w.n} m={w.m}; oracle with {THREADS} threads")
    g = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=P))
    o = O.OracleState(levels=P, mode=O.PAR, threads=THREADS)
    g.load(hl.GraphInput(w.n, w.u, w.v, w.w))
    o.load(w.n, w.u, w.v, w.w)
    qs = [(g.modularity(hl.PART_COMMUNITY), o.modularity(o.hierarchy()[-1]))]
    for (u, v, a) in w.batches:
        g.hit_leiden(hl.GraphInput(w.n, u, v, a))
        o.apply(u, v, a)
        qs.append((g.modularity(hl.PART_COMMUNITY), o.modularity(o.hierarchy()[-1])))
    say(f"mid Q (gpu, oracle) after load and each batch: {qs}")
    yield w, g, qs
    g.close()


@pytest.mark.timeout(900)
def test_mid_modularity_tracks_the_oracle(mid):
    w, g, qs = mid
    assert len(qs) == 4
    for i, (qg, qo) in enumerate(qs):
        assert abs(qg - qo) <= 1e-3, (i, qg, qo)


@pytest.mark.timeout(600)
def test_mid_structure(mid):
    w, g, _ = mid
    assert_hierarchy_consistent(g, P, "mid")
    refined_connected(g, w.n, P)
    for p in range(P - 1):
        check_contraction(g, p, tol=1e-9)


