"""Throughput empty-community ids (parallel_frontier.rs:204-216) in the oracle.

The reference gives vertex v the empty community n + v.  That id is free only
while every label in [n, 2n) was made by the empty moves of the same level:
with several levels a level also holds labels pulled down from the level above
(def_update, hit_leiden.rs:1733-1793), fresh ids of vertices added by growth
and labels a caller writes, and n + v can then name a populated community.
The oracle's rule (D10) keeps n + v where the reference defines it and places
the range above every label in use elsewhere.  These tests count, inside the
oracle, moves to an empty id whose community total was not zero when the
round's moves were selected.
"""
import numpy as np
import pytest

from hit_leiden_amd import gen
from oracle import hl_oracle as O


@pytest.fixture
def collisions():
    L = O.lib()
    L.or_set_reference_empty_ids(0)
    L.or_empty_collisions(1)
    yield L
    L.or_set_reference_empty_ids(0)
    L.or_empty_collisions(1)


def two_triangles_state(gamma):
    """Two disjoint triangles, labels [7, 7, 7, 8, 8, 8] written by the caller
    (n = 6: label 8 = n + 2 is vertex 2's empty id under n + v).  At gamma = 4
    leaving a triangle for an empty community gains -2/12 + 4*2*(6-2)/144 > 0
    and no neighbour community is better, so every vertex proposes its empty id;
    at gamma = 2 the same gain is negative and nothing moves."""
    u = np.array([0, 1, 0, 3, 4, 3], np.uint32)
    v = np.array([1, 2, 2, 4, 5, 5], np.uint32)
    o = O.OracleState(levels=1, gamma=gamma, mode=O.PAR)
    o.load_graph(6, u, v)
    o.write_map(0, O.MAP_F, np.array([7, 7, 7, 8, 8, 8], np.uint32))
    return o


def test_reference_ids_join_a_populated_community(collisions):
    """With n + v, vertex 2's move to "empty" 8 lands in triangle {3,4,5}."""
    collisions.or_set_reference_empty_ids(1)
    o = two_triangles_state(4.0)
    o.apply(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    assert collisions.or_empty_collisions(1) >= 1
    f = o.read_map(0, O.MAP_F)
    assert f.tolist() == [6, 7, 8, 8, 8, 8]  # vertex 2 joined the other triangle's community


def test_placed_range_stays_empty(collisions):
    """D10: caller-written labels move the empty range above them (ebase = 9)."""
    o = two_triangles_state(4.0)
    o.apply(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    assert collisions.or_empty_collisions(1) == 0
    assert collisions.or_empty_base(o.h, 0) == 9
    # every vertex left for a community of its own (ids 9 + v)
    assert o.read_map(0, O.MAP_F).tolist() == [9, 10, 11, 12, 13, 14]


@pytest.mark.parametrize("levels", [1, 2, 4])
def test_multilevel_stream_never_collides(collisions, levels):
    """Growth, pulls and deletions over 20 batches: no empty move lands in a
    populated community, one level keeps n + v, and the label space stays
    within a fixed multiple of the level size (no creep from batch to batch)."""
    w = gen.lfr(1500, 8.0, 0.3, seed=5)
    w.make_batches(20, 150, 60, mode=0, seed=6)
    o = O.OracleState(levels=levels, gamma=2.0, mode=O.PAR)
    o.load(w.n, w.u, w.v)
    bases = []
    for (u, v, a) in w.batches:
        o.apply(u, v, a)
        bases.append([collisions.or_empty_base(o.h, p) for p in range(levels)])
    assert collisions.or_empty_collisions(1) == 0
    if levels == 1:
        assert all(b == [w.n] for b in bases)  # the reference's n + v
    top = max(o.level_n(p) for p in range(levels))
    assert max(max(b) for b in bases) <= 4 * levels * top
    assert bases[-1] == bases[len(bases) // 2]  # placed once, then stable
