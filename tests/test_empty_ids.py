"""Throughput empty-community ids (parallel_frontier.rs:204-216) in the oracle.

The reference gives vertex v the empty community n + v.  That id is free only
while every label in [n, 2n) was made by the empty moves of the same level:
with several levels a level also holds labels pulled down from the level above
(def_update, hit_leiden.rs:1733-1793), fresh ids of vertices added by growth
and labels a caller writes, and n + v can then name a populated community.
The oracle (D10) relabels a level's labels by rank before its throughput
movement: the order of labels, and so every tie decision, is kept, every label
ends below n, and n + v is always free.  Mode bit REF_IDS skips the relabel
(the reference).  These tests count, inside the oracle, moves to an empty id
whose community total was not zero when the round's moves were selected.
"""
import numpy as np
import pytest

from hit_leiden_amd import gen
from oracle import hl_oracle as O


@pytest.fixture
def collisions():
    L = O.lib()
    L.or_empty_collisions(1)
    yield L
    L.or_empty_collisions(1)


def two_triangles_state(gamma, mode=O.PAR):
    """Two disjoint triangles, labels [7, 7, 7, 8, 8, 8] written by the caller
    (n = 6: label 8 = n + 2 is vertex 2's empty id under n + v).  At gamma = 4
    leaving a triangle for an empty community gains -2/12 + 4*2*(6-2)/144 > 0
    and no neighbour community is better, so every vertex proposes its empty id;
    at gamma = 2 the same gain is negative and nothing moves."""
    u = np.array([0, 1, 0, 3, 4, 3], np.uint32)
    v = np.array([1, 2, 2, 4, 5, 5], np.uint32)
    o = O.OracleState(levels=1, gamma=gamma, mode=mode)
    o.load_graph(6, u, v)
    o.write_map(0, O.MAP_F, np.array([7, 7, 7, 8, 8, 8], np.uint32))
    return o


def test_reference_ids_join_a_populated_community(collisions):
    """With n + v, vertex 2's move to "empty" 8 lands in triangle {3,4,5}."""
    o = two_triangles_state(4.0, mode=O.PAR | O.REF_IDS)
    o.apply(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    assert collisions.or_empty_collisions(1) >= 1
    f = o.read_map(0, O.MAP_F)
    assert f.tolist() == [6, 7, 8, 8, 8, 8]  # vertex 2 joined the other triangle's community


def test_relabelled_ids_stay_empty(collisions):
    """D10: the written labels become ranks [0,0,0,1,1,1] first, so n + v is free."""
    o = two_triangles_state(4.0)
    o.apply(np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    assert collisions.or_empty_collisions(1) == 0
    # every vertex left for a community of its own (ids 6 + v)
    assert o.read_map(0, O.MAP_F).tolist() == [6, 7, 8, 9, 10, 11]


@pytest.mark.parametrize("levels", [1, 2, 4])
def test_multilevel_stream_never_collides(collisions, levels):
    """Growth, pulls and deletions over 20 batches at gamma 2: no empty move
    lands in a populated community, and every level's labels stay below twice
    its size (the relabel keeps the label space from growing)."""
    w = gen.lfr(1500, 8.0, 0.3, seed=5)
    w.make_batches(20, 150, 60, mode=0, seed=6)
    o = O.OracleState(levels=levels, gamma=2.0, mode=O.PAR)
    o.load(w.n, w.u, w.v)
    for (u, v, a) in w.batches:
        o.apply(u, v, a)
        for p in range(levels):
            assert o.read_map(p, O.MAP_F).max() < 2 * o.level_n(p)
    assert collisions.or_empty_collisions(1) == 0


def test_relabel_keeps_every_decision():
    """The relabel only renames communities: with unit weights the partitions
    (first-seen relabelled) equal those of the run without it at one level,
    where labels never come from elsewhere."""
    from hit_leiden_amd import _first_seen
    w = gen.lfr(2000, 8.0, 0.3, seed=8)
    w.make_batches(6, 200, 60, mode=0, seed=9)
    a = O.OracleState(levels=1, mode=O.PAR)
    b = O.OracleState(levels=1, mode=O.PAR | O.REF_IDS)
    a.load(w.n, w.u, w.v)
    b.load(w.n, w.u, w.v)
    for (u, v, x) in w.batches:
        a.apply(u, v, x)
        b.apply(u, v, x)
        np.testing.assert_array_equal(_first_seen(a.read_map(0, O.MAP_F)), _first_seen(b.read_map(0, O.MAP_F)))
        np.testing.assert_array_equal(a.read_map(0, O.MAP_S_CUR), b.read_map(0, O.MAP_S_CUR))
