"""The oracle's static path: run() / multilevel_leiden (hit_leiden.rs:30-118,
167-323) with the queue shuffle seeded.

* The reference's own tests of this path, restated: the multilevel unit tests
  (hit_leiden.rs:2354-2446: two cliques joined by a bridge give two
  communities, a single vertex, four isolated vertices, hierarchy rows of n),
  tests/integration/test_connected_graph_not_all_singletons.rs:21-48,
  test_deterministic_identity.rs:19-29 and
  tests/property/test_partition_invariants.rs:20-31.
* Exact-arithmetic simulations of its two steps, scoring every decision by
  differences of the resolution-gamma modularity of whole partitions (not by
  the oracle's formulas): the queue movement (single_level_movement,
  :764-932: active endpoints of cross-community edges in shuffled order, FIFO,
  candidates in ascending id with strict '>', no empty community, requeue of
  the neighbours outside the new community in row order) and the refinement
  (refine_singleton_merge, :487-710: components split with the largest kept
  (ties to the smaller minimum vertex), then the split-off singletons merged
  with the T-filter and the EPS tie rules).  The queue order comes from the
  product's shuffle, which test_graphio pins against the oracle's own.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

from hit_leiden_amd import gen
from oracle import hl_oracle as O
from test_refine_bruteforce import modularity_g, random_case


def two_cliques():
    return 6, [0, 1, 0, 3, 4, 3, 2], [1, 2, 2, 4, 5, 5, 3]


def test_two_cliques_two_communities():
    n, u, v = two_cliques()
    lab, H, it = O.static_run(n, u, v)
    assert it > 0 and len(H) >= 1
    assert lab.tolist() == [0, 0, 0, 1, 1, 1]
    assert all(len(r) == n for r in H)


def test_single_vertex_and_isolated_vertices():
    lab, H, it = O.static_run(1, [], [])
    assert lab.tolist() == [0] and it >= 1 and len(H) >= 1
    lab, _, _ = O.static_run(4, [], [])
    assert len(np.unique(lab)) == 4


def test_connected_graph_not_all_singletons():
    n, u, v = two_cliques()
    w = [1.0] * 6 + [0.05]
    lab, _, _ = O.static_run(n, u, v, w)
    assert len(np.unique(lab)) < n


@pytest.mark.parametrize("n", [0, 1, 7, 49])
def test_partition_length(n):
    lab, _, _ = O.static_run(n, [], [])
    assert len(lab) == n


def test_deterministic_replay():
    w = gen.small_config()
    a = O.static_run(w.n, w.u, w.v, seed=5)
    b = O.static_run(w.n, w.u, w.v, seed=5)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    # labels are first-seen contiguous (canonicalize_community_ids_in_place)
    lab = a[0]
    first = [lab[0]]
    for x in lab[1:]:
        if x not in first and x == max(first) + 1:
            first.append(x)
    assert lab.max() + 1 == len(np.unique(lab))


# ---- exact simulations -------------------------------------------------------

def neighbours(n, edges):
    nbr = [[] for _ in range(n)]
    for a, b in edges:
        nbr[a].append(b)
        nbr[b].append(a)
    return [sorted(x) for x in nbr]


def sim_movement(n, edges, f, gamma, seed):
    g = Fr(gamma)
    nbr = neighbours(n, edges)
    f = list(f)
    active = sorted({x for a, b in edges if f[a] != f[b] for x in (a, b)})
    if not edges:
        active = list(range(n))
    order = gen.shuffle_order(seed, len(active)) if active else []
    queue = [active[i] for i in order]
    inq = set(queue)
    h = 0
    while h < len(queue):
        v = queue[h]
        h += 1
        inq.discard(v)
        q0 = modularity_g(n, edges, f, g)
        best, bg = f[v], Fr(0)
        for c in sorted({f[u] for u in nbr[v]} - {f[v]}):
            trial = list(f)
            trial[v] = c
            gain = (modularity_g(n, edges, trial, g) - q0) / 2
            if gain > bg:
                best, bg = c, gain
        if bg <= 0:
            continue
        f[v] = best
        for u in nbr[v]:
            if f[u] != best and u not in inq:
                queue.append(u)
                inq.add(u)
    return f


def sim_refine(n, edges, f, gamma):
    g = Fr(gamma)
    two_m = Fr(2 * len(edges))
    nbr = neighbours(n, edges)
    deg = [len(x) for x in nbr]
    sub = list(f)
    nxt = max(sub) + 1
    refined = []
    for sc in sorted(set(sub)):
        members = [x for x in range(n) if sub[x] == sc]
        seen, comps = set(), []
        for st in members:
            if st in seen:
                continue
            comp, stack = [], [st]
            seen.add(st)
            while stack:
                x = stack.pop()
                comp.append(x)
                for y in nbr[x]:
                    if sub[y] == sc and y not in seen:
                        seen.add(y)
                        stack.append(y)
            comps.append(comp)
        if len(comps) <= 1:
            continue
        keep = max(range(len(comps)), key=lambda i: (len(comps[i]), -min(comps[i])))
        for i, comp in enumerate(comps):
            if i == keep:
                continue
            for x in comp:
                sub[x] = nxt
                refined.append(x)
            nxt += 1
    for v in sorted(refined, key=lambda x: (deg[x], x)):
        if sum(1 for x in sub if x == sub[v]) != 1:
            continue
        C = f[v]
        cands = sorted({sub[u] for u in nbr[v] if f[u] == C} - {sub[v]})
        q0 = modularity_g(n, edges, sub, g)
        best, bg, bw = sub[v], Fr(0), Fr(-1)
        for S in cands:
            members = [x for x in range(n) if sub[x] == S]
            fl = list(f)
            for x in members:
                fl[x] = n + 1000
            dq = modularity_g(n, edges, fl, g) - modularity_g(n, edges, f, g)
            w_int = sum(1 for a, b in edges if sub[a] == S and sub[b] == S)
            if dq / 2 - Fr(2 * w_int) / two_m > Fr(1, 10**12):
                continue
            trial = list(sub)
            trial[v] = S
            gain = (modularity_g(n, edges, trial, g) - q0) / 2
            w = Fr(sum(1 for u in nbr[v] if sub[u] == S))
            if gain > bg or (gain == bg and (w > bw or (w == bw and S < best))):
                best, bg, bw = S, gain, w
        if bg > 0:
            sub[v] = best
    return sub


def test_static_movement_matches_simulation():
    rng = np.random.default_rng(11)
    moved = 0
    for case in range(200):
        n, edges, f, gamma = random_case(rng)
        seed = int(rng.integers(0, 2**32))
        want = sim_movement(n, edges, f, gamma, seed)
        got = O.static_movement(n, edges, None, f, gamma, seed)
        assert got.tolist() == want, (case, n, edges, f, gamma, seed)
        moved += sum(1 for a, b in zip(want, f) if a != b)
    assert moved > 100


def test_static_refine_matches_simulation():
    rng = np.random.default_rng(12)
    changed = 0
    for case in range(200):
        n, edges, f, gamma = random_case(rng)
        want = sim_refine(n, edges, f, gamma)
        got = O.static_refine(n, edges, None, f, gamma)
        assert got.tolist() == want, (case, n, edges, f, gamma)
        changed += sum(1 for a, b in zip(want, f) if a != b)
    assert changed > 50
