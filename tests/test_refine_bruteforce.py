"""Brute-force pins of the oracle's singleton-merge refinement (deterministic
mode, try_merge_singleton_refinement + should_skip_candidate_t_filter,
hit_leiden.rs:1359-1548).

The simulation below does not use the oracle's gain or filter formulas.  It
scores every decision by differences of the resolution-gamma modularity
Q_g(P) = sum_c [2 w_int(c) / 2m - g (d_c / 2m)^2] of whole partitions, in exact
rational arithmetic:

* merging singleton {v} into subcommunity S changes Q_g of the subcommunity
  partition by exactly twice the reference's gain (w(v,S)/2m - g k_v d_S/(2m)^2
  for a self-loop-free v), so the argmax and the "> 0" test use that difference;
* the T-filter value of S inside community C is half the change of Q_g of the
  community partition when S leaves C for a community of its own, minus
  2 w_int(S) / 2m: the reference's w(S, C) counts the edges inside S from both
  ends (hit_leiden.rs:1534-1545).

Visit order (degree ascending, then id), the singleton test, candidates from
same-community neighbours in ascending id with strict '>' and the 1e-9
threshold follow the reference.  A planted mistake (the resolution term of the
gain dropped) must disagree with the oracle on some of the same graphs.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import hl_oracle as O


def modularity_g(n, edges, labels, gamma):
    two_m = Fr(2 * len(edges))
    deg = [0] * n
    w_int = {}
    for a, b in edges:
        deg[a] += 1
        deg[b] += 1
        if labels[a] == labels[b]:
            w_int[labels[a]] = w_int.get(labels[a], 0) + 1
    d = {}
    for v in range(n):
        d[labels[v]] = d.get(labels[v], 0) + deg[v]
    return sum(Fr(2 * w_int.get(c, 0)) / two_m - gamma * (Fr(dc) / two_m) ** 2 for c, dc in d.items())


def simulate(n, edges, f, gamma, drop_gamma=False):
    """s starts as the identity, so every vertex is refined (hit_leiden.rs:1333-1339)."""
    g = Fr(gamma)
    two_m = Fr(2 * len(edges))
    nbr = [[] for _ in range(n)]
    for a, b in edges:
        nbr[a].append(b)
        nbr[b].append(a)
    deg = [len(x) for x in nbr]
    s = list(range(n))
    for v in sorted(range(n), key=lambda x: (deg[x], x)):
        if sum(1 for x in s if x == s[v]) != 1:
            continue
        C = f[v]
        cands = sorted({s[u] for u in nbr[v] if f[u] == C and s[u] != s[v]})
        q0 = modularity_g(n, edges, s, g if not drop_gamma else Fr(0))
        best, best_gain = s[v], Fr(0)
        for S in cands:
            members = [x for x in range(n) if s[x] == S]
            # T-filter: half the change of the community partition when S leaves C, minus the
            # internal S edges counted from their second end
            fl = list(f)
            for x in members:
                fl[x] = n + 1000
            dq = modularity_g(n, edges, fl, g) - modularity_g(n, edges, f, g)
            w_int = sum(1 for a, b in edges if s[a] == S and s[b] == S)
            if dq / 2 - Fr(2 * w_int) / two_m > Fr(1, 10**9):
                continue
            trial = list(s)
            trial[v] = S
            gain = (modularity_g(n, edges, trial, g if not drop_gamma else Fr(0)) - q0) / 2
            if gain > best_gain:
                best, best_gain = S, gain
        if best_gain > 0:
            s[v] = best
    return s


def random_case(rng):
    n = int(rng.integers(5, 10))
    pairs = [(a, b) for a in range(n) for b in range(a + 1, n)]
    k = int(rng.integers(n, min(len(pairs), 3 * n) + 1))
    pick = rng.choice(len(pairs), size=k, replace=False)
    edges = [pairs[i] for i in sorted(pick)]
    f = rng.integers(0, int(rng.integers(1, 4)), n).tolist()
    gamma = float(rng.choice([0.5, 1.0, 2.0]))
    return n, edges, f, gamma


def oracle_refine(n, edges, f, gamma):
    s, R = O.inc_refinement(n, edges, None, f, list(range(n)), [0] * n, gamma, O.DET)
    assert R.all()
    return s.tolist()


def test_singleton_merges_match_brute_force():
    rng = np.random.default_rng(2024)
    merged = 0
    for _ in range(300):
        n, edges, f, gamma = random_case(rng)
        want = simulate(n, edges, f, gamma)
        got = oracle_refine(n, edges, f, gamma)
        assert got == want, (n, edges, f, gamma)
        merged += sum(1 for v in range(n) if got[v] != v)
    assert merged > 100  # the cases exercise merges, not only the no-op path


def test_planted_gain_mistake_is_caught():
    """Dropping the resolution term of the gain changes the outcome on some graphs."""
    rng = np.random.default_rng(2024)
    differ = 0
    for _ in range(300):
        n, edges, f, gamma = random_case(rng)
        if simulate(n, edges, f, gamma, drop_gamma=True) != oracle_refine(n, edges, f, gamma):
            differ += 1
    assert differ > 0


@pytest.mark.parametrize("gamma", [0.5, 1.0, 2.0])
def test_each_merge_raises_subcommunity_modularity(gamma):
    """Every merge the oracle makes raises Q_g of the subcommunity partition
    (the gain is positive and equals half the change)."""
    rng = np.random.default_rng(7)
    for _ in range(100):
        n, edges, f, _ = random_case(rng)
        s = oracle_refine(n, edges, f, gamma)
        assert modularity_g(n, edges, s, Fr(gamma)) >= modularity_g(n, edges, list(range(n)), Fr(gamma))


def simulate_throughput(n, edges, f, gamma, connected):
    """Snapshot refinement (inc_refinement_parallel + compute_blocked_subcommunities,
    throughput.rs:239-359, hit_leiden.rs:1554-1637) from the identity s: every
    vertex proposes against the snapshot, a candidate S is blocked when its
    T-filter value (as above) exceeds 1e-9, and all proposals apply
    (throughput.rs:278-287) -- or, with `connected`, all but those whose own
    subcommunity another proposal targets."""
    g = Fr(gamma)
    two_m = Fr(2 * len(edges))
    nbr = [[] for _ in range(n)]
    for a, b in edges:
        nbr[a].append(b)
        nbr[b].append(a)
    s = list(range(n))
    q0 = modularity_g(n, edges, s, g)
    blocked = {}

    def is_blocked(S):
        if S not in blocked:
            fl = list(f)
            fl[S] = n + 1000  # s is the identity: subcommunity S is vertex S alone
            dq = modularity_g(n, edges, fl, g) - modularity_g(n, edges, f, g)
            blocked[S] = dq / 2 > Fr(1, 10**9)
        return blocked[S]

    prop = {}
    for v in range(n):
        cands = sorted({u for u in nbr[v] if f[u] == f[v]})
        best, best_gain = v, Fr(0)
        for S in cands:
            if is_blocked(S):
                continue
            trial = list(s)
            trial[v] = S
            gain = (modularity_g(n, edges, trial, g) - q0) / 2
            if gain > best_gain:
                best, best_gain = S, gain
        if best_gain > 0:
            prop[v] = best
    targets = set(prop.values())
    for v, S in prop.items():
        if connected and v in targets:
            continue
        s[v] = S
    return s


@pytest.mark.parametrize("mode", [O.THR, O.PAR])
def test_throughput_refinement_matches_brute_force(mode):
    rng = np.random.default_rng(99)
    moved = 0
    for _ in range(300):
        n, edges, f, gamma = random_case(rng)
        want = simulate_throughput(n, edges, f, gamma, connected=mode == O.PAR)
        got, _ = O.inc_refinement(n, edges, None, f, list(range(n)), [0] * n, gamma, mode)
        assert got.tolist() == want, (n, edges, f, gamma)
        moved += sum(1 for v in range(n) if want[v] != v)
    assert moved > 100
