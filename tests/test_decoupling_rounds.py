"""The decoupled selection of the throughput movement (throughput.rs:204-236)
is a sequential greedy over the candidates in priority order: a move is
accepted unless its source community is the target of an accepted move or its
target is the source of one.  The multi-kernel path on the device reaches the
same decisions with priority rounds (rejected by an earlier accepted
conflicting move, waiting while an earlier undecided one could conflict,
accepted otherwise) and, past a round cap, settles the undecided rest in order
against flags of every accepted move, later ones included.  These tests check
by brute force, on random candidate lists with heavy sharing of communities
and on long conflict chains, that rounds plus settle give the greedy's
decisions for every cap."""
import numpy as np
import pytest

UND, ACC, REJ = 0, 1, 2


def greedy(frm, to):
    """throughput.rs:204-236: one pass in priority order."""
    emitters, acceptors, out = set(), set(), []
    for f, t in zip(frm, to):
        if f in acceptors or t in emitters:
            out.append(REJ)
        else:
            out.append(ACC)
            emitters.add(f)
            acceptors.add(t)
    return out


def rounds_then_settle(frm, to, cap, reverse_settle=False):
    """Priority rounds as k_mark4/k_decide4, then k_settle_prep/k_settle_seq.
    reverse_settle plants a mistake (the rest settled in reverse order)."""
    n = len(frm)
    state = [UND] * n
    rounds = 0
    while UND in state:
        if rounds == cap:
            emit = {frm[p] for p in range(n) if state[p] == ACC}
            recv = {to[p] for p in range(n) if state[p] == ACC}
            for p in (reversed(range(n)) if reverse_settle else range(n)):
                if state[p] != UND:
                    continue
                if frm[p] in recv or to[p] in emit:
                    state[p] = REJ
                else:
                    state[p] = ACC
                    emit.add(frm[p])
                    recv.add(to[p])
            break
        # minima over the state at the start of the round
        min_u_to, min_u_from, min_a_to, min_a_from = {}, {}, {}, {}
        for p in range(n):
            if state[p] == UND:
                min_u_to.setdefault(to[p], p)
                min_u_from.setdefault(frm[p], p)
            elif state[p] == ACC:
                min_a_to.setdefault(to[p], p)
                min_a_from.setdefault(frm[p], p)
        big = n + 1
        new = list(state)
        for p in range(n):
            if state[p] != UND:
                continue
            f, t = frm[p], to[p]
            if min_a_to.get(f, big) < p or min_a_from.get(t, big) < p:
                new[p] = REJ
            elif min_u_to.get(f, big) < p or min_u_from.get(t, big) < p:
                pass
            else:
                new[p] = ACC
        assert new != state, "a round decides at least the first undecided candidate"
        state = new
        rounds += 1
    return state


def random_lists(seed, n, comms):
    rng = np.random.default_rng(seed)
    frm = rng.integers(0, comms, n).tolist()
    to = rng.integers(0, comms, n).tolist()
    to = [t if t != f else (t + 1) % comms for f, t in zip(frm, to)]
    return frm, to


@pytest.mark.parametrize("cap", [0, 1, 2, 3, 48])
def test_random_candidates(cap):
    for seed in range(60):
        frm, to = random_lists(seed, 40 + seed, 3 + seed % 9)
        assert rounds_then_settle(frm, to, cap) == greedy(frm, to), seed


@pytest.mark.parametrize("cap", [0, 1, 5, 48])
def test_conflict_chain(cap):
    """c0 -> c1, c1 -> c2, ...: each move waits for the one before it, so the
    rounds alone need one round per link."""
    n = 120
    frm = list(range(n))
    to = list(range(1, n + 1))
    want = greedy(frm, to)
    assert want == [ACC if p % 2 == 0 else REJ for p in range(n)]
    assert rounds_then_settle(frm, to, cap) == want


def test_settle_flags_from_later_moves_are_harmless():
    """A later accepted move never conflicts with an earlier undecided one:
    with the rounds stopped after one, the flags of every accepted move give
    the greedy's decisions."""
    frm = [0, 1, 5, 2]
    to = [1, 2, 6, 3]
    assert rounds_then_settle(frm, to, 1) == greedy(frm, to) == [ACC, REJ, ACC, ACC]


def test_order_of_the_settle_matters():
    """The comparison has teeth: settling the rest in reverse order differs
    from the greedy on the chain and on some random lists."""
    frm, to = list(range(40)), list(range(1, 41))
    assert rounds_then_settle(frm, to, 1, reverse_settle=True) != greedy(frm, to)
    wrong = 0
    for seed in range(60):
        frm, to = random_lists(seed, 40 + seed, 3 + seed % 9)
        wrong += rounds_then_settle(frm, to, 0, reverse_settle=True) != greedy(frm, to)
    assert wrong > 0
