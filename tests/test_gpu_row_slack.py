"""Rows with slack: a row that fits its capacity is merged in place (its old
entries staged aside first), one that outgrows it moves to new pool space with
an eighth more room, and the pool is compacted (capacities kept) when it runs
out.  Many batches that keep growing a few hub rows and deleting entries force
all three paths, and the state must stay bit-exact with the oracle after every
batch."""
import numpy as np
import pytest

import hit_leiden_amd as hl
from parity_util import assert_same_state, pair

pytestmark = pytest.mark.gpu

P = 3


def growing_batches(n, hubs, batches, per_hub, rng, weighted):
    """Each batch: per_hub new edges at every hub (new neighbours), a few
    random edges, and deletions of edges inserted two batches before."""
    out, hist = [], []
    seen = set()
    for b in range(batches):
        u, v = [], []
        for h in hubs:
            cand = rng.choice(n, size=per_hub * 2, replace=False)
            took = 0
            for x in cand:
                key = (min(h, int(x)), max(h, int(x)))
                if int(x) != h and key not in seen:
                    seen.add(key)
                    u.append(h)
                    v.append(int(x))
                    took += 1
                    if took == per_hub:
                        break
        for _ in range(40):
            a, c = rng.integers(0, n, size=2)
            key = (min(a, c), max(a, c))
            if a != c and key not in seen:
                seen.add(key)
                u.append(int(a))
                v.append(int(c))
        w = (rng.uniform(0.5, 1.5, len(u)) if weighted else np.ones(len(u))).tolist()
        hist.append((list(u), list(v), list(w)))
        du, dv, dw = [], [], []
        if b >= 2:  # delete a quarter of the edges inserted two batches before
            ou, ov, ow = hist[b - 2]
            for i in range(0, len(ou), 4):
                du.append(ou[i])
                dv.append(ov[i])
                dw.append(-ow[i])
                seen.discard((min(ou[i], ov[i]), max(ou[i], ov[i])))
        out.append((np.array(u + du, np.uint32), np.array(v + dv, np.uint32), np.array(w + dw, np.float64)))
    return out


@pytest.mark.parametrize("mode", [hl.RunMode.DETERMINISTIC, hl.RunMode.THROUGHPUT])
def test_growing_rows_stay_exact(mode):
    rng = np.random.default_rng(21)
    n = 3000
    base = set()
    while len(base) < 6000:
        a, c = rng.integers(0, n, size=2)
        if a != c:
            base.add((min(a, c), max(a, c)))
    bu = np.array([e[0] for e in base], np.uint32)
    bv = np.array([e[1] for e in base], np.uint32)
    g, o = pair(mode, P)
    g.load(hl.GraphInput(n, bu, bv))
    o.load(n, bu, bv)
    for i, (u, v, a) in enumerate(growing_batches(n, [0, 1, 2, 7, 50], 24, 40, rng, weighted=False)):
        g.hit_leiden(hl.GraphInput(n, u, v, a))
        o.apply(u, v, a)
        assert_same_state(g, o, P, tag=f"batch {i}")
    g.close()


def test_growing_rows_real_weights_det():
    """Non-integer weights: deterministic mode keeps every canonical order,
    so the state stays bit-exact through the moves and compactions too."""
    rng = np.random.default_rng(5)
    n = 2000
    base = set()
    while len(base) < 4000:
        a, c = rng.integers(0, n, size=2)
        if a != c:
            base.add((min(a, c), max(a, c)))
    bu = np.array([e[0] for e in base], np.uint32)
    bv = np.array([e[1] for e in base], np.uint32)
    bw = rng.uniform(0.5, 1.5, len(bu))
    g, o = pair(hl.RunMode.DETERMINISTIC, P)
    g.load(hl.GraphInput(n, bu, bv, bw))
    o.load(n, bu, bv, bw)
    for i, (u, v, a) in enumerate(growing_batches(n, [3, 4, 900], 20, 30, rng, weighted=True)):
        g.hit_leiden(hl.GraphInput(n, u, v, a))
        o.apply(u, v, a)
        assert_same_state(g, o, P, tag=f"batch {i}")
    g.close()
