"""hl_modularity_labels_noalloc (the allocation-free variant of
hl_modularity_labels, compute_modularity hit_leiden.rs:121-161) against the
allocating entry point, through the C ABI.

Two calls in a row on one GPU share the fixed workspace: the second, on a
smaller graph with fewer labels, must not see the first call's sums.  The
modularity is a sum of f64 atomics, so the two paths agree to rounding
(bar 1e-9, the deterministic-mode modularity bar); with unit weights every
partial sum is an exact integer and they agree exactly.
"""
import ctypes as C

import numpy as np
import pytest

import hit_leiden_amd as hl
from hit_leiden_amd import gen

pytestmark = pytest.mark.gpu


def noalloc(g, labels):
    a = np.ascontiguousarray(labels, dtype=np.uint32)
    q = C.c_double()
    st = hl.lib().hl_modularity_labels_noalloc(g._h, hl._ptr(a), len(a), C.byref(q))
    return st, q.value


def engine(n, mean_deg, seed, weighted):
    w = gen.planted(n, max(n // 100, 2), mean_deg, 0.3, seed=seed) if weighted else gen.lfr(n, mean_deg, 0.2,
                                                                                            seed=seed)
    g = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=3))
    g.load(hl.GraphInput(w.n, w.u, w.v, w.w))
    return g


@pytest.mark.parametrize("weighted", [False, True])
def test_noalloc_matches_allocating_entry_point(weighted):
    big = engine(20000, 10.0, 3, weighted)
    small = engine(3000, 6.0, 4, weighted)
    tol = 1e-12
    for g, tag in ((big, "first call"), (small, "second call, smaller input")):
        for labels in (g.hierarchy(canonical=True)[-1], g.hierarchy(canonical=True)[0],
                       np.arange(g.level_size(0), dtype=np.uint32)):
            st, q = noalloc(g, labels)
            assert st == hl.HL_OK, hl.lib().hl_last_error(g._h)
            want = g.modularity_labels(labels)
            assert abs(q - want) <= tol, (tag, q, want)
    # the big graph again after the small one: the workspace holds the small call's sums
    labels = big.hierarchy(canonical=True)[-1]
    st, q = noalloc(big, labels)
    assert st == hl.HL_OK and abs(q - big.modularity_labels(labels)) <= tol
    big.close()
    small.close()


def test_noalloc_rejects_what_does_not_fit():
    g = engine(3000, 6.0, 5, False)
    labels = g.hierarchy(canonical=True)[-1].copy()
    labels[7] = 1 << 21
    st, _ = noalloc(g, labels)
    assert st == hl.HL_INVALID_INPUT and "2^21" in hl.lib().hl_last_error(g._h).decode()
    st, _ = noalloc(g, labels[:100])
    assert st == hl.HL_INVALID_INPUT
    g.close()
    h = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT, levels=1))
    n = (1 << 21) + 1
    h.load_graph(hl.GraphInput(n, np.array([0], np.uint32), np.array([1], np.uint32)))
    st, _ = noalloc(h, np.zeros(n, np.uint32))
    assert st == hl.HL_INVALID_INPUT and "2^21" in hl.lib().hl_last_error(h._h).decode()
    h.close()
