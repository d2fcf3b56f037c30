"""Helpers of the full-size configuration tests (level graphs read back from the
device, checked with numpy and scipy)."""
import os

import numpy as np

import hit_leiden_amd as hl
from parity_util import assert_classes_connected

P = 4
THREADS = min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "16")))


def say(*a):
    print(*a, flush=True)


def level_edges(g, p):
    """G^p as (rows, cols, weights) with each undirected edge once (self-loops once)."""
    off, col, w = g.supergraph(p)
    rows = np.repeat(np.arange(len(off) - 1, dtype=np.int64), np.diff(off.astype(np.int64)))
    keep = rows <= col
    return rows[keep], col[keep].astype(np.int64), w[keep], off, col, w


def check_degrees_add_up(g, levels, rtol):
    """k^{p+1}(a) = sum of k^p(v) over the children v of a (self-loops count twice)."""
    prev = None
    for p in range(levels):
        off, col, w = g.supergraph(p)
        rows = np.repeat(np.arange(len(off) - 1, dtype=np.int64), np.diff(off.astype(np.int64)))
        k = np.bincount(rows, weights=w, minlength=len(off) - 1)
        self_ = rows == col
        k += np.bincount(rows[self_], weights=w[self_], minlength=len(off) - 1)
        if prev is not None:
            kp, s = prev
            want = np.bincount(s, weights=kp, minlength=len(k))
            np.testing.assert_allclose(k, want[:len(k)], rtol=rtol, atol=rtol, err_msg=f"level {p} degrees")
            assert len(want) <= len(k) or not np.any(want[len(k):]), f"level {p}: children of missing vertices"
        prev = (k, g.partition(p, hl.MAP_S_CUR).astype(np.int64))


def check_contraction(g, p, tol):
    """G^{p+1} equals the contraction of G^p under s_cur^p, every super-edge."""
    r, c, x, *_ = level_edges(g, p)
    s = g.partition(p, hl.MAP_S_CUR).astype(np.int64)
    n1 = g.level_size(p + 1)
    a, b = s[r], s[c]
    key = np.minimum(a, b) * n1 + np.maximum(a, b)
    uk, inv = np.unique(key, return_inverse=True)
    sums = np.bincount(inv, weights=x)
    live = np.abs(sums) > 1e-9
    r1, c1, x1, *_ = level_edges(g, p + 1)
    k1 = r1 * n1 + c1
    o = np.argsort(k1)
    assert np.array_equal(uk[live], k1[o]), f"level {p + 1}: super-edges differ"
    np.testing.assert_allclose(x1[o], sums[live], rtol=tol, atol=tol, err_msg=f"level {p + 1} weights")


def refined_connected(g, n, levels):
    off, col, _ = g.supergraph(0)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(off.astype(np.int64)))
    keep = rows < col
    H = g.hierarchy()
    assert_classes_connected(n, rows[keep], col[keep].astype(np.int64), H[levels - 2])


