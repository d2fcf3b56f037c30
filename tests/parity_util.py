"""Helpers comparing the GPU engine's state with the CPU oracle's."""
import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

import hit_leiden_amd as hl
from oracle import hl_oracle as O

MAPS = (hl.MAP_F, hl.MAP_G, hl.MAP_S_CUR, hl.MAP_S_PRE)
SETS = (hl.SET_ACTIVE, hl.SET_CHANGED, hl.SET_AFFECTED, hl.SET_REFINED)
MAP_NAMES = {hl.MAP_F: "f", hl.MAP_G: "g", hl.MAP_S_CUR: "s_cur", hl.MAP_S_PRE: "s_pre"}
SET_NAMES = {hl.SET_ACTIVE: "A", hl.SET_CHANGED: "B", hl.SET_AFFECTED: "K", hl.SET_REFINED: "R"}


def oracle_mode(mode, refine_reference=False):
    """The oracle mode a GPU configuration is compared with: DET, PAR (the
    parallel mode with a consistent hierarchy) or THR (the reference's own
    throughput rules, RunConfig.refine_reference)."""
    if mode == hl.RunMode.DETERMINISTIC:
        return O.DET
    return O.THR if refine_reference else O.PAR


def pair(mode, levels, gamma=1.0, **cfg):
    g = hl.HitLeiden(hl.RunConfig(resolution=gamma, mode=mode, levels=levels, **cfg))
    o = O.OracleState(levels=levels, gamma=gamma, mode=oracle_mode(mode, cfg.get("refine_reference", False)))
    return g, o


def assert_same_state(g, o, levels, tag="", weights_exact=True):
    """Bit-exact comparison of every per-level map, set, CSR and delta (and, in
    the parallel mode, the hierarchy invariants)."""
    if g.config.mode == hl.RunMode.THROUGHPUT and not g.config.refine_reference:
        assert_hierarchy_consistent(g, levels, tag)
    for p in range(levels):
        n = g.level_size(p)
        assert n == o.level_n(p), f"{tag} level {p} size {n} vs {o.level_n(p)}"
        for which in MAPS:
            a, b = g.partition(p, which), o.read_map(p, which)
            if not np.array_equal(a, b):
                bad = np.flatnonzero(a != b)
                raise AssertionError(f"{tag} level {p} map {MAP_NAMES[which]}: {len(bad)} differ, first {bad[:5]} "
                                     f"gpu {a[bad[:5]]} oracle {b[bad[:5]]}")
        for which in SETS:
            a, b = g.frontier(p, which), o.read_set(p, which).astype(bool)
            if not np.array_equal(a, b):
                bad = np.flatnonzero(a != b)
                raise AssertionError(f"{tag} level {p} set {SET_NAMES[which]}: {len(bad)} differ, first {bad[:5]}")
        off, col, w = g.supergraph(p)
        ooff, ocol, ow = o.read_graph(p)
        assert np.array_equal(off, ooff), f"{tag} level {p} CSR offsets differ"
        assert np.array_equal(col, ocol), f"{tag} level {p} CSR columns differ"
        if weights_exact:
            assert np.array_equal(w, ow), f"{tag} level {p} CSR weights differ"
        else:
            np.testing.assert_allclose(w, ow, rtol=1e-12, atol=1e-12)
        du, dv, dw = g.level_delta(p)
        ou, ov, ow2 = o.read_delta(p)
        assert np.array_equal(du, ou) and np.array_equal(dv, ov), f"{tag} level {p} delta differs"
        if weights_exact:
            assert np.array_equal(dw, ow2), f"{tag} level {p} delta weights differ"
        if weights_exact:
            assert g.twice_m(p) == o.twice_m(p), f"{tag} level {p} 2m"


def csr_to_edges(off, col, w):
    n = len(off) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(off.astype(np.int64)))
    return rows, col.astype(np.int64), w


def assert_supergraph_consistent(g, levels, tol=0.0):
    """G^{p+1} equals the supergraph of G^p under s_cur^p (aggregation exactness)."""
    for p in range(levels - 1):
        off, col, w = g.supergraph(p)
        r, c, x = csr_to_edges(off, col, w)
        s = g.partition(p, hl.MAP_S_CUR).astype(np.int64)
        n1 = g.level_size(p + 1)
        # each undirected edge once: r < c, plus self-loops (stored once)
        keep = r <= c
        a, b, ww = s[r[keep]], s[c[keep]], x[keep]
        lo, hi = np.minimum(a, b), np.maximum(a, b)
        want = {}
        for i, j, y in zip(lo.tolist(), hi.tolist(), ww.tolist()):
            want[(i, j)] = want.get((i, j), 0.0) + y
        off1, col1, w1 = g.supergraph(p + 1)
        r1, c1, x1 = csr_to_edges(off1, col1, w1)
        k1 = r1 <= c1
        got = {(int(i), int(j)): float(y) for i, j, y in zip(r1[k1], c1[k1], x1[k1])}
        want = {k: v for k, v in want.items() if v > 1e-12}
        assert set(got) == set(want), f"level {p + 1}: {len(set(got) ^ set(want))} super-edges differ"
        for k, v in want.items():
            assert abs(got[k] - v) <= tol * max(1.0, abs(v)), (p, k, got[k], v)
        assert n1 >= (s.max() + 1 if len(s) else 0)


def assert_classes_connected(n, u, v, labels):
    """Every label class induces a connected subgraph (isolated vertices alone)."""
    labels = np.asarray(labels)
    same = labels[u] == labels[v]
    A = sp.coo_matrix((np.ones(same.sum()), (u[same], v[same])), shape=(n, n)).tocsr()
    ncomp, comp = connected_components(A, directed=False)
    nlab = len(np.unique(labels))
    assert ncomp == nlab, f"{ncomp} components for {nlab} classes"


def hierarchy_violations(maps_f, maps_s):
    """Per level p: (s^p-classes spanning several f^p-classes, vertices with
    f^p(v) != f^{p+1}(s^p(v))); the second is None at the top level."""
    out = []
    P = len(maps_f)
    for p in range(P):
        f, s = np.asarray(maps_f[p], np.int64), np.asarray(maps_s[p], np.int64)
        if len(s):
            pairs = np.unique(np.stack([s, f], 1), axis=0)
            straddle = len(pairs) - len(np.unique(s))
        else:
            straddle = 0
        cons = None
        if p + 1 < P:
            fu = np.asarray(maps_f[p + 1], np.int64)
            cons = int(np.sum(f != fu[s])) if len(s) else 0
        out.append((straddle, cons))
    return out


def assert_hierarchy_consistent(g, levels, tag=""):
    """Parallel mode: every s^p-class lies inside one f^p-class, and
    f^p = f^{p+1} o s^p below the top level (hierarchy consistency)."""
    fs = [g.partition(p, hl.MAP_F) for p in range(levels)]
    ss = [g.partition(p, hl.MAP_S_CUR) for p in range(levels)]
    for p, (straddle, cons) in enumerate(hierarchy_violations(fs, ss)):
        assert straddle == 0, f"{tag} level {p}: {straddle} subcommunities span two communities"
        assert not cons, f"{tag} level {p}: {cons} vertices with f^p != f^(p+1) o s^p"
