"""Checkpoint and restore of a one-graph engine's hierarchy.

`save` writes, per level, the level graph (each undirected edge once, a
self-loop once with its record weight) and the maps f, g, s_pre, s_cur, plus
the run configuration, to one `.npz` file (numpy arrays only; `restore` reads
it with allow_pickle=False).  `restore` rebuilds every level on the device
through hl_restore_level; the restored engine continues exactly as the saved
one would (community totals are recomputed from the maps).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import MAP_F, MAP_G, MAP_S_CUR, MAP_S_PRE, HitLeiden, InvalidInput, RunConfig, RunMode, lib, _ptr

FORMAT = 1


def _level_records(engine, p):
    off, col, w = engine.supergraph(p)
    n = len(off) - 1
    rows = np.repeat(np.arange(n, dtype=np.uint32), np.diff(off).astype(np.int64))
    keep = rows <= col  # each undirected edge once; a self-loop is stored once
    return rows[keep], col[keep], w[keep]


def save(engine: HitLeiden, path):
    if getattr(engine, "graph_offsets", None) is not None and engine.graph_count:
        raise ValueError("checkpoints cover one-graph engines")
    c = engine.config
    arrays = {"format": np.array([FORMAT]), "levels": np.array([engine.levels]),
              "config": np.array([c.resolution, int(c.mode), c.max_rounds, c.small_degree, c.wave_degree],
                                 dtype=np.float64)}
    for p in range(engine.levels):
        u, v, w = _level_records(engine, p)
        arrays[f"n{p}"] = np.array([engine.level_size(p)], dtype=np.int64)
        arrays[f"u{p}"], arrays[f"v{p}"], arrays[f"w{p}"] = u, v, w
        for name, which in (("f", MAP_F), ("g", MAP_G), ("spre", MAP_S_PRE), ("scur", MAP_S_CUR)):
            arrays[f"{name}{p}"] = engine.partition(p, which)
    np.savez(path, **arrays)


MAX_LABEL = 2**31 - 1  # labels the engine accepts (hl_write_partition, hl_restore_level)


def _read_levels(z, path):
    """The levels of a checkpoint, checked before any device work: ids below
    n, finite weights, maps of n labels below 2^31 (InvalidInput otherwise)."""
    P = int(z["levels"][0])
    if not 1 <= P <= 8:
        raise InvalidInput(f"{path}: {P} levels")
    out = []
    for p in range(P):
        n = int(z[f"n{p}"][0])
        if not 0 <= n <= MAX_LABEL or (p == 0 and n == 0):
            raise InvalidInput(f"{path}: level {p} has {n} vertices")
        u, v, w = (z[f"u{p}"], z[f"v{p}"], z[f"w{p}"])
        if not (len(u) == len(v) == len(w)):
            raise InvalidInput(f"{path}: level {p} edge arrays differ in length")
        if len(u) and (int(u.max()) >= n or int(v.max()) >= n or int(min(u.min(), v.min())) < 0):
            raise InvalidInput(f"{path}: level {p} edge endpoint out of range")
        if not np.all(np.isfinite(w)):
            raise InvalidInput(f"{path}: level {p} non-finite weight")
        maps = []
        for k in ("f", "g", "spre", "scur"):
            a = z[f"{k}{p}"]
            if len(a) != n:
                raise InvalidInput(f"{path}: level {p} map {k} has {len(a)} labels for {n} vertices")
            if n and (int(a.max()) > MAX_LABEL or int(a.min()) < 0):
                raise InvalidInput(f"{path}: level {p} map {k} holds a label outside [0, 2^31)")
            maps.append(np.ascontiguousarray(a, dtype=np.uint32))
        out.append((n, np.ascontiguousarray(u, dtype=np.uint32), np.ascontiguousarray(v, dtype=np.uint32),
                    np.ascontiguousarray(w, dtype=np.float64), maps))
    return out


def restore(path, device=0) -> HitLeiden:
    with np.load(path, allow_pickle=False) as z:
        if int(z["format"][0]) != FORMAT:
            raise ValueError(f"{path}: checkpoint format {int(z['format'][0])}, expected {FORMAT}")
        levels = _read_levels(z, path)
        gamma, mode, max_rounds, sd, wd = z["config"].tolist()
    e = HitLeiden(RunConfig(resolution=gamma, mode=RunMode(int(mode)), levels=len(levels),
                            max_rounds=int(max_rounds), device=device, small_degree=int(sd), wave_degree=int(wd)))
    for p, (n, u, v, w, maps) in enumerate(levels):
        e._check(lib().hl_restore_level(e._h, p, n, len(u), _ptr(u), _ptr(v), _ptr(w), *[_ptr(a) for a in maps]))
    return e
