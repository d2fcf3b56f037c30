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

from . import MAP_F, MAP_G, MAP_S_CUR, MAP_S_PRE, HitLeiden, RunConfig, RunMode, lib, _ptr

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


def restore(path, device=0) -> HitLeiden:
    with np.load(path, allow_pickle=False) as z:
        if int(z["format"][0]) != FORMAT:
            raise ValueError(f"{path}: checkpoint format {int(z['format'][0])}, expected {FORMAT}")
        P = int(z["levels"][0])
        gamma, mode, max_rounds, sd, wd = z["config"].tolist()
        e = HitLeiden(RunConfig(resolution=gamma, mode=RunMode(int(mode)), levels=P, max_rounds=int(max_rounds),
                                device=device, small_degree=int(sd), wave_degree=int(wd)))
        for p in range(P):
            n = int(z[f"n{p}"][0])
            u = np.ascontiguousarray(z[f"u{p}"], dtype=np.uint32)
            v = np.ascontiguousarray(z[f"v{p}"], dtype=np.uint32)
            w = np.ascontiguousarray(z[f"w{p}"], dtype=np.float64)
            maps = [np.ascontiguousarray(z[f"{k}{p}"], dtype=np.uint32) for k in ("f", "g", "spre", "scur")]
            e._check(lib().hl_restore_level(e._h, p, n, len(u), _ptr(u), _ptr(v), _ptr(w), *[_ptr(a) for a in maps]))
    return e
