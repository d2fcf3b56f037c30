"""Autotuner for the launch shapes of the throughput path.

Two stages, each timing a few candidate configurations on a sample of the
caller's workload (a graph and some of its batches, device-resident; or, for
the batched form, a set of graphs and one set of their deltas) and keeping the
fastest:

1. the degree-bucket thresholds: the local-moving proposals give one lane per
   vertex up to `small_degree` (at most 16, the private LDS slice), one
   wavefront per vertex up to `wave_degree` (at most 512, the wavefront's LDS
   table), then one workgroup per vertex;
2. the workgroup shape of the grid-wide neighbour marks (the rows of the moved
   vertices, the largest kernel on power-law graphs): `marks_block` threads
   per workgroup and `marks_grid` workgroups.

Results are cached per workload signature in a JSON file when `cache` is given.
"""
from __future__ import annotations

import json
import os
import time
from typing import NamedTuple

import numpy as np

from . import GraphInput, HitLeiden, RunConfig, RunMode, hip

SMALL = (4, 8, 16)
WAVE = (64, 128, 256, 512)
MARKS_BLOCK = (128, 256, 512)
MARKS_GRID = (512, 1024, 2048)


class Tuned(NamedTuple):
    small_degree: int
    wave_degree: int
    marks_block: int
    marks_grid: int

    def config(self, **kw):
        """A RunConfig with these launch shapes (kw: the other fields)."""
        return RunConfig(small_degree=self.small_degree, wave_degree=self.wave_degree,
                         marks_block=self.marks_block, marks_grid=self.marks_grid, **kw)


def signature(n, u, v):
    """Workload key: size and a coarse degree histogram."""
    deg = np.bincount(np.concatenate([u, v]), minlength=n) if len(u) else np.zeros(n, np.int64)
    edges = [0, 4, 8, 16, 32, 64, 128, 256, 512, 4096, 1 << 30]
    hist = np.histogram(deg, bins=edges)[0]
    return f"n{n}_m{len(u)}_" + "_".join(str(int(x)) for x in (hist * 1000 // max(n, 1)))


def _time(config, graph, batches, repeat):
    e = HitLeiden(config)
    try:
        e.load(graph)
        dev = [(len(b[0]), hip.DeviceArray(b[0]), hip.DeviceArray(b[1]), hip.DeviceArray(b[2])) for b in batches]
        hip.synchronize()
        best = float("inf")
        for _ in range(repeat):
            t = time.perf_counter()
            for k, du, dv, da in dev:
                e.apply_batch_device(k, du.address, dv.address, da.address)
            best = min(best, time.perf_counter() - t)
        for _, a, b, c in dev:
            a.free(), b.free(), c.free()
        return best
    finally:
        e.close()


# a candidate replaces the incumbent only when it is this much faster: timings of
# near-equal shapes differ by run-to-run noise, not by their shape
MARGIN = 0.03


def _search(time_of, base, log):
    """The two stages over time_of(Tuned) -> seconds, starting from base."""
    results = {}

    def timed(t):
        if t not in results:
            results[t] = time_of(t)
            if log:
                log(f"[autotune] {t._asdict()}: {results[t] * 1e3:.2f} ms")
        return results[t]

    def pick(incumbent, candidates):
        best = incumbent
        for t in candidates:
            if timed(t) < timed(best) * (1.0 - MARGIN):
                best = t
        return best

    best = pick(base, [base._replace(small_degree=sd, wave_degree=wd) for sd in SMALL for wd in WAVE])
    best = pick(best, [best._replace(marks_block=mb, marks_grid=mg) for mb in MARKS_BLOCK for mg in MARKS_GRID])
    return best


def _cached(cache, key, compute):
    table = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
        if key in table and len(table[key]) == 4:
            return Tuned(*table[key])
    best = compute()
    if cache:
        table[key] = list(best)
        with open(cache, "w") as f:
            json.dump(table, f, indent=1)
    return best


def tune(graph: GraphInput, batches, levels=4, gamma=1.0, repeat=1, cache=None, log=None) -> Tuned:
    """The launch shapes with the lowest time for the batches [(src, dst, alpha)]
    applied after loading graph, in throughput mode."""
    key = signature(graph.node_count, graph.src, graph.dst)

    def time_of(t):
        cfg = t.config(mode=RunMode.THROUGHPUT, levels=levels, resolution=gamma)
        return _time(cfg, graph, batches, repeat)

    return _cached(cache, key, lambda: _search(time_of, Tuned(16, 512, 256, 1024), log))


def tune_graphs(graphs, deltas, levels=4, gamma=1.0, cache=None, log=None) -> Tuned:
    """The same for the batched form: graphs loaded into one engine
    (hl_load_graphs), then one pass over deltas {graph index: GraphInput}."""
    sig = signature(sum(g.node_count for g in graphs[:1]), graphs[0].src, graphs[0].dst) if graphs else "empty"
    key = f"graphs{len(graphs)}_{sum(len(d) for d in deltas.values())}_{sig}"

    def time_of(t, repeat=2):
        best = float("inf")
        for _ in range(repeat):  # fresh engines: the pass starts from the same state each time
            e = HitLeiden(t.config(mode=RunMode.THROUGHPUT, levels=levels, resolution=gamma))
            try:
                e.load_graphs(graphs)
                hip.synchronize()
                t0 = time.perf_counter()
                e.apply_graphs(deltas)
                best = min(best, time.perf_counter() - t0)
            finally:
                e.close()
        return best

    return _cached(cache, key, lambda: _search(time_of, Tuned(16, 512, 256, 1024), log))
