"""Autotuner for the degree-bucket thresholds of the throughput path.

The local-moving proposals split the active vertices by degree: one lane per
vertex up to `small_degree` (at most 16, the private LDS slice), one wavefront
per vertex up to `wave_degree` (at most 512, the wavefront's LDS table), then
one workgroup per vertex.  The best split depends on the degree distribution,
so `tune` times a few candidate pairs on a sample of the caller's workload
(a graph and some of its batches, device-resident) and returns the fastest.
Results are cached per workload signature in a JSON file when `cache` is given.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import GraphInput, HitLeiden, RunConfig, RunMode, hip

SMALL = (4, 8, 16)
WAVE = (64, 128, 256, 512)


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


def tune(graph: GraphInput, batches, levels=4, gamma=1.0, repeat=1, cache=None, log=None):
    """Return (small_degree, wave_degree) with the lowest time for the batches
    [(src, dst, alpha)] applied after loading graph, in throughput mode."""
    key = signature(graph.node_count, graph.src, graph.dst)
    table = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
        if key in table:
            return tuple(table[key])
    results = {}
    for sd in SMALL:
        for wd in WAVE:
            cfg = RunConfig(mode=RunMode.THROUGHPUT, levels=levels, resolution=gamma, small_degree=sd,
                            wave_degree=wd)
            results[(sd, wd)] = _time(cfg, graph, batches, repeat)
            if log:
                log(f"[autotune] small={sd} wave={wd}: {results[(sd, wd)] * 1e3:.2f} ms")
    best = min(results, key=results.get)
    if cache:
        table[key] = list(best)
        with open(cache, "w") as f:
            json.dump(table, f, indent=1)
    return best
