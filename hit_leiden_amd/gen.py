"""Seeded synthetic workloads (libhl_gen.so): the graphs and edge-delta
streams of the five configurations, plus small variants for tests.

    small    LFR-style, n=9,877, tau1=2.5, tau2=1.5 in [5,200], mu=0.2, m=21,778,
             5 batches of 1,000 degree-proportional new edges
    mid      planted partition, n=1e6, 1,000 log-normal blocks (sigma=1), mean degree 32,
             mu=0.3, w ~ U[0.5,1.5], 20 batches of 9,000 inserts + 1,000 deletes
    large    R-MAT scale 24, edge factor 16, (0.57,0.19,0.19,0.05), 10 batches of
             95,000 R-MAT inserts + 5,000 deletes
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_GEN_SO = os.path.join(_HERE, "libhl_gen.so")
_g = None


def glib():
    global _g
    if _g is None:
        if not os.path.exists(_GEN_SO):
            from . import build
            build.build_gen()
        L = C.CDLL(_GEN_SO)
        vp = C.c_void_p
        L.hg_lfr.restype = vp
        L.hg_lfr.argtypes = [C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint32, C.c_uint32, C.c_double,
                             C.c_uint64, C.c_uint64]
        L.hg_planted.restype = vp
        L.hg_planted.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_uint64]
        L.hg_rmat.restype = vp
        L.hg_rmat.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int]
        L.hg_sbm.restype = vp
        L.hg_sbm.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_uint64]
        L.hg_n.restype = C.c_uint32
        L.hg_n.argtypes = [vp]
        L.hg_m.restype = C.c_uint64
        L.hg_m.argtypes = [vp]
        L.hg_edges.argtypes = [vp, vp, vp, vp]
        L.hg_labels.argtypes = [vp, vp]
        L.hg_make_batches.argtypes = [vp, C.c_uint32, C.c_uint64, C.c_uint64, C.c_int, C.c_double, C.c_double,
                                      C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64]
        L.hg_num_batches.restype = C.c_uint32
        L.hg_num_batches.argtypes = [vp]
        L.hg_batch_len.restype = C.c_uint64
        L.hg_batch_len.argtypes = [vp, C.c_uint32]
        L.hg_batch.argtypes = [vp, C.c_uint32, vp, vp, vp]
        L.hg_free.argtypes = [vp]
        L.hg_shuffle_order.argtypes = [C.c_uint64, C.c_uint64, vp]
        L.hg_stdrng_u32.argtypes = [C.c_uint64, C.c_uint64, vp]
        L.hg_chacha_block.argtypes = [vp, C.c_int, C.c_uint64, C.c_uint64, vp]
        _g = L
    return _g


class Workload:
    """A generated graph (n, u, v, w[, labels]) and its batches [(u, v, alpha)]."""

    def __init__(self, h, weighted):
        L = glib()
        self.n = int(L.hg_n(h))
        m = int(L.hg_m(h))
        self.u = np.zeros(m, dtype=np.uint32)
        self.v = np.zeros(m, dtype=np.uint32)
        self.w = np.zeros(m, dtype=np.float64)
        L.hg_edges(h, self.u.ctypes.data, self.v.ctypes.data, self.w.ctypes.data)
        if not weighted:
            self.w = None
        lab = np.zeros(max(self.n, 1), dtype=np.uint32)
        self.labels = lab[:self.n] if L.hg_labels(h, lab.ctypes.data) else None
        self.batches = []
        self._h = h

    @property
    def m(self):
        return len(self.u)

    def make_batches(self, nb, ins, dels, mode, wlo=1.0, whi=1.0, scale=0, abc=(0.57, 0.19, 0.19), seed=1):
        L = glib()
        L.hg_make_batches(self._h, nb, ins, dels, mode, wlo, whi, scale, abc[0], abc[1], abc[2], seed)
        self.batches = []
        for i in range(L.hg_num_batches(self._h)):
            k = int(L.hg_batch_len(self._h, i))
            u = np.zeros(k, dtype=np.uint32)
            v = np.zeros(k, dtype=np.uint32)
            a = np.zeros(k, dtype=np.float64)
            L.hg_batch(self._h, i, u.ctypes.data, v.ctypes.data, a.ctypes.data)
            self.batches.append((u, v, a))
        return self.batches

    def close(self):
        if self._h:
            glib().hg_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


def lfr(n, mean_deg, mu, target_m=0, tau1=2.5, tau2=1.5, cmin=5, cmax=200, seed=1):
    return Workload(glib().hg_lfr(n, tau1, mean_deg, tau2, cmin, cmax, mu, target_m, seed), weighted=False)


def planted(n, blocks, mean_deg, mu, sigma=1.0, wlo=0.5, whi=1.5, seed=1):
    return Workload(glib().hg_planted(n, blocks, sigma, mean_deg, mu, wlo, whi, seed), weighted=True)


def rmat(scale, ef=16, abc=(0.57, 0.19, 0.19), seed=1, threads=0):
    return Workload(glib().hg_rmat(scale, ef, abc[0], abc[1], abc[2], seed, threads), weighted=False)


def sbm(n, blocks, mean_deg, mu, seed=1):
    return Workload(glib().hg_sbm(n, blocks, mean_deg, mu, seed), weighted=False)


# ---- the reference loaders' seeded shuffle (rand 0.8.5 StdRng) ----

def shuffle_order(seed, n):
    """`StdRng::seed_from_u64(seed)` + `SliceRandom::shuffle` of n elements
    (dynamic_graph.rs:85-87): out[i] = the original index of position i."""
    out = np.zeros(max(int(n), 1), dtype=np.uint64)
    glib().hg_shuffle_order(int(seed), int(n), out.ctypes.data)
    return out[:int(n)]


def stdrng_u32(seed, count):
    out = np.zeros(max(int(count), 1), dtype=np.uint32)
    glib().hg_stdrng_u32(int(seed), int(count), out.ctypes.data)
    return out[:int(count)]


def chacha_block(key_words, rounds, counter=0, stream=0):
    k = np.ascontiguousarray(key_words, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint32)
    glib().hg_chacha_block(k.ctypes.data, int(rounds), int(counter), int(stream), out.ctypes.data)
    return out


# ---- the named configurations ----

def small_config(seed=7):
    """LFR-style n=9,877, m=21,778, 5 x 1,000 degree-proportional new edges."""
    w = lfr(9877, 4.6, 0.2, target_m=21778, seed=seed)
    w.make_batches(5, 1000, 0, mode=0, seed=seed + 1)
    return w


def mid_config(seed=11, n=1_000_000, blocks=1000, batches=20):
    """planted partition n=1e6, mean degree 32, mu=0.3, w~U[0.5,1.5];
    20 x (9,000 new inserts + 1,000 deletes)."""
    w = planted(n, blocks, 32.0, 0.3, seed=seed)
    w.make_batches(batches, 9000, 1000, mode=1, wlo=0.5, whi=1.5, seed=seed + 1)
    return w


def large_config(seed=13, scale=24, batches=10, threads=0):
    """R-MAT scale 24, ef 16; 10 x (95,000 R-MAT inserts + 5,000 deletes)."""
    w = rmat(scale, 16, seed=seed, threads=threads)
    w.make_batches(batches, 95000, 5000, mode=2, scale=scale, seed=seed + 1)
    return w


def batched_config(count=4096, steps=1, seed=19, threads=16):
    """SBM graphs with n log-uniform in [200, 50,000], 4-64 blocks, mean degree
    8, mu = 0.25, unit weights; each gets `steps` successive batches of k new-edge
    insertions with uniform endpoints, k log-uniform in [16, 2,000] per graph."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(200), np.log(50000), count)).astype(np.int64)
    blocks = rng.integers(4, 65, count)
    ks = np.exp(rng.uniform(np.log(16), np.log(2000), count)).astype(np.int64)

    def one(i):
        w = sbm(int(ns[i]), int(blocks[i]), 8.0, 0.25, seed=seed * 100003 + i)
        w.make_batches(steps, int(ks[i]), 0, mode=1, seed=seed * 7919 + i)
        return w

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(one, range(count)))


def served_config(count=256, seed=23, threads=16):
    """LFR-style graphs with n log-uniform in [10,000, 2,000,000], degree
    exponent 2.5, mean degree 10, mu = 0.2, unit weights."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(10000), np.log(2000000), count)).astype(np.int64)
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(lambda i: lfr(int(ns[i]), 10.0, 0.2, seed=seed * 100003 + i), range(count)))


def served_requests(sizes, rate, count, seed=29, kmin=10, kmax=5000, zipf_s=1.1):
    """Poisson arrivals at `rate` per second; each request inserts k uniform
    random vertex pairs (k log-uniform in [kmin, kmax]) into a graph drawn
    Zipf(zipf_s) by index.  Returns [(graph, src, dst, arrival)] with
    graph-local ids."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    z = 1.0 / np.arange(1, len(sizes) + 1) ** zipf_s
    pick = rng.choice(len(sizes), size=count, p=z / z.sum())
    ks = np.exp(rng.uniform(np.log(kmin), np.log(kmax), count)).astype(np.int64)
    out = []
    for i in range(count):
        n = int(sizes[pick[i]])
        a = rng.integers(0, n, ks[i])
        b = rng.integers(0, n - 1, ks[i])
        b = b + (b >= a)  # no self-loops
        out.append((int(pick[i]), a.astype(np.uint32), b.astype(np.uint32), float(t[i])))
    return out

