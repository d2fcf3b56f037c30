"""hit_leiden_amd -- MI355X (gfx950) engine for HIT-Leiden's incremental
hierarchical community update.

Python mirror of the reference crate's API for this path (nameearly/hit-leiden,
Rust): `RunMode`/`RunConfig` (src/core/config.rs:19-51), `GraphInput`
(src/core/types.rs:133-148), the `HitLeidenError` variants (src/core/error.rs:19-27),
`PartitionState` + `hit_leiden()` (src/core/partition/state.rs:19-64,
src/core/algorithm/hit_leiden.rs:936-1018) and `compute_modularity`
(hit_leiden.rs:121-161).  Every call goes through the C ABI in
include/hitleiden.h (libhitleiden.so); there is no CPU fallback -- importing
this package on a machine without the built library or a gfx950 device
raises AccelerationError.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhitleiden.so")

HL_OK, HL_INVALID_INPUT, HL_BACKEND, HL_ACCELERATION, HL_OOM = range(5)
MAP_F, MAP_G, MAP_S_CUR, MAP_S_PRE = range(4)
SET_ACTIVE, SET_CHANGED, SET_AFFECTED, SET_REFINED = range(4)
PART_COMMUNITY, PART_REFINED, PART_F0 = range(3)
MAX_LEVELS = 8


class RunMode(enum.IntEnum):
    """RunMode (src/core/config.rs:19-23)."""
    DETERMINISTIC = 0
    THROUGHPUT = 1


class HitLeidenError(Exception):
    """HitLeidenError (src/core/error.rs:19-27)."""


class InvalidInput(HitLeidenError):
    pass


class BackendError(HitLeidenError):
    pass


class AccelerationError(HitLeidenError):
    pass


class OutOfMemory(BackendError):
    pass


_ERRORS = {HL_INVALID_INPUT: InvalidInput, HL_BACKEND: BackendError, HL_ACCELERATION: AccelerationError,
           HL_OOM: OutOfMemory}


class _Config(C.Structure):
    _fields_ = [("gamma", C.c_double), ("mode", C.c_int32), ("levels", C.c_int32), ("max_rounds", C.c_int32),
                ("device", C.c_int32), ("small_degree", C.c_uint32), ("wave_degree", C.c_uint32),
                ("load_pass_records", C.c_uint32), ("key_cache", C.c_uint32), ("hub_tables", C.c_uint32),
                ("decide_rounds", C.c_uint32), ("refine_reference", C.c_uint32), ("marks_block", C.c_uint32),
                ("marks_grid", C.c_uint32), ("persist_rounds", C.c_uint32)]


class BatchStats(C.Structure):
    """hl_batch_stats: per-batch counters of one load/apply call."""
    _fields_ = [("iterations", C.c_uint64), ("levels", C.c_uint32), ("level_n", C.c_uint32 * MAX_LEVELS),
                ("delta_len", C.c_uint64 * MAX_LEVELS), ("active", C.c_uint64 * MAX_LEVELS),
                ("evaluated", C.c_uint64 * MAX_LEVELS), ("moved", C.c_uint64 * MAX_LEVELS),
                ("affected", C.c_uint64 * MAX_LEVELS), ("refined", C.c_uint64 * MAX_LEVELS),
                ("rounds", C.c_uint64 * MAX_LEVELS), ("bytes_alg", C.c_uint64), ("seconds", C.c_double),
                ("decide_rounds", C.c_uint64 * MAX_LEVELS), ("settled", C.c_uint64),
                ("alloc_count", C.c_uint64), ("alloc_bytes", C.c_uint64), ("alloc_seconds", C.c_double),
                ("host_syncs", C.c_uint64), ("persist_rounds", C.c_uint64)]

    def as_dict(self):
        P = self.levels
        return {
            "iterations": self.iterations, "levels": P, "level_n": list(self.level_n[:P]),
            "delta_len": list(self.delta_len[:P]), "active": list(self.active[:P]),
            "evaluated": list(self.evaluated[:P]), "moved": list(self.moved[:P]),
            "affected": list(self.affected[:P]), "refined": list(self.refined[:P]),
            "rounds": list(self.rounds[:P]), "bytes_alg": self.bytes_alg, "seconds": self.seconds,
            "decide_rounds": list(self.decide_rounds[:P]), "settled": self.settled,
            "alloc_count": self.alloc_count, "alloc_bytes": self.alloc_bytes, "alloc_seconds": self.alloc_seconds,
            "host_syncs": self.host_syncs, "persist_rounds": self.persist_rounds,
        }


_lib = None


def library_path():
    return _LIB_PATH


class _Request(C.Structure):
    _fields_ = [("graph", C.c_uint32), ("k", C.c_uint64), ("src", C.c_void_p), ("dst", C.c_void_p),
                ("alpha", C.c_void_p), ("arrival", C.c_double)]


class ServeStats(C.Structure):
    """hl_serve_stats: passes, records, largest pass (requests), wall and busy seconds,
    the modelled device footprint of the largest pass and the passes free memory cut short."""
    _fields_ = [("passes", C.c_uint64), ("records", C.c_uint64), ("max_pass_requests", C.c_uint64),
                ("seconds", C.c_double), ("busy_seconds", C.c_double), ("max_pass_bytes", C.c_uint64),
                ("footprint_cut_passes", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_EMIT = C.CFUNCTYPE(None, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint32), C.c_uint32)


def lib():
    """Load libhitleiden.so; fails loudly when it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise AccelerationError(f"{_LIB_PATH} is not built (run hit_leiden_amd.build.build_all())")
        L = C.CDLL(_LIB_PATH)
        vp, u32p, u64p, f64p, u8p = C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), \
            C.POINTER(C.c_double), C.POINTER(C.c_uint8)
        sig = {
            "hl_create": ([C.POINTER(_Config), C.POINTER(vp)], C.c_int),
            "hl_destroy": ([vp], None),
            "hl_load": ([vp, C.c_uint32, C.c_uint64, vp, vp, vp, C.POINTER(BatchStats)], C.c_int),
            "hl_load_graph": ([vp, C.c_uint32, C.c_uint64, vp, vp, vp], C.c_int),
            "hl_apply_batch": ([vp, C.c_uint64, vp, vp, vp, C.POINTER(BatchStats)], C.c_int),
            "hl_apply_batch_device": ([vp, C.c_uint64, vp, vp, vp, C.POINTER(BatchStats)], C.c_int),
            "hl_level_count": ([vp, C.POINTER(C.c_int32)], C.c_int),
            "hl_level_size": ([vp, C.c_int32, u32p], C.c_int),
            "hl_read_partition": ([vp, C.c_int32, C.c_int32, vp, C.c_uint64], C.c_int),
            "hl_write_partition": ([vp, C.c_int32, C.c_int32, vp, C.c_uint64], C.c_int),
            "hl_read_frontier": ([vp, C.c_int32, C.c_int32, vp, C.c_uint64], C.c_int),
            "hl_read_level_graph": ([vp, C.c_int32, vp, vp, vp, C.c_uint64, u64p], C.c_int),
            "hl_read_level_delta": ([vp, C.c_int32, vp, vp, vp, C.c_uint64, u64p], C.c_int),
            "hl_twice_m": ([vp, C.c_int32, f64p], C.c_int),
            "hl_read_hierarchy": ([vp, vp, C.c_uint64, C.c_int32], C.c_int),
            "hl_modularity": ([vp, C.c_int32, f64p], C.c_int),
            "hl_modularity_labels": ([vp, vp, C.c_uint64, f64p], C.c_int),
            "hl_modularity_labels_noalloc": ([vp, vp, C.c_uint64, f64p], C.c_int),
            "hl_run_static": ([vp, C.c_uint32, C.c_uint64, vp, vp, vp, C.c_uint32, C.c_uint64, vp, vp, C.c_uint32,
                               u32p, C.POINTER(C.c_int64)], C.c_int),
            "hl_last_error": ([vp], C.c_char_p),
            "hl_create_error": ([], C.c_char_p),
            "hl_profile_slots": ([], C.c_int32),
            "hl_profile_enable": ([vp, C.c_int32], C.c_int),
            "hl_profile_read": ([vp, C.c_int32, C.c_char_p, C.c_uint64, u64p, f64p, u64p], C.c_int),
            "hl_load_graphs": ([vp, C.c_uint32, vp, C.c_uint64, vp, vp, vp, C.POINTER(BatchStats)], C.c_int),
            "hl_serve": ([vp, C.POINTER(_Request), C.c_uint64, C.c_uint64, C.c_int32, vp, vp, vp, _EMIT, vp,
                          C.POINTER(ServeStats)], C.c_int),
            "hl_graph_count": ([vp, u32p], C.c_int),
            "hl_compare_partitions": ([vp, vp, vp, C.c_uint64, C.c_uint32, C.c_uint32, vp], C.c_int),
            "hl_restore_level": ([vp, C.c_int32, C.c_uint32, C.c_uint64, vp, vp, vp, vp, vp, vp, vp], C.c_int),
            "hl_modularity_graphs": ([vp, vp, C.c_uint64], C.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _first_seen(labels):
    """canonicalize_community_ids_in_place (hit_leiden.rs:358-370) on a numpy row."""
    _, first, inv = np.unique(labels, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    return rank[inv].astype(np.uint32)


@dataclass
class RunConfig:
    """RunConfig (src/core/config.rs:25-51) restricted to this path."""
    resolution: float = 1.0
    mode: RunMode = RunMode.THROUGHPUT
    max_iterations: int = 10  # config.rs:31,46: the static path's level cap (run(), multilevel_leiden)
    levels: int = 1
    max_rounds: int = 20
    device: int = 0
    small_degree: int = 0
    wave_degree: int = 0
    load_pass_records: int = 0
    key_cache: int = 0  # 1: cache each row entry's neighbour community in the throughput movement
    hub_tables: bool = True  # evaluate hub rows from community tables kept current by the moves
    decide_rounds: int = 0  # priority rounds of the multi-kernel decoupling before the in-order settle (0: 48)
    # throughput mode: False keeps the hierarchy consistent (refined communities inside their
    # communities, f^p = f^{p+1} o s^p); True applies the reference's own refinement and
    # def_update rules (throughput.rs:278-287, hit_leiden.rs:1233-1240,1733-1793)
    refine_reference: bool = False
    marks_block: int = 0  # workgroup size of the grid-wide neighbour marks (0: 256; 128, 256, 512)
    marks_grid: int = 0  # their workgroups (0: 1,024)
    # a level's throughput rounds in one cooperative kernel while they fit it (same results)
    persist_rounds: int = True  # True / False, or the kernel's workgroups (2..1024)


class GraphInput:
    """GraphInput (src/core/types.rs:133-148): node_count + (u, v, w) edges;
    w None means unit weights (Option<f64>::unwrap_or(1.0))."""

    def __init__(self, node_count, src, dst, weights=None):
        self.node_count = int(node_count)
        self.src = np.ascontiguousarray(src, dtype=np.uint32)
        self.dst = np.ascontiguousarray(dst, dtype=np.uint32)
        self.weights = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        if self.src.shape != self.dst.shape or (self.weights is not None and self.weights.shape != self.src.shape):
            raise InvalidInput("src, dst and weights must have the same length")

    @classmethod
    def from_edges(cls, node_count, edges):
        """edges: iterable of (u, v) or (u, v, w)."""
        e = list(edges)
        if not e:
            return cls(node_count, [], [], [])
        u = [x[0] for x in e]
        v = [x[1] for x in e]
        w = [x[2] if len(x) > 2 and x[2] is not None else 1.0 for x in e]
        return cls(node_count, u, v, w)

    def __len__(self):
        return len(self.src)


class HitLeiden:
    """Device-resident PartitionState + graph hierarchy on one gfx950 GPU.

    load() runs the initial pass (Algorithm 6 with dG = all edges on the
    identity state); hit_leiden(delta) applies one edge-delta batch
    (hit_leiden.rs:936-1018) and returns the iteration count.
    """

    def __init__(self, config: RunConfig | None = None):
        self.config = config or RunConfig()
        c = _Config()
        c.gamma = float(self.config.resolution)
        c.mode = int(self.config.mode)
        c.levels = int(self.config.levels)
        c.max_rounds = int(self.config.max_rounds)
        c.device = int(self.config.device)
        c.small_degree = int(self.config.small_degree)
        c.wave_degree = int(self.config.wave_degree)
        c.load_pass_records = int(self.config.load_pass_records)
        c.key_cache = int(self.config.key_cache)
        c.hub_tables = 0 if self.config.hub_tables else 1
        c.decide_rounds = int(self.config.decide_rounds)
        c.refine_reference = 1 if self.config.refine_reference else 0
        c.marks_block = int(self.config.marks_block)
        c.marks_grid = int(self.config.marks_grid)
        pr = self.config.persist_rounds
        c.persist_rounds = (0 if pr else 1) if isinstance(pr, bool) else int(pr)
        h = C.c_void_p()
        st = lib().hl_create(C.byref(c), C.byref(h))
        if st != HL_OK:
            raise _ERRORS.get(st, HitLeidenError)(lib().hl_create_error().decode())
        self._h = h
        self.last_stats = BatchStats()

    def close(self):
        if getattr(self, "_h", None):
            lib().hl_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st):
        if st != HL_OK:
            msg = lib().hl_last_error(self._h)
            raise _ERRORS.get(st, HitLeidenError)(msg.decode() if msg else "error")

    # ---- load / update ----
    def load(self, graph: GraphInput) -> BatchStats:
        s = BatchStats()
        self._check(lib().hl_load(self._h, graph.node_count, len(graph), _ptr(graph.src), _ptr(graph.dst),
                                  _ptr(graph.weights), C.byref(s)))
        self.last_stats = s
        return s

    def run_static(self, graph: GraphInput, max_passes=10, tol=1e-6):
        """Static clustering with the incremental kernels: load, then empty-delta
        passes (each re-evaluates every vertex, hit_leiden.rs:1054-1057) until the
        hierarchy stops changing, the community modularity gains less than tol, or
        max_passes.  Synchronous throughput rounds can swap vertices back and
        forth: a pass that lowers the modularity is undone (the engine returns to
        the state before it, kept as an in-memory checkpoint) and ends the run.
        The reference's own static order, multilevel_leiden
        (hit_leiden.rs:167-323), is `static_leiden`; this returns (passes,
        modularity of the communities)."""
        import io
        from . import checkpoint
        self.load(graph)
        empty = GraphInput(graph.node_count, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
        prev = self.hierarchy()
        q = self.modularity(PART_COMMUNITY)
        passes = 0
        for passes in range(1, max_passes + 1):
            saved = io.BytesIO()
            checkpoint.save(self, saved)
            self.hit_leiden(empty)
            cur = self.hierarchy()
            q_new = self.modularity(PART_COMMUNITY)
            if q_new < q:
                saved.seek(0)
                back = checkpoint.restore(saved, device=self.config.device)
                self._h, back._h = back._h, self._h
                back.close()
                break
            if all(np.array_equal(x, y) for x, y in zip(prev, cur)) or q_new - q < tol:
                q = q_new
                break
            prev, q = cur, q_new
        return passes, q

    def static_leiden(self, graph: GraphInput, max_levels=10, seed=0):
        """run() / multilevel_leiden (hit_leiden.rs:30-118, 167-323) on the
        device: movement with the shuffled FIFO queue, Leiden refinement,
        aggregation of the input edges, up to max_levels levels, then a final
        polish on the input graph.  The queue shuffle draws from one
        StdRng::seed_from_u64(seed) for the whole run (the reference draws an
        unseeded thread_rng).  The handle's loaded state is untouched.  Returns
        (labels numbered first-seen, hierarchy rows, movement calls)."""
        n = graph.node_count
        rows_cap = int(max_levels) + 2
        labels = np.zeros(max(n, 1), dtype=np.uint32)
        hier = np.zeros(max(n, 1) * rows_cap, dtype=np.uint32)
        rows = C.c_uint32()
        calls = C.c_int64()
        self._check(lib().hl_run_static(self._h, n, len(graph), _ptr(graph.src), _ptr(graph.dst),
                                        _ptr(graph.weights), int(max_levels), int(seed), _ptr(labels), _ptr(hier),
                                        rows_cap, C.byref(rows), C.byref(calls)))
        r = min(rows.value, rows_cap)
        h = hier[:r * n].reshape(r, n) if n else np.zeros((r, 0), np.uint32)
        return labels[:n], h, int(calls.value)

    def load_graph(self, graph: GraphInput):
        self._check(lib().hl_load_graph(self._h, graph.node_count, len(graph), _ptr(graph.src), _ptr(graph.dst),
                                        _ptr(graph.weights)))

    def hit_leiden(self, delta: GraphInput) -> int:
        """Apply one edge-delta batch; alpha > 0 inserts/strengthens, alpha < 0 deletes/weakens."""
        s = BatchStats()
        self._check(lib().hl_apply_batch(self._h, len(delta), _ptr(delta.src), _ptr(delta.dst),
                                         _ptr(delta.weights), C.byref(s)))
        self.last_stats = s
        return int(s.iterations)

    apply_batch = hit_leiden

    def apply_batch_device(self, k, d_src, d_dst, d_alpha=None) -> BatchStats:
        """Batch already in device memory: raw device pointers (ints)."""
        s = BatchStats()
        self._check(lib().hl_apply_batch_device(self._h, int(k), C.c_void_p(d_src), C.c_void_p(d_dst),
                                                None if d_alpha is None else C.c_void_p(d_alpha), C.byref(s)))
        self.last_stats = s
        return s

    # ---- several independent graphs in one engine (batched form) ----
    def load_graphs(self, graphs) -> BatchStats:
        """Load independent graphs (a list of GraphInput) into one engine; graph i
        owns the union vertex ids [graph_offsets[i], graph_offsets[i+1])."""
        sizes = np.array([g.node_count for g in graphs], dtype=np.uint64)
        off = np.zeros(len(graphs) + 1, dtype=np.uint64)
        np.cumsum(sizes, out=off[1:])
        if off[-1] >= 2**32:
            raise InvalidInput("too many vertices for u32 ids")
        self.graph_offsets = off.astype(np.uint32)
        src = np.concatenate([g.src.astype(np.uint64) + off[i] for i, g in enumerate(graphs)] or [np.zeros(0)])
        dst = np.concatenate([g.dst.astype(np.uint64) + off[i] for i, g in enumerate(graphs)] or [np.zeros(0)])
        weighted = any(g.weights is not None for g in graphs)
        w = np.concatenate([g.weights if g.weights is not None else np.ones(len(g), np.float64)
                            for g in graphs]) if weighted else None
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        s = BatchStats()
        self._check(lib().hl_load_graphs(self._h, len(graphs), _ptr(self.graph_offsets), len(src), _ptr(src),
                                         _ptr(dst), _ptr(w), C.byref(s)))
        self.last_stats = s
        return s

    def graphs_delta(self, deltas):
        """Union-id arrays of {graph index: GraphInput batch}; every id is graph-local
        (below that graph's node_count), or InvalidInput."""
        off = self.graph_offsets.astype(np.uint64)
        items = sorted(deltas.items())
        for i, d in items:
            if not 0 <= i < len(off) - 1:
                raise InvalidInput(f"graph index {i} out of range")
            size = int(off[i + 1] - off[i])
            if len(d) and (int(d.src.max()) >= size or int(d.dst.max()) >= size):
                raise InvalidInput(f"a record of graph {i} names a vertex outside its {size} vertices")
        src = np.concatenate([d.src.astype(np.uint64) + off[i] for i, d in items] or [np.zeros(0)])
        dst = np.concatenate([d.dst.astype(np.uint64) + off[i] for i, d in items] or [np.zeros(0)])
        a = np.concatenate([d.weights if d.weights is not None else np.ones(len(d), np.float64)
                            for _, d in items] or [np.zeros(0)])
        return (np.ascontiguousarray(src, dtype=np.uint32), np.ascontiguousarray(dst, dtype=np.uint32),
                np.ascontiguousarray(a, dtype=np.float64))

    def apply_graphs(self, deltas) -> BatchStats:
        """One pass over the batches of several graphs ({graph index: GraphInput})."""
        src, dst, a = self.graphs_delta(deltas)
        s = BatchStats()
        self._check(lib().hl_apply_batch(self._h, len(src), _ptr(src), _ptr(dst), _ptr(a), C.byref(s)))
        self.last_stats = s
        return s

    def serve(self, requests, max_pass_records=1 << 20, emit=None, coalesce=False):
        """Open-loop serving of [(graph, GraphInput batch with graph-local ids, arrival seconds)]
        (hl_serve).  emit(request, labels) receives each request's community labels
        as soon as they are final: at the round of its pass where its graph's
        frontier at the top level empties, or at the pass's end.  Returns (done
        seconds, pass index, ServeStats); the seconds at which each request's
        communities were final are left in self.last_emitted."""
        keep, reqs = [], (_Request * max(len(requests), 1))()
        for i, (gi, b, t) in enumerate(requests):
            src = np.ascontiguousarray(b.src, dtype=np.uint32)
            dst = np.ascontiguousarray(b.dst, dtype=np.uint32)
            a = None if b.weights is None else np.ascontiguousarray(b.weights, dtype=np.float64)
            keep.append((src, dst, a))
            reqs[i] = _Request(int(gi), len(src), src.ctypes.data, dst.ctypes.data,
                               None if a is None else a.ctypes.data, float(t))
        done = np.zeros(max(len(requests), 1), dtype=np.float64)
        pass_of = np.zeros(max(len(requests), 1), dtype=np.uint64)
        emitted = np.zeros(max(len(requests), 1), dtype=np.float64)
        st = ServeStats()
        cb = _EMIT(0)
        if emit is not None:
            def _cb(_ctx, r, labels, n):
                emit(int(r), np.ctypeslib.as_array(labels, shape=(n,)).copy() if n else np.zeros(0, np.uint32))
            cb = _EMIT(_cb)
        self._check(lib().hl_serve(self._h, reqs, len(requests), int(max_pass_records), int(bool(coalesce)),
                                   _ptr(done), _ptr(pass_of), _ptr(emitted), cb, None, C.byref(st)))
        self.last_emitted = emitted[:len(requests)]
        return done[:len(requests)], pass_of[:len(requests)], st

    def compare_partitions(self, a, b, a_largest=None, b_largest=None):
        """contingency_metrics (hit_leiden_incremental.rs:577-666) on the device:
        (nmi, purity a->b, purity b->a, Jaccard of the largest communities; the
        largest are found here unless given)."""
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        if len(a) != len(b):
            raise InvalidInput("labelings differ in length")
        if a_largest is None:
            a_largest = int(np.bincount(a).argmax()) if len(a) else 0
        if b_largest is None:
            b_largest = int(np.bincount(b).argmax()) if len(b) else 0
        out = np.zeros(4, dtype=np.float64)
        self._check(lib().hl_compare_partitions(self._h, _ptr(a), _ptr(b), len(a), a_largest, b_largest, _ptr(out)))
        return tuple(float(x) for x in out)

    @property
    def graph_count(self):
        x = C.c_uint32()
        self._check(lib().hl_graph_count(self._h, C.byref(x)))
        return x.value

    def modularity_graphs(self):
        q = np.zeros(max(self.graph_count, 1), dtype=np.float64)
        self._check(lib().hl_modularity_graphs(self._h, _ptr(q), len(q)))
        return q[:self.graph_count]

    def graph_hierarchy(self, i, canonical=True):
        """Hierarchy rows restricted to graph i (first-seen relabelled per row when canonical)."""
        h = self.hierarchy(canonical=False)
        lo, hi = int(self.graph_offsets[i]), int(self.graph_offsets[i + 1])
        rows = [r[lo:hi].copy() for r in h]
        if canonical:
            rows = [_first_seen(r) for r in rows]
        return rows

    # ---- state readout (PartitionState fields) ----
    @property
    def levels(self):
        x = C.c_int32()
        self._check(lib().hl_level_count(self._h, C.byref(x)))
        return x.value

    def level_size(self, level):
        x = C.c_uint32()
        self._check(lib().hl_level_size(self._h, level, C.byref(x)))
        return x.value

    def partition(self, level, which):
        n = self.level_size(level)
        out = np.zeros(max(n, 1), dtype=np.uint32)
        self._check(lib().hl_read_partition(self._h, level, which, _ptr(out), len(out)))
        return out[:n]

    def write_partition(self, level, which, values):
        a = np.ascontiguousarray(values, dtype=np.uint32)
        self._check(lib().hl_write_partition(self._h, level, which, _ptr(a), len(a)))

    @property
    def community_mapping_per_level(self):
        return [self.partition(p, MAP_F) for p in range(self.levels)]

    @property
    def refined_community_mapping_per_level(self):
        return [self.partition(p, MAP_G) for p in range(self.levels)]

    @property
    def current_subcommunity_mapping_per_level(self):
        return [self.partition(p, MAP_S_CUR) for p in range(self.levels)]

    @property
    def previous_subcommunity_mapping_per_level(self):
        return [self.partition(p, MAP_S_PRE) for p in range(self.levels)]

    @property
    def node_to_comm(self):
        """g^1, the reference's output (hit_leiden.rs:1016)."""
        return self.partition(0, MAP_G)

    def frontier(self, level, which):
        n = self.level_size(level)
        out = np.zeros(max(n, 1), dtype=np.uint8)
        self._check(lib().hl_read_frontier(self._h, level, which, _ptr(out), len(out)))
        return out[:n].astype(bool)

    def supergraph(self, level):
        """G^p as (offsets, cols, weights), rows sorted, self-loop = one entry."""
        nnz = C.c_uint64()
        self._check(lib().hl_read_level_graph(self._h, level, None, None, None, 0, C.byref(nnz)))
        n = self.level_size(level)
        off = np.zeros(n + 1, dtype=np.uint64)
        col = np.zeros(max(nnz.value, 1), dtype=np.uint32)
        w = np.zeros(max(nnz.value, 1), dtype=np.float64)
        self._check(lib().hl_read_level_graph(self._h, level, _ptr(off), _ptr(col), _ptr(w), len(col),
                                              C.byref(nnz)))
        return off, col[:nnz.value], w[:nnz.value]

    def level_delta(self, level):
        k = C.c_uint64()
        self._check(lib().hl_read_level_delta(self._h, level, None, None, None, 0, C.byref(k)))
        u = np.zeros(max(k.value, 1), dtype=np.uint32)
        v = np.zeros(max(k.value, 1), dtype=np.uint32)
        w = np.zeros(max(k.value, 1), dtype=np.float64)
        self._check(lib().hl_read_level_delta(self._h, level, _ptr(u), _ptr(v), _ptr(w), len(u), C.byref(k)))
        return u[:k.value], v[:k.value], w[:k.value]

    def twice_m(self, level=0):
        x = C.c_double()
        self._check(lib().hl_twice_m(self._h, level, C.byref(x)))
        return x.value

    def hierarchy(self, canonical=False):
        """PartitionResult.hierarchy_levels (types.rs:107-110): P rows of n labels."""
        n = self.level_size(0)
        P = self.levels
        out = np.zeros(max(n * P, 1), dtype=np.uint32)
        self._check(lib().hl_read_hierarchy(self._h, _ptr(out), len(out), 1 if canonical else 0))
        return out[:n * P].reshape(P, n)

    def modularity(self, which=PART_COMMUNITY):
        q = C.c_double()
        self._check(lib().hl_modularity(self._h, which, C.byref(q)))
        return q.value

    # ---- device timing ----
    def profile(self, on=True, only=None):
        """Start (reset) or stop per-kernel HIP-event timing on the engine stream;
        only=<kernel name> records that kernel family alone."""
        code = 1 if on else 0
        if on and only is not None:
            code = 2 + self.profile_slot(only)
        self._check(lib().hl_profile_enable(self._h, code))

    def profile_slot(self, name):
        buf = C.create_string_buffer(64)
        for s in range(lib().hl_profile_slots()):
            self._check(lib().hl_profile_read(self._h, s, buf, 64, None, None, None))
            if buf.value.decode() == name:
                return s
        raise KeyError(name)

    def profile_read(self):
        """{kernel: (launches, device_ms, algorithmic_bytes)} since profile(True)."""
        out = {}
        buf = C.create_string_buffer(64)
        for s in range(lib().hl_profile_slots()):
            n, ms, b = C.c_uint64(), C.c_double(), C.c_uint64()
            self._check(lib().hl_profile_read(self._h, s, buf, 64, C.byref(n), C.byref(ms), C.byref(b)))
            if n.value:
                out[buf.value.decode()] = (n.value, ms.value, b.value)
        return out

    def modularity_labels(self, labels):
        a = np.ascontiguousarray(labels, dtype=np.uint32)
        q = C.c_double()
        self._check(lib().hl_modularity_labels(self._h, _ptr(a), len(a), C.byref(q)))
        return q.value


def compute_modularity(graph: GraphInput, labels, device=0) -> float:
    """compute_modularity (hit_leiden.rs:121-161) on the GPU."""
    h = HitLeiden(RunConfig(device=device))
    try:
        h.load_graph(graph)
        return h.modularity_labels(labels)
    finally:
        h.close()


from .validation import PartitionResult, RunOutcome, ValidationOutcome, check, run, validate  # noqa: E402

__all__ = ["ServeStats", "RunMode", "RunConfig", "GraphInput", "HitLeiden", "BatchStats", "compute_modularity",
           "PartitionResult", "RunOutcome", "ValidationOutcome", "check", "run", "validate",
           "HitLeidenError", "InvalidInput", "BackendError", "AccelerationError", "OutOfMemory", "lib",
           "library_path", "MAP_F", "MAP_G", "MAP_S_CUR", "MAP_S_PRE", "SET_ACTIVE", "SET_CHANGED",
           "SET_AFFECTED", "SET_REFINED", "PART_COMMUNITY", "PART_REFINED", "PART_F0"]
