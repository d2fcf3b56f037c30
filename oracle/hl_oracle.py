"""ctypes binding of the CPU oracle (libhl_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker or the reported CPU
baseline.  The product path (hit_leiden_amd) never imports it.

The oracle restates the reference's incremental path
(src/core/algorithm/hit_leiden.rs:120-161,934-1793; throughput.rs:79-359;
parallel_frontier.rs:133-239; in_memory.rs:30-138) with canonical orders.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhl_oracle.so")
_lib = None

DET, THR, CONNECTED, REF_IDS, AS_WRITTEN = 0, 1, 2, 4, 8
PAR = THR | CONNECTED
MAP_F, MAP_G, MAP_S_CUR, MAP_S_PRE = 0, 1, 2, 3
SET_A, SET_B, SET_K, SET_R = 0, 1, 2, 3

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int]
        L.or_destroy.argtypes = [C.c_void_p]
        for name in ("or_load", "or_load_graph"):
            fn = getattr(L, name)
            fn.restype = C.c_int
            fn.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64, _u32p, _u32p, C.c_void_p]
        L.or_apply.restype = C.c_int64
        L.or_apply.argtypes = [C.c_void_p, C.c_uint64, _u32p, _u32p, C.c_void_p]
        L.or_level_n.restype = C.c_uint32
        L.or_level_n.argtypes = [C.c_void_p, C.c_int]
        L.or_levels.argtypes = [C.c_void_p]
        L.or_read_map.argtypes = [C.c_void_p, C.c_int, C.c_int, _u32p]
        L.or_write_map.argtypes = [C.c_void_p, C.c_int, C.c_int, _u32p]
        L.or_read_set.argtypes = [C.c_void_p, C.c_int, C.c_int, _u8p]
        L.or_level_iters.restype = C.c_uint32
        L.or_level_iters.argtypes = [C.c_void_p, C.c_int]
        L.or_read_graph.restype = C.c_uint64
        L.or_read_graph.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_read_delta.restype = C.c_uint64
        L.or_read_delta.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_twice_m.restype = C.c_double
        L.or_twice_m.argtypes = [C.c_void_p, C.c_int]
        L.or_hierarchy.argtypes = [C.c_void_p, _u32p]
        L.or_modularity_labels.restype = C.c_double
        L.or_modularity_labels.argtypes = [C.c_void_p, _u32p]
        L.or_apply_delta_fn.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, _f64p, C.c_uint64, _u32p, _u32p, _f64p,
                                        C.c_uint32, C.POINTER(C.c_uint64), C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_inc_movement_fn.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, _f64p, _u32p, _u32p, C.c_uint64,
                                         _u32p, _u32p, _f64p, C.c_double, C.c_int, _u8p, _u8p]
        L.or_inc_refinement_fn.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, _f64p, _u32p, _u32p, _u8p,
                                           C.c_double, C.c_int, _u8p]
        L.or_inc_aggregation_fn.restype = C.c_int64
        L.or_inc_aggregation_fn.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, _f64p, C.c_uint64, _u32p, _u32p,
                                            _f64p, _u32p, _u32p, _u8p, _u32p, _u32p, _f64p, C.c_uint64,
                                            C.POINTER(C.c_uint32)]
        L.or_def_update_fn.argtypes = [C.c_int, _u32p, _u32p, _u32p, _u8p]
        L.or_modularity_fn.restype = C.c_double
        L.or_modularity_fn.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, _f64p, _u32p]
        L.or_select_decoupled_fn.restype = C.c_uint64
        L.or_select_decoupled_fn.argtypes = [C.c_uint64, _u32p, _u32p, _u32p, _f64p, _u32p]
        L.or_canonicalize_fn.argtypes = [C.c_uint32, _u32p]
        L.or_should_skip_aggregation_fn.argtypes = [C.c_uint64, C.c_uint32, _u8p]
        L.or_empty_collisions.argtypes = [C.c_int]
        L.or_empty_collisions.restype = C.c_uint64
        L.or_stdrng_shuffle.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p]
        L.or_stdrng_shuffle.restype = None
        L.or_chacha_block.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64, C.c_void_p]
        L.or_chacha_block.restype = None
        L.or_static_run.restype = C.c_int64
        L.or_static_run.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, C.c_void_p, C.c_double, C.c_uint32,
                                    C.c_uint64, _u32p, _u32p, C.c_uint32, C.POINTER(C.c_uint32)]
        L.or_static_movement_fn.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, C.c_void_p, _u32p, C.c_double,
                                            C.c_uint64]
        L.or_static_refine_fn.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, C.c_void_p, _u32p, C.c_double,
                                          _u32p]
        _lib = L
    return _lib


def _edges(edges, weights=None):
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    u = np.ascontiguousarray(e[:, 0].astype(np.uint32))
    v = np.ascontiguousarray(e[:, 1].astype(np.uint32))
    if weights is None:
        w = np.ones(len(u), dtype=np.float64)
    else:
        w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
    return u, v, w


def _wptr(w):
    return None if w is None else np.ascontiguousarray(w, dtype=np.float64).ctypes.data_as(C.c_void_p)


class OracleState:
    """Mirror of PartitionState + hit_leiden() with P levels (state.rs:19-64)."""

    def __init__(self, levels=1, gamma=1.0, mode=DET, threads=0):
        self.h = lib().or_create(levels, gamma, mode, threads)
        if not self.h:
            raise ValueError("invalid oracle config")
        self.levels = levels

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def load(self, n, u, v, w=None):
        u = np.ascontiguousarray(u, dtype=np.uint32)
        v = np.ascontiguousarray(v, dtype=np.uint32)
        wa = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
        r = lib().or_load(self.h, n, len(u), u, v, _wptr(wa))
        if r < 0:
            raise ValueError("invalid edges")
        return r

    def load_graph(self, n, u, v, w=None):
        u = np.ascontiguousarray(u, dtype=np.uint32)
        v = np.ascontiguousarray(v, dtype=np.uint32)
        wa = None if w is None else np.ascontiguousarray(w, dtype=np.float64)
        if lib().or_load_graph(self.h, n, len(u), u, v, _wptr(wa)) < 0:
            raise ValueError("invalid edges")

    def apply(self, u, v, alpha=None):
        u = np.ascontiguousarray(u, dtype=np.uint32)
        v = np.ascontiguousarray(v, dtype=np.uint32)
        a = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
        r = lib().or_apply(self.h, len(u), u, v, _wptr(a))
        if r < 0:
            raise ValueError("invalid batch")
        return r

    def level_n(self, p):
        return lib().or_level_n(self.h, p)

    def read_map(self, p, which):
        out = np.zeros(self.level_n(p), dtype=np.uint32)
        lib().or_read_map(self.h, p, which, out)
        return out

    def write_map(self, p, which, arr):
        lib().or_write_map(self.h, p, which, np.ascontiguousarray(arr, dtype=np.uint32))

    def read_set(self, p, which):
        out = np.zeros(self.level_n(p), dtype=np.uint8)
        lib().or_read_set(self.h, p, which, out)
        return out

    def level_iters(self, p):
        return lib().or_level_iters(self.h, p)

    def read_graph(self, p):
        n = self.level_n(p)
        nnz = lib().or_read_graph(self.h, p, None, None, None)
        off = np.zeros(n + 1, dtype=np.uint64)
        col = np.zeros(max(nnz, 1), dtype=np.uint32)
        w = np.zeros(max(nnz, 1), dtype=np.float64)
        lib().or_read_graph(self.h, p, off.ctypes.data_as(C.c_void_p), col.ctypes.data_as(C.c_void_p),
                            w.ctypes.data_as(C.c_void_p))
        return off, col[:nnz], w[:nnz]

    def read_delta(self, p):
        k = lib().or_read_delta(self.h, p, None, None, None)
        u = np.zeros(max(k, 1), dtype=np.uint32)
        v = np.zeros(max(k, 1), dtype=np.uint32)
        w = np.zeros(max(k, 1), dtype=np.float64)
        lib().or_read_delta(self.h, p, u.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p),
                            w.ctypes.data_as(C.c_void_p))
        return u[:k], v[:k], w[:k]

    def twice_m(self, p=0):
        return lib().or_twice_m(self.h, p)

    def hierarchy(self):
        n0 = self.level_n(0)
        out = np.zeros(self.levels * n0, dtype=np.uint32)
        lib().or_hierarchy(self.h, out)
        return out.reshape(self.levels, n0)

    def modularity(self, labels):
        return lib().or_modularity_labels(self.h, np.ascontiguousarray(labels, dtype=np.uint32))


# ---- function-level wrappers (reference unit-test entry points) ----

def apply_delta(n, edges, weights, delta, delta_w, n_new=None):
    u, v, w = _edges(edges, weights)
    du, dv, dw = _edges(delta, delta_w)
    n_new = n if n_new is None else n_new
    m = C.c_uint64(0)
    lib().or_apply_delta_fn(n, len(u), u, v, w, len(du), du, dv, dw, n_new, C.byref(m), None, None, None)
    ou = np.zeros(max(m.value, 1), dtype=np.uint32)
    ov = np.zeros(max(m.value, 1), dtype=np.uint32)
    ow = np.zeros(max(m.value, 1), dtype=np.float64)
    lib().or_apply_delta_fn(n, len(u), u, v, w, len(du), du, dv, dw, n_new, C.byref(m),
                            ou.ctypes.data_as(C.c_void_p), ov.ctypes.data_as(C.c_void_p),
                            ow.ctypes.data_as(C.c_void_p))
    k = m.value
    return list(zip(ou[:k].tolist(), ov[:k].tolist(), ow[:k].tolist()))


def inc_movement(n, edges, weights, f, s, delta, delta_w, gamma, mode):
    u, v, w = _edges(edges, weights)
    du, dv, dw = _edges(delta, delta_w)
    fa = np.ascontiguousarray(f, dtype=np.uint32).copy()
    sa = np.ascontiguousarray(s, dtype=np.uint32)
    B = np.zeros(n, dtype=np.uint8)
    K = np.zeros(n, dtype=np.uint8)
    it = lib().or_inc_movement_fn(n, len(u), u, v, w, fa, sa, len(du), du, dv, dw, gamma, mode, B, K)
    return fa, B.astype(bool), K.astype(bool), it


def inc_refinement(n, edges, weights, f, s, K, gamma, mode):
    u, v, w = _edges(edges, weights)
    fa = np.ascontiguousarray(f, dtype=np.uint32)
    sa = np.ascontiguousarray(s, dtype=np.uint32).copy()
    Ka = np.ascontiguousarray(K, dtype=np.uint8)
    R = np.zeros(n, dtype=np.uint8)
    lib().or_inc_refinement_fn(n, len(u), u, v, w, fa, sa, Ka, gamma, mode, R)
    return sa, R.astype(bool)


def inc_aggregation(n, edges, weights, delta, delta_w, s_pre, s_cur, R, cap=4096):
    u, v, w = _edges(edges, weights)
    du, dv, dw = _edges(delta, delta_w)
    sp = np.ascontiguousarray(s_pre, dtype=np.uint32).copy()
    sc = np.ascontiguousarray(s_cur, dtype=np.uint32)
    Ra = np.ascontiguousarray(R, dtype=np.uint8)
    ou = np.zeros(cap, dtype=np.uint32)
    ov = np.zeros(cap, dtype=np.uint32)
    ow = np.zeros(cap, dtype=np.float64)
    nn = C.c_uint32(0)
    k = lib().or_inc_aggregation_fn(n, len(u), u, v, w, len(du), du, dv, dw, sp, sc, Ra, ou, ov, ow, cap,
                                    C.byref(nn))
    if k < 0:
        raise ValueError("capacity")
    return list(zip(ou[:k].tolist(), ov[:k].tolist(), ow[:k].tolist())), sp, nn.value


def def_update(f_levels, s_levels, B_levels):
    P = len(f_levels)
    sizes = np.array([len(x) for x in f_levels], dtype=np.uint32)
    f = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.uint32) for x in f_levels]))
    s = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.uint32) for x in s_levels]))
    B = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.uint8) for x in B_levels]))
    lib().or_def_update_fn(P, sizes, f, s, B)
    out_f, out_b, off = [], [], 0
    for z in sizes:
        out_f.append(f[off:off + z].copy())
        out_b.append(B[off:off + z].astype(bool))
        off += z
    return out_f, out_b


def modularity(n, edges, weights, labels):
    u, v, w = _edges(edges, weights)
    return lib().or_modularity_fn(n, len(u), u, v, w, np.ascontiguousarray(labels, dtype=np.uint32))


def select_decoupled_moves(cands):
    """cands: list of (node, from, to, gain); returns selected nodes in order."""
    k = len(cands)
    node = np.array([c[0] for c in cands], dtype=np.uint32)
    frm = np.array([c[1] for c in cands], dtype=np.uint32)
    to = np.array([c[2] for c in cands], dtype=np.uint32)
    gain = np.array([c[3] for c in cands], dtype=np.float64)
    out = np.zeros(max(k, 1), dtype=np.uint32)
    c = lib().or_select_decoupled_fn(k, node, frm, to, gain, out)
    return out[:c].tolist()


def stdrng_shuffle(seed, n):
    out = np.zeros(max(int(n), 1), dtype=np.uint64)
    lib().or_stdrng_shuffle(int(seed), int(n), out.ctypes.data_as(C.c_void_p))
    return out[:int(n)]


def chacha_block(key_words, rounds, counter=0, stream=0):
    k = np.ascontiguousarray(key_words, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint32)
    lib().or_chacha_block(k.ctypes.data_as(C.c_void_p), int(rounds), int(counter), int(stream),
                          out.ctypes.data_as(C.c_void_p))
    return out


def static_run(n, u, v, w=None, gamma=1.0, max_levels=10, seed=0):
    """run() (hit_leiden.rs:30-118): (labels, hierarchy rows, movement calls)."""
    u = np.ascontiguousarray(u, dtype=np.uint32)
    v = np.ascontiguousarray(v, dtype=np.uint32)
    cap = max_levels + 2
    labels = np.zeros(max(n, 1), dtype=np.uint32)
    hier = np.zeros(max(n, 1) * cap, dtype=np.uint32)
    rows = C.c_uint32(0)
    it = lib().or_static_run(n, len(u), u, v, _wptr(w), gamma, max_levels, seed, labels, hier, cap, C.byref(rows))
    if it < 0:
        raise ValueError("invalid edges")
    r = min(rows.value, cap)
    return labels[:n], hier[:r * n].reshape(r, n) if n else np.zeros((r, 0), np.uint32), int(it)


def static_movement(n, edges, weights, f, gamma, seed):
    u, v, w = _edges(edges, weights)
    fa = np.ascontiguousarray(f, dtype=np.uint32).copy()
    lib().or_static_movement_fn(n, len(u), u, v, _wptr(w), fa, gamma, seed)
    return fa


def static_refine(n, edges, weights, f, gamma):
    u, v, w = _edges(edges, weights)
    fa = np.ascontiguousarray(f, dtype=np.uint32)
    out = np.zeros(n, dtype=np.uint32)
    lib().or_static_refine_fn(n, len(u), u, v, _wptr(w), fa, gamma, out)
    return out


def canonicalize(labels):
    a = np.ascontiguousarray(labels, dtype=np.uint32).copy()
    lib().or_canonicalize_fn(len(a), a)
    return a


def should_skip_aggregation(delta_len, R):
    Ra = np.ascontiguousarray(R, dtype=np.uint8)
    return bool(lib().or_should_skip_aggregation_fn(delta_len, len(Ra), Ra))
