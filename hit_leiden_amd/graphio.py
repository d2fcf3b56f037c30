"""Graph files and update streams for the engine.

* Edge-list text: one edge per line, `src dst [weight]`, whitespace separated,
  `#` comments (the format the reference's README.md:36-37 describes; the
  reference ships no loader for it).
* Binary edge list, the server's bulk format: a 32-byte header
  (magic `HLEB`, version u32 = 1, n u32, flags u32 (bit 0: weighted), m u64,
  reserved u64) followed by src u32[m], dst u32[m] and, when weighted, w f64[m],
  little-endian.  `read_binary` memory-maps the arrays.
* `paper_split` (DynamicGraphBuilder::paper_split, dynamic_graph.rs:78-131):
  shuffle the edges, load the first `initial_ratio` of them, then `rounds`
  batches of `batch_size` edges.  The reference returns the cumulative graph
  after each round; the incremental path consumes the difference, so this
  returns the batches themselves (`cumulative()` rebuilds the states).  The
  shuffle is the reference's own: rand 0.8.5 `StdRng::seed_from_u64(seed)`
  (ChaCha12) and `SliceRandom::shuffle`, restated in hl_rng.h.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import GraphInput

MAGIC = b"HLEB"
VERSION = 1
_HEADER = struct.Struct("<4sIIIQQ")


def read_edge_list(path, node_count=None, one_based=False) -> GraphInput:
    """`src dst [weight]` lines; n = max id + 1 unless node_count is given."""
    import pandas as pd
    df = pd.read_csv(path, sep=r"\s+", comment="#", header=None, engine="c")
    if df.shape[1] < 2:
        raise ValueError(f"{path}: need at least two columns")
    src = df.iloc[:, 0].to_numpy(np.int64)
    dst = df.iloc[:, 1].to_numpy(np.int64)
    if one_based:
        src, dst = src - 1, dst - 1
    if len(src) and (src.min() < 0 or dst.min() < 0):
        raise ValueError(f"{path}: negative vertex id")
    n = int(max(src.max(initial=-1), dst.max(initial=-1)) + 1) if node_count is None else int(node_count)
    if len(src) and max(src.max(), dst.max()) >= n:
        raise ValueError(f"{path}: vertex id >= node_count")
    w = df.iloc[:, 2].to_numpy(np.float64) if df.shape[1] > 2 else None
    return GraphInput(n, src.astype(np.uint32), dst.astype(np.uint32), w)


def write_edge_list(path, g: GraphInput):
    with open(path, "w") as f:
        if g.weights is None:
            for a, b in zip(g.src.tolist(), g.dst.tolist()):
                f.write(f"{a} {b}\n")
        else:
            for a, b, x in zip(g.src.tolist(), g.dst.tolist(), g.weights.tolist()):
                f.write(f"{a} {b} {x!r}\n")


def write_binary(path, g: GraphInput):
    m = len(g.src)
    flags = 1 if g.weights is not None else 0
    with open(path, "wb") as f:
        f.write(_HEADER.pack(MAGIC, VERSION, g.node_count, flags, m, 0))
        np.ascontiguousarray(g.src, dtype="<u4").tofile(f)
        np.ascontiguousarray(g.dst, dtype="<u4").tofile(f)
        if flags:
            np.ascontiguousarray(g.weights, dtype="<f8").tofile(f)


def read_binary(path) -> GraphInput:
    with open(path, "rb") as f:
        head = f.read(_HEADER.size)
    if len(head) != _HEADER.size:
        raise ValueError(f"{path}: truncated header")
    magic, version, n, flags, m, _ = _HEADER.unpack(head)
    if magic != MAGIC or version != VERSION:
        raise ValueError(f"{path}: not a version-{VERSION} HLEB file")
    off = _HEADER.size
    src = np.memmap(path, dtype="<u4", mode="r", offset=off, shape=(m,))
    dst = np.memmap(path, dtype="<u4", mode="r", offset=off + 4 * m, shape=(m,))
    w = np.memmap(path, dtype="<f8", mode="r", offset=off + 8 * m, shape=(m,)) if flags & 1 else None
    if m and (int(src.max()) >= n or int(dst.max()) >= n):
        raise ValueError(f"{path}: vertex id >= node_count")
    return GraphInput(n, np.asarray(src), np.asarray(dst), None if w is None else np.asarray(w))


@dataclass
class IncrementalSplit:
    """IncrementalSplit (dynamic_graph.rs:21-27) with update batches as deltas."""
    initial_graph: GraphInput
    update_batches: List[GraphInput] = field(default_factory=list)
    batch_size: int = 0
    rounds: int = 0

    def cumulative(self, i):
        """The reference's update_batches[i]: the initial edges plus batches 0..i."""
        parts = [self.initial_graph] + self.update_batches[:i + 1]
        return GraphInput(self.initial_graph.node_count, np.concatenate([p.src for p in parts]),
                          np.concatenate([p.dst for p in parts]))


def paper_split(graph: GraphInput, initial_ratio, batch_size, rounds, seed) -> IncrementalSplit:
    """Shuffle, take floor(ratio * m) edges as the initial graph, then at most
    `rounds` batches of `batch_size` (dynamic_graph.rs:78-131); unit weights as
    the reference drops them."""
    from .gen import shuffle_order
    m = len(graph.src)
    order = shuffle_order(seed, m)
    src, dst = graph.src[order], graph.dst[order]
    ratio = min(max(float(initial_ratio), 0.0), 1.0)
    init = int(np.floor(m * ratio))
    avail = m - init
    eff = 0 if batch_size == 0 else min(rounds, avail // batch_size)
    n = graph.node_count
    batches = [GraphInput(n, src[init + r * batch_size:init + (r + 1) * batch_size].copy(),
                          dst[init + r * batch_size:init + (r + 1) * batch_size].copy()) for r in range(eff)]
    return IncrementalSplit(GraphInput(n, src[:init].copy(), dst[:init].copy()), batches, batch_size, eff)
