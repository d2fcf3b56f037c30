"""Bad caller input is rejected before any device work (the reference panics on
out-of-range ids, hit_leiden.rs:1060; `run` returns InvalidInput,
hit_leiden.rs:31-43).

CPU tests cover the host-side checks of the Python mirror (checkpoint
contents, graph-local batch ids); the GPU tests at the end call the C ABI
with labels the kernels cannot hold (0xFFFFFFFF is every hash table's empty
key) and expect HL_INVALID_INPUT with the handle unchanged.
"""
import numpy as np
import pytest

import hit_leiden_amd as hl
from hit_leiden_amd import checkpoint
from conftest import gpu_available


def write_checkpoint(path, n=5, f=None, u=None, v=None, w=None):
    f = np.arange(n, dtype=np.uint32) if f is None else np.asarray(f, np.uint32)
    u = np.array([0, 1], np.uint32) if u is None else np.asarray(u, np.uint32)
    v = np.array([1, 2], np.uint32) if v is None else np.asarray(v, np.uint32)
    w = np.ones(len(u)) if w is None else np.asarray(w, np.float64)
    ident = np.arange(n, dtype=np.uint32)
    np.savez(path, format=np.array([checkpoint.FORMAT]), levels=np.array([1]),
             config=np.array([1.0, 1, 20, 0, 0], np.float64), n0=np.array([n]), u0=u, v0=v, w0=w,
             f0=f, g0=ident, spre0=ident, scur0=ident)


@pytest.mark.parametrize("label", [0xFFFFFFFF, 2**31])
def test_checkpoint_label_out_of_range(tmp_path, label):
    p = tmp_path / "bad.npz"
    write_checkpoint(p, f=[0, 1, label, 3, 4])
    with pytest.raises(hl.InvalidInput, match="2\\^31"):
        checkpoint.restore(p)


def test_checkpoint_map_length(tmp_path):
    p = tmp_path / "bad.npz"
    write_checkpoint(p, f=[0, 1, 2])
    with pytest.raises(hl.InvalidInput, match="labels for 5 vertices"):
        checkpoint.restore(p)


def test_checkpoint_endpoint_and_weight(tmp_path):
    p = tmp_path / "bad.npz"
    write_checkpoint(p, u=[0, 7], v=[1, 2])
    with pytest.raises(hl.InvalidInput, match="endpoint"):
        checkpoint.restore(p)
    write_checkpoint(p, w=[1.0, np.nan])
    with pytest.raises(hl.InvalidInput, match="non-finite"):
        checkpoint.restore(p)


def test_graphs_delta_rejects_ids_outside_their_graph():
    """Two endpoints that overflow into the next graph would pass the engine's
    'joins two graphs' check and edit the wrong graph: graphs_delta rejects them."""
    e = hl.HitLeiden.__new__(hl.HitLeiden)  # the union-id mapping needs no device
    e.graph_offsets = np.array([0, 10, 30], np.uint32)
    ok = e.graphs_delta({1: hl.GraphInput(20, [0, 19], [5, 3])})
    assert ok[0].tolist() == [10, 29]
    with pytest.raises(hl.InvalidInput, match="outside its 10 vertices"):
        e.graphs_delta({0: hl.GraphInput(10, [12], [15])})
    with pytest.raises(hl.InvalidInput, match="graph index"):
        e.graphs_delta({2: hl.GraphInput(1, [0], [0])})


gpu = pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")


@pytest.mark.gpu
@gpu
def test_write_partition_rejects_sentinel_labels():
    g = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT))
    g.load(hl.GraphInput(5, [0, 1, 3], [1, 2, 4]))
    before = g.partition(0, hl.MAP_F).copy()
    for bad in (0xFFFFFFFF, 2**31):
        with pytest.raises(hl.InvalidInput, match="2\\^31"):
            g.write_partition(0, hl.MAP_F, np.array([0, 1, bad, 3, 4], np.uint32))
        np.testing.assert_array_equal(g.partition(0, hl.MAP_F), before)
    g.close()


@pytest.mark.gpu
@gpu
def test_restore_level_rejects_sentinel_labels():
    import ctypes as C
    g = hl.HitLeiden(hl.RunConfig(mode=hl.RunMode.THROUGHPUT))
    u = np.array([0, 1], np.uint32)
    v = np.array([1, 2], np.uint32)
    w = np.ones(2)
    bad = np.array([0, 1, 0xFFFFFFFF], np.uint32)
    st = hl.lib().hl_restore_level(g._h, 0, 3, 2, hl._ptr(u), hl._ptr(v), hl._ptr(w), None, None, None,
                                   hl._ptr(bad))
    assert st == hl.HL_INVALID_INPUT
    assert "2^31" in hl.lib().hl_last_error(g._h).decode()
    g.close()
