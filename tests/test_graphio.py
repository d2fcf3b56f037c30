"""Edge-list files, the binary bulk format and paper_split (CPU only)."""
import numpy as np
import pytest

import hit_leiden_amd as hl
from hit_leiden_amd import graphio


def test_text_round_trip(tmp_path):
    g = hl.GraphInput(7, [0, 1, 5], [1, 6, 2], [1.5, 0.25, 3.0])
    p = tmp_path / "g.txt"
    graphio.write_edge_list(p, g)
    r = graphio.read_edge_list(p)
    assert r.node_count == 7
    np.testing.assert_array_equal(r.src, g.src)
    np.testing.assert_array_equal(r.dst, g.dst)
    np.testing.assert_array_equal(r.weights, g.weights)


def test_text_comments_unweighted_one_based(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("# header\n1 2\n2 3\n\n3 1\n")
    r = graphio.read_edge_list(p, one_based=True)
    assert r.node_count == 3 and r.weights is None
    np.testing.assert_array_equal(r.src, [0, 1, 2])
    np.testing.assert_array_equal(r.dst, [1, 2, 0])
    with pytest.raises(ValueError):
        graphio.read_edge_list(p, node_count=2)


@pytest.mark.parametrize("weighted", [False, True])
def test_binary_round_trip(tmp_path, weighted):
    rng = np.random.default_rng(1)
    n, m = 1000, 5000
    g = hl.GraphInput(n, rng.integers(0, n, m), rng.integers(0, n, m), rng.random(m) if weighted else None)
    p = tmp_path / "g.hleb"
    graphio.write_binary(p, g)
    assert p.stat().st_size == 32 + m * (16 if weighted else 8)
    r = graphio.read_binary(p)
    assert r.node_count == n
    np.testing.assert_array_equal(r.src, g.src)
    np.testing.assert_array_equal(r.dst, g.dst)
    if weighted:
        np.testing.assert_array_equal(r.weights, g.weights)
    else:
        assert r.weights is None
    with open(p, "r+b") as f:
        f.write(b"XXXX")
    with pytest.raises(ValueError):
        graphio.read_binary(p)


def test_paper_split_matches_reference_shape():
    # dynamic_graph.rs:138-156: ring of 100, ratio 0.8, batches of 5, 4 rounds
    g = hl.GraphInput(100, np.arange(100), (np.arange(100) + 1) % 100)
    s = graphio.paper_split(g, 0.8, 5, 4, 42)
    assert len(s.initial_graph) == 80
    assert len(s.update_batches) == 4 and s.rounds == 4 and s.batch_size == 5
    assert len(s.cumulative(0)) == 85
    assert len(s.cumulative(3)) == 100
    # every edge exactly once across the initial graph and the batches
    parts = [s.initial_graph] + s.update_batches
    key = np.sort(np.concatenate([p.src.astype(np.int64) * 100 + p.dst for p in parts]))
    np.testing.assert_array_equal(key, np.sort(g.src.astype(np.int64) * 100 + g.dst))
    # rounds are capped by the edges left
    assert graphio.paper_split(g, 0.9, 5, 10, 1).rounds == 2
    assert graphio.paper_split(g, 0.5, 0, 3, 1).rounds == 0
