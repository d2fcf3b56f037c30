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


# ---- the seeded shuffle: rand 0.8.5 StdRng (ChaCha12) + SliceRandom::shuffle ----

def words_hex(words):
    return b"".join(int(w).to_bytes(4, "little") for w in words).hex()


# Published ChaCha keystream vectors: the all-zero 256-bit key and nonce at 20,
# 12 and 8 rounds (the first 32 bytes; RFC 7539 A.1 #1 for 20 rounds, the
# ChaCha test-vector draft for 12 and 8), and RFC 8439 section 2.3.2 (key
# 00..1f, counter 1, nonce 00:00:00:09:00:00:00:4a:00:00:00:00).
ZERO_KEY = {20: "76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7",
            12: "9bf49a6a0755f953811fce125f2683d50429c3bb49e074147e0089a52eae155f",
            8: "3e00ef2f895f40d67f5bb8e81f09a5a12c840ec3ce9a7f3b181be188ef711a1e"}
RFC8439_232 = ("10f1e7e4d13b5915500fdd1fa32071c4c7d1f4c733c068030422aa9ac3d46c4e"
               "d2826446079faa0914c2d705d98b02a2b5129cd1de164eb9cbd083e8a2503c4e")


@pytest.mark.parametrize("rounds", [20, 12, 8])
def test_chacha_zero_key_vectors(rounds):
    from hit_leiden_amd import gen
    from oracle import hl_oracle as O
    for impl in (gen.chacha_block, O.chacha_block):
        assert words_hex(impl(np.zeros(8, np.uint32), rounds))[:64] == ZERO_KEY[rounds]


def test_chacha_rfc8439_block():
    from hit_leiden_amd import gen
    from oracle import hl_oracle as O
    key = np.frombuffer(bytes(range(32)), dtype="<u4")
    # the 96-bit nonce in the 64-bit counter / 64-bit stream layout: word 13 is the counter's high half
    counter = (0x09000000 << 32) | 1
    for impl in (gen.chacha_block, O.chacha_block):
        assert words_hex(impl(key, 20, counter, 0x4a000000)) == RFC8439_232


@pytest.mark.parametrize("seed,n", [(0, 1), (42, 10), (7, 1000), (2**63 + 5, 4097), (123456789, 200000)])
def test_shuffle_two_restatements_agree(seed, n):
    from hit_leiden_amd import gen
    from oracle import hl_oracle as O
    a, b = gen.shuffle_order(seed, n), O.stdrng_shuffle(seed, n)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(np.sort(a), np.arange(n))


def test_shuffle_is_uniform():
    """Position of each element over 20,000 seeds of a 5-element shuffle: a
    chi-square test against the uniform distribution (a wrong bound in the
    Fisher-Yates loop or the rejection zone skews it)."""
    from hit_leiden_amd import gen
    n, trials = 5, 20000
    counts = np.zeros((n, n))
    for s in range(trials):
        o = gen.shuffle_order(s, n)
        counts[np.arange(n), o] += 1
    exp = trials / n
    chi2 = ((counts - exp) ** 2 / exp).sum()
    assert chi2 < 60.0, chi2  # 16 degrees of freedom: p ~ 1e-6 at 60


def test_paper_split_uses_the_reference_shuffle():
    """The edges land in StdRng::seed_from_u64(seed) + shuffle order
    (dynamic_graph.rs:85-87)."""
    from hit_leiden_amd import gen
    g = hl.GraphInput(100, np.arange(100), (np.arange(100) + 1) % 100)
    s = graphio.paper_split(g, 0.8, 5, 4, 42)
    order = gen.shuffle_order(42, 100)
    np.testing.assert_array_equal(s.initial_graph.src, g.src[order[:80]])
    np.testing.assert_array_equal(s.update_batches[0].dst, g.dst[order[80:85]])
