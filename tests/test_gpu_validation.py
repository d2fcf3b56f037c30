"""The reference's run-level tests on the device: run() (the static
multilevel_leiden, HitLeiden.static_leiden) judged with validate and check, as
tests/contract/test_run_validate.rs:20-30,
tests/integration/test_throughput_equivalence.rs:20-34 and
tests/property/test_partition_invariants.rs:20-31 do."""
import pytest

import hit_leiden_amd as hl
from hit_leiden_amd.validation import check, validate
from test_gpu_parity import say

pytestmark = pytest.mark.gpu

DET, THR = hl.RunMode.DETERMINISTIC, hl.RunMode.THROUGHPUT


def static_result(n, edges, mode):
    return hl.run(hl.GraphInput.from_edges(n, edges), hl.RunConfig(mode=mode)).partition


def test_run_and_validate_contract():
    """RunConfig::default() is the deterministic mode (config.rs:39-50)."""
    say("validate: path of three")
    out = hl.run(hl.GraphInput.from_edges(3, [(0, 1), (1, 2)]), hl.RunConfig(mode=DET, levels=1))
    assert check(out)
    assert len(out.partition.node_to_community) == 3
    assert validate(out, out, DET).equivalence_passed


def test_throughput_equivalence_bounds():
    say("validate: one edge, two throughput runs")
    a = static_result(2, [(0, 1, 1.0)], THR)
    b = static_result(2, [(0, 1, 1.0)], THR)
    v = validate(a, b, THR)
    assert v.hard_invariants_passed
    assert v.equivalence_passed


@pytest.mark.parametrize("n", [0, 1, 2, 7, 49])
def test_partition_len_matches_nodes(n):
    """Edgeless graphs: one community per vertex, modularity 0."""
    say(f"validate: {n} isolated vertices")
    r = static_result(n, [], DET)
    assert len(r.node_to_community) == n
    assert r.community_count == n
    assert r.quality_score == 0.0
    assert check(r)


def test_empty_vertex_set_rejected():
    """run() takes node_count 0 (above); the incremental engine's load
    rejects an empty vertex set before any device work."""
    e = hl.HitLeiden(hl.RunConfig(mode=DET, levels=1))
    try:
        with pytest.raises(hl.InvalidInput):
            e.load(hl.GraphInput.from_edges(0, []))
    finally:
        e.close()
