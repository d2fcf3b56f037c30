"""The validation layer (hit_leiden_amd.validation) against the rules of
src/core/validation/equivalence.rs:21-44 and invariants.rs:19-28, at their
boundaries: equal and unequal partitions in deterministic mode, quality
deltas on either side of 0.001 in throughput mode, ids at the node count."""
import numpy as np
import pytest

import hit_leiden_amd as hl
from hit_leiden_amd.validation import PartitionResult, check, validate


def part(labels, q):
    return PartitionResult(node_to_community=np.asarray(labels, np.int64), quality_score=q)


def test_deterministic_identity():
    a = part([0, 0, 1], 0.25)
    v = validate(a, part([0, 0, 1], 0.25), hl.RunMode.DETERMINISTIC)
    assert v.hard_invariants_passed and v.deterministic_identity_passed and v.equivalence_passed
    assert v.quality_delta_vs_reference == 0.0
    v = validate(a, part([0, 1, 1], 0.25), hl.RunMode.DETERMINISTIC)
    assert v.deterministic_identity_passed is False and v.equivalence_passed is False
    # equal quality does not make different partitions equivalent, nor does a length change
    assert not validate(a, part([0, 0], 0.25), hl.RunMode.DETERMINISTIC).equivalence_passed


def test_throughput_quality_bound():
    a = part([0, 0, 1], 0.5)
    v = validate(a, part([1, 1, 0], 0.5 - 0.00075), hl.RunMode.THROUGHPUT)
    assert v.deterministic_identity_passed is None and v.equivalence_passed
    assert v.quality_delta_vs_reference == pytest.approx(0.00075)
    assert not validate(a, part([0, 0, 1], 0.5 + 0.0011), hl.RunMode.THROUGHPUT).equivalence_passed
    # different partitions of equal quality pass in throughput mode
    assert validate(a, part([0, 1, 2], 0.5), hl.RunMode.THROUGHPUT).equivalence_passed


def test_self_validation_passes():
    """tests/contract/test_run_validate.rs:20-30: a result validated against
    itself is equivalent, in both modes."""
    a = part([0, 1, 1, 2], 0.1)
    for mode in hl.RunMode:
        assert validate(a, a, mode).equivalence_passed


def test_run_outcomes_accepted():
    """validate and check take RunOutcomes as the reference's do; a missing
    partition fails the check and cannot be validated."""
    a = hl.RunOutcome(partition=part([0, 0, 1], 0.3))
    assert validate(a, a, hl.RunMode.DETERMINISTIC).equivalence_passed
    assert check(a)
    assert not check(hl.RunOutcome(partition=None))
    with pytest.raises(ValueError):
        validate(a, hl.RunOutcome(partition=None), hl.RunMode.THROUGHPUT)


def test_check_ids_below_node_count():
    assert check(part([0, 1, 2], 0.0))
    assert check(part([], 0.0))
    assert not check(part([0, 3, 1], 0.0))
    assert not check(None)


class _Engine:
    """Stands in for HitLeiden: a two-level hierarchy and its modularity."""

    def __init__(self, rows, q):
        self.rows, self.q, self.asked = rows, q, []

    def hierarchy(self, canonical=False):
        self.asked.append(canonical)
        return np.asarray(self.rows, np.uint32)

    def modularity(self, which):
        assert which == hl.PART_COMMUNITY
        return self.q


def test_partition_result_takes_the_communities():
    """node_to_community is the last hierarchy row (the communities pulled down
    to the vertices), canonical, with its modularity and community count."""
    from hit_leiden_amd.validation import partition_result
    e = _Engine([[0, 1, 1, 2, 3], [0, 0, 0, 1, 1]], 0.42)
    r = partition_result(e, iteration_count=3, run_id="x")
    assert e.asked == [True]
    assert r.node_to_community.tolist() == [0, 0, 0, 1, 1]
    assert r.community_count == 2 and r.quality_score == 0.42 and r.iteration_count == 3
    assert [row.tolist() for row in r.hierarchy_levels] == [[0, 1, 1, 2, 3], [0, 0, 0, 1, 1]]
    assert check(r)
