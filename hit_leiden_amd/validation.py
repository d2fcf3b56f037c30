"""The reference's validation layer over the engine's results.

* ``PartitionResult`` (src/core/types.rs:104-114) and ``partition_result``,
  which builds one from an engine: ``node_to_community`` is the community
  partition (row P-1 of the hierarchy, the top level pulled down to the
  vertices, relabelled first-seen contiguous as
  canonicalize_community_ids_in_place does, hit_leiden.rs:358-370) and
  ``quality_score`` its modularity.
* ``validate`` (src/core/validation/equivalence.rs:21-44, exported as
  ``hit_leiden::validate``, src/lib.rs:38-44): deterministic runs are
  equivalent when their node-to-community vectors are equal, throughput runs
  when their quality scores differ by at most 0.001.  Hard invariants are
  reported as passed, as the reference reports them.
* ``check`` (src/core/validation/invariants.rs:19-28): a partition exists and
  every community id is below its node count.
* ``run`` (hit_leiden::run, src/lib.rs:27-29; hit_leiden.rs:30-118): the
  reference's static multilevel_leiden on the device
  (``HitLeiden.static_leiden``) returned as a ``RunOutcome``.

The benchmark's contingency metrics (NMI, purity, largest-community Jaccard,
hit_leiden_incremental.rs:577-666) run on the device:
``HitLeiden.compare_partitions``.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import PART_COMMUNITY, RunMode

QUALITY_TOLERANCE = 0.001  # equivalence.rs:41 (quality_tolerance of the config is unused there)


@dataclass
class PartitionResult:
    node_to_community: np.ndarray
    quality_score: float
    hierarchy_levels: List[np.ndarray] = field(default_factory=list)
    community_count: int = 0
    iteration_count: int = 0
    run_id: str = ""


@dataclass
class RunOutcome:
    """src/core/types.rs:127-131, with the partition only."""
    partition: Optional[PartitionResult]


def run(graph, config=None, seed=0) -> RunOutcome:
    """hit_leiden::run (src/lib.rs:27-29, hit_leiden.rs:30-118) on the device:
    multilevel_leiden (HitLeiden.static_leiden) with max_levels =
    config.max_iterations (config.rs:31, > 0) and the queue shuffle drawn from
    StdRng::seed_from_u64(seed); node_to_community numbered first-seen
    (canonicalize_community_ids_in_place), hierarchy_levels the rows,
    iteration_count the movement calls and quality_score compute_modularity
    of the communities (hit_leiden.rs:121-161: 0 without edge weight).  The
    static path does not depend on the mode (single_level_movement always
    takes the queue, :806)."""
    from . import HitLeiden, RunConfig
    cfg = config or RunConfig()
    e = HitLeiden(cfg)
    try:
        labels, rows, calls = e.static_leiden(graph, max_levels=cfg.max_iterations, seed=seed)
        q = 0.0
        if graph.node_count and len(graph):
            e.load_graph(graph)
            q = e.modularity_labels(labels)
        part = PartitionResult(node_to_community=labels.astype(np.int64), quality_score=float(q),
                               hierarchy_levels=[r.astype(np.int64) for r in rows],
                               community_count=int(len(np.unique(labels))), iteration_count=int(calls),
                               run_id=f"run:{getattr(graph, 'dataset_id', '')}")
        return RunOutcome(partition=part)
    finally:
        e.close()


def _partition(x):
    return x.partition if isinstance(x, RunOutcome) else x


@dataclass
class ValidationOutcome:
    """src/core/report.rs:28-33"""
    hard_invariants_passed: bool
    deterministic_identity_passed: Optional[bool]
    quality_delta_vs_reference: Optional[float]
    equivalence_passed: bool


def partition_result(engine, iteration_count=0, run_id="") -> PartitionResult:
    """The engine's current communities as a PartitionResult."""
    levels = engine.hierarchy(canonical=True)
    labels = levels[-1].astype(np.int64) if len(levels) else np.zeros(0, np.int64)
    return PartitionResult(node_to_community=labels, quality_score=engine.modularity(PART_COMMUNITY),
                           hierarchy_levels=[row.astype(np.int64) for row in levels],
                           community_count=int(len(np.unique(labels))), iteration_count=int(iteration_count),
                           run_id=run_id)


def validate(reference, candidate, mode: RunMode) -> ValidationOutcome:
    """equivalence.rs:21-44 over two RunOutcomes or PartitionResults.  Both
    partitions must exist (the reference unwraps them)."""
    reference, candidate = _partition(reference), _partition(candidate)
    if reference is None or candidate is None:
        raise ValueError("validate needs two partitions")
    a = np.asarray(reference.node_to_community)
    b = np.asarray(candidate.node_to_community)
    same = a.shape == b.shape and bool(np.array_equal(a, b))
    delta = abs(float(reference.quality_score) - float(candidate.quality_score))
    if RunMode(mode) == RunMode.DETERMINISTIC:
        return ValidationOutcome(True, same, delta, same)
    return ValidationOutcome(True, None, delta, delta <= QUALITY_TOLERANCE)


def check(run_or_partition) -> bool:
    """invariants.rs:19-28: ids below the node count (an absent partition fails)."""
    partition = _partition(run_or_partition)
    if partition is None:
        return False
    labels = np.asarray(partition.node_to_community)
    return bool(labels.size == 0 or (labels.min() >= 0 and labels.max() < labels.size))
