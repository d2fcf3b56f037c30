"""The autotuner's search and cache on the CPU, over a stand-in timing
function (the device timings are exercised in test_gpu_parity)."""
import json

from hit_leiden_amd import autotune


def fake(calls):
    def time_of(t):
        calls.append(t)
        return (abs(t.small_degree - 8) + abs(t.wave_degree - 256) / 64 + abs(t.marks_block - 512) / 128 +
                abs(t.marks_grid - 2048) / 512)
    return time_of


def test_two_stage_search_finds_the_separable_minimum():
    calls = []
    best = autotune._search(fake(calls), autotune.Tuned(16, 512, 256, 1024), None)
    assert best == autotune.Tuned(8, 256, 512, 2048)
    # stage 1 times every threshold pair, stage 2 every marks shape; the shared point once
    assert len(calls) == len(set(calls)) == (len(autotune.SMALL) * len(autotune.WAVE) +
                                             len(autotune.MARKS_BLOCK) * len(autotune.MARKS_GRID) - 1)


def test_candidates_are_valid_launch_shapes():
    assert all(0 < s <= 16 for s in autotune.SMALL)
    assert all(0 < w <= 512 for w in autotune.WAVE)
    assert set(autotune.MARKS_BLOCK) <= {128, 256, 512}
    assert all(0 < g <= 65536 for g in autotune.MARKS_GRID)


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tune.json")
    calls = []
    t1 = autotune._cached(path, "k", lambda: autotune._search(fake(calls), autotune.Tuned(16, 512, 256, 1024), None))
    n = len(calls)
    t2 = autotune._cached(path, "k", lambda: autotune._search(fake(calls), autotune.Tuned(16, 512, 256, 1024), None))
    assert t1 == t2 and len(calls) == n  # the second call reads the cache
    assert json.load(open(path))["k"] == list(t1)
    cfg = t1.config(levels=3)
    assert (cfg.small_degree, cfg.wave_degree, cfg.marks_block, cfg.marks_grid, cfg.levels) == (8, 256, 512, 2048, 3)


def test_noise_sized_differences_keep_the_built_in_shape():
    """Candidates within the margin of the incumbent (run-to-run noise) do not
    replace it; one clearly faster does."""
    base = autotune.Tuned(16, 512, 256, 1024)

    def flat(t):
        return 1.0 - 0.01 * (t != base)  # every other shape 1 % faster
    assert autotune._search(flat, base, None) == base

    def one_better(t):
        return 0.9 if t == base._replace(marks_grid=2048) else 1.0
    assert autotune._search(one_better, base, None) == base._replace(marks_grid=2048)
