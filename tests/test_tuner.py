"""Launch-shape tuner: candidate lists from the native library and the
coordinate-descent selection (no GPU needed)."""
from __future__ import annotations

import pytest

from quantem.gpu_amd.ssb.compute.hip import tuner


@pytest.mark.parametrize("n", [128, 256, 512, 1024])
def test_candidates_start_with_the_default(n):
    assert tuner.candidates(n, "stream_wg")[0] == 1536
    pairs = tuner.candidates(n, "gather_pairs")
    assert pairs[0] == {128: 16, 256: 8, 512: 8, 1024: 4}[n] and len(pairs) == 2
    # at 256 the wave-local kernel comes first, then the three tile shapes
    assert tuner.candidates(n, "trotter") == list(range(1 if n == 1024 else (5 if n == 256 else 3)))


def test_candidates_reject_unsupported_scan_sizes():
    with pytest.raises(NotImplementedError, match="scan size"):
        tuner.candidates(192, "trotter")


def test_coordinate_descent_picks_each_knobs_fastest_value():
    cost = {"a": {1: 5.0, 2: 3.0, 3: 3.0}, "b": {10: 1.0, 20: 2.0}}
    seen = []

    def measure(knob, value, cfg):
        seen.append(dict(cfg))
        return cost[knob][value] + (0.5 if cfg["a"] == 3 else 0.0)

    best, table = tuner.coordinate_descent(measure, {"a": [1, 2, 3], "b": [10, 20]}, {"a": 1, "b": 10})
    assert best == {"a": 2, "b": 10}          # tie 3.0 vs 3.5 resolved by the penalty; earlier wins ties
    assert table["a"] == {1: 5.0, 2: 3.0, 3: 3.5}
    assert all(c["a"] == 2 for c in seen[3:])  # later knobs see the best earlier choice


def test_cache_key_buckets_bf_counts():
    k1 = tuner.cache_key("MI355X", 256, (128, 128), 793)
    assert k1 == tuner.cache_key("MI355X", 256, (128, 128), 1024) == "MI355X|scan256|det128x128|bf1024"
    assert k1 != tuner.cache_key("MI355X", 256, (128, 128), 1025)
