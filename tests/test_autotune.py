"""The autotuner: its crossover rule on synthetic measurements (CPU), and a
short tuning run on the device whose choices come from the measured ladder
and leave decodes bit-exact (GPU)."""

from __future__ import annotations

import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def autotune():
    return importlib.import_module("nd-image-codecs_amd.autotune")


def test_crossover_rule(autotune):
    sizes = [100, 200, 400, 800]
    # the small form wins up to 200, loses at 400: the crossing is interpolated
    assert autotune.crossover(sizes, [1, 2, 5, 9], [4, 4, 4, 4]) == 200 + int(200 * 2 / (2 + 1))
    # loses at once
    assert autotune.crossover(sizes, [5, 6, 7, 8], [1, 1, 1, 1]) == 0
    # wins everywhere: up to the largest size measured, not beyond
    assert autotune.crossover(sizes, [1, 1, 1, 1], [2, 2, 2, 2]) == 800
    # a tie counts as a win; with the tie just before the loss the crossing is the tie
    assert autotune.crossover(sizes, [1, 2, 3, 5], [1, 2, 3, 4]) == 400



def test_knob_registry(ndic):
    """The knobs are host state: defaults, the accepted ranges, set / reset
    round trips and the errors for unknown keys and values out of range."""
    defaults = {"ht_decode_wave_max_blocks": 8192, "parse_wave_max_planes": 4096, "ht_wave_grain": 1,
                "fuse_dwt_lift": 4096, "pipeline_windows": 1, "window_elems": 0, "ht_lane_waves": 0,
                "graph_windows": 0, "ht_decode_sort": 0}
    keys = ndic.tune_keys()
    assert set(defaults) <= set(keys)
    before = {k: ndic.tune_get(k) for k in keys}
    try:
        for k in keys:
            ndic.tune_reset(k)
        for k, v in defaults.items():
            assert ndic.tune_get(k) == v, k
        for k, lo, hi in [("pipeline_windows", 0, 3), ("ht_wave_grain", 1, 64), ("ht_lane_waves", 0, 4),
                          ("graph_windows", 0, 1), ("ht_decode_sort", 0, 1)]:
            for v in (lo, hi):
                ndic.tune_set(k, v)
                assert ndic.tune_get(k) == v
            for v in (lo - 1, hi + 1):
                with pytest.raises(ValueError):
                    ndic.tune_set(k, v)
            assert ndic.tune_get(k) == hi  # a refused value leaves the knob as it was
            ndic.tune_reset(k)
            assert ndic.tune_get(k) == defaults[k]
        with pytest.raises(KeyError):
            ndic.tune_set("no_such_knob", 1)
        with pytest.raises(KeyError):
            ndic.tune_get("no_such_knob")
    finally:
        for k, v in before.items():
            ndic.tune_reset(k)
            if ndic.tune_get(k) != v:
                ndic.tune_set(k, v)

@pytest.mark.gpu
def test_short_tuning_run(autotune, ndic):
    keys = autotune.KNOBS
    before = {k: ndic.tune_get(k) for k in keys}
    try:
        rec = autotune.tune(quick=True, reps=1, ladder=(200, 1600, 6400))
        m = rec["measurements"]
        assert set(rec["knobs"]) == set(keys)
        assert all(len(m[k]) == 3 for k in ("decode_wave_ms", "decode_lane_ms", "encode_wave_ms", "encode_lane_ms"))
        assert all(v > 0 for v in m["decode_wave_ms"] + m["decode_lane_ms"])
        assert rec["knobs"]["ht_wave_grain"] in autotune.GRAINS
        for k, v in rec["knobs"].items():
            assert ndic.tune_get(k) == v
        # the tuned knobs decode a small batch bit-exactly: the decoded window
        # (zeroed before) equals the samples it was encoded from, which are
        # not all zero
        w = autotune._Window(400, seed=5)
        w.decode_stages(1)
        for (z, y, x), o in zip(w.shapes, w.offs):
            n = z * y * x * 2
            assert w.out.download(o, n) == w.raw.download(o, n)
            assert np.frombuffer(w.raw.download(o, n), np.uint16).any()
    finally:
        for k, v in before.items():
            ndic.tune_reset(k)
            if ndic.tune_get(k) != v:
                ndic.tune_set(k, v)
