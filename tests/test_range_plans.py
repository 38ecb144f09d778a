"""Byte-range plans over an ndht chunk's coefficient-plane index (CPU).

The oracle restatement (oracle/range_oracle.py) is pinned first: the
reference's low_pass_indices vectors (crates/ndic-lift/src/lib.rs:297-315),
the container sample of crates/ndic-codestream/src/container.rs:349-364
(header_len 116) and plans derived by hand for it.  The library's planner
(ndic_range_plan_chunk, RangeIndex of crates/ndic-codestream/src/
range.rs:141-369) must then agree with the oracle on random containers and
targets, errors included.
"""

from __future__ import annotations

import struct

import numpy as np
import pytest

import range_oracle as ro


def container(dims, lv, planes=None) -> bytes:
    """ndht [header | index] (container.rs:249-280); planes = [(offset, len, prefix)]."""
    b = bytearray(b"ndht") + bytes([1, 1 if planes is not None else 0, lv, len(dims)])
    for d in dims:
        b += struct.pack("<I", d)
    for off, ln, prefix in planes or []:
        b += struct.pack("<QI", off, ln) + struct.pack(f"<{lv + 1}I", *prefix)
    return bytes(b)


SAMPLE = container([4, 16, 32], 2, [(116 + 100 * i, 100, [40, 70, 100]) for i in range(4)])


def test_oracle_low_pass_vectors():
    step = lambda kind, levels, group: {"dimension": 0, "kind": kind, "levels": levels, "group": group}
    assert ro.low_pass_indices(step("lift53", 2, 0), 32) == list(range(8))
    assert ro.low_pass_indices(step("haar", 1, 4), 8) == [0, 1, 4, 5]
    assert ro.low_pass_indices(step("delta", 0, 4), 10) == [0, 4, 8]
    assert ro.low_pass_indices(step("lift53", 8, 0), 5) == [0]
    assert ro.low_pass_indices(step("lift53", 8, 0), 1) == [0]
    assert ro.approx_len(1, 6, 1) == 3


def test_oracle_sample_plans():
    h = ro.parse_chunk_prefix(SAMPLE)
    assert h["header_len"] == 116 and len(SAMPLE) == 116
    # resolutions of a 32 x 16 plane at 2 levels: (8, 4), (16, 8), (32, 16)
    p = ro.plan(SAMPLE, "thumbnail", 16)
    assert (p["max_res"], p["decoded_size"], p["ranges"], p["total_bytes"]) == (1, [8, 16], [(0, 185)], 186)
    p = ro.plan(SAMPLE, "thumbnail", 4)  # nothing fits: resolution 0
    assert (p["max_res"], p["decoded_size"], p["ranges"]) == (0, [4, 8], [(0, 155)])
    p = ro.plan(SAMPLE, "plane", 2)
    assert (p["ranges"], p["total_bytes"], p["decoded_size"]) == ([(0, 115), (316, 415)], 216, [16, 32])
    lift = [{"dimension": 0, "kind": "lift53", "levels": 2, "group": 0}]
    p = ro.plan(SAMPLE, "thumbnail_3d", 32, lift)  # 4 -> 1 low-pass plane
    assert (p["planes"], p["decoded_size"], p["ranges"], p["max_res"]) == ([0], [1, 16, 32], [(0, 215)], 2)
    haar = [{"dimension": 0, "kind": "haar", "levels": 1, "group": 0}]
    p = ro.plan(SAMPLE, "thumbnail_3d", 16, haar)  # 4 -> 2 low-pass planes
    assert (p["planes"], p["decoded_size"], p["ranges"], p["total_bytes"]) == (
        [0, 1], [2, 8, 16], [(0, 185), (216, 285)], 256)
    p = ro.plan(SAMPLE, "thumbnail_3d", 16, [])  # untransformed: every plane
    assert p["planes"] == [0, 1, 2, 3] and p["decoded_size"] == [4, 8, 16]


def random_container(rng):
    ndim = int(rng.integers(2, 6))
    dims = [int(rng.integers(1, 6)) for _ in range(ndim - 2)] + [int(rng.integers(1, 300)), int(rng.integers(1, 300))]
    lv = int(rng.integers(0, 6))
    n = int(np.prod(dims[:-2])) if ndim > 2 else 1
    fixed = 8 + 4 * ndim + n * (12 + 4 * (lv + 1))
    off, planes = fixed + int(rng.integers(0, 3)), []
    for _ in range(n):
        ln = int(rng.integers(1, 500))
        prefix = sorted(int(x) for x in rng.integers(1, ln + 1, lv + 1))
        planes.append((off, ln, prefix))
        off += ln + int(rng.integers(0, 3))
    return dims, lv, container(dims, lv, planes)


def random_steps(rng, lead):
    steps = []
    for d in range(lead):
        if rng.uniform() < 0.6:
            steps.append({"dimension": d, "kind": ["delta", "haar", "lift53"][int(rng.integers(0, 3))],
                          "levels": int(rng.integers(1, 4)), "group": int(rng.choice([0, 2, 3, 4]))})
    return steps


def test_planner_matches_oracle(ndic):
    rng = np.random.default_rng(17)
    for _ in range(300):
        dims, lv, pre = random_container(rng)
        target = ["thumbnail", "plane", "thumbnail_3d"][int(rng.integers(0, 3))]
        steps = random_steps(rng, len(dims) - 2)
        arg = int(rng.integers(0, int(np.prod(dims[:-2])) + 1)) if target == "plane" else int(rng.integers(1, 400))
        try:
            exp = ro.plan(pre, target, arg, steps)
        except ro.PlanError as e:
            with pytest.raises(ndic.NdicError) as got:
                ndic.range_plan(pre, target, arg, steps)
            assert (got.value.kind, e.message in str(got.value)) == (e.kind, True), (target, arg, dims, e.message)
            continue
        got = ndic.range_plan(pre, target, arg, steps)
        assert got == exp, (dims, lv, target, arg, steps)


@pytest.mark.parametrize("mutate", ["short", "magic", "version", "flags", "levels", "ndim", "zero", "trunc_hdr",
                                    "trunc_index", "overlap", "prefix", "prefix_len", "no_index", "step_dim",
                                    "plane_range"])
def test_planner_errors_match_oracle(ndic, mutate):
    b = bytearray(SAMPLE)
    target, arg, steps = "thumbnail", 16, []
    if mutate == "short":
        b = b[:7]
    elif mutate == "magic":
        b[0] = ord("x")
    elif mutate == "version":
        b[4] = 2
    elif mutate == "flags":
        b[5] = 3
    elif mutate == "levels":
        b[6] = 33
    elif mutate == "ndim":
        b[7] = 1
    elif mutate == "zero":
        b[12:16] = b"\0\0\0\0"
    elif mutate == "trunc_hdr":
        b = b[:14]
    elif mutate == "trunc_index":
        b = b[:100]
    elif mutate == "overlap":
        struct.pack_into("<Q", b, 20 + 24, 150)
    elif mutate == "prefix":
        struct.pack_into("<I", b, 20 + 12, 90)
    elif mutate == "prefix_len":
        struct.pack_into("<I", b, 20 + 20, 101)
    elif mutate == "no_index":
        b = bytearray(container([4, 16, 32], 2))
    elif mutate == "step_dim":
        target, steps = "thumbnail_3d", [{"dimension": 1, "kind": "haar", "levels": 1, "group": 0}]
    elif mutate == "plane_range":
        target, arg = "plane", 4
    with pytest.raises(ro.PlanError) as e:
        ro.plan(bytes(b), target, arg, steps)
    with pytest.raises(ndic.NdicError) as got:
        ndic.range_plan(bytes(b), target, arg, steps)
    assert got.value.kind == e.value.kind, (mutate, str(got.value))
    assert e.value.message in str(got.value), (mutate, str(got.value))
    if e.value.kind == "Codestream":
        assert got.value.offset == e.value.offset, mutate
