"""nd_lift on the GPU against the oracle, bit-exact.

Covers the reference's lift test matrix (crates/ndic-lift/tests/roundtrip.rs,
crates/ndic-lift/src/chunk.rs:320-472): every kind, levels, groups (including
segments longer than the LDS budget), odd and singleton extents, multi-step
configs, u8..i64 dtypes, and the refusal paths.
"""

from __future__ import annotations

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NP = {"u8": np.uint8, "i8": np.int8, "u16": np.uint16, "i16": np.int16, "u32": np.uint32, "i32": np.int32,
      "u64": np.uint64, "i64": np.int64}
RANGES = {"u8": (0, 255), "i8": (-128, 127), "u16": (0, 65535), "i16": (-32768, 32767), "u32": (0, 2**20),
          "i32": (-2**20, 2**20), "u64": (0, 2**40), "i64": (-2**40, 2**40)}


def _roundtrip(ndic, oracle, arr, dt, steps):
    raw = arr.tobytes()
    exp = oracle.lift_chunk(True, raw, arr.shape, dt, steps)
    got = ndic.lift_forward_bytes(raw, arr.shape, ndic.DTYPE_CODES[np.dtype(NP[dt]).name], steps)
    assert got == exp
    back_exp = oracle.lift_chunk(False, exp, arr.shape, dt, steps)
    back = ndic.lift_inverse_bytes(exp, arr.shape, ndic.DTYPE_CODES[np.dtype(NP[dt]).name], steps)
    assert back == back_exp == raw


def test_golden_vectors_on_gpu(ndic, golden):
    cases = json.loads((golden / "nd_lift_vectors.json").read_text())["cases"]
    for c in cases:
        steps = c["configuration"]["transforms"]
        t = np.int32 if c["plane"] == "i32" else np.int64
        arr = np.array(c["input"], t).reshape(c["shape"])
        out = ndic.forward(arr, steps)
        assert out.ravel().tolist() == c["expected"], c["name"]
        assert np.array_equal(ndic.inverse(out, steps, t), arr), c["name"]


@pytest.mark.parametrize("kind", ["delta", "haar", "lift53"])
@pytest.mark.parametrize("levels", [1, 2, 5])
@pytest.mark.parametrize("group", [0, 3, 8])
def test_kinds_levels_groups(ndic, oracle, kind, levels, group):
    rng = np.random.default_rng(levels * 100 + group)
    for shape in [(33, 5, 7), (2, 130), (64, 3), (1, 9, 9), (7,)]:
        arr = rng.integers(0, 4096, size=shape).astype(np.uint16)
        steps = [{"dimension": 0, "kind": kind, "levels": levels, "group": group}]
        _roundtrip(ndic, oracle, arr, "u16", steps)


@pytest.mark.parametrize("dt", list(NP))
def test_dtypes_multistep(ndic, oracle, dt):
    rng = np.random.default_rng(7)
    lo, hi = RANGES[dt]
    arr = rng.integers(lo, hi, size=(6, 5, 11), dtype=np.int64).astype(NP[dt])
    steps = [{"dimension": 0, "kind": "lift53", "levels": 2}, {"dimension": 2, "kind": "haar", "levels": 3},
             {"dimension": 1, "kind": "delta", "group": 2}]
    _roundtrip(ndic, oracle, arr, dt, steps)


def test_long_segments_global_path(ndic, oracle):
    rng = np.random.default_rng(9)
    for kind in ["delta", "haar", "lift53"]:
        arr = rng.integers(0, 1000, size=(300, 70)).astype(np.uint16)
        _roundtrip(ndic, oracle, arr, "u16", [{"dimension": 0, "kind": kind, "levels": 4}])
        arr64 = rng.integers(-2**40, 2**40, size=(150, 3)).astype(np.int64)
        _roundtrip(ndic, oracle, arr64, "i64", [{"dimension": 0, "kind": kind, "levels": 4}])


def test_config1_shape(ndic, oracle):
    rng = np.random.default_rng(1)
    arr = rng.integers(0, 65536, size=(1, 8, 32, 64, 64)).astype(np.uint16)
    steps = [{"dimension": 1, "kind": "lift53", "levels": 2}, {"dimension": 2, "kind": "haar", "levels": 2}]
    _roundtrip(ndic, oracle, arr, "u16", steps)


def test_narrow_error_reports_lifted_range(ndic, oracle):
    """The out-of-range message gives the range of the lifted plane that was
    narrowed (crates/ndic-zarr/src/lift.rs:95-101), also when the last step
    narrows straight into the output."""
    rng = np.random.default_rng(4)
    vol = rng.integers(-3000, 3000, size=(6, 9, 70)).astype(np.int16)
    for steps in ([{"dimension": 0, "kind": "lift53", "levels": 2}],
                  [{"dimension": 0, "kind": "delta"}, {"dimension": 2, "kind": "haar", "levels": 2}]):
        coeffs = oracle.lift_chunk(True, vol.tobytes(), vol.shape, "i16", steps)
        lo, hi = int(vol.min()), int(vol.max())
        with pytest.raises(ndic.NdicError, match=rf"coefficient range \[{lo}, {hi}\] does not narrow back to u8"):
            ndic.lift_inverse_bytes(coeffs, vol.shape, "uint8", steps)


def test_refusals_match(ndic, oracle):
    with pytest.raises(ndic.NdicError, match="leave the 32-bit plane") as e:
        ndic.lift_forward_bytes(np.array([-2**31, 2**31 - 1], np.int32).tobytes(), [2], "int32",
                                [{"dimension": 0, "kind": "delta"}])
    assert e.value.kind == "InvalidArgument"
    with pytest.raises(ndic.NdicError, match=r"coefficient range \[70000, 70000\] does not narrow back to u16"):
        ndic.lift_inverse_bytes(np.array([70000], np.int32).tobytes(), [1], "uint16", [])
    with pytest.raises(ndic.NdicError, match="overflow budget"):
        ndic.lift_forward_bytes(np.array([2**32 - 1], np.uint32).tobytes(), [1], "uint32", [])
    with pytest.raises(ndic.NdicError, match="not supported") as e:
        ndic.lift_forward_bytes(b"\0\0", [1], "uint16", [], version="1.0")
    assert e.value.kind == "Unsupported"
    with pytest.raises(ndic.NdicError, match="needs levels >= 1"):
        ndic.lift_forward_bytes(b"\0\0", [1], "uint16", [{"dimension": 0, "kind": "haar", "levels": 0}])
    with pytest.raises(ndic.NdicError, match="out of range"):
        ndic.lift_forward_bytes(b"\0\0", [1], "uint16", [{"dimension": 3, "kind": "delta"}])
    # hostile planes decode to garbage, not an error (wrapping inverse)
    hostile = np.array([-2**31, 2**31 - 1, -2**31 + 1, 2**31 - 2, 0, -1, -2**31, 2**31 - 1], np.int32)
    for kind in ["delta", "haar", "lift53"]:
        steps = [{"dimension": 0, "kind": kind, "levels": 3}]
        assert ndic.lift_inverse_bytes(hostile.tobytes(), [8], "int32", steps) == \
            oracle.lift_chunk(False, hostile.tobytes(), [8], "i32", steps)


def test_empty_chunk(ndic):
    assert ndic.lift_forward_bytes(b"", [0, 4], "uint16", [{"dimension": 0, "kind": "lift53", "levels": 2}]) == b""
