"""Plane-selective decodes of device-resident chunks (ndic_htj2k_decode_select_batch).

Partial decodes (crates/ndic-zarr/src/htj2k_codec.rs:319-411) must equal the
matching slice of the oracle's full decode, for the sub-chunk ranges of
crates/ndic-zarr/tests/htj2k_zarrs.rs:155-189 and random ones, every dtype,
with and without the coefficient-plane index.  thumbnail_3d decodes
(crates/ndic-codestream/src/range.rs:308-369) must return, for the planes
the oracle planner selects, what the oracle's codestream decoder gives for
each of them at the planned resolution.
"""

from __future__ import annotations

import importlib

import numpy as np
import pytest

import range_oracle as ro

pytestmark = pytest.mark.gpu

NPT = {"u8": np.uint8, "i8": np.int8, "u16": np.uint16, "i16": np.int16, "u32": np.uint32, "i32": np.int32}
NAME = {"u8": "uint8", "i8": "int8", "u16": "uint16", "i16": "int16", "u32": "uint32", "i32": "int32"}


@pytest.fixture(scope="module")
def dev():
    return importlib.import_module("nd-image-codecs_amd.device")


def run(ndic, dev, chunks, descs_of):
    """Uploads the chunks, plans (d_out 0), decodes, returns (descs, outputs)."""
    blob = b"".join(chunks)
    d_in = dev.DeviceBuffer.from_bytes(blob)
    offs = np.cumsum([0] + [len(c) for c in chunks])[:-1]
    descs = descs_of(offs)
    arr = ndic.decode_select_batch(d_in.ptr, 0, descs)
    al = lambda n: (n + 255) // 256 * 256
    out_off, total = [], 0
    for d in arr:
        out_off.append(total)
        d.out_offset = total
        total += al(d.out_bytes)
    d_out = dev.DeviceBuffer(total + 256, zero=True)
    ndic.decode_select_batch(d_in.ptr, d_out.ptr, arr)
    dev.synchronize()
    raw = d_out.download()
    return arr, [raw[o:o + d.out_bytes] for o, d in zip(out_off, arr)]


def volume(rng, shape, dt):
    info = np.iinfo(NPT[dt])
    lo, hi = max(int(info.min), -(2 ** 20)), min(int(info.max), 2 ** 20)
    return rng.integers(lo, hi, size=shape, endpoint=True).astype(NPT[dt])


def test_partial_decode_reference_ranges(ndic, oracle, dev):
    """The sub-chunk reads of htj2k_zarrs.rs:155-189 on an (8, 2, 8, 8) u16 chunk."""
    i = np.arange(8 * 2 * 8 * 8)
    data = ((np.abs(((i * 37) % 101) - 50)) * 400).astype(np.uint16).reshape(8, 2, 8, 8)
    chunk = oracle.htj2k_encode_chunk(data.tobytes(), data.shape, "u16", xy_levels=2)
    ranges = [[(3, 1), (1, 1), (5, 1), (6, 1)], [(2, 1), (0, 1), (4, 1), (0, 8)], [(2, 4), (0, 2), (1, 2), (1, 2)]]
    arr, outs = run(ndic, dev, [chunk] * len(ranges), lambda offs: [
        ndic.make_select_desc(data.shape, "uint16", 2, offs[k], len(chunk), subset=r) for k, r in enumerate(ranges)])
    for r, d, got in zip(ranges, arr, outs):
        sl = tuple(slice(a, a + n) for a, n in r)
        assert list(d.out_shape[:d.out_ndim]) == [n for _, n in r]
        assert got == data[sl].tobytes(), r


@pytest.mark.parametrize("index", [True, False])
def test_partial_decode_random_subsets(ndic, oracle, dev, index):
    rng = np.random.default_rng(31 + index)
    chunks, specs = [], []
    for k, (shape, dt, lv) in enumerate([((3, 33, 61), "u8", 3), ((2, 3, 40, 20), "i16", 2), ((5, 17, 9), "u16", 1),
                                         ((4, 64, 70), "i32", 4), ((2, 2, 2, 31, 7), "i8", 2), ((6, 1, 9), "u32", 0)]):
        arr = volume(rng, shape, dt)
        ch = oracle.htj2k_encode_chunk(arr.tobytes(), shape, dt, xy_levels=lv, index=index)
        for _ in range(3):
            start = [int(rng.integers(0, s)) for s in shape]
            count = [int(rng.integers(1, s - a + 1)) for a, s in zip(start, shape)]
            chunks.append(ch)
            specs.append((arr, dt, lv, list(zip(start, count))))
    arrs, outs = run(ndic, dev, chunks, lambda offs: [
        ndic.make_select_desc(a.shape, NAME[dt], lv, offs[k], len(chunks[k]), subset=sub)
        for k, (a, dt, lv, sub) in enumerate(specs)])
    for (a, dt, lv, sub), got in zip(specs, outs):
        sl = tuple(slice(s, s + n) for s, n in sub)
        assert got == a[sl].tobytes(), (a.shape, dt, lv, sub)


def test_partial_decode_of_lift_ht_coefficients(ndic, oracle, dev):
    """Under nd_lift the htj2k codec holds int32 coefficients: the subset is
    of those (the lift's own partial decoder asks for whole chunks)."""
    rng = np.random.default_rng(5)
    vol = volume(rng, (9, 40, 36), "u16")
    steps = [{"dimension": 0, "kind": "lift53", "levels": 2, "group": 0}]
    chunk = oracle.lift_ht_encode(vol, "u16", steps, xy_levels=3)
    coeffs = np.frombuffer(oracle.htj2k_decode_chunk(chunk, vol.shape, "i32"), np.int32).reshape(vol.shape)
    sub = [(1, 5), (3, 30), (2, 20)]
    _, outs = run(ndic, dev, [chunk], lambda offs: [
        ndic.make_select_desc(vol.shape, "int32", 3, offs[0], len(chunk), subset=sub)])
    assert outs[0] == coeffs[1:6, 3:33, 2:22].tobytes()


def test_thumbnail_3d_matches_oracle(ndic, oracle, dev):
    rng = np.random.default_rng(9)
    cases = []
    for shape, steps, lv, max_px in [
            ((16, 64, 80), [{"dimension": 0, "kind": "lift53", "levels": 2, "group": 0}], 3, 20),
            ((9, 33, 47), [{"dimension": 0, "kind": "haar", "levels": 1, "group": 4}], 2, 64),
            ((10, 50, 50), [{"dimension": 0, "kind": "delta", "levels": 1, "group": 4}], 4, 8),
            ((3, 4, 40, 30), [{"dimension": 1, "kind": "lift53", "levels": 1, "group": 0}], 2, 15),
            ((2, 6, 24, 24), [{"dimension": 0, "kind": "haar", "levels": 1, "group": 0},
                              {"dimension": 1, "kind": "lift53", "levels": 2, "group": 3}], 3, 3),
            ((5, 20, 20), [], 2, 1000)]:
        vol = volume(rng, shape, "u16")
        chunk = oracle.lift_ht_encode(vol, "u16", steps, xy_levels=lv)
        cases.append((chunk, shape, steps, lv, max_px))
    arr, outs = run(ndic, dev, [c[0] for c in cases], lambda offs: [
        ndic.make_select_desc(shape, "int32", lv, offs[k], len(ch), max_px=mp, transforms=steps)
        for k, (ch, shape, steps, lv, mp) in enumerate(cases)])
    for (chunk, shape, steps, lv, max_px), d, got in zip(cases, arr, outs):
        plan = ro.plan(chunk, "thumbnail_3d", max_px, steps)
        assert d.max_res == plan["max_res"]
        assert list(d.out_shape[:d.out_ndim]) == plan["decoded_size"], (shape, steps, max_px)
        exp = b"".join(oracle.decode_codestream(ro.plane_codestream(chunk, p), plan["max_res"]).astype(np.int32)
                       .tobytes() for p in plan["planes"])
        assert got == exp, (shape, steps, max_px)


def test_thumbnail_3d_without_index_and_prefix_only(ndic, oracle, dev):
    """An index-less chunk decodes the same planes (offsets walked on the
    device); the plan-only call gives the sizes without decoding."""
    rng = np.random.default_rng(2)
    vol = volume(rng, (8, 30, 30), "i16")
    steps = [{"dimension": 0, "kind": "haar", "levels": 2, "group": 0}]
    coeffs = oracle.lift_chunk(True, vol.tobytes(), vol.shape, "i16", steps)
    with_idx = oracle.htj2k_encode_chunk(coeffs, vol.shape, "i32", xy_levels=2)
    no_idx = oracle.htj2k_encode_chunk(coeffs, vol.shape, "i32", xy_levels=2, index=False)
    _, outs = run(ndic, dev, [with_idx, no_idx], lambda offs: [
        ndic.make_select_desc(vol.shape, "int32", 2, offs[k], len(c), max_px=8, transforms=steps)
        for k, c in enumerate([with_idx, no_idx])])
    assert outs[0] == outs[1] and len(outs[0]) == 2 * 8 * 8 * 4


def test_select_errors(ndic, oracle, dev):
    rng = np.random.default_rng(4)
    vol = volume(rng, (3, 20, 20), "i16")
    chunk = oracle.htj2k_encode_chunk(vol.tobytes(), vol.shape, "i16", xy_levels=2)
    d_in = dev.DeviceBuffer.from_bytes(chunk)
    with pytest.raises(ndic.NdicError, match="subset exceeds the chunk"):
        ndic.decode_select_batch(d_in.ptr, 0, [ndic.make_select_desc(vol.shape, "int16", 2, 0, len(chunk),
                                                                     subset=[(0, 1), (15, 6), (0, 4)])])
    with pytest.raises(ndic.NdicError, match="not a leading chunk axis"):
        ndic.decode_select_batch(d_in.ptr, 0, [ndic.make_select_desc(
            vol.shape, "int32", 2, 0, len(chunk), max_px=8,
            transforms=[{"dimension": 1, "kind": "haar", "levels": 1, "group": 0}])])
    # a window that does not fit the requested dtype: the error names the window's range
    d_out = dev.DeviceBuffer(4096, zero=True)
    sub = [(1, 1), (2, 5), (3, 6)]
    win = vol[1:2, 2:7, 3:9]
    arr = (ndic.SelectDesc * 1)(ndic.make_select_desc(vol.shape, "uint8", 2, 0, len(chunk), subset=sub))
    with pytest.raises(ndic.NdicError) as e:
        ndic.decode_select_batch(d_in.ptr, d_out.ptr, arr)
    assert arr[0].status != 0
    assert f"[{int(win.min())}, {int(win.max())}]" in str(e.value), str(e.value)
