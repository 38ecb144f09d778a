"""The fused finest-DWT-level + inverse nd_lift + narrowing pass against the
oracle and against the unfused passes (``fuse_dwt_lift`` = 0).

Every chunk is encoded by the oracle (nd_lift forward, then htj2k of the
int32 plane: crates/ndic-zarr/src/htj2k.rs:113-441), decoded in one batch with
the fused pass on and off, and must come back bit-exact.  The cases cover
what decides the fused kernel's work: odd, tiny and tile-straddling plane
sizes (the 64 x 32 tiles, degenerate 1-wide / 1-tall planes as in
crates/ndic-htj2k/src/dwt/mod.rs:430-494), lift axes of every length up to
past 64, one and two synthesis levels of haar / 5-3 and delta
(crates/ndic-lift/src/kernel.rs:186-272), group segments, lift axes that are
not the outermost one, every dtype of at most 32 bits, and chunks the fused
pass must leave to the separate passes (three lift levels on long axes,
no DWT levels).  A chunk whose lifted samples do not narrow reports the same
status and message either way.
"""

from __future__ import annotations

import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CODE = {"uint8": "u8", "int8": "i8", "uint16": "u16", "int16": "i16", "uint32": "u32", "int32": "i32"}


def _volume(rng, shape, dt):
    info = np.iinfo(np.dtype(dt))
    lo, hi = max(int(info.min), -(2 ** 20)), min(int(info.max), 2 ** 20)
    z = np.indices(shape).sum(axis=0).astype(np.float64)
    smooth = lo + (hi - lo) * (0.5 + 0.4 * np.sin(z / 7.0))
    noise = rng.integers(-(hi - lo) // 50, (hi - lo) // 50 + 1, size=shape)
    return np.clip(smooth + noise, lo, hi).astype(np.dtype(dt))


def _cases():
    rng = np.random.default_rng(5)
    cases = []
    planes = [(1, 1), (1, 7), (9, 1), (2, 3), (31, 33), (32, 64), (33, 65), (64, 97), (130, 66)]
    zs = [1, 2, 3, 4, 5, 8, 17, 33, 64, 70]
    k = 0
    for kind, levels in [("delta", 1), ("haar", 1), ("haar", 2), ("lift53", 1), ("lift53", 2)]:
        for z in zs:
            y, x = planes[k % len(planes)]
            dt = list(CODE)[k % len(CODE)]
            cases.append(((z, y, x), dt, [{"dimension": 0, "kind": kind, "levels": levels, "group": 0}], 1 + k % 5))
            k += 1
    # group segments, inner lift axes, 4-d chunks
    cases.append(((13, 40, 50), "uint16", [{"dimension": 0, "kind": "lift53", "levels": 2, "group": 5}], 3))
    cases.append(((9, 35, 20), "int16", [{"dimension": 0, "kind": "haar", "levels": 2, "group": 4}], 2))
    cases.append(((2, 6, 33, 40), "uint8", [{"dimension": 1, "kind": "lift53", "levels": 2, "group": 0}], 2))
    cases.append(((3, 5, 17, 70), "int16", [{"dimension": 0, "kind": "delta", "levels": 1, "group": 0}], 4))
    cases.append(((2, 3, 4, 24, 24), "uint16", [{"dimension": 1, "kind": "haar", "levels": 1, "group": 2}], 1))
    # three levels: fused on short axes, separate passes on long ones
    cases.append(((4, 48, 48), "uint16", [{"dimension": 0, "kind": "lift53", "levels": 3, "group": 0}], 3))
    cases.append(((19, 48, 48), "uint16", [{"dimension": 0, "kind": "lift53", "levels": 3, "group": 0}], 3))
    # no DWT level, lift along y, two steps: never fused
    cases.append(((6, 40, 40), "uint16", [{"dimension": 0, "kind": "lift53", "levels": 2, "group": 0}], 0))
    cases.append(((6, 40, 40), "int16", [{"dimension": 1, "kind": "haar", "levels": 2, "group": 0}], 2))
    cases.append(((6, 40, 40), "uint16", [{"dimension": 0, "kind": "lift53", "levels": 2, "group": 0},
                                          {"dimension": 2, "kind": "haar", "levels": 1, "group": 0}], 2))
    out = []
    for shape, dt, steps, lv in cases:
        out.append((_volume(rng, shape, dt), dt, steps, lv))
    return out


def _batch(ndic, oracle, specs, out_dtypes=None):
    dev = importlib.import_module("nd-image-codecs_amd.device")
    chunks = [oracle.lift_ht_encode(vol, CODE[dt], steps, xy_levels=lv) for vol, dt, steps, lv in specs]
    d_in = dev.DeviceBuffer.from_bytes(b"".join(chunks))
    al = lambda n: (n + 255) // 256 * 256
    d_out = dev.DeviceBuffer(sum(al(v.nbytes) for v, *_ in specs) + 256)
    descs, ioff, ooff = [], 0, 0
    for i, ((vol, dt, steps, lv), ch) in enumerate(zip(specs, chunks)):
        odt = out_dtypes[i] if out_dtypes else dt
        descs.append(ndic.make_desc(vol.shape, odt, steps, lv, ioff, len(ch), ooff))
        ioff += len(ch)
        ooff += al(vol.nbytes)
    return dev, d_in, d_out, descs, al


def _decode(ndic, dev, d_in, d_out, descs, fuse):
    ndic.tune_set("fuse_dwt_lift", fuse)
    try:
        arr = (ndic.ChunkDesc * len(descs))(*descs)
        d_out.memset(0)
        err = None
        try:
            ndic.decode_batch(d_in.ptr, d_out.ptr, arr)
        except ndic.NdicError as e:
            err = str(e)
        dev.synchronize()
        return d_out.download(), [d.status for d in arr], err
    finally:
        ndic.tune_reset("fuse_dwt_lift")


def test_fused_matches_oracle_and_separate_passes(ndic, oracle):
    specs = _cases()
    dev, d_in, d_out, descs, al = _batch(ndic, oracle, specs)
    got = {}
    for fuse in (1, 0):
        out, status, err = _decode(ndic, dev, d_in, d_out, descs, fuse)
        assert err is None and not any(status), (fuse, err, status)
        o = 0
        for i, (vol, dt, steps, lv) in enumerate(specs):
            assert out[o:o + vol.nbytes] == vol.tobytes(), (fuse, i, vol.shape, dt, steps, lv)
            o += al(vol.nbytes)
        got[fuse] = out
    assert got[0] == got[1]


def test_fused_range_error_matches_separate_passes(ndic, oracle):
    rng = np.random.default_rng(8)
    good = (_volume(rng, (5, 40, 70), "uint16"), "uint16",
            [{"dimension": 0, "kind": "lift53", "levels": 2, "group": 0}], 3)
    vol = rng.integers(-3000, 3000, size=(6, 33, 47)).astype(np.int16)
    bad = (vol, "int16", [{"dimension": 0, "kind": "haar", "levels": 2, "group": 0}], 2)
    specs = [good, bad]
    dev, d_in, d_out, descs, al = _batch(ndic, oracle, specs, out_dtypes=["uint16", "uint8"])
    res = {fuse: _decode(ndic, dev, d_in, d_out, descs, fuse) for fuse in (1, 0)}
    for fuse, (out, status, err) in res.items():
        assert status[0] == 0 and status[1] != 0, (fuse, status)
        lo, hi = int(vol.min()), int(vol.max())
        assert f"coefficient range [{lo}, {hi}] does not narrow back to u8" in err, (fuse, err)
        assert out[:good[0].nbytes] == good[0].tobytes()
    assert res[1][1] == res[0][1] and res[1][2] == res[0][2]


@pytest.mark.parametrize("window", [60_000, 250_000])
def test_pipelined_windows_match(ndic, oracle, window):
    """Small windows so that the batch spans many: two, three or four windows
    in flight (pipeline_windows = 1, 2, 3) and one window after the other give the same
    bit-exact samples and the same per-chunk statuses, a failing chunk in the
    middle included."""
    rng = np.random.default_rng(9)
    specs = _cases()
    vol = rng.integers(-3000, 3000, size=(6, 33, 47)).astype(np.int16)
    bad_at = len(specs) // 2
    specs.insert(bad_at, (vol, "int16", [{"dimension": 0, "kind": "haar", "levels": 2, "group": 0}], 2))
    out_dt = [dt for _, dt, _, _ in specs]
    out_dt[bad_at] = "uint8"
    dev, d_in, d_out, descs, al = _batch(ndic, oracle, specs, out_dtypes=out_dt)
    ndic.tune_set("window_elems", window)
    try:
        res = {}
        for pipe in (1, 2, 3, 0):
            ndic.tune_set("pipeline_windows", pipe)
            res[pipe] = _decode(ndic, dev, d_in, d_out, descs, 1)
    finally:
        ndic.tune_reset("pipeline_windows")
        ndic.tune_reset("window_elems")
    for pipe, (out, status, err) in res.items():
        assert [i for i, st in enumerate(status) if st] == [bad_at], (pipe, status)
        assert "does not narrow back to u8" in err, (pipe, err)
        o = 0
        for i, (v, dt, steps, lv) in enumerate(specs):
            if i != bad_at:
                assert out[o:o + v.nbytes] == v.tobytes(), (pipe, i, v.shape, dt, steps, lv)
            o += al(v.nbytes)
    assert res[1] == res[0] and res[2] == res[0] and res[3] == res[0]
