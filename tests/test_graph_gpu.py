"""Decode windows captured into HIP graphs (graph_windows = 1).

A batch spanning several windows is decoded eagerly once, then through
graphs four times in a row; between the graph runs the same process decodes
another batch eagerly and synchronises.  Every graph run must equal the
eager run byte for byte (the first instantiates the cached graphs, the later
ones update them in place with the new windows' parameters).  Plane-selective
decodes take the same path and are checked the same way.
"""

from __future__ import annotations

import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CODE = {"uint8": "u8", "uint16": "u16", "int16": "i16"}


def _batch(ndic, oracle, dev, seed, n):
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(n):
        shape = (int(rng.integers(1, 12)), int(rng.integers(8, 90)), int(rng.integers(8, 90)))
        dt = ["uint8", "uint16", "int16"][k % 3]
        info = np.iinfo(np.dtype(dt))
        vol = rng.integers(max(info.min, -3000), min(info.max, 3000), size=shape).astype(dt)
        steps = [{"dimension": 0, "kind": ["delta", "haar", "lift53"][k % 3], "levels": 2, "group": 0}]
        lv = int(rng.integers(0, 5))
        specs.append((vol, dt, steps, lv, oracle.lift_ht_encode(vol, CODE[dt], steps, xy_levels=lv)))
    d_in = dev.DeviceBuffer.from_bytes(b"".join(s[4] for s in specs))
    al = lambda x: (x + 255) // 256 * 256
    d_out = dev.DeviceBuffer(sum(al(s[0].nbytes) for s in specs) + 256, zero=True)
    descs, ioff, ooff = [], 0, 0
    for vol, dt, steps, lv, ch in specs:
        descs.append(ndic.make_desc(vol.shape, dt, steps, lv, ioff, len(ch), ooff))
        ioff += len(ch)
        ooff += al(vol.nbytes)
    return specs, d_in, d_out, descs, al


def _decode(ndic, dev, d_in, d_out, descs):
    d_out.memset(0)
    ndic.decode_batch(d_in.ptr, d_out.ptr, descs)
    dev.synchronize()
    return d_out.download()


@pytest.fixture
def knobs(ndic):
    yield ndic
    for k in ("graph_windows", "window_elems", "pipeline_windows"):
        ndic.tune_reset(k)


@pytest.mark.parametrize("pipeline", [0, 1, 3])
def test_graph_windows_replay_equal_eager(knobs, oracle, pipeline):
    ndic = knobs
    dev = importlib.import_module("nd-image-codecs_amd.device")
    specs, d_in, d_out, descs, al = _batch(ndic, oracle, dev, 3, 18)
    other = _batch(ndic, oracle, dev, 4, 6)
    ndic.tune_set("window_elems", 40_000)
    ndic.tune_set("pipeline_windows", pipeline)
    ndic.tune_set("graph_windows", 0)
    eager = _decode(ndic, dev, d_in, d_out, descs)
    o = 0
    for vol, *_ in specs:
        assert eager[o:o + vol.nbytes] == vol.tobytes()
        o += al(vol.nbytes)
    for rep in range(4):
        ndic.tune_set("graph_windows", 1)
        got = _decode(ndic, dev, d_in, d_out, descs)
        assert got == eager, rep
        ndic.tune_set("graph_windows", 0)
        _decode(ndic, dev, *other[1:4])  # eager work between the replays


def test_graph_select_decodes_equal_eager(knobs, oracle):
    ndic = knobs
    dev = importlib.import_module("nd-image-codecs_amd.device")
    rng = np.random.default_rng(8)
    chunks, descs_in = [], []
    for k in range(5):
        vol = rng.integers(0, 4000, size=(6 + k, 40, 33)).astype(np.uint16)
        steps = [{"dimension": 0, "kind": "haar", "levels": 1, "group": 0}]
        chunks.append(oracle.lift_ht_encode(vol, "u16", steps, xy_levels=3))
        descs_in.append((vol.shape, steps))
    d_in = dev.DeviceBuffer.from_bytes(b"".join(chunks))
    offs = np.cumsum([0] + [len(c) for c in chunks])[:-1]

    def descs():
        ds = [ndic.make_select_desc(sh, "int32", 3, offs[k], len(chunks[k]), max_px=10, transforms=st)
              for k, (sh, st) in enumerate(descs_in)]
        ds += [ndic.make_select_desc(sh, "int32", 3, offs[k], len(chunks[k]), subset=[(1, 2), (3, 20), (0, 33)])
               for k, (sh, st) in enumerate(descs_in)]
        arr = ndic.decode_select_batch(d_in.ptr, 0, ds)
        o = 0
        for d in arr:
            d.out_offset = o
            o += (d.out_bytes + 255) // 256 * 256
        return arr, o

    arr, total = descs()
    d_out = dev.DeviceBuffer(total + 256, zero=True)
    ndic.tune_set("graph_windows", 0)
    ndic.decode_select_batch(d_in.ptr, d_out.ptr, arr)
    dev.synchronize()
    eager = d_out.download()
    for rep in range(3):
        ndic.tune_set("graph_windows", 1)
        d_out.memset(0)
        ndic.decode_select_batch(d_in.ptr, d_out.ptr, descs()[0])
        dev.synchronize()
        assert d_out.download() == eager, rep
