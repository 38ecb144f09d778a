"""The block index of variable-size ZFP streams on the device.

The indexed encode writes the same stream bytes as the plain encode (and
the oracle), and beside each variable-size stream an index whose entries are
the oracle's block positions (a walk of the stream's blocks) less the
header; the indexed decode, which finds every block through the index
instead of walking the stream, equals the oracle's decode.  An index that
disagrees with its stream fails that chunk alone (Codestream), one of the
wrong length is refused (InvalidArgument).
"""

from __future__ import annotations

import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TYPES = {"float32": ("f32", np.float32), "float64": ("f64", np.float64), "int32": ("i32", np.int32),
         "int64": ("i64", np.int64), "uint16": ("u16", np.uint16)}
MODES = [{"mode": "reversible"}, {"mode": "fixed_accuracy", "tolerance": 1e-3},
         {"mode": "fixed_precision", "precision": 19}, {"mode": "fixed_rate", "rate": 10.0}]


@pytest.fixture(scope="module")
def dev():
    return importlib.import_module("nd-image-codecs_amd.device")


def field(rng, shape, dt):
    npt = np.dtype(TYPES[dt][1])
    if npt.kind == "f":
        return np.cumsum(rng.standard_normal(shape), axis=-1).astype(npt)
    info = np.iinfo(npt)
    lo, hi = max(info.min, -(2 ** 40)), min(info.max, 2 ** 40)
    return rng.integers(lo, hi, size=shape, endpoint=True).astype(npt)


def zcfg(ndic, cfg):
    return ndic.zfp_config(cfg["mode"], rate=cfg.get("rate"), tolerance=cfg.get("tolerance"),
                           precision=cfg.get("precision"))


def cases(rng):
    out = []
    for shape in [(70,), (9, 13), (6, 10, 11), (3, 5, 6, 7), (16, 16, 16)]:
        for dt in TYPES:
            for cfg in MODES:
                if cfg["mode"] == "fixed_accuracy" and dt in ("int32", "int64", "uint16"):
                    continue
                out.append((field(rng, shape, dt), dt, cfg))
    return out


def encode(ndic, dev, items, indexed=True):
    offs, o = [], 0
    for a, _, _ in items:
        offs.append(o)
        o += (a.nbytes + 255) // 256 * 256
    d_raw = dev.DeviceBuffer(o, zero=True)
    for (a, _, _), off in zip(items, offs):
        d_raw.upload(a.tobytes(), off)
    cap = 3 * o + (1 << 20)
    d_enc = dev.DeviceBuffer(cap, zero=True)
    icap = sum(4 * (int(np.prod([(d + 3) // 4 for d in a.shape])) + 1) + 8 for a, _, _ in items)
    d_idx = dev.DeviceBuffer(icap, zero=True)
    descs = [ndic.make_zfp_desc(a.shape, dt, zcfg(ndic, cfg), off, a.nbytes) for (a, dt, cfg), off in zip(items, offs)]
    arr = ndic.zfp_encode_batch(d_raw.ptr, d_enc.ptr, cap, descs, d_index_ptr=d_idx.ptr if indexed else 0,
                                index_capacity=icap if indexed else 0)
    return d_enc, d_idx, arr


def decode(ndic, dev, items, d_enc, d_idx, arr, descs_fix=None):
    offs, o = [], 0
    for a, _, _ in items:
        offs.append(o)
        o += (a.nbytes + 255) // 256 * 256
    d_out = dev.DeviceBuffer(o)
    d_out.memset(0x77)
    descs = (ndic.ZfpDesc * len(items))(*[
        ndic.make_zfp_desc(a.shape, dt, zcfg(ndic, cfg), e.out_offset, e.out_len, off, e.index_offset, e.index_len)
        for (a, dt, cfg), e, off in zip(items, arr, offs)])
    if descs_fix:
        descs_fix(descs)
    err = ndic.Error()
    import ctypes as C
    rc = ndic.lib().ndic_zfp_decode_batch_indexed(C.c_void_p(d_enc.ptr), C.c_void_p(d_idx.ptr), C.c_void_p(d_out.ptr),
                                                  descs, len(descs), None, C.byref(err))
    out = d_out.download()
    return rc, [int(d.status) for d in descs], [out[off:off + a.nbytes] for (a, _, _), off in zip(items, offs)]


def test_indexed_encode_and_decode_match_oracle(ndic, oracle, dev):
    rng = np.random.default_rng(12)
    items = cases(rng)
    d_enc, d_idx, arr = encode(ndic, dev, items)
    plain_enc, _, parr = encode(ndic, dev, items, indexed=False)
    idx_all = d_idx.download()
    for (a, dt, cfg), e, p in zip(items, arr, parr):
        odt, npt = TYPES[dt]
        stream = d_enc.download(e.out_offset, e.out_len)
        assert stream == plain_enc.download(p.out_offset, p.out_len)
        assert stream == oracle.zfp_encode_chunk(a.tobytes(), a.shape, odt, cfg), (a.shape, dt, cfg)
        if cfg["mode"] == "fixed_rate":
            assert e.index_len == 0
            continue
        nblocks = int(np.prod([(d + 3) // 4 for d in a.shape]))
        assert e.index_len == 4 * (nblocks + 1) and e.index_offset % 8 == 0
        idx = np.frombuffer(idx_all[e.index_offset:e.index_offset + e.index_len], np.uint32)
        sc = np.int32 if npt in (np.uint16,) else npt  # narrow integers ride the int32 stream
        starts = oracle.zfp_block_starts(stream, a.shape, sc, cfg)
        assert np.array_equal(idx.astype(np.uint64), starts - starts[0]), (a.shape, dt, cfg)
    rc, st, outs = decode(ndic, dev, items, d_enc, d_idx, arr)
    assert rc == 0 and not any(st)
    for (a, dt, cfg), o in zip(items, outs):
        odt = TYPES[dt][0]
        e_stream = oracle.zfp_encode_chunk(a.tobytes(), a.shape, odt, cfg)
        assert o == oracle.zfp_decode_chunk(e_stream, a.shape, odt, cfg), (a.shape, dt, cfg)


def test_index_that_disagrees_fails_its_chunk_alone(ndic, dev):
    rng = np.random.default_rng(5)
    items = [(field(rng, (12, 17, 9), "float32"), "float32", {"mode": "reversible"}) for _ in range(5)]
    d_enc, d_idx, arr = encode(ndic, dev, items)
    idx = bytearray(d_idx.download())

    def poke(i, k, delta):
        at = arr[i].index_offset + 4 * k
        v = int.from_bytes(idx[at:at + 4], "little") + delta
        idx[at:at + 4] = (v & 0xFFFFFFFF).to_bytes(4, "little")

    poke(1, 7, 1)            # an inner block offset one bit off
    poke(2, 0, 64)           # the first entry not at the first block
    nb = arr[3].index_len // 4 - 1
    poke(3, nb, 1 << 20)     # the end past the stream
    d_idx.upload(bytes(idx))
    rc, st, outs = decode(ndic, dev, items, d_enc, d_idx, arr)
    assert st == [0, 2, 2, 2, 0] and rc == 2
    for i in (0, 4):
        assert outs[i] == items[i][0].tobytes()  # reversible: lossless

    def short(descs):
        descs[4].index_len -= 4

    rc, st, _ = decode(ndic, dev, items, d_enc, d_idx, arr, short)
    assert st[4] == 1 and st[0] == 0 and rc in (1, 2)


def test_c3_reversible_fields_indexed(ndic, oracle, dev):
    import bench
    rng = np.random.default_rng(4)
    f32 = bench.sine_field(rng, (64, 256, 256)).astype(np.float32)
    f64 = bench.sine_field(rng, (4, 64, 128, 128), drift_axis=0)
    items = [(f32, "float32", {"mode": "reversible"}), (f64, "float64", {"mode": "reversible"})]
    d_enc, d_idx, arr = encode(ndic, dev, items)
    rc, st, outs = decode(ndic, dev, items, d_enc, d_idx, arr)
    assert rc == 0 and st == [0, 0]
    assert outs[0] == f32.tobytes() and outs[1] == f64.tobytes()
