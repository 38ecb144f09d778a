"""The device path at the benchmark configurations' own shapes, both
directions against the oracle:

* the 32x256x256 u16 chunk (nd_lift 5/3 along z, 5 DWT levels): the whole
  chain encoded on the device byte-identical to the oracle's chunk, decoded
  back bit-exact, batched and one chunk per launch sequence;
* the 8x1x32x512x512 u16 5D series, transposed to c,t,z,y,x
  (crates/ndic-zarr/src/series.rs:418-442), nd_lift 5/3 along t and haar
  along z, 5 DWT levels: the same checks;
* a sample of the uneven batch drawn with the benchmark's own recipe
  (including a 64x1024x1024 chunk, an all-zero chunk and a uniform-random
  chunk): every chunk's device bytes equal the oracle's, the device decode
  equals the oracle's decode and the original samples;
* the ZFP fields f32 64x256x256 and f64 4x64x128x128 at fixed rate 8 and 16
  and reversible: device streams byte-identical to the oracle's, decodes
  equal;
* 256 independent ZFP fields of 16^3..64^3 samples in one batch, and 600
  bricks of them read through the brick batch, against the oracle's chunk
  decode and brick reader.
"""

from __future__ import annotations

import importlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import bench
import datagen

pytestmark = pytest.mark.gpu

SHORT = {"uint8": "u8", "uint16": "u16", "int16": "i16", "float32": "f32", "float64": "f64"}
POOL = 8  # host threads for the oracle (its ctypes calls release the GIL)


@pytest.fixture(scope="module")
def dev():
    return importlib.import_module("nd-image-codecs_amd.device")


def al(n, a=256):
    return (n + a - 1) // a * a


def device_chain(ndic, dev, vols, specs, plain=False):
    """Encodes the chunks on the device in one batch, then decodes that
    arena; returns (chunk bytes per item, decoded bytes per item)."""
    offs, o = [], 0
    for v in vols:
        offs.append(o)
        o += al(v.nbytes)
    d_raw = dev.DeviceBuffer(max(o, 1), zero=True)
    for v, off in zip(vols, offs):
        d_raw.upload(v.tobytes(), off)
    cap = 3 * o + (64 << 20)
    d_enc = dev.DeviceBuffer(cap, zero=True)
    edescs = (ndic.ChunkDesc * len(vols))(*[
        ndic.make_desc(v.shape, dt, steps, lv, offs[i], v.nbytes, 0) for i, (v, (dt, steps, lv)) in
        enumerate(zip(vols, specs))])
    ndic.encode_batch(d_raw.ptr, d_enc.ptr, cap, edescs, plain=plain)
    chunks = [d_enc.download(d.out_offset, d.out_len) for d in edescs]
    d_out = dev.DeviceBuffer(max(o, 1))
    d_out.memset(0xA5)
    ddescs = [ndic.make_desc(v.shape, dt, steps, lv, edescs[i].out_offset, edescs[i].out_len, offs[i])
              for i, (v, (dt, steps, lv)) in enumerate(zip(vols, specs))]
    ndic.decode_batch(d_enc.ptr, d_out.ptr, ddescs, plain=plain)
    return chunks, [d_out.download(off, v.nbytes) for v, off in zip(vols, offs)]


def test_c0_chain(ndic, oracle, dev):
    vol = datagen.c0_volume((32, 256, 256), seed=0)
    steps = [{"dimension": 0, "kind": "lift53", "levels": 2, "group": 0}]
    exp = oracle.lift_ht_encode(vol, "u16", steps, xy_levels=5)
    for plain in (False, True):
        chunks, decs = device_chain(ndic, dev, [vol], [("uint16", steps, 5)], plain=plain)
        assert chunks[0] == exp, plain
        assert decs[0] == vol.tobytes(), plain
    # the single-chunk codec calls (host buffers) give the same bytes
    coeffs = ndic.lift_forward_bytes(vol, vol.shape, "uint16", steps)
    assert ndic.htj2k_encode(coeffs, vol.shape, "int32", xy_levels=5) == exp


def test_c1_series_chunk(ndic, oracle, dev):
    series = datagen.series_5d((8, 1, 32, 512, 512), seed=1)
    vol = np.ascontiguousarray(series.transpose(1, 0, 2, 3, 4))  # c, t, z, y, x
    steps = [{"dimension": 1, "kind": "lift53", "levels": 2, "group": 0},
             {"dimension": 2, "kind": "haar", "levels": 2, "group": 0}]
    assert (vol == 65535).mean() > 0.009 and vol.max() == 65535
    exp = oracle.lift_ht_encode(vol, "u16", steps, xy_levels=5)
    chunks, decs = device_chain(ndic, dev, [vol], [("uint16", steps, 5)])
    assert chunks[0] == exp
    assert decs[0] == vol.tobytes()
    # one chunk per launch sequence gives the same bytes and samples
    chunks, decs = device_chain(ndic, dev, [vol], [("uint16", steps, 5)], plain=True)
    assert chunks[0] == exp and decs[0] == vol.tobytes()


def test_c2_sample_both_directions(ndic, oracle, dev):
    rng = np.random.default_rng(2024)
    specs = bench.draw_specs(dev, rng, "c2", 36)
    shape0, dt0, steps0, lv0, _, blobs0, seed0 = specs[0]
    specs[0] = ((64, 1024, 1024), "uint16", [{"dimension": 0, "kind": "lift53", "levels": 2, "group": 0}], 5,
                dev.SYNTH_BLOBS, [(32.0, 500.0, 300.0, 2400.0, 19.0)] + list(blobs0[1:]), seed0)
    specs[1] = ((7, 97, 311), "int16", steps0, 4, dev.SYNTH_ZERO, blobs0, 11)
    specs[2] = ((5, 333, 129), "uint8", steps0, 3, dev.SYNTH_RANDOM, blobs0, 12)
    specs[3] = ((9, 200, 64), "uint16", [{"dimension": 0, "kind": "delta", "levels": 2, "group": 0}], 0,
                dev.SYNTH_RANDOM, blobs0, 13)
    vols = []
    for shape, dt, steps, lv, content, blobs, seed in specs:
        n = int(np.prod(shape)) * np.dtype(dt).itemsize
        buf = dev.DeviceBuffer(al(n), zero=True)
        dev.synth_chunk(buf.ptr, shape, dt, content, blobs, seed)
        vols.append(np.frombuffer(buf.download(0, n), dt).reshape(shape))
    assert not vols[1].any() and vols[2].std() > 50
    chunks, decs = device_chain(ndic, dev, vols, [(s[1], s[2], s[3]) for s in specs])

    def check(i):
        shape, dt, steps, lv = specs[i][:4]
        exp = oracle.lift_ht_encode(vols[i], SHORT[dt], steps, xy_levels=lv)
        back = oracle.lift_ht_decode(exp, shape, SHORT[dt], steps)
        return chunks[i] == exp, back.tobytes() == vols[i].tobytes(), decs[i] == vols[i].tobytes()

    with ThreadPoolExecutor(POOL) as ex:
        res = list(ex.map(check, range(len(specs))))
    bad = [(i, specs[i][:4], r) for i, r in enumerate(res) if not all(r)]
    assert not bad, bad


ZFP_MODES = [("fixed_rate", {"rate": 8.0}), ("fixed_rate", {"rate": 16.0}), ("reversible", {})]


@pytest.mark.parametrize("which", ["f32 64x256x256", "f64 4x64x128x128"])
def test_c3_fields(ndic, oracle, dev, which):
    rng = np.random.default_rng(4)
    f32 = bench.sine_field(rng, (64, 256, 256)).astype(np.float32)
    fld = f32 if which.startswith("f32") else bench.sine_field(rng, (4, 64, 128, 128), drift_axis=0)
    odt = SHORT[fld.dtype.name]
    d_in = dev.DeviceBuffer.from_bytes(fld.tobytes())
    cap = 3 * fld.nbytes + (1 << 20)
    d_enc = dev.DeviceBuffer(cap, zero=True)
    d_out = dev.DeviceBuffer(fld.nbytes)
    for mode, extra in ZFP_MODES:
        cfg = {"mode": mode, **extra}
        exp = oracle.zfp_encode_chunk(fld.tobytes(), fld.shape, odt, cfg)
        zc = ndic.zfp_config(mode, rate=extra.get("rate"))
        arr = ndic.zfp_encode_batch(d_in.ptr, d_enc.ptr, cap, [ndic.make_zfp_desc(fld.shape, fld.dtype.name, zc, 0,
                                                                                  fld.nbytes)])
        assert d_enc.download(arr[0].out_offset, arr[0].out_len) == exp, mode
        d_out.memset(0x5A)
        ndic.zfp_decode_batch(d_enc.ptr, d_out.ptr, [ndic.make_zfp_desc(fld.shape, fld.dtype.name, zc,
                                                                        arr[0].out_offset, arr[0].out_len, 0)])
        want = oracle.zfp_decode_chunk(exp, fld.shape, odt, cfg)
        assert d_out.download(0, fld.nbytes) == want, mode
        if mode == "reversible":
            assert want == fld.tobytes()


def test_zfp_field_batch_and_bricks(ndic, oracle, dev):
    serve = importlib.import_module("nd-image-codecs_amd.serve")
    rng = np.random.default_rng(8)
    base = bench.sine_field(rng, (64, 128, 128)).astype(np.float32)
    fields = []
    for i in range(256):
        shape = tuple(int(v) for v in np.exp(rng.uniform(np.log(16), np.log(64.99), 3)).astype(int))
        z0, y0, x0 = (int(rng.integers(0, b - s + 1)) for b, s in zip(base.shape, shape))
        f = np.ascontiguousarray(base[z0:z0 + shape[0], y0:y0 + shape[1], x0:x0 + shape[2]])
        if i % 4 == 3:
            f = f.astype(np.float64)
        fields.append((f, 8.0 if i % 2 else 16.0))
    offs, o = [], 0
    for f, _ in fields:
        offs.append(o)
        o += al(f.nbytes)
    d_raw = dev.DeviceBuffer(o, zero=True)
    for (f, _), off in zip(fields, offs):
        d_raw.upload(f.tobytes(), off)
    cap = o + 256 * 4096
    d_enc = dev.DeviceBuffer(cap, zero=True)
    cfgs = [ndic.zfp_config("fixed_rate", rate=r) for _, r in fields]
    arr = ndic.zfp_encode_batch(d_raw.ptr, d_enc.ptr, cap, [
        ndic.make_zfp_desc(f.shape, f.dtype.name, c, off, f.nbytes) for (f, _), c, off in zip(fields, cfgs, offs)])
    d_out = dev.DeviceBuffer(o)
    ndic.zfp_decode_batch(d_enc.ptr, d_out.ptr, [
        ndic.make_zfp_desc(f.shape, f.dtype.name, c, a.out_offset, a.out_len, off)
        for (f, _), c, a, off in zip(fields, cfgs, arr, offs)])
    out = d_out.download()
    streams = [d_enc.download(a.out_offset, a.out_len) for a in arr]

    def check(i):
        (f, r), off = fields[i], offs[i]
        cfg = {"mode": "fixed_rate", "rate": r}
        exp = oracle.zfp_encode_chunk(f.tobytes(), f.shape, SHORT[f.dtype.name], cfg)
        want = oracle.zfp_decode_chunk(exp, f.shape, SHORT[f.dtype.name], cfg)
        return streams[i] == exp, out[off:off + f.nbytes] == want

    with ThreadPoolExecutor(POOL) as ex:
        res = list(ex.map(check, range(len(fields))))
    assert all(all(r) for r in res), [i for i, r in enumerate(res) if not all(r)][:10]
    # bricks of those streams through the brick batch, against the oracle brick reader
    n = 600
    bdesc = (serve.BrickDesc * n)()
    picks, boff = [], 0
    for k in range(n):
        i = int(rng.integers(0, len(fields)))
        f, r = fields[i]
        b = [int(rng.integers(0, (d + 3) // 4)) for d in f.shape]
        d = bdesc[k]
        d.in_offset, d.in_len, d.out_offset, d.ndim = arr[i].out_offset, arr[i].out_len, boff, 3
        for a in range(3):
            d.shape[a], d.brick[a] = f.shape[a], b[a]
        d.dtype = ndic._zfp_dtype(f.dtype.name)
        d.rate = r
        picks.append((i, b, boff))
        boff += 64 * f.dtype.itemsize
    d_b = dev.DeviceBuffer(boff, zero=True)
    serve._lib()
    err = ndic.Error()
    import ctypes as C
    rc = ndic.lib().ndic_zfp_decode_bricks(C.c_void_p(d_enc.ptr), C.c_void_p(d_b.ptr), bdesc, n, None, C.byref(err))
    ndic._check(rc, err)
    got = d_b.download()
    for i, b, off in picks:
        f, r = fields[i]
        exp, bshape = oracle.zfp_decode_brick(streams[i], f.shape, SHORT[f.dtype.name], r, b)
        assert got[off:off + len(exp)] == exp, (i, b)
