"""The serving loop: every response equals the same request served alone and
the oracle's answer; Poisson arrivals are all served with latency reported."""

from __future__ import annotations

import importlib

import numpy as np
import pytest

import datagen

pytestmark = pytest.mark.gpu

STEPS = {"u8": "uint8", "u16": "uint16", "i16": "int16"}


@pytest.fixture(scope="module")
def world(ndic, oracle):
    serve = importlib.import_module("nd-image-codecs_amd.serve")
    dev = importlib.import_module("nd-image-codecs_amd.device")
    rng = np.random.default_rng(5)
    specs, vols = [], []
    for i in range(12):
        z = int(rng.integers(1, 6))
        y, x = int(rng.integers(20, 150)), int(rng.integers(20, 150))
        dt = ["uint8", "uint16", "int16"][i % 3]
        info = np.iinfo(dt)
        vol = datagen.blobs_volume((z, y, x), n_blobs=3, base=info.max / 8, amp=(info.max / 10, info.max / 3),
                                   lo=info.min, hi=info.max, seed=40 + i, dtype=np.dtype(dt))
        steps = [{"dimension": 0, "kind": ["delta", "haar", "lift53"][i % 3], "levels": 2, "group": 0}]
        lv = [0, 3, 4, 5][i % 4]
        specs.append(((z, y, x), dt, steps, lv))
        vols.append(vol)
    al = lambda n: (n + 15) // 16 * 16
    raw = b"".join(v.tobytes() + bytes(al(v.nbytes) - v.nbytes) for v in vols)
    d_raw = dev.DeviceBuffer.from_bytes(raw)
    descs, off, raw_offs = [], 0, []
    for (shape, dt, steps, lv), v in zip(specs, vols):
        descs.append(ndic.make_desc(shape, dt, steps, lv, off, v.nbytes, 0))
        raw_offs.append(off)
        off += al(v.nbytes)
    cap = 3 * len(raw) + (1 << 20)
    arena = dev.DeviceBuffer(cap, zero=True)
    arr = (ndic.ChunkDesc * len(descs))(*descs)
    ndic.encode_batch(d_raw.ptr, arena.ptr, cap, arr)
    store = serve.Store()
    store.add_htj2k_arena(arena, arr, specs, d_raw, raw_offs)
    # ZFP fields at fixed rate, streams at 8-byte aligned offsets
    fields, blob = [], b""
    for k, (shape, dt, rate) in enumerate([((9, 10, 11), "float32", 8.0), ((3, 5, 6, 7), "float64", 16.0),
                                           ((13, 17), "int32", 12.0)]):
        a = np.cumsum(rng.standard_normal(shape), axis=-1).astype(dt) if dt.startswith("float") else \
            rng.integers(-(2**20), 2**20, size=shape).astype(dt)
        s = ndic.nd_zfp_encode(a.tobytes(), shape, dt, mode="fixed_rate", rate=rate)
        fields.append(serve.ZfpField(len(blob), len(s), shape, dt, rate))
        blob += s
    store.add_zfp_arena(dev.DeviceBuffer.from_bytes(blob), fields)
    zfp_blob = blob
    return serve, dev, store, vols, zfp_blob


def make_requests(serve, store, n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        k = i % 3
        if k == 0:
            reqs.append(serve.Request(serve.THUMB, 0.0, key=int(rng.integers(0, len(store.ht))), res=3))
        elif k == 1:
            reqs.append(serve.Request(serve.DECODE, 0.0, key=int(rng.integers(0, len(store.ht)))))
        else:
            key = int(rng.integers(0, len(store.zfp)))
            f = store.zfp[key]
            reqs.append(serve.Request(serve.BRICK, 0.0, key=key,
                                      brick=tuple(int(rng.integers(0, (d + 3) // 4)) for d in f.shape)))
    return reqs


def test_mix_equals_alone_and_oracle(world, ndic, oracle):
    serve, dev, store, vols, zfp_blob = world
    server = serve.Server(store, out_budget=64 << 20)
    reqs = make_requests(serve, store, 30, 1)
    results = {}

    def keep(srv, batch):
        for r in batch:
            results[id(r)] = srv.result(r)

    stats = server.run(reqs, verify=keep)
    assert stats.passes >= 1 and sum(len(v) for v in stats.latencies_ms.values()) == len(reqs)
    mixed = [results[id(r)] for r in reqs]
    for r, got in zip(reqs, mixed):
        alone = serve.Request(r.kind, 0.0, key=r.key, brick=r.brick, res=r.res)
        server.run([alone])
        assert server.result(alone) == got, (r.kind, r.key)
        c = store.ht[r.key] if r.kind != serve.BRICK else None
        if r.kind == serve.DECODE:
            assert got == vols[r.key].tobytes()
        elif r.kind == serve.THUMB:
            chunk = store.ht_arena.download(c.offset, c.length)
            off, ln, pre = c.plane0
            res = min(r.res, c.levels)
            exp = oracle.decode_codestream(chunk[off:off + pre[res]], res, prefix=True)
            assert got == exp.tobytes(), (r.key, res)
        else:
            f = store.zfp[r.key]
            s = zfp_blob[f.offset:f.offset + f.length]
            full = np.frombuffer(ndic.nd_zfp_decode(s, f.shape, f.dtype, mode="fixed_rate", rate=f.rate),
                                 f.dtype).reshape(f.shape)
            sl = tuple(slice(4 * b, 4 * b + 4) for b in r.brick)
            assert got == np.ascontiguousarray(full[sl]).tobytes()


def test_poisson_arrivals_all_served(world):
    serve, dev, store, vols, _ = world
    rng = np.random.default_rng(2)
    reqs = serve.poisson_requests(rng, rate=400.0, duration=0.5, store=store, n_encode_sources=0)
    server = serve.Server(store, out_budget=32 << 20)
    stats = server.run(reqs)
    summ = stats.summary()
    assert summ["requests"] == len(reqs)
    assert summ["p50_ms"] is not None and summ["p99_ms"] >= summ["p50_ms"]
    assert all(r.done >= r.arrival for r in reqs)


def test_admission_splits_passes_by_memory(world):
    serve, dev, store, vols, _ = world
    big = max(range(len(store.ht)), key=lambda k: int(np.prod(store.ht[k].shape)) * 2)
    server = serve.Server(store, out_budget=int(np.prod(store.ht[big].shape)) * 2 + 4096)
    reqs = [serve.Request(serve.DECODE, 0.0, key=big) for _ in range(3)]
    outs = []
    stats = server.run(reqs, verify=lambda srv, batch: outs.extend(srv.result(r) for r in batch))
    assert stats.passes == 3
    assert outs == [vols[big].tobytes()] * 3
