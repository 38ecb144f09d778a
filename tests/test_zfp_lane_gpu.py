"""The ZFP block coders on the GPU at sizes where many waves run at once:
one lane per block (1D-3D fields) and four lanes per block (4D fields).

Streams must equal the oracle's byte for byte and decode to the oracle's
samples: every dtype and mode on shapes with partial blocks, fields large
enough that waves decode from their LDS stage and streams whose blocks are too
large for it (they read global memory), walked variable-size streams, bricks,
and 4D fields of thousands of blocks encoded repeatedly (a stream that differs
between two encodes of the same field fails).
"""

from __future__ import annotations

import numpy as np
import pytest


pytestmark = pytest.mark.gpu

NP = {"u8": np.uint8, "i8": np.int8, "u16": np.uint16, "i16": np.int16, "i32": np.int32, "i64": np.int64,
      "f32": np.float32, "f64": np.float64}
LONG = {k: np.dtype(v).name for k, v in NP.items()}
MODES = [{"mode": "reversible"}, {"mode": "fixed_rate", "rate": 8.0}, {"mode": "fixed_rate", "rate": 16.0},
         {"mode": "fixed_rate", "rate": 3.25}, {"mode": "fixed_precision", "precision": 14},
         {"mode": "fixed_accuracy", "tolerance": 1e-4}]


def kw(cfg: dict) -> dict:
    return {k: v for k, v in cfg.items()}


def field(rng, shape, dtype, kind="smooth"):
    npt = NP[dtype]
    if np.dtype(npt).kind == "f":
        if kind == "noise":
            return rng.standard_normal(shape).astype(npt) * npt(1e3)
        a = np.cumsum(rng.standard_normal(shape), axis=-1).astype(npt) * npt(2.5)
        return a
    info = np.iinfo(npt)
    lo, hi = max(info.min, -(2**40)), min(info.max, 2**40)
    if kind == "noise":
        return rng.integers(lo, hi, size=shape, endpoint=True).astype(npt)
    base = np.cumsum(rng.integers(-40, 41, size=shape), axis=-1)
    return np.clip(base + (lo + hi) // 2, lo, hi).astype(npt)


def check(ndic, oracle, a, dtype, cfg, encodes=1):
    raw = a.tobytes()
    exp = oracle.zfp_encode_chunk(raw, a.shape, dtype, cfg)
    for _ in range(encodes):
        got = ndic.nd_zfp_encode(raw, a.shape, LONG[dtype], **kw(cfg))
        assert got == exp, (a.shape, dtype, cfg)
    out = ndic.nd_zfp_decode(exp, a.shape, LONG[dtype], **kw(cfg))
    assert out == oracle.zfp_decode_chunk(exp, a.shape, dtype, cfg), (a.shape, dtype, cfg)


@pytest.mark.parametrize("dtype", list(NP))
def test_small_fields_match_oracle(ndic, oracle, dtype):
    rng = np.random.default_rng(sum(map(ord, dtype)))
    for shape in ([6], [70], [3, 5], [9, 21], [2, 3, 5], [7, 9, 19], [2, 3, 5, 6], [5, 2, 9, 7]):
        for kind in ("smooth", "noise"):
            a = field(rng, shape, dtype, kind)
            for cfg in MODES:
                check(ndic, oracle, a, dtype, cfg)


@pytest.mark.parametrize("dtype", ["f32", "f64", "i32", "i64", "u16"])
def test_large_fields_staged_and_unstaged(ndic, oracle, dtype):
    """Fields of a few thousand blocks: fixed-rate waves decode from their
    LDS stage; noisy 64-bit reversible blocks are too large for it."""
    rng = np.random.default_rng(7)
    for shape, kind in (((13, 45, 130), "smooth"), ((9, 31, 66), "noise")):
        a = field(rng, shape, dtype, kind)
        for cfg in ({"mode": "fixed_rate", "rate": 8.0}, {"mode": "fixed_rate", "rate": 21.5}, {"mode": "reversible"},
                    {"mode": "fixed_precision", "precision": 20}):
            check(ndic, oracle, a, dtype, cfg, encodes=2)


@pytest.mark.parametrize("dtype", ["i32", "u16", "f32", "f64", "i64"])
def test_large_4d_fields_repeatable(ndic, oracle, dtype):
    rng = np.random.default_rng(13)
    for shape, kind in (((3, 13, 45, 66), "smooth"), ((4, 16, 32, 64), "noise")):
        a = field(rng, shape, dtype, kind)
        for cfg in ({"mode": "fixed_rate", "rate": 8.0}, {"mode": "reversible"}):
            check(ndic, oracle, a, dtype, cfg, encodes=3)


def test_bricks_match_oracle(ndic, oracle):
    """Every brick of four fixed-rate fields in random order against the
    oracle's brick reader (the batch form over several streams runs in the
    configuration and allocation-free tests)."""
    rng = np.random.default_rng(9)
    shapes = [(9, 10, 11), (5, 17, 8), (12, 4, 9), (6, 13)]
    cfg = {"mode": "fixed_rate", "rate": 12.0}
    fields = [field(rng, s, "f32") for s in shapes]
    streams = [oracle.zfp_encode_chunk(a.tobytes(), a.shape, "f32", cfg) for a in fields]
    want = []
    for i, a in enumerate(fields):
        grid = [(d + 3) // 4 for d in a.shape]
        want += [(i, [int(c) for c in np.unravel_index(k, grid)]) for k in range(int(np.prod(grid)))]
    order = rng.permutation(len(want))
    for j in order:
        i, br = want[j]
        got = ndic.nd_zfp_decode_brick(streams[i], fields[i].shape, "float32", br, rate=12.0)
        exp_bytes, exp_shape = oracle.zfp_decode_brick(streams[i], fields[i].shape, "f32", 12.0, br)
        assert got[0] == exp_bytes and got[1] == exp_shape, (i, br)


def test_reference_bench_fixture(ndic, oracle):
    """The reference's ZFP chunk benchmark fixture (f32 32x64x64): the device
    streams have the lengths the reference's baseline records give (347,456
    bytes lossless, 131,088 at rate 8), equal the oracle's, and decode --
    the lossless one by walking the stream -- to the oracle's samples."""
    import datagen

    a = datagen.zfp_bench_fixture()
    for cfg, length in (({"mode": "reversible"}, 347456), ({"mode": "fixed_rate", "rate": 8.0}, 131088)):
        got = ndic.nd_zfp_encode(a.tobytes(), a.shape, "float32", **kw(cfg))
        assert len(got) == length, cfg
        check(ndic, oracle, a, "f32", cfg)
    s = ndic.nd_zfp_encode(a.tobytes(), a.shape, "float32", mode="reversible")
    assert ndic.nd_zfp_decode(s, a.shape, "float32", mode="reversible") == a.tobytes()


@pytest.mark.parametrize("dims", [1, 2, 3, 4])
def test_walk_of_random_streams_matches_oracle(ndic, oracle, dims):
    """Streams whose blocks are random bits (ones at densities 1/2, 1/10 and
    1/50) behind a valid header, decoded by walking them: the samples equal
    the oracle's decode of the same bits.  Sparse ones make long tokens, zero
    runs and late last coefficients, dense ones short tokens and full
    budgets -- the cases the walk's ring turns, last-coefficient turns and
    general turns divide between them."""
    rng = np.random.default_rng(40 + dims)
    shape = {1: [300], 2: [20, 24], 3: [8, 12, 16], 4: [4, 8, 8, 8]}[dims]
    nblocks = int(np.prod([(n + 3) // 4 for n in shape]))
    modes = [{"mode": "reversible"}, {"mode": "fixed_precision", "precision": 20},
             {"mode": "fixed_accuracy", "tolerance": 1e-3}]
    for dtype in ("f32", "f64", "i32", "i64"):
        a = np.zeros(shape, NP[dtype])
        for cfg in modes:
            if cfg["mode"] == "fixed_accuracy" and dtype[0] == "i":
                continue
            head = oracle.zfp_encode_chunk(a.tobytes(), shape, dtype, cfg)[:12]  # 96 header bits
            # the longest stream the configuration allows (zfp_stream_maximum_size,
            # crates/ndic-zfp/src/lib.rs:380-384): longer ones are refused
            prec = 32 if dtype in ("f32", "i32") else 64
            values = 4 ** dims
            mb = 1 + {"f32": 8, "f64": 11}.get(dtype, 0) + values - 1
            mb += (5 if prec == 32 else 6) if cfg["mode"] == "reversible" and dtype[0] == "f" else 0
            mb += values * min(cfg.get("precision", 64), prec)
            size = ((148 + nblocks * min(mb, 16658) + 63) // 64) * 8
            for density in (0.5, 0.1, 0.02):
                bits = rng.random((size - 12) * 8) < density
                s = head + np.packbits(bits, bitorder="little").tobytes()
                got = ndic.nd_zfp_decode(s, shape, LONG[dtype], **kw(cfg))
                assert got == oracle.zfp_decode_chunk(s, shape, dtype, cfg), (dtype, cfg, density)


@pytest.mark.parametrize("dims", [1, 2, 3, 4])
def test_fixed_rate_random_streams_match_oracle(ndic, oracle, dims):
    """Fixed-rate streams whose blocks are random bits (the header and the
    length of a real stream): raw-bit blocks, lowest exponents, negative
    zeros and full plane budgets decode as the oracle decodes them."""
    rng = np.random.default_rng(50 + dims)
    shape = {1: [300], 2: [20, 24], 3: [8, 12, 16], 4: [4, 8, 8, 8]}[dims]
    for dtype in ("f32", "f64", "i32", "i64"):
        a = np.zeros(shape, NP[dtype])
        for rate in (3.25, 8.0, 16.0, 31.0):
            cfg = {"mode": "fixed_rate", "rate": rate}
            real = oracle.zfp_encode_chunk(a.tobytes(), shape, dtype, cfg)
            for density in (0.5, 0.1):
                # keep 19 bytes: the 96- or 148-bit header (a rate above 2,048
                # bits per block takes the long mode word)
                bits = rng.random((len(real) - 19) * 8) < density
                s = real[:19] + np.packbits(bits, bitorder="little").tobytes()
                got = ndic.nd_zfp_decode(s, shape, LONG[dtype], **kw(cfg))
                assert got == oracle.zfp_decode_chunk(s, shape, dtype, cfg), (dtype, rate, density)
