"""The oracle's fixed-rate brick reader (BrickIndex, BrickReader,
decompress_brick, NdZfpBrickDecoder / decode_chunk_brick:
crates/ndic-zfp/src/lib.rs:386-599, crates/ndic-zfp/src/chunk.rs:419-540),
pinned on the CPU by the reference's own invariants:

* the computed stream length is the length compress writes
  (lib.rs:681-689);
* every brick equals the full decode's slice, edge bricks clipped
  (lib.rs:691-721), in any visit order (lib.rs:724-749);
* brick k's bits start at header_bits + k * bits_per_brick and its byte
  range tiles the payload (lib.rs:752-764) -- checked by decoding each brick
  from a stream whose bytes outside the header and the brick's byte range
  are overwritten: a wrong offset reads the garbage and differs;
* the chunk form over narrow integers demotes like the full chunk decode.

The full decode these compare with is pinned by the reference's 104 stream
hashes and the C-library vectors (test_zfp_oracle.py).
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest

ODT = {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64", np.dtype(np.int32): "i32",
       np.dtype(np.int64): "i64"}


def smooth(shape, dtype, seed=0):
    rng = np.random.default_rng(seed)
    idx = np.indices(shape).astype(np.float64)
    v = sum(np.sin(0.37 * (a + 1) * idx[a] + a) for a in range(len(shape))) + rng.normal(0, 1e-3, shape)
    if np.dtype(dtype).kind == "f":
        return v.astype(dtype)
    return np.round(v * 2 ** 20).astype(dtype)


def all_bricks(grid):
    return list(itertools.product(*[range(g) for g in grid]))


@pytest.mark.parametrize("shape", [[16, 16], [9, 10, 11], [4, 4, 4, 4]])
def test_fixed_rate_stream_has_the_computed_length(oracle, shape):
    data = smooth(shape, np.float32)
    stream, hb = oracle.zfp_compress(data, {"mode": "fixed_rate", "rate": 8.0})
    idx = oracle.zfp_brick_index(shape, np.float32, 8.0)
    assert len(stream) == idx["stream_len"] and hb == idx["header_bits"]


CASES = [([9, 10, 11], np.float32, 8.0), ([9, 10, 7], np.float32, 16.0), ([17], np.float64, 12.0),
         ([6, 13], np.int32, 10.0), ([3, 5, 6, 7], np.float64, 8.0), ([5, 9, 4], np.int64, 24.0),
         ([16, 16], np.float64, 8.0), ([2, 3, 5], np.int32, 32.0)]


@pytest.mark.parametrize("shape,dtype,rate", CASES)
def test_every_brick_is_the_full_decode_slice(oracle, shape, dtype, rate):
    data = smooth(shape, dtype, seed=len(shape))
    cfg = {"mode": "fixed_rate", "rate": rate}
    stream, _ = oracle.zfp_compress(data, cfg)
    full = oracle.zfp_decompress(stream, shape, dtype, cfg)
    idx = oracle.zfp_brick_index(shape, dtype, rate)
    assert idx["grid"] == [(d + 3) // 4 for d in shape]
    bricks = all_bricks(idx["grid"])
    for b in reversed(bricks):  # out-of-stream-order seeks
        got, bshape = oracle.zfp_decode_brick(stream, shape, ODT[np.dtype(dtype)], rate, b)
        sl = tuple(slice(4 * c, 4 * c + 4) for c in b)
        assert bshape == list(full[sl].shape), b
        assert got == np.ascontiguousarray(full[sl]).tobytes(), b


@pytest.mark.parametrize("shape,dtype,rate", CASES[:5])
def test_brick_byte_ranges_locate_the_brick(oracle, shape, dtype, rate):
    data = smooth(shape, dtype, seed=7)
    cfg = {"mode": "fixed_rate", "rate": rate}
    stream, _ = oracle.zfp_compress(data, cfg)
    idx = oracle.zfp_brick_index(shape, dtype, rate)
    n = int(np.prod(idx["grid"]))
    first = oracle.zfp_byte_range(idx, 0)
    last = oracle.zfp_byte_range(idx, n - 1)
    assert first[0] == idx["header_bits"] // 8
    assert last[0] + last[1] <= idx["stream_len"]
    head = (idx["header_bits"] + 7) // 8
    for k, b in enumerate(all_bricks(idx["grid"])):
        off, ln = oracle.zfp_byte_range(idx, k)
        assert ln >= idx["bits_per_brick"] // 8
        masked = bytearray(b"\xa5" * len(stream))
        masked[:head] = stream[:head]
        masked[off:off + ln] = stream[off:off + ln]
        want, _ = oracle.zfp_decode_brick(stream, shape, ODT[np.dtype(dtype)], rate, b)
        got, _ = oracle.zfp_decode_brick(bytes(masked), shape, ODT[np.dtype(dtype)], rate, b)
        assert got == want, (k, b)


@pytest.mark.parametrize("dt", ["u8", "i8", "u16", "i16", "i32", "i64", "f32", "f64"])
def test_chunk_bricks_match_the_chunk_decode(oracle, dt):
    np_dt = {"u8": np.uint8, "i8": np.int8, "u16": np.uint16, "i16": np.int16, "i32": np.int32,
             "i64": np.int64, "f32": np.float32, "f64": np.float64}[dt]
    shape = [7, 9, 6]
    rng = np.random.default_rng(3)
    if np.dtype(np_dt).kind == "f":
        vals = np.cumsum(rng.standard_normal(shape), axis=-1).astype(np_dt)
    else:
        info = np.iinfo(np_dt)
        vals = rng.integers(max(info.min, -(2 ** 30)), min(info.max, 2 ** 30), size=shape, endpoint=True).astype(np_dt)
    cfg = {"mode": "fixed_rate", "rate": 12.0}
    stream = oracle.zfp_encode_chunk(vals.tobytes(), shape, dt, cfg)
    full = np.frombuffer(oracle.zfp_decode_chunk(stream, shape, dt, cfg), np_dt).reshape(shape)
    for b in all_bricks([(d + 3) // 4 for d in shape]):
        got, bshape = oracle.zfp_decode_brick(stream, shape, dt, 12.0, b)
        sl = tuple(slice(4 * c, 4 * c + 4) for c in b)
        assert got == np.ascontiguousarray(full[sl]).tobytes() and bshape == list(full[sl].shape), b


def test_brick_reader_refusals(oracle):
    shape = [8, 8]
    data = smooth(shape, np.float32)
    stream, _ = oracle.zfp_compress(data, {"mode": "fixed_rate", "rate": 8.0})
    with pytest.raises(oracle.OracleError):  # out of the grid
        oracle.zfp_decode_brick(stream, shape, "f32", 8.0, [2, 0])
    with pytest.raises(oracle.OracleError):  # length differs from the computed size
        oracle.zfp_decode_brick(stream + bytes(8), shape, "f32", 8.0, [0, 0])
    with pytest.raises(oracle.OracleError):  # another rate: the mode word differs
        oracle.zfp_decode_brick(stream, shape, "f32", 16.0, [0, 0])
    with pytest.raises(oracle.OracleError):  # another field: the header differs
        oracle.zfp_decode_brick(stream, [8, 8], "f64", 8.0, [0, 0])
    bad = bytearray(stream)
    bad[0] ^= 0xFF
    with pytest.raises(oracle.OracleError):  # magic
        oracle.zfp_decode_brick(bytes(bad), shape, "f32", 8.0, [0, 0])


@pytest.mark.parametrize("shape,dtype,rate", CASES[:4])
def test_block_starts_walk(oracle, shape, dtype, rate):
    """The oracle's block walk (what a block index records): at fixed rate it
    reproduces the computed addressing header_bits + k * bits_per_brick; in
    every mode it starts at the header's end, increases, and ends where the
    stream's 64-bit padding begins."""
    data = smooth(shape, dtype, seed=5)
    idx = oracle.zfp_brick_index(shape, dtype, rate)
    stream, hb = oracle.zfp_compress(data, {"mode": "fixed_rate", "rate": rate})
    starts = oracle.zfp_block_starts(stream, shape, dtype, {"mode": "fixed_rate", "rate": rate})
    n = int(np.prod(idx["grid"]))
    assert list(starts) == [hb + k * idx["bits_per_brick"] for k in range(n + 1)]
    for cfg in ({"mode": "reversible"}, {"mode": "fixed_precision", "precision": 17}):
        stream, hb = oracle.zfp_compress(data, cfg)
        starts = oracle.zfp_block_starts(stream, shape, dtype, cfg)
        assert starts[0] == hb and len(starts) == n + 1
        assert np.all(np.diff(starts.astype(np.int64)) > 0)
        assert len(stream) == (int(starts[-1]) + 63) // 64 * 8
