"""The ZFP oracle against the reference's ZFP pins (CPU only).

Pins:
- fixtures/zfp/checksums.json: 104 stream hashes over dims x dtype x mode
  (crates/ndic-zfp/tests/checksums.rs:21-60, inputs from
  crates/ndic-zfp/tests/common/mod.rs:17-101);
- fixtures/zfp/tiny-chunk-4x8x8-rate8.zfp byte identity
  (crates/ndic-zfp/tests/checksums.rs:62-86);
- 384 stream and decode hashes written by the LLNL zfp 0.5.5 C library
  (tests/golden/make_zfp_vectors.py), which reproduces the 104 reference
  hashes too, plus that library's measured 4D coefficient order;
- closed forms: fixed-rate stream length (crates/ndic-zfp/src/lib.rs:380-384),
  reversible identity, error bound of fixed accuracy, promotion/demotion of
  narrow integers (crates/ndic-zfp/src/chunk.rs:309-356);
- the stream lengths the reference's ZFP chunk benchmarks record for their
  f32 32x64x64 fixture: 347,456 bytes lossless, 131,088 bytes at rate 8
  (bench/baselines/main/zfp-reversible/zfp__chunk_encode_f32_zyx_32x64x64.json:32,
  bench/baselines/main/zfp-rate8/zfp__chunk_encode_f32_zyx_32x64x64.json:32).
"""

from __future__ import annotations

import json
import math

import numpy as np
import pytest

from zfp_fixtures import SHORT, ZFP_EXTRA_TYPES, case_values, fnv1a64, ramp_chunk, tiny_chunk_f32


def _checksum_cases(golden):
    return json.loads((golden / "zfp_checksums.json").read_text())["cases"]


def test_reference_checksums(oracle, golden):
    cases = _checksum_cases(golden)
    assert len(cases) == 104
    for c in cases:
        n = int(np.prod(c["shape"]))
        out = oracle.zfp_encode_chunk(ramp_chunk(n, c["dtype"]), c["shape"], SHORT[c["dtype"]], c["configuration"])
        assert len(out) == c["bytes"], c
        assert "0x%016x" % fnv1a64(out) == c["fnv1a64"], c


def test_reference_checksum_streams_decode(oracle, golden):
    for c in _checksum_cases(golden):
        n = int(np.prod(c["shape"]))
        raw = ramp_chunk(n, c["dtype"])
        dt = SHORT[c["dtype"]]
        stream = oracle.zfp_encode_chunk(raw, c["shape"], dt, c["configuration"])
        back = oracle.zfp_decode_chunk(stream, c["shape"], dt, c["configuration"])
        assert len(back) == len(raw)
        if c["configuration"]["mode"] == "reversible":
            assert back == raw, c


def test_tiny_chunk_fixture(oracle, golden):
    fixture = (golden / "tiny-chunk-4x8x8-rate8.zfp").read_bytes()
    cfg = {"mode": "fixed_rate", "rate": 8.0}
    raw = tiny_chunk_f32()
    assert oracle.zfp_encode_chunk(raw, [4, 8, 8], "f32", cfg) == fixture
    back = np.frombuffer(oracle.zfp_decode_chunk(fixture, [4, 8, 8], "f32", cfg), np.float32)
    src = np.frombuffer(raw, np.float32)
    assert np.max(np.abs(back - src)) < 16.0


def test_libzfp_vectors(oracle, golden):
    doc = json.loads((golden / "zfp_libzfp_vectors.json").read_text())
    assert oracle.zfp_perm(4) == doc["perm4"]
    assert len(doc["cases"]) == 384
    for c in doc["cases"]:
        dt = ZFP_EXTRA_TYPES[c["type"]]
        a = case_values(c["kind"], dt, c["shape"], c["seed"])
        s, _ = oracle.zfp_compress(a, c["mode"])
        assert len(s) == c["bytes"], c["name"]
        assert "0x%016x" % fnv1a64(s) == c["stream_fnv1a64"], c["name"]
        b = oracle.zfp_decompress(s, c["shape"], dt, c["mode"])
        assert "0x%016x" % fnv1a64(b.tobytes()) == c["decoded_fnv1a64"], c["name"]


def test_coefficient_orders_are_permutations_by_sequency(oracle):
    for d in (1, 2, 3, 4):
        p = oracle.zfp_perm(d)
        assert sorted(p) == list(range(4 ** d))
        seq = [sum((v >> (2 * a)) & 3 for a in range(d)) for v in p]
        assert seq == sorted(seq)
        sq = [sum(((v >> (2 * a)) & 3) ** 2 for a in range(d)) for v in p]
        for i in range(1, len(p)):
            if seq[i] == seq[i - 1]:
                assert sq[i] >= sq[i - 1]


@pytest.mark.parametrize("rate", [1.0, 2.5, 8.0, 16.0, 31.0])
@pytest.mark.parametrize("tname", ["int32", "int64", "float32", "float64"])
def test_fixed_rate_length_closed_form(oracle, rate, tname):
    dt = ZFP_EXTRA_TYPES[tname]
    for shape in ([13], [7, 9], [5, 6, 7], [3, 5, 4, 6]):
        d = len(shape)
        a = case_values("wild", dt, shape, 7)
        s, hb = oracle.zfp_compress(a, {"mode": "fixed_rate", "rate": rate})
        bits = math.floor((4 ** d) * rate + 0.5)
        if tname == "float32":
            bits = max(bits, 9)
        if tname == "float64":
            bits = max(bits, 12)
        assert hb == 96 + (52 if bits > 2048 else 0)
        blocks = int(np.prod([(n + 3) // 4 for n in shape]))
        assert len(s) == (hb + blocks * bits + 63) // 64 * 8


@pytest.mark.parametrize("dtype", ["u8", "i8", "u16", "i16", "i32", "i64", "f32", "f64"])
def test_reversible_identity(oracle, dtype):
    np_t = {"u8": np.uint8, "i8": np.int8, "u16": np.uint16, "i16": np.int16, "i32": np.int32, "i64": np.int64,
            "f32": np.float32, "f64": np.float64}[dtype]
    rng = np.random.default_rng(5)
    for shape in ([1], [5], [4, 4], [3, 9, 2], [2, 3, 5, 7]):
        if np.dtype(np_t).kind == "f":
            a = (rng.standard_normal(shape) * np.exp(rng.uniform(-30, 30, shape))).astype(np_t)
        else:
            info = np.iinfo(np_t)
            a = rng.integers(info.min, info.max, size=shape, endpoint=True, dtype=np_t)
        cfg = {"mode": "reversible"}
        s = oracle.zfp_encode_chunk(a.tobytes(), shape, dtype, cfg)
        assert oracle.zfp_decode_chunk(s, shape, dtype, cfg) == a.tobytes()


def test_fixed_accuracy_bound(oracle):
    rng = np.random.default_rng(9)
    a = np.cumsum(rng.standard_normal((6, 9, 10)), axis=2)
    for tol in (1e-3, 0.1, 2.0):
        cfg = {"mode": "fixed_accuracy", "tolerance": tol}
        s, _ = oracle.zfp_compress(a, cfg)
        b = oracle.zfp_decompress(s, a.shape, np.float64, cfg)
        assert np.max(np.abs(a - b)) <= tol


def test_narrow_promotion_closed_form(oracle):
    # u8 v -> (v - 0x80) << 23: a reversible stream of the promoted int32
    # field equals the chunk-level stream of the u8 chunk
    v = np.arange(40, dtype=np.uint8).reshape(5, 8) * 6
    cfg = {"mode": "reversible"}
    promoted = ((v.astype(np.int32) - 0x80) << 23).astype(np.int32)
    s, _ = oracle.zfp_compress(promoted, cfg)
    assert oracle.zfp_encode_chunk(v.tobytes(), [5, 8], "u8", cfg) == s
    i16 = (np.arange(40, dtype=np.int16) * 801 - 16000).reshape(5, 8)
    s, _ = oracle.zfp_compress((i16.astype(np.int32) << 15).astype(np.int32), cfg)
    assert oracle.zfp_encode_chunk(i16.tobytes(), [5, 8], "i16", cfg) == s


def test_legacy_dims_squeeze(oracle):
    # chunk.rs:610-633: singleton axes are squeezed, then left-padded to dims
    a = np.arange(256, dtype=np.float32).reshape(4, 1, 8, 8)
    cfg = {"mode": "fixed_rate", "rate": 8.0, "dims": 3}
    s1 = oracle.zfp_encode_chunk(a.tobytes(), [4, 1, 8, 8], "f32", cfg)
    s2 = oracle.zfp_encode_chunk(a.tobytes(), [4, 8, 8], "f32", cfg)
    assert s1 == s2
    b = np.arange(64, dtype=np.float32)
    s3 = oracle.zfp_encode_chunk(b.tobytes(), [64], "f32", {"mode": "reversible", "dims": 2})
    s4, _ = oracle.zfp_compress(b.reshape(1, 64), {"mode": "reversible"})
    assert s3 == s4


def test_decode_rejects_mismatched_streams(oracle):
    a = np.linspace(-1, 1, 64, dtype=np.float32).reshape(4, 16)
    cfg = {"mode": "fixed_rate", "rate": 8.0}
    s, _ = oracle.zfp_compress(a, cfg)
    with pytest.raises(oracle.OracleError):
        oracle.zfp_decompress(b"xfp" + s[3:], a.shape, np.float32, cfg)
    with pytest.raises(oracle.OracleError):
        oracle.zfp_decompress(s, (16, 4), np.float32, cfg)
    with pytest.raises(oracle.OracleError):
        oracle.zfp_decompress(s, a.shape, np.float64, cfg)
    with pytest.raises(oracle.OracleError):
        oracle.zfp_decompress(s, a.shape, np.float32, {"mode": "fixed_rate", "rate": 4.0})
    with pytest.raises(oracle.OracleError):
        oracle.zfp_decompress(s + bytes(8), a.shape, np.float32, cfg)


def test_reference_bench_fixture_stream_lengths(oracle):
    """The reference's ZFP chunk benchmarks encode datagen.zfp_bench_fixture()
    (bench/rs/ndic-bench-cli/src/workloads/zfp.rs:58-81) to 347,456 bytes
    losslessly and to 131,088 bytes at 8 bits per value (their baseline
    records' bytes_out); a lossless stream of that exact length leaves little
    room for a wrong plane, group test or header.  The lossless stream decodes
    to the fixture."""
    import datagen

    a = datagen.zfp_bench_fixture()
    assert a.nbytes == 524288
    s = oracle.zfp_encode_chunk(a.tobytes(), a.shape, "f32", {"mode": "reversible"})
    assert len(s) == 347456
    assert oracle.zfp_decode_chunk(s, a.shape, "f32", {"mode": "reversible"}) == a.tobytes()
    r = oracle.zfp_encode_chunk(a.tobytes(), a.shape, "f32", {"mode": "fixed_rate", "rate": 8.0})
    assert len(r) == 131088
