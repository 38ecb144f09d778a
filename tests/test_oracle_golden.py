"""The oracle against the reference's own golden data (CPU only).

Pins: fixtures/nd-lift/vectors.json (crates/ndic-lift/tests/vectors.rs:20),
the kernel closed forms (crates/ndic-lift/src/kernel.rs:386-432), the derived
HT tables (crates/ndic-htj2k/src/block/tables.rs:116-122), the byte-stable
codestream fixtures (crates/ndic-codestream/tests/fixtures.rs:25-73,
crates/ndic-zarr/tests/htj2k_zarrs.rs:373-398) and the 5/3 DWT closed forms
(crates/ndic-htj2k/src/dwt/mod.rs:430-494).
"""

from __future__ import annotations

import json

import numpy as np
import pytest


def test_ht_tables_match_reference(oracle, golden):
    g = json.loads((golden / "ht_tables.json").read_text())
    t = oracle.ht_tables()
    for name, vals in g.items():
        assert np.array_equal(np.array(vals, np.uint16), t[name]), name


def test_nd_lift_vectors(oracle, golden):
    cases = json.loads((golden / "nd_lift_vectors.json").read_text())["cases"]
    assert len(cases) == 13
    for c in cases:
        steps = c["configuration"]["transforms"]
        if c["plane"] == "i32":
            raw = oracle.lift_chunk(True, np.array(c["input"], np.int32).tobytes(), c["shape"], "i32", steps)
            got = np.frombuffer(raw, np.int32).tolist()
            back = oracle.lift_chunk(False, raw, c["shape"], "i32", steps)
            assert np.frombuffer(back, np.int32).tolist() == c["input"], c["name"]
        else:
            got = oracle.lift_plane_i64(True, np.array(c["input"], np.int64), c["shape"], steps).tolist()
        assert got == c["expected"], c["name"]


@pytest.mark.parametrize("kind,levels,inp,exp", [
    ("haar", 1, [2, 7, -2, -7], [4, -5, 5, -5]),
    ("lift53", 1, [0, 4, 0, 0], [2, 1, 4, 0]),
    ("lift53", 1, list(range(8)), [0, 2, 4, 6, 0, 0, 0, 1]),
    ("lift53", 2, [5] * 7, [5, 5, 0, 0, 0, 0, 0]),
    ("haar", 1, [9] * 6, [9, 9, 9, 0, 0, 0]),
    ("delta", 0, [3, 4, 5, 6, 7], [3, 1, 1, 1, 1]),
])
def test_lift_closed_forms(oracle, kind, levels, inp, exp):
    steps = [{"dimension": 0, "kind": kind, "levels": levels}]
    raw = oracle.lift_chunk(True, np.array(inp, np.int32).tobytes(), [len(inp)], "i32", steps)
    assert np.frombuffer(raw, np.int32).tolist() == exp


def test_lift_budget_refusals(oracle):
    steps = [{"dimension": 0, "kind": "delta"}]
    with pytest.raises(oracle.OracleError, match="leave the 32-bit plane"):
        oracle.lift_chunk(True, np.array([-2**31, 2**31 - 1], np.int32).tobytes(), [2], "i32", steps)
    with pytest.raises(oracle.OracleError, match="does not narrow back to u16"):
        oracle.lift_chunk(False, np.array([70000], np.int32).tobytes(), [1], "u16", [])
    with pytest.raises(oracle.OracleError, match="not supported"):
        oracle.lift_chunk(True, b"\0\0", [1], "u16", [], version="0.2")


def test_lift_delta_volume(oracle):
    # crates/ndic-lift/src/chunk.rs:320-330: [2,2,3] delta along z
    raw = oracle.lift_chunk(True, np.arange(12, dtype=np.int32).tobytes(), [2, 2, 3], "i32",
                            [{"dimension": 0, "kind": "delta"}])
    assert np.frombuffer(raw, np.int32).tolist() == [0, 1, 2, 3, 4, 5, 6, 6, 6, 6, 6, 6]


def test_dwt_closed_forms(oracle):
    # DC plane: all energy lands in LL; impulse round trip; odd/degenerate shapes
    dc = np.full((16, 16), 7, np.int32)
    f = oracle.dwt53(True, dc, 2)
    assert np.all(f[:4, :4] == 7) and np.count_nonzero(f) == 16
    rng = np.random.default_rng(3)
    for (h, w) in [(1, 1), (1, 7), (7, 1), (2, 2), (33, 65), (5, 3)]:
        for lv in range(0, 6):
            p = rng.integers(-5000, 5000, size=(h, w)).astype(np.int32)
            assert np.array_equal(oracle.dwt53(False, oracle.dwt53(True, p, lv), lv), p)


def test_dwt_matches_textbook_1d(oracle):
    # one level on a single row is T.800 F-9 / F-10 with symmetric extension
    x = np.array([[3, 9, -4, 7, 1, 0, 5]], np.int32)
    n = x.shape[1]
    ext = lambda k: x[0, k if k < n else 2 * (n - 1) - k]
    d = [int(ext(2 * i + 1)) - ((int(ext(2 * i)) + int(ext(2 * i + 2))) >> 1) for i in range(n // 2)]
    dd = lambda i: d[min(max(i, 0), len(d) - 1)] if i >= 0 else d[0]
    s = [int(x[0, 2 * i]) + ((dd(i - 1) + dd(i) + 2) >> 2) for i in range((n + 1) // 2)]
    assert oracle.dwt53(True, x, 1)[0].tolist() == s + d


def test_codestream_fixtures_byte_stable(oracle, golden):
    j2c = (golden / "tiny-gradient-8x8.j2c").read_bytes()
    img = np.array([(7 * (i % 8) + 13 * (i // 8)) % 256 for i in range(64)], np.int32).reshape(8, 8)
    assert oracle.encode_image(img, 8, False, 1) == j2c
    assert np.array_equal(oracle.decode_codestream(j2c), img)
    # the .jph wraps the same codestream in boxes: its jp2c payload must decode too
    jph = (golden / "tiny-gradient-8x8.jph").read_bytes()
    k = jph.find(b"jp2c")
    assert np.array_equal(oracle.decode_codestream(jph[k + 4:]), img)


def test_ndht_fixture_byte_stable(oracle, golden):
    nd = (golden / "tiny-chunk-4x8x8.ndht").read_bytes()
    vals = np.array([(i * 7) % 4096 for i in range(256)], np.uint16)
    assert oracle.htj2k_encode_chunk(vals.tobytes(), (4, 8, 8), "u16", xy_levels=2) == nd
    assert oracle.htj2k_decode_chunk(nd, (4, 8, 8), "u16") == vals.tobytes()
    # index table of fixtures/codestreams/tiny-chunk-4x8x8.ndht.md
    import struct
    entries = [struct.unpack_from("<QI3I", nd, 20 + 24 * p) for p in range(4)]
    assert entries == [(116, 200, 161, 177, 198), (316, 200, 161, 177, 198), (516, 201, 162, 178, 199),
                       (717, 202, 163, 179, 200)]


def test_ht_block_roundtrip_random(oracle):
    # crates/ndic-htj2k/tests/block_roundtrip.rs:110-140 as an invariant
    rng = np.random.default_rng(11)
    for _ in range(60):
        w, h = int(rng.integers(1, 65)), int(rng.integers(1, 65))
        kmax = int(rng.integers(2, 29))
        dens = rng.uniform(0, 1)
        mag = rng.integers(0, 2 ** kmax, size=(h, w)).astype(np.int64)
        mag[rng.uniform(size=(h, w)) > dens] = 0
        if mag.max() == 0:
            mag[0, 0] = 1
        sign = rng.integers(0, 2, size=(h, w)).astype(np.uint32) << 31
        shift = 31 - kmax
        sm = (sign | (mag.astype(np.uint32) << shift)).astype(np.uint32)
        sm[mag == 0] = 0
        seg = oracle.ht_encode_block(sm, kmax - 1)
        dec = oracle.ht_decode_block(seg, w, h, kmax - 1)
        got = ((dec & 0x7FFFFFFF) >> shift).astype(np.int64)
        assert np.array_equal(got, mag)
        assert np.array_equal(dec[mag > 0] >> 31, sm[mag > 0] >> 31)
