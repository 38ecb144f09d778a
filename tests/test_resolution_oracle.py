"""The oracle's reduced-resolution decode, pinned on the CPU by the
reference's own invariants (so the GPU thumbnail tests, which compare with
it, rest on a checked oracle):

* crates/ndic-codestream/tests/codestream_roundtrip.rs:103-128
  (partial_decode_matches_wavelet_pyramid): decode_to_resolution(r) equals
  the forward 5/3 DWT of the level-shifted samples over all L levels, cut to
  the LL region of L - r levels and inverted over r levels only, shifted
  back and clamped to the sample range;
* crates/ndic-codestream/tests/range_plans.rs:33-64
  (prefix_decode_equals_downsampled_full_decode): the codestream cut after
  the last resolution-r packet decodes, as a prefix, to the same samples as
  the whole codestream decoded to resolution r;
* the ndht index's prefix[r] is exactly that cut (container.rs:199-340: the
  largest end of a packet of resolution <= r).

The forward and inverse DWT used here are the oracle's, pinned by the T.800
closed forms in test_oracle_golden.py; a wrong inverse level count, a wrong
LL region, a missing shift or clamp, or a cut at the wrong packet fails these.
"""

from __future__ import annotations

import struct

import numpy as np
import pytest

import cs_tools


def level_dims(w, h, levels):
    for _ in range(levels):
        w, h = (w + 1) // 2, (h + 1) // 2
    return w, h


def pyramid(oracle, samples, depth, signed, levels, r):
    """The reference pyramid of codestream_roundtrip.rs:112-125."""
    h, w = samples.shape
    half = 0 if signed else 1 << (depth - 1)
    x = samples.astype(np.int32) - half
    f = oracle.dwt53(True, x, levels)
    ew, eh = level_dims(w, h, levels - r)
    region = np.ascontiguousarray(f[:eh, :ew])
    out = oracle.dwt53(False, region, r)
    if signed:
        return out
    return np.clip(out + half, 0, (1 << depth) - 1)


CASES = [
    # (width, height, levels, depth, signed)
    (120, 68, 3, 8, False),   # the reference's own case
    (97, 311, 5, 12, False),
    (33, 7, 4, 16, False),
    (9, 1, 2, 8, False),
    (130, 65, 5, 15, True),
    (1, 1, 0, 8, False),
    (64, 64, 6, 10, True),
]


@pytest.mark.parametrize("w,h,levels,depth,signed", CASES)
def test_partial_decode_matches_wavelet_pyramid(oracle, w, h, levels, depth, signed):
    rng = np.random.default_rng(w * 1000 + h)
    lo, hi = (-(1 << (depth - 1)), (1 << (depth - 1)) - 1) if signed else (0, (1 << depth) - 1)
    yy, xx = np.mgrid[0:h, 0:w]
    smooth = (lo + hi) / 2 + (hi - lo) / 3 * np.sin(xx / 7.0) * np.cos(yy / 5.0)
    samples = np.clip(np.round(smooth + rng.normal(0, (hi - lo) / 40, (h, w))), lo, hi).astype(np.int32)
    stream = oracle.encode_image(samples, depth, signed, levels)
    assert np.array_equal(oracle.decode_codestream(stream), samples)
    for r in range(levels + 1):
        dec = oracle.decode_codestream(stream, r)
        assert dec.shape == level_dims(w, h, levels - r)[::-1], r
        assert np.array_equal(dec, pyramid(oracle, samples, depth, signed, levels, r)), r
    # a resolution past the levels is the full decode
    assert np.array_equal(oracle.decode_codestream(stream, levels + 3), samples)


def test_prefix_decode_equals_downsampled_full_decode(oracle):
    rng = np.random.default_rng(4)
    samples = rng.integers(0, 256, size=(121, 190)).astype(np.int32)
    stream = oracle.encode_image(samples, 8, False, 4)
    lay = cs_tools.layout(stream)
    assert len(lay.packets) == 5  # one packet per resolution (one component, maximal precincts)
    for r in range(5):
        off, n = lay.packets[r]
        cut = stream[:off + n]
        full = oracle.decode_codestream(stream, r)
        assert np.array_equal(oracle.decode_codestream(cut, r, True), full), r
        assert full.shape == level_dims(190, 121, 4 - r)[::-1]
        if r < 4:
            # one byte short of the resolution's last packet is an error
            with pytest.raises(oracle.OracleError):
                oracle.decode_codestream(stream[:off + n - 1], r, True)
    # the resolution-0 prefix is a small part of the stream (range_plans.rs:69-82)
    assert lay.packets[0][0] + lay.packets[0][1] < len(stream) // 4


def test_index_prefixes_are_the_resolution_cuts(oracle):
    rng = np.random.default_rng(9)
    shape, lv = (3, 77, 140), 4
    vol = rng.integers(0, 4096, size=shape).astype(np.uint16)
    chunk = oracle.htj2k_encode_chunk(vol.tobytes(), shape, "u16", xy_levels=lv)
    entry = 12 + 4 * (lv + 1)
    for p in range(shape[0]):
        off, ln, *prefix = struct.unpack_from("<QI%dI" % (lv + 1), chunk, 8 + 4 * 3 + p * entry)
        cs = chunk[off:off + ln]
        lay = cs_tools.layout(cs)
        assert prefix == [o + n for o, n in lay.packets], p
        for r in range(lv + 1):
            got = oracle.decode_codestream(cs[:prefix[r]], r, True)
            assert np.array_equal(got, pyramid(oracle, vol[p].astype(np.int32), 12, False, lv, r)), (p, r)
