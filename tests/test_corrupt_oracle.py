"""Malformed input on the oracle (CPU): every single-byte mutation and every
truncation of a small codestream and of a small ndht chunk ends in a clean
error or an image, never a crash, as the reference requires
(crates/ndic-codestream/tests/malformed.rs:30-85,
crates/ndic-zarr/tests/htj2k_zarrs.rs:257).  The GPU corruption tests hold the
device to these same oracle outcomes, so the targeted cases are pinned here
first: an inflated cleanup length in the second code-block of a packet, a COD
that announces EPH markers the packets lack, and prefix cuts inside an
included packet all fail.
"""

from __future__ import annotations

import struct

import numpy as np
import pytest

import cs_tools

MUTATION_VALUES = (0x00, 0xFF, 0x01, 0x80)


def small_stream(oracle) -> bytes:
    """malformed.rs small_stream: 32 x 20 u8 samples (i * 7) mod 256, 2 levels."""
    s = np.array([(i * 7) % 256 for i in range(32 * 20)], np.int32).reshape(20, 32)
    return oracle.encode_image(s, 8, False, 2)


def small_image() -> np.ndarray:
    return np.array([(i * 7) % 256 for i in range(32 * 20)], np.int32).reshape(20, 32)


def mutations(data: bytes, positions=None):
    for pos in (range(len(data)) if positions is None else positions):
        for v in MUTATION_VALUES:
            if data[pos] != v:
                m = bytearray(data)
                m[pos] = v
                yield pos, v, bytes(m)


def inflate_second_block(cs: bytes) -> bytes:
    """The resolution-1 packet's second code-block (its LH band) with its
    cleanup length raised so that the packet's segments overrun its body:
    the header includes all three blocks, the body then runs out before the
    last segment, after the earlier blocks' segments were found."""
    lay = cs_tools.layout(cs)
    off, n = lay.packets[1]
    blocks = cs_tools.one_block_packet(cs, off, 3)
    assert len(blocks) == 3 and blocks[1]["band"] == 1
    lens = [b["len1"] for b in blocks]
    body = sum(lens)  # the packet's bytes after its header
    m = bytearray(cs)
    # set the length field's zero bits from the top until the segments overrun
    val = lens[1]
    for k, (byte, bit) in enumerate(blocks[1]["len1_bits"]):
        weight = 1 << (len(blocks[1]["len1_bits"]) - 1 - k)
        if not (val & weight) and lens[0] + val + lens[2] <= body:
            m[byte] |= 1 << bit
            val |= weight
    assert lens[0] + val + lens[2] > body and lens[0] + val <= body
    return bytes(m)


def set_eph(cs: bytes) -> bytes:
    """COD's Scod with the EPH bit set; the packets carry no EPH markers."""
    at = cs_tools.layout(cs).markers[cs_tools.COD]
    m = bytearray(cs)
    m[at + 4] |= 0x04
    return bytes(m)


def test_small_stream_decodes(oracle):
    assert np.array_equal(oracle.decode_codestream(small_stream(oracle)), small_image())


def test_single_byte_mutations_never_crash(oracle):
    cs = small_stream(oracle)
    outcomes = {"error": 0, "image": 0}
    for pos, v, m in mutations(cs):
        try:
            oracle.decode_codestream(m)
            outcomes["image"] += 1
        except oracle.OracleError:
            outcomes["error"] += 1
    assert outcomes["error"] > 0 and outcomes["image"] > 0


def test_truncations_error(oracle):
    """Every cut of the stream fails in strict mode; in prefix mode only the
    two cuts that drop (part of) the final EOC still decode the full image."""
    cs = small_stream(oracle)
    for cut in range(len(cs)):
        with pytest.raises(oracle.OracleError):
            oracle.decode_codestream(cs[:cut])
        if cut < len(cs) - 2:
            with pytest.raises(oracle.OracleError):
                oracle.decode_codestream(cs[:cut], -1, True)
        else:
            assert np.array_equal(oracle.decode_codestream(cs[:cut], -1, True), small_image())


def test_hostile_header_values_rejected(oracle):
    """malformed.rs:55-85: XTsiz = 0 is refused; a 0xFFFFFFFF canvas does not decode."""
    cs = small_stream(oracle)
    siz = 6  # SOC + SIZ marker + Lsiz
    m = bytearray(cs)
    m[siz + 18:siz + 22] = bytes(4)
    with pytest.raises(oracle.OracleError):
        oracle.decode_codestream(bytes(m))
    m = bytearray(cs)
    m[siz + 2:siz + 10] = b"\xff" * 8
    with pytest.raises(oracle.OracleError):
        oracle.decode_codestream(bytes(m))


def test_targeted_packet_corruptions_error(oracle):
    cs = small_stream(oracle)
    with pytest.raises(oracle.OracleError):
        oracle.decode_codestream(inflate_second_block(cs))
    with pytest.raises(oracle.OracleError):
        oracle.decode_codestream(set_eph(cs))
    lay = cs_tools.layout(cs)
    for r, (off, n) in enumerate(lay.packets):
        for cut in (off + 1, off + n // 2, off + n - 1):
            for max_res in range(r, len(lay.packets)):
                with pytest.raises(oracle.OracleError):
                    oracle.decode_codestream(cs[:cut], max_res, True)
        # a cut at the packet's end leaves every resolution up to r decodable
        oracle.decode_codestream(cs[:off + n], r, True)


def test_chunk_mutations_and_truncations_never_crash(oracle):
    """An ndht chunk (4 x 8 x 8 u16, 2 levels): every mutation of its header
    and index, a sample of the codestream bytes, and every truncation."""
    vals = np.array([(i * 7) % 4096 for i in range(256)], np.uint16).reshape(4, 8, 8)
    chunk = oracle.htj2k_encode_chunk(vals.tobytes(), vals.shape, "u16", xy_levels=2)
    ndim, lv = 3, 2
    head = 8 + 4 * ndim + 4 * (12 + 4 * (lv + 1))
    errors = 0
    for pos, v, m in mutations(chunk, list(range(head)) + list(range(head, len(chunk), 7))):
        try:
            oracle.htj2k_decode_chunk(m, vals.shape, "u16")
        except oracle.OracleError:
            errors += 1
    assert errors > 0
    for cut in range(len(chunk)):
        with pytest.raises(oracle.OracleError):
            oracle.htj2k_decode_chunk(chunk[:cut], vals.shape, "u16")
    assert oracle.htj2k_decode_chunk(chunk, vals.shape, "u16") == vals.tobytes()
    # the index's first prefix entry points inside plane 0's codestream
    off, ln, p0 = struct.unpack_from("<QII", chunk, 8 + 4 * ndim)
    assert 0 < p0 <= ln and off >= head
