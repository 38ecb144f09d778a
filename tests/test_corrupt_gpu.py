"""Corrupted and truncated input on the GPU, held to the oracle's outcome.

The reference requires a clean error or an image for every single-byte
mutation (0x00, 0xFF, 0x01, 0x80 at every position) and every truncation of
a codestream (crates/ndic-codestream/tests/malformed.rs:30-85), and a clean
error for malformed chunks (crates/ndic-zarr/tests/htj2k_zarrs.rs:257).  Here
every mutated or cut stream goes through one device batch and each item's
own outcome must be the oracle's: both fail, or both decode the same
samples.  Bytes of the main and tile-part headers may in addition be refused
as Unsupported by the device where the oracle decodes them, since the device
path's scope is one HT component, reversible 5/3, one layer, maximal
precincts and 64 x 64 code-blocks; and an oracle image of another size than
the chunk declares is an error (htj2k.rs decode_plane's geometry check).

Both forms of the packet parser (wavefront and thread per plane) and of the
block decoder (wavefront and lane per block) run every sweep.
"""

from __future__ import annotations

import struct

import numpy as np
import pytest

import batch_tools
import cs_tools
from test_corrupt_oracle import inflate_second_block, mutations, set_eph, small_image, small_stream

pytestmark = pytest.mark.gpu

FORMS = {"wave": 1 << 62, "lane": 0}


@pytest.fixture(params=list(FORMS))
def form(request, ndic):
    keys = ("parse_wave_max_planes", "ht_decode_wave_max_blocks")
    before = {k: ndic.tune_get(k) for k in keys}
    for k in keys:
        ndic.tune_set(k, FORMS[request.param])
    yield request.param
    for k, v in before.items():
        ndic.tune_set(k, v)


def oracle_outcome(oracle, data, max_res, prefix, shape):
    try:
        a = oracle.decode_codestream(data, max_res, prefix)
    except oracle.OracleError:
        return None
    return a if a.shape == shape else None


def compare(oracle, ndic, streams, labels, width, height, levels, max_res=-1, prefix=False, header_end=0):
    r = levels if max_res < 0 else min(max_res, levels)
    shape = (batch_tools.reduced(height, levels - r), batch_tools.reduced(width, levels - r))
    got = batch_tools.plane_batch(ndic, streams, width, height, levels, max_res, prefix)
    bad = []
    errors = 0
    for (st, arr), data, lab in zip(got, streams, labels):
        exp = oracle_outcome(oracle, data, max_res, prefix, shape)
        errors += exp is None
        if exp is None and st == 0:
            bad.append((lab, "device decoded, oracle failed"))
        elif exp is not None and st != 0:
            in_header = isinstance(lab, tuple) and lab[0] < header_end
            if not (in_header and st == batch_tools.NDIC_UNSUPPORTED):
                bad.append((lab, f"device status {st}, oracle decoded"))
        elif exp is not None and not np.array_equal(arr, exp):
            bad.append((lab, "samples differ"))
    assert not bad, f"{len(bad)} of {len(streams)} outcomes differ from the oracle: {bad[:12]}"
    return errors


def test_small_stream_mutations_match_oracle(ndic, oracle, form):
    cs = small_stream(oracle)
    lay = cs_tools.layout(cs)
    labels, streams = [], []
    for pos, v, m in mutations(cs):
        labels.append((pos, v))
        streams.append(m)
    errors = compare(oracle, ndic, streams, labels, 32, 20, 2, header_end=lay.body_start)
    assert errors > 100  # the sweep reaches the error paths, packet headers included


def test_small_stream_truncations_match_oracle(ndic, oracle, form):
    cs = small_stream(oracle)
    cuts = list(range(len(cs)))
    streams = [cs[:c] for c in cuts]
    compare(oracle, ndic, streams, cuts, 32, 20, 2)
    for r in range(3):
        compare(oracle, ndic, streams, cuts, 32, 20, 2, max_res=r, prefix=True)


def test_targeted_packet_corruptions_error(ndic, oracle, form):
    """The three cases whose earlier handling left included code-blocks
    without a segment: each is an error, alone and among good streams (which
    still decode)."""
    cs = small_stream(oracle)
    lay = cs_tools.layout(cs)
    off1, n1 = lay.packets[1]
    bad = [inflate_second_block(cs), set_eph(cs)]
    for b in bad:
        with pytest.raises(ndic.NdicError):
            ndic.htj2k_decode_plane(b)
    with pytest.raises(ndic.NdicError):
        ndic.htj2k_decode_plane(cs[:off1 + n1 // 2], max_res=1, prefix=True)
    with pytest.raises(ndic.NdicError):
        ndic.htj2k_decode_plane(cs[:off1 + n1 - 1], max_res=2, prefix=True)
    mixed = [cs, bad[0], cs, bad[1], cs]
    res = batch_tools.plane_batch(ndic, mixed, 32, 20, 2)
    assert [st != 0 for st, _ in res] == [False, True, False, True, False]
    for st, a in res:
        if st == 0:
            assert np.array_equal(a, small_image())
    cut = batch_tools.plane_batch(ndic, [cs[:off1 + n1 // 2], cs[:off1 + n1]], 32, 20, 2, max_res=1, prefix=True)
    assert cut[0][0] != 0 and cut[1][0] == 0
    assert np.array_equal(cut[1][1], oracle.decode_codestream(cs[:off1 + n1], 1, True))


def test_multiblock_packet_headers_match_oracle(ndic, oracle, form):
    """A 150 x 200 u16 plane with 3 levels (several code-blocks per band at
    the finest resolutions): every byte of every packet header region (the
    first 24 bytes of each packet) mutated, and cuts at and inside every
    packet, strict and as prefixes at each resolution."""
    rng = np.random.default_rng(31)
    img = rng.integers(0, 4096, size=(150, 200)).astype(np.int32)
    img[40:90, 60:150] = 0
    cs = oracle.encode_image(img, 12, False, 3)
    lay = cs_tools.layout(cs)
    positions = sorted({p for off, n in lay.packets for p in range(off, min(off + 24, off + n))})
    labels, streams = [], []
    for pos, v, m in mutations(cs, positions):
        labels.append((pos, v))
        streams.append(m)
    compare(oracle, ndic, streams, labels, 200, 150, 3)
    cuts = sorted({c for off, n in lay.packets for c in (off, off + 1, off + 5, off + n // 2, off + n - 1)})
    trunc = [cs[:c] for c in cuts]
    compare(oracle, ndic, trunc, cuts, 200, 150, 3)
    for r in range(4):
        compare(oracle, ndic, trunc, cuts, 200, 150, 3, max_res=r, prefix=True)


def test_chunk_mutations_match_oracle(ndic, oracle, form):
    """ndht chunks through the batch decoder (no lift steps, u16): every
    mutation of the container header and index, every seventh codestream
    byte, and every truncation; each chunk's status and samples against the
    oracle's chunk decode."""
    vals = np.array([(i * 7) % 4096 for i in range(256)], np.uint16).reshape(4, 8, 8)
    chunk = oracle.htj2k_encode_chunk(vals.tobytes(), vals.shape, "u16", xy_levels=2)
    ndim, lv = 3, 2
    head = 8 + 4 * ndim + 4 * (12 + 4 * (lv + 1))
    # codestream header bytes: from each plane's offset to its first packet
    planes = [struct.unpack_from("<QI", chunk, 8 + 4 * ndim + (12 + 4 * (lv + 1)) * p) for p in range(4)]
    in_header = set()
    for off, ln in planes:
        in_header.update(range(off, off + cs_tools.layout(chunk[off:off + ln]).body_start))
    muts = list(mutations(chunk, list(range(head)) + list(range(head, len(chunk), 7))))
    cands = [m for _, _, m in muts] + [chunk[:c] for c in range(len(chunk))]
    where = [pos for pos, _, _ in muts] + [-1] * len(chunk)
    got = batch_tools.chunk_batch(ndic, cands, vals.shape, "uint16", [], lv)
    bad = []
    for i, ((st, out), data) in enumerate(zip(got, cands)):
        try:
            exp = oracle.htj2k_decode_chunk(data, vals.shape, "u16")
        except oracle.OracleError:
            exp = None
        if exp is not None and st == batch_tools.NDIC_UNSUPPORTED and where[i] in in_header:
            continue  # outside the device path's codestream scope (e.g. other code-block sizes)
        if (exp is None) != (st != 0) or (exp is not None and out != exp):
            bad.append((i, st, exp is None))
    assert not bad, f"{len(bad)} of {len(cands)} chunk outcomes differ from the oracle: {bad[:12]}"
    # the good chunk among corrupted ones still decodes
    mixed = batch_tools.chunk_batch(ndic, [cands[0], chunk, chunk[:50]], vals.shape, "uint16", [], lv)
    assert mixed[1] == (0, vals.tobytes()) and mixed[2][0] != 0
    off, ln = struct.unpack_from("<QI", chunk, 8 + 4 * ndim)
    assert off >= head and ln > 0
