"""The HT cleanup coders on the device against each other and the oracle:
one code-block per wavefront (lanes over columns, wave-parallel stream
unstuffing / bit packing and MagSgn offsets), one code-block per lane, and
the two-pass decoder (VLC / MEL one block per lane into per-quad records,
then MagSgn one block per wavefront).

Covers what the reference's block tests cover (crates/ndic-htj2k/tests/
block_roundtrip.rs:110-140: every block size 1..64, K_max 2..28, densities
from empty to full), the refinement passes after the cleanup pass
(decode.rs:398-600), segments whose MagSgn part is too large for the wave
decoder's LDS bit array (decoded serially by lane 0), corrupted segments
(both decoders give the same samples and the same error), and whole chunk
batches decoded with each decoder forced; and the encoders' byte output
against the oracle over the same range of blocks and chunks.
"""

from __future__ import annotations

import ctypes as C
import importlib

import numpy as np
import pytest

import datagen
import devhooks

pytestmark = pytest.mark.gpu
KNOB = "ht_decode_wave_max_blocks"
ENC_KNOB = "ht_encode_wave_max_blocks"


@pytest.fixture
def force(ndic):
    """Selects the decoder: "wave" (wavefront per block), "lane" (lane per
    block, workgroup shape chosen by the library), "lane1" / "lane4" (lane
    per block with one / four waves per workgroup sharing the tables),
    "lane_sorted" (lane per block, blocks taken heaviest first), "split"
    (the two-pass decoder)."""
    def set_(mode: str) -> None:
        ndic.tune_set(KNOB, 1 << 40 if mode == "wave" else 0)
        ndic.tune_set("ht_lane_waves", {"lane1": 1, "lane4": 4}.get(mode, 0))
        ndic.tune_set("ht_decode_sort", 1 if mode == "lane_sorted" else 0)
        ndic.tune_set("ht_decode_split", 1 if mode == "split" else 0)

    yield set_
    for k in (KNOB, "ht_lane_waves", "ht_decode_sort", "ht_decode_split"):
        ndic.tune_reset(k)


@pytest.fixture
def force_enc(ndic):
    def set_(mode: str) -> None:
        ndic.tune_set(ENC_KNOB, 1 << 40 if mode == "wave" else 0)

    yield set_
    ndic.tune_reset(ENC_KNOB)


def decode_rc(ndic, coded: bytes, w: int, h: int, mmsb: int, passes: int = 1, len1=None, len2: int = 0,
              causal: bool = False):
    """(status, samples) of the device block decode; samples are kept on error."""
    lib = devhooks.bind(ndic)
    out = np.zeros((h, w), np.uint32)
    err = ndic.Error()
    rc = lib.ndic_dev_ht_decode_block(coded, len(coded), out.ctypes.data, w, h, mmsb, passes,
                                      len(coded) if len1 is None else len1, len2, int(causal), C.byref(err))
    return rc, out


def random_block(rng, i):
    w, h = int(rng.integers(1, 65)), int(rng.integers(1, 65))
    if i % 7 == 0:
        w, h = 64, 64
    kmax = int(rng.integers(2, 29))
    dens = [0.0, 0.02, 0.3, 0.8, 1.0][i % 5]
    return w, h, kmax, datagen.sm_block(rng, w, h, kmax, dens)


def test_wave_decoder_matches_oracle(ndic, oracle, force):
    force("wave")
    rng = np.random.default_rng(31)
    for i in range(240):
        w, h, kmax, sm = random_block(rng, i)
        coded = oracle.ht_encode_block(sm, kmax - 1)
        got = devhooks.ht_decode_block(ndic, coded, w, h, kmax - 1)
        assert np.array_equal(got, oracle.ht_decode_block(coded, w, h, kmax - 1)), (i, w, h, kmax)


@pytest.mark.parametrize("mode", ["wave", "split"])
def test_large_magsgn_segments(ndic, oracle, force, mode):
    """Dense 64x64 blocks at K_max 20-30: the wavefront decoder hands MagSgn
    parts of more than 12,280 bytes to lane 0's serial decode, the two-pass
    decoder streams every size through its LDS ring."""
    force(mode)
    rng = np.random.default_rng(32)
    sizes = []
    for kmax in (20, 24, 27, 28, 29, 30):
        sm = datagen.sm_block(rng, 64, 64, kmax, 1.0)
        coded = oracle.ht_encode_block(sm, kmax - 1)
        sizes.append(len(coded))
        got = devhooks.ht_decode_block(ndic, coded, 64, 64, kmax - 1)
        assert np.array_equal(got, oracle.ht_decode_block(coded, 64, 64, kmax - 1)), kmax
    assert min(sizes) < 12280 < max(sizes), sizes


def test_wave_decoder_refinement_passes(ndic, oracle, force):
    force("wave")
    rng = np.random.default_rng(33)
    for i in range(60):
        w, h = int(rng.integers(1, 65)), int(rng.integers(1, 65))
        kmax = int(rng.integers(3, 29))
        sm = datagen.sm_block(rng, w, h, kmax, float(rng.uniform(0.02, 0.6)))
        cleanup = oracle.ht_encode_block(sm, kmax - 1)
        for passes in (2, 3):
            n2 = int(rng.integers(1, 400))
            seg2 = [bytes(n2), b"\xff" * n2, rng.integers(0, 256, size=n2, dtype=np.uint8).tobytes()][i % 3]
            coded = cleanup + seg2
            causal = bool(i & 1)
            exp = oracle.ht_decode_block(coded, w, h, kmax - 1, passes, len(cleanup), len(seg2), causal)
            got = devhooks.ht_decode_block(ndic, coded, w, h, kmax - 1, passes, len(cleanup), len(seg2), causal)
            assert np.array_equal(got, exp), (i, w, h, kmax, passes)


def test_decoders_agree_on_corrupt_segments(ndic, oracle, force):
    """Flipped, zeroed and 0xFF bytes in every part of the segment, wrong
    Scup nibbles, truncations: the two device decoders return the same
    samples and the same status; where the oracle accepts the segment the
    samples are the oracle's."""
    rng = np.random.default_rng(34)
    checked = accepted = 0
    for i in range(300):
        w, h, kmax, sm = random_block(rng, i)
        if i % 5 == 0:
            continue  # empty blocks have no segment to corrupt
        coded = bytearray(oracle.ht_encode_block(sm, kmax - 1))
        kind = i % 6
        n = len(coded)
        if kind == 0:
            for _ in range(int(rng.integers(1, 4))):
                coded[int(rng.integers(0, n))] ^= 1 << int(rng.integers(0, 8))
        elif kind == 1:
            a = int(rng.integers(0, n))
            coded[a:a + 8] = b"\xff" * len(coded[a:a + 8])
        elif kind == 2:
            a = int(rng.integers(0, n))
            coded[a:a + 8] = bytes(len(coded[a:a + 8]))
        elif kind == 3:
            coded[-2] = (coded[-2] & 0xF0) | int(rng.integers(0, 16))
        elif kind == 4:
            coded = coded[:max(2, n - int(rng.integers(1, 16)))]
        else:
            coded[-1] = int(rng.integers(0, 256))
        coded = bytes(coded)
        mmsb = kmax - 1
        force("lane")
        rc_l, out_l = decode_rc(ndic, coded, w, h, mmsb)
        force("wave")
        rc_w, out_w = decode_rc(ndic, coded, w, h, mmsb)
        assert rc_l == rc_w and np.array_equal(out_l, out_w), (i, kind, w, h, kmax)
        force("split")
        rc_s, out_s = decode_rc(ndic, coded, w, h, mmsb)
        assert rc_s == rc_w and np.array_equal(out_s, out_w), (i, kind, w, h, kmax)
        checked += 1
        try:
            exp = oracle.ht_decode_block(coded, w, h, mmsb)
        except Exception:
            continue
        accepted += 1
        if rc_w == 0:
            assert np.array_equal(out_w, exp), (i, kind)
    assert checked > 200 and accepted > 20, (checked, accepted)


@pytest.mark.parametrize("mode", ["wave", "lane1", "lane4", "lane_sorted", "split"])
def test_chunks_decode_with_either_decoder(ndic, oracle, force, mode):
    force(mode)
    rng = np.random.default_rng(35)
    for k, (shape, dt, lv) in enumerate([((3, 33, 61), "u16", 3), ((1, 64, 64), "i16", 0), ((2, 97, 131), "u8", 5),
                                         ((5, 1, 9), "u16", 2), ((2, 200, 70), "i32", 4), ((1, 129, 257), "u16", 5)]):
        npt = {"u8": np.uint8, "u16": np.uint16, "i16": np.int16, "i32": np.int32}[dt]
        info = np.iinfo(npt)
        lo, hi = max(info.min, -(2 ** 20)), min(info.max, 2 ** 20)
        arr = rng.integers(lo, hi, size=shape, endpoint=True).astype(npt)
        if k == 1:
            arr[:] = 0
        raw = arr.tobytes()
        chunk = oracle.htj2k_encode_chunk(raw, shape, dt, xy_levels=lv)
        name = {"u8": "uint8", "u16": "uint16", "i16": "int16", "i32": "int32"}[dt]
        assert ndic.htj2k_decode(chunk, shape, name) == raw, (mode, shape, dt, lv)


@pytest.mark.parametrize("mode", ["wave", "lane", "lane1", "lane4", "lane_sorted", "split"])
def test_window_of_9217_blocks_with_either_decoder(ndic, force, mode):
    """One window of 9,217 code-blocks (an odd count) of blob content,
    encoded on the device and decoded back bit-exactly by each decoder."""
    force(mode)
    dev = importlib.import_module("nd-image-codecs_amd.device")
    shapes = [(64, 200, 200)] * 9 + [(1, 64, 64)]
    al = lambda n: (n + 15) // 16 * 16
    sizes = [int(np.prod(s)) * 2 for s in shapes]
    d_raw = dev.DeviceBuffer(sum(al(n) for n in sizes), zero=True)
    offs, off = [], 0
    for k, (s, n) in enumerate(zip(shapes, sizes)):
        dev.synth_chunk(d_raw.ptr + off, s, "uint16", dev.SYNTH_BLOBS,
                        [(s[0] / 2, s[1] / 3, s[2] / 2, 3000.0, 9.0), (s[0] / 4, s[1] / 2, s[2] / 5, 900.0, 20.0)],
                        100 + k)
        offs.append(off)
        off += al(n)
    steps = [{"dimension": 0, "kind": "lift53", "levels": 2, "group": 0}]
    cap = 3 * off + (1 << 20)
    d_enc = dev.DeviceBuffer(cap, zero=True)
    arr = (ndic.ChunkDesc * len(shapes))(*[ndic.make_desc(s, "uint16", steps, 0, offs[i], 0, 0)
                                             for i, s in enumerate(shapes)])
    ndic.encode_batch(d_raw.ptr, d_enc.ptr, cap, arr)
    d_out = dev.DeviceBuffer(off, zero=True)
    descs = [ndic.make_desc(s, "uint16", steps, 0, arr[i].out_offset, arr[i].out_len, offs[i])
             for i, s in enumerate(shapes)]
    ndic.decode_batch(d_enc.ptr, d_out.ptr, descs)
    assert dev.count_mismatch(d_out.ptr, d_raw.ptr, off) == 0


@pytest.mark.parametrize("mode", ["wave", "lane"])
def test_encoders_match_oracle(ndic, oracle, force_enc, mode):
    """Every block size, K_max 2..30, densities from one sample to full,
    plus blocks of all-ones magnitudes (0xFF runs in MagSgn: stuffing after
    every byte) and checkerboards."""
    force_enc(mode)
    rng = np.random.default_rng(41)
    cases = []
    for i in range(260):
        w, h = int(rng.integers(1, 65)), int(rng.integers(1, 65))
        if i % 9 == 0:
            w, h = 64, 64
        kmax = int(rng.integers(2, 31))
        dens = [0.0, 0.01, 0.3, 0.8, 1.0][i % 5]
        cases.append((w, h, kmax, datagen.sm_block(rng, w, h, kmax, dens)))
    for kmax in (8, 16, 30):
        ones = np.full((64, 64), ((1 << kmax) - 1) << (31 - kmax), np.uint32)
        ones[::3, ::2] |= np.uint32(1 << 31)
        cases.append((64, 64, kmax, ones))
        chk = np.zeros((64, 64), np.uint32)
        chk[::2, ::2] = np.uint32(5 << (31 - kmax))
        cases.append((64, 64, kmax, chk))
    for w, h, kmax, sm in cases:
        exp = oracle.ht_encode_block(sm, kmax - 1)
        got = devhooks.ht_encode_block(ndic, sm, kmax - 1)
        assert got == exp, (mode, w, h, kmax)
    # encode_block's callers skip all-zero blocks (crates/ndic-htj2k/src/block/
    # encode.rs:103-105; the packet header signals non-inclusion): the device
    # coders mark such a block not included
    assert devhooks.ht_encode_block(ndic, np.zeros((5, 17), np.uint32), 3) == b""


@pytest.mark.parametrize("mode", ["wave", "lane"])
def test_chunks_encode_with_either_encoder(ndic, oracle, force_enc, mode):
    force_enc(mode)
    rng = np.random.default_rng(42)
    for k, (shape, dt, lv) in enumerate([((3, 33, 61), "u16", 3), ((1, 64, 64), "i16", 0), ((2, 97, 131), "u8", 5),
                                         ((5, 1, 9), "u16", 2), ((2, 200, 70), "i32", 4), ((1, 129, 257), "u16", 5)]):
        npt = {"u8": np.uint8, "u16": np.uint16, "i16": np.int16, "i32": np.int32}[dt]
        info = np.iinfo(npt)
        lo, hi = max(info.min, -(2 ** 20)), min(info.max, 2 ** 20)
        arr = rng.integers(lo, hi, size=shape, endpoint=True).astype(npt)
        if k == 1:
            arr[:] = 0
        name = {"u8": "uint8", "u16": "uint16", "i16": "int16", "i32": "int32"}[dt]
        exp = oracle.htj2k_encode_chunk(arr.tobytes(), shape, dt, xy_levels=lv)
        assert ndic.htj2k_encode(arr, shape, name, xy_levels=lv) == exp, (mode, shape, dt, lv)
    vol = datagen.c0_volume((4, 256, 256), seed=3)
    exp = oracle.htj2k_encode_chunk(vol.tobytes(), vol.shape, "u16", xy_levels=5)
    assert ndic.htj2k_encode(vol, vol.shape, "uint16", xy_levels=5) == exp


@pytest.mark.parametrize("mode", ["split", "lane"])
def test_decoders_match_oracle_per_block(ndic, oracle, force, mode):
    """Every block size and density of the block round-trip test through the
    lane-per-block and the two-pass decoders, and the refinement passes
    after them."""
    force(mode)
    rng = np.random.default_rng(36)
    for i in range(160):
        w, h, kmax, sm = random_block(rng, i)
        coded = oracle.ht_encode_block(sm, kmax - 1)
        got = devhooks.ht_decode_block(ndic, coded, w, h, kmax - 1)
        assert np.array_equal(got, oracle.ht_decode_block(coded, w, h, kmax - 1)), (mode, i, w, h, kmax)
    for i in range(40):
        w, h = int(rng.integers(1, 65)), int(rng.integers(1, 65))
        kmax = int(rng.integers(3, 29))
        sm = datagen.sm_block(rng, w, h, kmax, float(rng.uniform(0.02, 0.6)))
        cleanup = oracle.ht_encode_block(sm, kmax - 1)
        seg2 = rng.integers(0, 256, size=int(rng.integers(1, 300)), dtype=np.uint8).tobytes()
        exp = oracle.ht_decode_block(cleanup + seg2, w, h, kmax - 1, 3, len(cleanup), len(seg2), bool(i & 1))
        got = devhooks.ht_decode_block(ndic, cleanup + seg2, w, h, kmax - 1, 3, len(cleanup), len(seg2), bool(i & 1))
        assert np.array_equal(got, exp), (mode, i)
