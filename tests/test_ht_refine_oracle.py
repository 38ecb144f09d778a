"""Pins the oracle's HT SigProp and MagRef decode (the two refinement passes
of crates/ndic-htj2k/src/block/decode.rs:398-600) with an independent
restatement of what the passes mean: a sample-by-sample scan of each 4-row
stripe, column by column, in which a not-yet-significant sample with a
significant 8-neighbour (the stripe below counted only from the cleanup pass,
and not at all in vertically causal mode; samples already visited counted with
what this pass found) reads one significance bit, the newly significant
samples of each group of four columns then read one sign bit each, and
MagRef reads one bit per cleanup-significant sample in the same scan order.
The bit streams are modelled byte by byte: SigProp forward from the start of
the refinement segment (a byte after 0xFF gives 7 bits), MagRef backward from
its end (a byte whose low 7 bits are all set after a byte above 0x8F gives 7
bits), both read as zero once exhausted (crates/ndic-htj2k/src/block/streams.rs:20-265).

The oracle's version is the reference's bit-mask formulation (a 16-bit
significance word per 4x4 group, propagation masks, the carry into the next
group); this one shares none of it, so a wrong mask, shift or carry fails here.
"""

from __future__ import annotations

import numpy as np
import pytest

from datagen import sm_block


def fwd_stream(seg: bytes) -> int:
    acc, n, stuffed = 0, 0, False
    for b in seg:
        acc |= b << n
        n += 7 if stuffed else 8
        stuffed = b == 0xFF
    return acc


def rev_stream(seg: bytes) -> int:
    acc, n, prev_big = 0, 0, True
    for b in reversed(seg):
        acc |= b << n
        n += 7 if (prev_big and (b & 0x7F) == 0x7F) else 8
        prev_big = b > 0x8F
    return acc


class Bits:
    def __init__(self, acc: int):
        self.acc, self.pos = acc, 0

    def take(self) -> int:
        b = (self.acc >> self.pos) & 1
        self.pos += 1
        return b


def scan_refine(cleanup: np.ndarray, missing_msbs: int, passes: int, seg2: bytes, causal: bool) -> np.ndarray:
    h, w = cleanup.shape
    out = cleanup.astype(np.uint32).copy()
    if passes < 2 or not seg2 or missing_msbs >= 29:
        return out
    p = 30 - missing_msbs
    sig0 = cleanup != 0
    sig = sig0.copy()
    sp = Bits(fwd_stream(seg2))
    for s0 in range(0, h, 4):
        rows = range(s0, min(s0 + 4, h))
        for g0 in range(0, w, 4):
            cols = range(g0, min(g0 + 4, w))
            newly = []
            for c in cols:
                for r in rows:
                    if sig0[r, c]:
                        continue
                    cand = False
                    for dr in (-1, 0, 1):
                        for dc in (-1, 0, 1):
                            rr, cc = r + dr, c + dc
                            if (dr, dc) == (0, 0) or not (0 <= rr < h and 0 <= cc < w):
                                continue
                            if rr >= s0 + 4:
                                known = sig0[rr, cc] and not causal
                            elif rr < s0 or cc < c or (cc == c and rr < r):
                                known = sig[rr, cc]  # visited earlier in this pass
                            else:
                                known = sig0[rr, cc]
                            cand = cand or bool(known)
                    if cand and sp.take():
                        sig[r, c] = True
                        newly.append((r, c))
            for (r, c) in newly:
                out[r, c] = (sp.take() << 31) | (1 << (p - 1)) | (1 << (p - 2))
    if passes > 2:
        mr = Bits(rev_stream(seg2))
        keep = ~np.uint32((3 << (p - 2)))
        for s0 in range(0, h, 4):
            for c in range(w):
                for r in range(s0, min(s0 + 4, h)):
                    if sig0[r, c]:
                        bit = mr.take()
                        out[r, c] = (int(out[r, c]) & int(keep)) | (bit << (p - 1)) | (1 << (p - 2))
    return out


def refinement_bytes(rng, n: int, kind: str) -> bytes:
    if kind == "zeros":
        return bytes(n)
    if kind == "ones":
        return b"\xff" * n
    return rng.integers(0, 256, size=n, dtype=np.uint8).tobytes()


@pytest.mark.parametrize("causal", [False, True])
def test_refinement_passes_match_scan_restatement(oracle, causal):
    rng = np.random.default_rng(11 + causal)
    cases = grew = 0
    for i in range(90):
        w, h = int(rng.integers(1, 65)), int(rng.integers(1, 65))
        if i < 6:
            w, h = [(64, 64), (1, 1), (4, 4), (5, 9), (64, 3), (2, 64)][i]
        kmax = int(rng.integers(3, 29))
        sm = sm_block(rng, w, h, kmax, float(rng.uniform(0.02, 0.6)))
        mmsb = kmax - 1
        cleanup = oracle.ht_encode_block(sm, mmsb)
        base = oracle.ht_decode_block(cleanup, w, h, mmsb)
        for passes in (2, 3):
            kind = ["random", "zeros", "ones"][i % 3] if i >= 6 else "random"
            seg2 = refinement_bytes(rng, int(rng.integers(1, 400)), kind)
            coded = cleanup + seg2
            got = oracle.ht_decode_block(coded, w, h, mmsb, passes, len(cleanup), len(seg2), causal)
            exp = scan_refine(base, mmsb, passes, seg2, causal)
            assert np.array_equal(got, exp), (i, w, h, kmax, passes, kind)
            cases += 1
            grew += int(np.count_nonzero(got) > np.count_nonzero(base))
    assert cases == 180 and grew > 60  # SigProp made samples significant in many of them


def test_zero_refinement_changes_only_magref_bits(oracle):
    """All-zero refinement bytes: SigProp finds nothing and MagRef clears the
    refined bit, so the samples follow from the cleanup decode in closed form."""
    rng = np.random.default_rng(3)
    w, h, kmax = 37, 29, 12
    sm = sm_block(rng, w, h, kmax, 0.3)
    cleanup = oracle.ht_encode_block(sm, kmax - 1)
    base = oracle.ht_decode_block(cleanup, w, h, kmax - 1)
    p = 30 - (kmax - 1)
    two = oracle.ht_decode_block(cleanup + bytes(64), w, h, kmax - 1, 2, len(cleanup), 64)
    assert np.array_equal(two, base)
    three = oracle.ht_decode_block(cleanup + bytes(64), w, h, kmax - 1, 3, len(cleanup), 64)
    nz = base != 0
    exp = base.copy()
    exp[nz] = (base[nz] - np.uint32(1 << (p - 1))) + np.uint32(1 << (p - 2))
    assert np.array_equal(three, exp)
