"""Codestream structure helpers for the corruption tests (host-side, test only).

A marker walk over a single-tile Part 1/15 codestream
(crates/ndic-codestream/src/reader.rs:120-291) that returns where the main
header markers, the tile-part's PLT packet lengths and its body lie, and a
reader of one-block-per-band packet headers
(crates/ndic-codestream/src/packet.rs:190-330, bitio.rs:104-161) that records
the bit position of each code-block's cleanup length, so a test can inflate
exactly one length field.
"""

from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import Dict, List, Tuple

SOC, SOT, SOD, EOC, COD, PLT = 0xFF4F, 0xFF90, 0xFF93, 0xFFD9, 0xFF52, 0xFF58


@dataclass
class Layout:
    markers: Dict[int, int] = field(default_factory=dict)  # marker -> offset of its first occurrence
    body_start: int = 0
    body_end: int = 0
    packets: List[Tuple[int, int]] = field(default_factory=list)  # (offset, length) per packet, in order


def layout(cs: bytes) -> Layout:
    """Main-header marker offsets, the tile-part body and its packets (from PLT)."""
    out = Layout()
    assert struct.unpack_from(">H", cs, 0)[0] == SOC
    pos = 2
    while True:
        m = struct.unpack_from(">H", cs, pos)[0]
        out.markers.setdefault(m, pos)
        if m == SOT:
            break
        pos += 2 + struct.unpack_from(">H", cs, pos + 2)[0]
    sot = pos
    psot = struct.unpack_from(">I", cs, sot + 6)[0]
    pos = sot + 12
    plt: List[int] = []
    while True:
        m = struct.unpack_from(">H", cs, pos)[0]
        out.markers.setdefault(m, pos)
        if m == SOD:
            pos += 2
            break
        ln = struct.unpack_from(">H", cs, pos + 2)[0]
        if m == PLT:
            acc = 0
            for b in cs[pos + 5:pos + 2 + ln]:
                acc = (acc << 7) | (b & 0x7F)
                if not b & 0x80:
                    plt.append(acc)
                    acc = 0
        pos += 2 + ln
    out.body_start = pos
    out.body_end = sot + psot if psot else len(cs) - 2
    off = pos
    for n in plt:
        out.packets.append((off, n))
        off += n
    assert off == out.body_end, "PLT lengths do not cover the body"
    return out


class _Bits:
    """bitio.rs HeaderBitReader: MSB first; a byte after 0xFF holds 7 bits."""

    def __init__(self, data: bytes, start: int):
        self.data, self.pos, self.avail, self.tmp, self.stuff = data, start, 0, 0, False

    def bit(self) -> Tuple[int, int, int]:
        """(bit, byte offset, bit index within the byte, 7 = MSB)."""
        if self.avail == 0:
            self.tmp = self.data[self.pos]
            self.pos += 1
            self.avail = 7 if self.stuff else 8
            self.stuff = self.tmp == 0xFF
        self.avail -= 1
        return (self.tmp >> self.avail) & 1, self.pos - 1, self.avail

    def bits(self, n: int) -> Tuple[int, List[Tuple[int, int]]]:
        v, where = 0, []
        for _ in range(n):
            b, byte, k = self.bit()
            v = (v << 1) | b
            where.append((byte, k))
        return v, where


def one_block_packet(cs: bytes, offset: int, nbands: int) -> List[dict]:
    """Parses a packet header whose bands hold one code-block each (tag trees
    of one node): per included block its band, missing MSBs, passes, Lblock
    and the (byte, bit) positions of its cleanup-length field, MSB first."""
    br = _Bits(cs, offset)
    if br.bit()[0] == 0:
        return []
    blocks = []
    for b in range(nbands):
        if br.bit()[0] == 0:  # inclusion tag tree, layer 0
            continue
        mm = 0
        while br.bit()[0] == 0:  # missing-MSB tag tree: zeros, then a one
            mm += 1
        np_ = 1
        if br.bit()[0]:
            np_ = 2
            if br.bit()[0]:
                v, _ = br.bits(2)
                np_ = 3 + v
                if v == 3:
                    v, _ = br.bits(5)
                    np_ = 6 + v
                    if v == 31:
                        np_ = 37 + br.bits(7)[0]
        lblock = 3
        while br.bit()[0]:
            lblock += 1
        assert np_ == 1, "the helper reads single-pass blocks only"
        len1, where = br.bits(lblock)
        blocks.append({"band": b, "mmsb": mm, "passes": np_, "lblock": lblock, "len1": len1, "len1_bits": where})
    return blocks
