"""Autotuner of the library's launch-parameter knobs (``ndic_tune_*``).

The HT block coder and the codestream parser each come in two forms whose
cost scales differently with the size of a batch window: one code-block
(plane) per wavefront, lowest latency, and one per lane (thread), highest
throughput.  The wavefront-per-block coders are persistent kernels fed by a
device work queue, handing out ``ht_wave_grain`` code-blocks at a time.  For
each shape class -- a window of N code-blocks of 256 x 256 planes, 5 DWT
levels, uint16 blob content -- the tuner times both forms on the device
(HIP events around each decode stage, wall time around an encode), and sets:

* ``ht_decode_wave_max_blocks``: the largest N at which the wavefront
  decoder is not slower (every smaller class measured agrees);
* ``ht_encode_wave_max_blocks``: the same for the encoder;
* ``parse_wave_max_planes``: the same for the parser, in planes;
* ``ht_wave_grain``: the queue grain with the fastest wavefront decode of
  the class nearest the decode crossover.

Every decode made while tuning is checked bit-exact against its source.
``tune()`` returns the choices with the measurements, ``apply()`` sets them,
``load_or_tune(path)`` reuses a saved result.
"""

from __future__ import annotations

import importlib
import json
import os
import time
from typing import Dict, List, Sequence

_pkg = importlib.import_module("nd-image-codecs_amd")
dev = importlib.import_module("nd-image-codecs_amd.device")

KNOBS = ("ht_decode_wave_max_blocks", "ht_encode_wave_max_blocks", "parse_wave_max_planes", "ht_wave_grain")
BLOCKS_PER_PLANE = 25  # 256 x 256 plane, 5 levels: 1 + 3 (1 + 1 + 1 + 1) + 3 x 4 code-blocks
PLANES_PER_CHUNK = 8
LADDER = (200, 800, 3200, 6400, 12800, 25600, 51200)
QUICK_LADDER = (200, 1600, 6400, 12800, 25600)
GRAINS = (1, 2, 4, 8)


def crossover(sizes: Sequence[int], fast_small: Sequence[float], fast_big: Sequence[float]) -> int:
    """Largest size up to which the small-batch form is never slower than the
    large-batch form (0 when it loses at the smallest size).  Between the last
    winning size and the first losing one the costs are interpolated
    linearly.  A form that wins at every measured size is chosen up to the
    largest one only: beyond it nothing was measured, and the large-batch form
    (the one built for throughput) stays in force."""
    for k, (n, a, b) in enumerate(zip(sizes, fast_small, fast_big)):
        if a > b:
            if k == 0:
                return 0
            n0, a0, b0 = sizes[k - 1], fast_small[k - 1], fast_big[k - 1]
            d0, d1 = b0 - a0, b - a  # >= 0 then < 0
            return int(n0 + (n - n0) * d0 / (d0 - d1))
    return int(sizes[-1])


class _Window:
    """A device batch of uint16 blob chunks (PLANES_PER_CHUNK planes of
    256 x 256 each) encoded on the device: one decode window of n_blocks."""

    def __init__(self, n_blocks: int, seed: int):
        planes = max(1, n_blocks // BLOCKS_PER_PLANE)
        self.shapes = []
        while planes > 0:
            z = min(PLANES_PER_CHUNK, planes)
            self.shapes.append((z, 256, 256))
            planes -= z
        self.steps = [{"dimension": 0, "kind": "lift53", "levels": 2, "group": 0}]
        sizes = [z * y * x * 2 for z, y, x in self.shapes]
        self.offs, off = [], 0
        for n in sizes:
            self.offs.append(off)
            off += (n + 255) // 256 * 256
        self.nbytes = off
        self.raw = dev.DeviceBuffer(off, zero=True)
        for k, (s, o) in enumerate(zip(self.shapes, self.offs)):
            dev.synth_chunk(self.raw.ptr + o, s, "uint16", dev.SYNTH_BLOBS,
                            [(s[0] / 2, 90.0, 120.0, 2500.0, 12.0), (s[0] / 3, 180.0, 60.0, 900.0, 20.0)], seed + k)
        self.cap = 3 * off + (16 << 20)
        self.enc = dev.DeviceBuffer(self.cap, zero=True)
        self.edescs = (_pkg.ChunkDesc * len(self.shapes))(*[
            _pkg.make_desc(s, "uint16", self.steps, 5, self.offs[i], 0, 0) for i, s in enumerate(self.shapes)])
        _pkg.encode_batch(self.raw.ptr, self.enc.ptr, self.cap, self.edescs)
        self.ddescs = [_pkg.make_desc(s, "uint16", self.steps, 5, self.edescs[i].out_offset, self.edescs[i].out_len,
                                      self.offs[i]) for i, s in enumerate(self.shapes)]
        self.out = dev.DeviceBuffer(off, zero=True)

    def decode_stages(self, reps: int) -> List[float]:
        """Median (over reps) per-stage device ms of one decode of the window."""
        runs = []
        for _ in range(reps):
            dev.profile(True)
            _pkg.decode_batch(self.enc.ptr, self.out.ptr, self.ddescs)
            dev.synchronize()
            ms, _, _ = dev.stage_times(5)
            dev.profile(False)
            runs.append(ms)
        if dev.count_mismatch(self.out.ptr, self.raw.ptr, self.nbytes) != 0:
            raise RuntimeError("autotune: a decode under trial knobs was not bit-exact")
        return [sorted(r[i] for r in runs)[len(runs) // 2] for i in range(5)]

    def encode_ms(self, reps: int) -> float:
        ts = []
        for _ in range(reps):
            d = (_pkg.ChunkDesc * len(self.shapes))(*self.edescs)
            dev.synchronize()
            t0 = time.perf_counter()
            _pkg.encode_batch(self.raw.ptr, self.enc.ptr, self.cap, d)
            dev.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
            if [x.out_len for x in d] != [x.out_len for x in self.edescs]:
                raise RuntimeError("autotune: an encode under trial knobs changed the stream lengths")
        return sorted(ts)[len(ts) // 2]


def _force(key: str, small_form: bool) -> None:
    _pkg.tune_set(key, 1 << 62 if small_form else 0)


def tune(quick: bool = True, reps: int = 3, seed: int = 11, ladder: Sequence[int] = ()) -> Dict:
    """Measures every knob's forms over the ladder of window sizes (code-blocks)
    and applies the choices."""
    ladder = tuple(ladder) or (QUICK_LADDER if quick else LADDER)
    saved = {k: _pkg.tune_get(k) for k in KNOBS}
    meas = {"blocks": list(ladder), "decode_wave_ms": [], "decode_lane_ms": [], "encode_wave_ms": [],
            "encode_lane_ms": [], "planes": [], "parse_wave_ms": [], "parse_thread_ms": [], "grain": {}}
    t_start = time.perf_counter()
    try:
        windows = {}
        for n in ladder:
            w = windows[n] = _Window(n, seed)
            meas["planes"].append(sum(s[0] for s in w.shapes))
            for small in (True, False):
                _force("ht_decode_wave_max_blocks", small)
                _force("parse_wave_max_planes", small)
                st = w.decode_stages(reps)
                meas["decode_wave_ms" if small else "decode_lane_ms"].append(round(st[1], 4))
                meas["parse_wave_ms" if small else "parse_thread_ms"].append(round(st[0], 4))
                _force("ht_encode_wave_max_blocks", small)
                meas["encode_wave_ms" if small else "encode_lane_ms"].append(round(w.encode_ms(reps), 4))
        chosen = {
            "ht_decode_wave_max_blocks": crossover(ladder, meas["decode_wave_ms"], meas["decode_lane_ms"]),
            "ht_encode_wave_max_blocks": crossover(ladder, meas["encode_wave_ms"], meas["encode_lane_ms"]),
            "parse_wave_max_planes": crossover(meas["planes"], meas["parse_wave_ms"], meas["parse_thread_ms"]),
        }
        # queue grain of the wavefront decoder, at the largest class it still wins (or the smallest)
        wins = [n for n in ladder if n <= chosen["ht_decode_wave_max_blocks"]]
        gw = windows[wins[-1] if wins else ladder[0]]
        _force("ht_decode_wave_max_blocks", True)
        _force("parse_wave_max_planes", False)
        for g in GRAINS:
            _pkg.tune_set("ht_wave_grain", g)
            meas["grain"][str(g)] = round(gw.decode_stages(reps)[1], 4)
        chosen["ht_wave_grain"] = int(min(GRAINS, key=lambda g: meas["grain"][str(g)]))
        del windows, gw
    except BaseException:
        for k, v in saved.items():
            _pkg.tune_set(k, v)
        raise
    apply(chosen)
    return {"knobs": chosen, "measurements": meas, "seconds": round(time.perf_counter() - t_start, 2),
            "library": _pkg.version()}


def apply(knobs: Dict) -> None:
    for k, v in knobs.items():
        _pkg.tune_set(k, int(v))


def load_or_tune(path: str, quick: bool = True) -> Dict:
    """Applies the tuning saved at path (same library version), else tunes and saves."""
    if os.path.exists(path):
        with open(path) as f:
            rec = json.load(f)
        if rec.get("library") == _pkg.version() and set(rec.get("knobs", {})) == set(KNOBS):
            apply(rec["knobs"])
            return rec
    rec = tune(quick=quick)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(rec, f, indent=1)
    return rec
