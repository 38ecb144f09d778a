"""Serving loop over an HBM-resident store of compressed chunks.

Requests arrive over time (thumbnail, full-chunk decode, ZFP brick, encode)
and go to one of three lanes by kind: small reads (thumbnails and bricks),
full-chunk decodes, encodes.  Each lane has one or more worker threads, each
of which the library gives its own HIP stream, so their work overlaps on the
device.  Each pass of a worker admits every request of its lane that has
arrived, as far as its output memory allows, and runs them as device
batches: thumbnails
through ``ndic_htj2k_decode_plane_batch`` (plane 0 cut at the ndht index's
``prefix[r]``, resolutions 0..r), bricks through ``ndic_zfp_decode_bricks``,
full decodes through ``ndic_lift_ht_decode_batch``, encodes through
``ndic_lift_ht_encode_batch``.  Requests that arrive while a pass is on the
GPU join the lane's next pass (continuous batching); latency is the time
from a request's arrival to the end of the pass that served it.

Failures stay with their request, as the reference's per-chunk error contract
has it (crates/ndic-zarr/tests/htj2k_zarrs.rs:257: a malformed chunk errors
cleanly): a request whose output could never fit its lane is refused alone,
and after each device batch every request takes its own descriptor's status,
so a corrupt chunk fails while the other requests of its pass are served.
Admission is by device memory: a pass starts only when the device memory it
may newly need (library scratch beyond what its worker already holds, encode
output growth) fits the free HBM measured through the library, less a
reserve and less what passes in flight on other workers have claimed;
otherwise it waits for memory to free up (back-pressure).

The thumbnail byte ranges follow ``RangeIndex::thumbnail``
(crates/ndic-codestream/src/range.rs:128-200): plane 0's codestream up to
the end of its resolution-r packets, read from the chunk's own index.
"""

from __future__ import annotations

import ctypes as C
import importlib
import queue
import struct
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

_pkg = importlib.import_module("nd-image-codecs_amd")
dev = importlib.import_module("nd-image-codecs_amd.device")

THUMB, DECODE, BRICK, ENCODE = "thumbnail", "decode", "brick", "encode"
STATUS_NAMES = {1: "InvalidArgument", 2: "Codestream", 3: "Unsupported", 4: "Io"}
# device scratch the library's windows take per decoded / encoded sample
# (int32 coefficient, temporary and reconstruction planes; encode adds the
# code-block slots), used to size a pass's claim on free HBM
DECODE_SCRATCH_PER_SAMPLE = 12
ENCODE_SCRATCH_PER_SAMPLE = 20


class PlaneDesc(C.Structure):
    _fields_ = [("in_offset", C.c_uint64), ("in_len", C.c_uint64), ("out_offset", C.c_uint64),
                ("width", C.c_uint32), ("height", C.c_uint32), ("levels", C.c_uint32), ("max_res", C.c_int32),
                ("prefix", C.c_int32), ("out_width", C.c_uint32), ("out_height", C.c_uint32), ("status", C.c_int32)]


class BrickDesc(C.Structure):
    _fields_ = [("in_offset", C.c_uint64), ("in_len", C.c_uint64), ("out_offset", C.c_uint64), ("ndim", C.c_uint32),
                ("shape", C.c_uint32 * 4), ("brick", C.c_uint32 * 4), ("dtype", C.c_int32), ("rate", C.c_double),
                ("status", C.c_int32)]


_bound = False


def _lib():
    global _bound
    lib = _pkg.lib()
    if not _bound:
        P, S, E = C.c_void_p, C.c_size_t, C.POINTER(_pkg.Error)
        lib.ndic_htj2k_decode_plane_batch.argtypes = [P, P, P, S, P, E]
        lib.ndic_htj2k_decode_plane_batch.restype = C.c_int
        lib.ndic_htj2k_decode_plane_batch_irreversible.argtypes = [P, P, P, S, P, E]
        lib.ndic_htj2k_decode_plane_batch_irreversible.restype = C.c_int
        lib.ndic_zfp_decode_bricks.argtypes = [P, P, P, S, P, E]
        lib.ndic_zfp_decode_bricks.restype = C.c_int
        _bound = True
    return lib


def reduced(n: int, levels: int) -> int:
    for _ in range(levels):
        n = (n + 1) // 2
    return n


@dataclass
class HtChunk:
    """A compressed nd-lift-ht chunk in the store."""
    offset: int
    length: int
    shape: tuple
    dtype: str
    steps: list
    levels: int
    raw_offset: int = -1           # where the original samples live (for checks), if kept
    plane0: tuple = (0, 0, [])     # (codestream offset in the chunk, length, prefix[r])


@dataclass
class ZfpField:
    offset: int
    length: int
    shape: tuple
    dtype: str
    rate: float


@dataclass
class Request:
    kind: str
    arrival: float
    key: int = 0
    brick: tuple = ()
    res: int = 3
    out_bytes: int = 0
    done: float = -1.0
    out_offset: int = -1
    out_buf: object = None
    out_len: int = 0              # encode: the chunk's length at out_offset
    status: int = 0               # ndic_status of the request (0: served)
    error: Optional[str] = None   # why it failed, when it did


def parse_plane0(head: bytes):
    """Plane 0's entry of an ndht coefficient-plane index
    (crates/ndic-codestream/src/container.rs:14-30): (offset, len, prefix[])."""
    if head[:4] != b"ndht":
        raise ValueError("not an ndht chunk")
    flags, levels, ndim = head[5], head[6], head[7]
    if not flags & 1:
        raise ValueError("ndht chunk has no coefficient-plane index")
    pos = 8 + 4 * ndim
    off, ln = struct.unpack_from("<QI", head, pos)
    pre = list(struct.unpack_from("<%dI" % (levels + 1), head, pos + 12))
    return off, ln, pre


class Store:
    """Compressed chunks kept in device memory for the server."""

    def __init__(self):
        self.ht: List[HtChunk] = []
        self.zfp: List[ZfpField] = []
        self.ht_arena: Optional[dev.DeviceBuffer] = None
        self.zfp_arena: Optional[dev.DeviceBuffer] = None
        self.raw_arena: Optional[dev.DeviceBuffer] = None

    def add_htj2k_arena(self, arena, descs, specs, raw_arena=None, raw_offsets=None):
        """Adopt an arena written by ndic_lift_ht_encode_batch (descs[i].out_offset/out_len);
        specs[i] = (shape, dtype, steps, levels)."""
        self.ht_arena = arena
        self.raw_arena = raw_arena
        for i, d in enumerate(descs):
            shape, dt, steps, lv = specs[i][:4]
            need = 8 + 4 * len(shape) + 12 + 4 * (lv + 1)
            head = arena.download(d.out_offset, min(d.out_len, need))
            self.ht.append(HtChunk(d.out_offset, d.out_len, tuple(shape), dt, steps, lv,
                                   raw_offsets[i] if raw_offsets is not None else -1, parse_plane0(head)))

    def add_zfp_arena(self, arena, fields: Sequence[ZfpField]):
        self.zfp_arena = arena
        self.zfp.extend(fields)


@dataclass
class Stats:
    latencies_ms: Dict[str, List[float]] = field(default_factory=dict)
    passes: int = 0
    bytes_out: int = 0
    wall_s: float = 0.0
    call_ms: Dict[str, float] = field(default_factory=dict)   # device-batch call time per kind
    calls: Dict[str, int] = field(default_factory=dict)
    items: Dict[str, int] = field(default_factory=dict)       # requests per kind over all calls
    pass_ms: List[float] = field(default_factory=list)        # wall time of each pass
    lane_passes: Dict[str, int] = field(default_factory=dict)
    failed: List[Request] = field(default_factory=list)       # requests that ended with an error
    waits: int = 0                                            # passes held back for device memory
    lock: threading.Lock = field(default_factory=threading.Lock)

    def record(self, kind: str, ms: float, n: int = 1):
        with self.lock:
            self.call_ms[kind] = self.call_ms.get(kind, 0.0) + ms
            self.calls[kind] = self.calls.get(kind, 0) + 1
            self.items[kind] = self.items.get(kind, 0) + n

    def summary(self):
        allv = sorted(v for vs in self.latencies_ms.values() for v in vs)

        def pct(v, q):
            if not v:
                return None
            return round(float(np.percentile(np.array(v), q)), 3)

        fails: Dict[str, int] = {}
        for r in self.failed:
            fails[r.kind] = fails.get(r.kind, 0) + 1
        out = {"requests": len(allv), "failed": fails, "memory_waits": self.waits,
               "passes": self.passes, "lane_passes": dict(self.lane_passes),
               "p50_ms": pct(allv, 50), "p99_ms": pct(allv, 99),
               "GBps": round(self.bytes_out / self.wall_s / 1e9, 3) if self.wall_s else None,
               "pass_ms": {"p50": pct(self.pass_ms, 50), "p99": pct(self.pass_ms, 99)},
               "by_kind": {k: {"n": len(v), "p50_ms": pct(sorted(v), 50), "p99_ms": pct(sorted(v), 99),
                               "mean_call_ms": round(self.call_ms.get(k, 0.0) / max(1, self.calls.get(k, 0)), 3),
                               "mean_batch": round(self.items.get(k, 0) / max(1, self.calls.get(k, 0)), 2)}
                           for k, v in self.latencies_ms.items()}}
        return out


class HbmGate:
    """Admission by device memory.  A pass claims the device memory it may
    newly need before it starts; the claim is granted when it fits the free
    HBM measured now (through the library), less a reserve kept for the rest
    of the process and less what passes in flight have claimed.  A claim is
    released when its pass ends (the memory it made the library allocate then
    shows in the measured free HBM instead)."""

    def __init__(self, reserve: int):
        self.reserve = int(reserve)
        self.claimed = 0
        self.cv = threading.Condition()
        self.total = _pkg.device_memory()[1]

    def free(self) -> int:
        return _pkg.device_memory()[0]

    def try_claim(self, need: int) -> bool:
        if need <= 0:
            return True
        with self.cv:
            if self.free() - self.reserve - self.claimed >= need:
                self.claimed += need
                return True
            return False

    def release(self, n: int) -> None:
        if n > 0:
            with self.cv:
                self.claimed -= n
                self.cv.notify_all()

    def wait(self, timeout: float) -> None:
        with self.cv:
            self.cv.wait(timeout)


class _Worker:
    """A lane worker's device state: its output memory and the high-water
    mark of the library scratch its thread context has grown to."""

    def __init__(self, lane: str, out: Optional["dev.DeviceBuffer"]):
        self.lane = lane
        self.out = out
        self.scratch_hw = 0
        self.enc_out: Optional["dev.DeviceBuffer"] = None


class Server:
    """Continuous-batching request loop on one GPU.

    Requests are routed by kind to three lanes: small reads (thumbnails and
    ZFP bricks), full-chunk decodes, and encodes.  A lane has `workers[lane]`
    host threads, each with its own library context (and so its own HIP
    stream) and output memory, taking requests from the lane's queue.  A
    worker runs passes: each admits every request of its lane that has
    arrived, as far as its output memory and the free device memory allow,
    and runs them as one device batch per kind.  All workers' work overlaps
    on the device, so a thumbnail never waits behind a large chunk decode or
    an encode, and an encode that arrives during another starts at once.
    """

    def __init__(self, store: Store, out_budget: int = 4 << 30, encode_sources=None, small_budget: int = 64 << 20,
                 workers: Optional[Dict[str, int]] = None, hbm_reserve: int = 2 << 30):
        self.store = store
        self.out_budget = int(out_budget)      # full-chunk decodes, per worker
        self.small_budget = int(small_budget)  # thumbnails and bricks, per worker
        # worker threads per lane: each has its own library context (stream) and output memory
        self.workers = dict(workers or {"small": 1, "bulk": 2, "encode": 2})
        self.outs = {"bulk": [dev.DeviceBuffer(self.out_budget) for _ in range(self.workers["bulk"])],
                     "small": [dev.DeviceBuffer(self.small_budget) for _ in range(self.workers["small"])]}
        self.out = self.outs["bulk"][0]
        # encode requests take samples from these (device buffer, shape, dtype, steps, levels) sources
        self.encode_sources = encode_sources or []
        self.gate = HbmGate(hbm_reserve)

    # -- request sizing
    def out_bytes(self, r: Request) -> int:
        s = self.store
        if r.kind == THUMB:
            c = s.ht[r.key]
            h, w = c.shape[-2], c.shape[-1]
            rl = c.levels - min(r.res, c.levels)
            return reduced(w, rl) * reduced(h, rl) * 4
        if r.kind == DECODE:
            c = s.ht[r.key]
            return int(np.prod(c.shape)) * np.dtype(c.dtype).itemsize
        if r.kind == BRICK:
            f = s.zfp[r.key]
            n = 1
            for d, b in zip(f.shape, r.brick):
                n *= max(0, min(4, d - 4 * b))
            return n * np.dtype(f.dtype).itemsize
        if r.kind == ENCODE:
            return r.out_len
        return 0

    def _encode_source(self, r: Request):
        return self.encode_sources[r.key % len(self.encode_sources)]

    def _scratch(self, r: Request) -> int:
        """Library device scratch a request's pass may need (an upper bound)."""
        s = self.store
        if r.kind == THUMB:
            c = s.ht[r.key]
            return c.shape[-1] * c.shape[-2] * DECODE_SCRATCH_PER_SAMPLE
        if r.kind == DECODE:
            return int(np.prod(s.ht[r.key].shape)) * DECODE_SCRATCH_PER_SAMPLE
        if r.kind == ENCODE and self.encode_sources:
            return int(np.prod(self._encode_source(r)[1])) * ENCODE_SCRATCH_PER_SAMPLE
        return 0

    def _lane_capacity(self, lane: str) -> int:
        return self.out_budget if lane == "bulk" else self.small_budget if lane == "small" else 0

    def buffer_of(self, r: Request) -> "dev.DeviceBuffer":
        return r.out_buf

    def result_ptr(self, r: Request) -> int:
        """Device address of a served request's output (valid until its lane's next pass)."""
        return self.buffer_of(r).ptr + r.out_offset

    def result(self, r: Request) -> bytes:
        if r.status != 0:
            raise _pkg.NdicError(r.status, 0, r.error or "request failed")
        return self.buffer_of(r).download(r.out_offset, self.out_bytes(r))

    # -- device batches; each request takes its own descriptor's status
    @staticmethod
    def _settle(reqs: List[Request], descs, exc: Optional[BaseException]) -> None:
        """Per-request outcome of one batch call.  Descriptors start at status
        -1: one still at -1 after a failed call was never reached (the call
        failed as a whole), so its request takes the call's error."""
        msg = str(exc) if exc is not None else ""
        for r, d in zip(reqs, descs):
            st = int(d.status)
            if st == 0:
                continue
            if st < 0:
                st = getattr(exc, "status", 4) or 4
            r.status = st
            r.error = f"{STATUS_NAMES.get(st, st)}: " + (msg if msg else "failed in its batch")

    @staticmethod
    def _call(fn) -> Optional[BaseException]:
        try:
            fn()
        except _pkg.NdicError as e:
            return e
        return None

    def _thumbs(self, thumbs: List[Request], out: "dev.DeviceBuffer", stats: Optional[Stats]):
        s, lib = self.store, _lib()
        t0 = time.perf_counter()
        arr = (PlaneDesc * len(thumbs))()
        for i, r in enumerate(thumbs):
            c = s.ht[r.key]
            off, ln, pre = c.plane0
            res = min(r.res, c.levels)
            d = arr[i]
            d.in_offset = c.offset + off
            d.in_len = pre[res] if pre else ln
            d.out_offset = r.out_offset
            d.width, d.height, d.levels = c.shape[-1], c.shape[-2], c.levels
            d.max_res, d.prefix = res, 1
            d.status = -1
        err = _pkg.Error()
        exc = self._call(lambda: _pkg._check(lib.ndic_htj2k_decode_plane_batch(
            C.c_void_p(s.ht_arena.ptr), C.c_void_p(out.ptr), arr, len(thumbs), None, C.byref(err)), err))
        self._settle(thumbs, arr, exc)
        if stats:
            stats.record(THUMB, (time.perf_counter() - t0) * 1e3, len(thumbs))

    def _bricks(self, bricks: List[Request], out: "dev.DeviceBuffer", stats: Optional[Stats]):
        s, lib = self.store, _lib()
        t0 = time.perf_counter()
        arr = (BrickDesc * len(bricks))()
        for i, r in enumerate(bricks):
            f = s.zfp[r.key]
            d = arr[i]
            d.in_offset, d.in_len, d.out_offset = f.offset, f.length, r.out_offset
            d.ndim = len(f.shape)
            for a in range(len(f.shape)):
                d.shape[a] = f.shape[a]
                d.brick[a] = r.brick[a]
            d.dtype = _pkg._zfp_dtype(f.dtype)
            d.rate = f.rate
            d.status = -1
        err = _pkg.Error()
        exc = self._call(lambda: _pkg._check(lib.ndic_zfp_decode_bricks(
            C.c_void_p(s.zfp_arena.ptr), C.c_void_p(out.ptr), arr, len(bricks), None, C.byref(err)), err))
        self._settle(bricks, arr, exc)
        if stats:
            stats.record(BRICK, (time.perf_counter() - t0) * 1e3, len(bricks))

    def _decodes(self, decs: List[Request], out: "dev.DeviceBuffer", stats: Optional[Stats]):
        s = self.store
        t0 = time.perf_counter()
        arr = (_pkg.ChunkDesc * len(decs))(*[
            _pkg.make_desc(s.ht[r.key].shape, s.ht[r.key].dtype, s.ht[r.key].steps, s.ht[r.key].levels,
                           s.ht[r.key].offset, s.ht[r.key].length, r.out_offset) for r in decs])
        for d in arr:
            d.status = -1
        exc = self._call(lambda: _pkg.decode_batch(s.ht_arena.ptr, out.ptr, arr))
        self._settle(decs, arr, exc)
        if stats:
            stats.record(DECODE, (time.perf_counter() - t0) * 1e3, len(decs))

    def _encode_cap(self, encs: List[Request]) -> int:
        total = sum(int(np.prod(self._encode_source(r)[1])) * np.dtype(self._encode_source(r)[2]).itemsize
                    for r in encs)
        return 2 * total + (64 << 20)

    def _run_encodes(self, encs: List[Request], w: _Worker, stats: Optional[Stats] = None):
        """Encodes each request's source chunk; the chunks land packed in the
        worker's encode output (one batch call per source)."""
        t0 = time.perf_counter()
        if not self.encode_sources:
            for r in encs:
                r.status, r.error = 1, "InvalidArgument: the server has no encode sources"
            return
        by_src: Dict[int, List[Request]] = {}
        for r in encs:
            by_src.setdefault(self._encode_source(r)[0].ptr, []).append(r)
        cap = self._encode_cap(encs)
        if w.enc_out is None or w.enc_out.nbytes < cap:
            w.enc_out = None
            w.enc_out = dev.DeviceBuffer(cap)
        cursor = 0
        for ptr, rs in by_src.items():
            _, shape, dt, steps, lv = self._encode_source(rs[0])
            n = int(np.prod(shape)) * np.dtype(dt).itemsize
            arr = (_pkg.ChunkDesc * len(rs))(*[_pkg.make_desc(shape, dt, steps, lv, 0, n, 0) for _ in rs])
            for d in arr:
                d.status = -1
            exc = self._call(lambda: _pkg.encode_batch(ptr, w.enc_out.ptr + cursor, w.enc_out.nbytes - cursor, arr))
            self._settle(rs, arr, exc)
            for r, d in zip(rs, arr):
                r.out_buf, r.out_offset, r.out_len = w.enc_out, cursor + int(d.out_offset), int(d.out_len)
            cursor += max((int(d.out_offset + d.out_len) for d in arr if d.status == 0), default=0)
            cursor = (cursor + 255) // 256 * 256
        if stats:
            stats.record(ENCODE, (time.perf_counter() - t0) * 1e3, len(encs))

    def _run_pass(self, w: _Worker, batch: List[Request], stats: Optional[Stats]):
        """One pass of a lane worker into its output memory; every call
        returns with its work done on the calling thread's stream."""
        cursor = 0
        if w.lane != "encode":
            for r in batch:
                r.out_offset = cursor
                r.out_buf = w.out
                cursor += (self.out_bytes(r) + 255) // 256 * 256
        if w.lane == "small":
            thumbs = [r for r in batch if r.kind == THUMB]
            bricks = [r for r in batch if r.kind == BRICK]
            if thumbs:
                self._thumbs(thumbs, w.out, stats)
            if bricks:
                self._bricks(bricks, w.out, stats)
        elif w.lane == "bulk":
            self._decodes(batch, w.out, stats)
        else:
            self._run_encodes(batch, w, stats)

    # -- admission
    def _admit(self, w: _Worker, pending: List[Request]):
        """-> (batch, refused, claim): the requests this pass serves (a prefix
        of the pending ones, in arrival order), the requests refused alone
        because they could never fit (their output exceeds the lane's memory,
        or they need more device memory than the GPU has to give), and the
        bytes of free HBM claimed for the pass.  An empty batch with nothing
        refused means the pass waits for device memory."""
        cap = self._lane_capacity(w.lane)
        batch: List[Request] = []
        refused: List[Request] = []
        used = 0
        for r in pending:
            if w.lane != "encode":
                b = (self.out_bytes(r) + 255) // 256 * 256
                if b > cap:
                    r.status, r.error = 1, "InvalidArgument: the request's output exceeds its lane's output memory"
                    refused.append(r)
                    continue
                if used + b > cap:
                    break
                used += b
            if self._scratch(r) > self.gate.total - self.gate.reserve:
                r.status, r.error = 4, "Io: the request needs more device memory than the GPU has"
                refused.append(r)
                continue
            batch.append(r)
        # the pass's claim on free HBM; shrink the pass until it fits
        while batch:
            need = max(0, sum(self._scratch(r) for r in batch) - w.scratch_hw)
            if w.lane == "encode":
                have = w.enc_out.nbytes if w.enc_out is not None else 0
                need += max(0, self._encode_cap(batch) - have)
            if self.gate.try_claim(need):
                return batch, refused, need
            if len(batch) == 1:
                break
            batch = batch[:max(1, len(batch) // 2)]
        return [], refused, 0

    # -- lanes: persistent host threads (each keeps its library context warm)
    _ROUTE = {THUMB: "small", BRICK: "small", DECODE: "bulk", ENCODE: "encode"}

    def _lanes(self):
        if getattr(self, "_lane_q", None) is None:
            self._lane_q = {k: queue.Queue() for k in ("small", "bulk", "encode")}
            self._lane_th = []
            for k, q in self._lane_q.items():
                for i in range(self.workers[k]):
                    w = _Worker(k, self.outs[k][i] if k in self.outs else None)
                    th = threading.Thread(target=self._lane_loop, args=(w, q), daemon=True)
                    th.start()
                    self._lane_th.append(th)
        return self._lane_q

    def close(self):
        """Stops the lane threads (their streams drain first)."""
        if getattr(self, "_lane_q", None) is not None:
            for k, q in self._lane_q.items():
                for _ in range(self.workers[k]):
                    q.put(("stop", None))
            for th in self._lane_th:
                th.join()
            self._lane_q = None

    def _finish(self, run, w: _Worker, batch: List[Request], refused: List[Request], t_pass: Optional[float]):
        stats = run["stats"]
        t = time.perf_counter() - run["t0"]
        with stats.lock:
            if t_pass is not None:
                stats.pass_ms.append((time.perf_counter() - t_pass) * 1e3)
                stats.passes += 1
                stats.lane_passes[w.lane] = stats.lane_passes.get(w.lane, 0) + 1
            for r in list(batch) + list(refused):
                r.done = t
                if r.status != 0:
                    stats.failed.append(r)
                    continue
                stats.latencies_ms.setdefault(r.kind, []).append((t - r.arrival) * 1e3)
                stats.bytes_out += self.out_bytes(r)

    def _lane_loop(self, w: _Worker, q: "queue.Queue"):
        """One worker of a lane.  Workers of a lane share its queue; a worker
        stops taking messages after the first end-of-run token it receives
        (every request of the run was queued before the tokens), finishes its
        pending requests, drains its stream and signals the token."""
        pending: List[Request] = []
        end_ev: Optional[threading.Event] = None

        def take(msg) -> bool:
            nonlocal end_ev
            kind, val = msg
            if kind == "req":
                pending.append(val)
            elif kind == "end":
                end_ev = val
            else:
                return False
            return True

        while True:
            if not pending and end_ev is None:
                if not take(q.get()):
                    return
            while end_ev is None:
                try:
                    msg = q.get_nowait()
                except queue.Empty:
                    break
                if not take(msg):
                    return
            if pending:
                run = self._run_ctx
                try:
                    batch, refused, claim = self._admit(w, pending)
                    gone = {id(r) for r in batch} | {id(r) for r in refused}
                    pending = [r for r in pending if id(r) not in gone]
                    if refused:
                        self._finish(run, w, [], refused, None)
                    if not batch:
                        if not refused:
                            with run["stats"].lock:
                                run["stats"].waits += 1
                            self.gate.wait(0.002)  # back-pressure: wait for device memory to free up
                        continue
                    try:
                        tp = time.perf_counter()
                        self._run_pass(w, batch, run["stats"])
                        w.scratch_hw = max(w.scratch_hw, sum(self._scratch(r) for r in batch))
                    finally:
                        self.gate.release(claim)
                    self._finish(run, w, batch, [], tp)
                    served = [r for r in batch if r.status == 0]
                    if run["verify"] is not None and served:
                        run["verify"](self, served)
                except BaseException as e:
                    run["errors"].append(e)
            if not pending and end_ev is not None:
                try:
                    dev.synchronize()  # work the verify hook queued on this worker's stream
                except BaseException as e:
                    self._run_ctx["errors"].append(e)
                end_ev.set()
                end_ev = None

    def run(self, requests: List[Request], verify=None) -> Stats:
        """Serves requests (sorted by arrival, seconds from start) in real time
        and returns when every one is done and the lanes' streams have
        drained.  Requests that fail carry their status and error and are
        listed in Stats.failed; the others are served.  verify(server,
        served) runs on the lane's thread after each pass, before the lane
        reuses its output memory.  Raises only for a failure of the loop
        itself (an exception that is not a request's own error)."""
        lanes = self._lanes()
        stats = Stats()
        for r in requests:
            r.status, r.error, r.done = 0, None, -1.0
        self._run_ctx = {"stats": stats, "verify": verify, "t0": time.perf_counter(), "errors": []}
        t0 = self._run_ctx["t0"]
        done = []
        try:
            for r in requests:
                wait = r.arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
                lanes[self._ROUTE[r.kind]].put(("req", r))
        finally:
            for k, q in lanes.items():
                for _ in range(self.workers[k]):
                    ev = threading.Event()
                    done.append(ev)
                    q.put(("end", ev))
            for ev in done:
                ev.wait()
        stats.wall_s = time.perf_counter() - t0
        if self._run_ctx["errors"]:
            raise self._run_ctx["errors"][0]
        return stats


def poisson_requests(rng: np.random.Generator, rate: float, duration: float, store: Store,
                     mix=(0.60, 0.25, 0.10, 0.05), thumb_res: int = 3, n_encode_sources: int = 1) -> List[Request]:
    """Poisson arrivals at `rate` req/s for `duration` s with the given kind mix."""
    n = int(rng.poisson(rate * duration))
    arrivals = np.sort(rng.uniform(0, duration, n))
    kinds = rng.choice(4, size=n, p=np.array(mix) / sum(mix))
    out = []
    for t, k in zip(arrivals, kinds):
        if k == 2 and not store.zfp:
            k = 0
        if k == 3 and n_encode_sources == 0:
            k = 0
        if k == 0:
            out.append(Request(THUMB, float(t), key=int(rng.integers(0, len(store.ht))), res=thumb_res))
        elif k == 1:
            out.append(Request(DECODE, float(t), key=int(rng.integers(0, len(store.ht)))))
        elif k == 2:
            key = int(rng.integers(0, len(store.zfp)))
            f = store.zfp[key]
            out.append(Request(BRICK, float(t), key=key,
                               brick=tuple(int(rng.integers(0, (d + 3) // 4)) for d in f.shape)))
        else:
            out.append(Request(ENCODE, float(t), key=int(rng.integers(0, max(1, n_encode_sources)))))
    return out
