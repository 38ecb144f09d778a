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

        out = {"requests": len(allv), "passes": self.passes, "lane_passes": dict(self.lane_passes),
               "p50_ms": pct(allv, 50), "p99_ms": pct(allv, 99),
               "GBps": round(self.bytes_out / self.wall_s / 1e9, 3) if self.wall_s else None,
               "pass_ms": {"p50": pct(self.pass_ms, 50), "p99": pct(self.pass_ms, 99)},
               "by_kind": {k: {"n": len(v), "p50_ms": pct(sorted(v), 50), "p99_ms": pct(sorted(v), 99),
                               "mean_call_ms": round(self.call_ms.get(k, 0.0) / max(1, self.calls.get(k, 0)), 3),
                               "mean_batch": round(self.items.get(k, 0) / max(1, self.calls.get(k, 0)), 2)}
                           for k, v in self.latencies_ms.items()}}
        return out


class Server:
    """Continuous-batching request loop on one GPU.

    Requests are routed by kind to three lanes: small reads (thumbnails and
    ZFP bricks), full-chunk decodes, and encodes.  A lane has `workers[lane]`
    host threads, each with its own library context (and so its own HIP
    stream) and output memory, taking requests from the lane's queue.  A
    worker runs passes: each admits every request of its lane that has
    arrived, as far as its output memory allows, and runs them as one device
    batch per kind.  All workers' work overlaps on the device, so a thumbnail
    never waits behind a large chunk decode or an encode, and an encode that
    arrives during another starts at once.
    """

    def __init__(self, store: Store, out_budget: int = 4 << 30, encode_sources=None, small_budget: int = 64 << 20,
                 workers: Optional[Dict[str, int]] = None):
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
        self.enc_out: Dict[int, dev.DeviceBuffer] = {}

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
                n *= min(4, d - 4 * b)
            return n * np.dtype(f.dtype).itemsize
        return 0

    def buffer_of(self, r: Request) -> "dev.DeviceBuffer":
        return r.out_buf

    def result_ptr(self, r: Request) -> int:
        """Device address of a served request's output (valid until its lane's next pass)."""
        return self.buffer_of(r).ptr + r.out_offset

    def result(self, r: Request) -> bytes:
        return self.buffer_of(r).download(r.out_offset, self.out_bytes(r))

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
        err = _pkg.Error()
        _pkg._check(lib.ndic_htj2k_decode_plane_batch(C.c_void_p(s.ht_arena.ptr), C.c_void_p(out.ptr), arr,
                                                      len(thumbs), None, C.byref(err)), err)
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
        err = _pkg.Error()
        _pkg._check(lib.ndic_zfp_decode_bricks(C.c_void_p(s.zfp_arena.ptr), C.c_void_p(out.ptr), arr,
                                               len(bricks), None, C.byref(err)), err)
        if stats:
            stats.record(BRICK, (time.perf_counter() - t0) * 1e3, len(bricks))

    def _decodes(self, decs: List[Request], out: "dev.DeviceBuffer", stats: Optional[Stats]):
        s = self.store
        t0 = time.perf_counter()
        descs = [_pkg.make_desc(s.ht[r.key].shape, s.ht[r.key].dtype, s.ht[r.key].steps, s.ht[r.key].levels,
                                s.ht[r.key].offset, s.ht[r.key].length, r.out_offset) for r in decs]
        _pkg.decode_batch(s.ht_arena.ptr, out.ptr, descs)
        if stats:
            stats.record(DECODE, (time.perf_counter() - t0) * 1e3, len(decs))

    def _run_encodes(self, encs: List[Request], stats: Optional[Stats] = None):
        if not self.encode_sources:
            return
        t0 = time.perf_counter()
        by_src: Dict[int, list] = {}
        total = 0
        for r in encs:
            buf, shape, dt, steps, lv = self.encode_sources[r.key % len(self.encode_sources)]
            n = int(np.prod(shape)) * np.dtype(dt).itemsize
            by_src.setdefault(buf.ptr, []).append(_pkg.make_desc(shape, dt, steps, lv, 0, n, 0))
            total += n
        cap = 2 * total + (64 << 20)
        me = threading.get_ident()
        buf = self.enc_out.get(me)
        if buf is None or buf.nbytes < cap:
            buf = self.enc_out[me] = dev.DeviceBuffer(cap)
        for ptr, ds in by_src.items():
            _pkg.encode_batch(ptr, buf.ptr, buf.nbytes, ds)
        if stats:
            stats.record(ENCODE, (time.perf_counter() - t0) * 1e3, len(encs))

    def _run_pass(self, lane: str, out: Optional["dev.DeviceBuffer"], batch: List[Request], stats: Optional[Stats]):
        """One pass of a lane worker into its output memory; every call
        returns with its work done on the calling thread's stream."""
        cursor = 0
        for r in batch:
            r.out_offset = cursor
            r.out_buf = out
            cursor += (self.out_bytes(r) + 255) // 256 * 256
        if lane == "small":
            thumbs = [r for r in batch if r.kind == THUMB]
            bricks = [r for r in batch if r.kind == BRICK]
            if thumbs:
                self._thumbs(thumbs, out, stats)
            if bricks:
                self._bricks(bricks, out, stats)
        elif lane == "bulk":
            self._decodes(batch, out, stats)
        else:
            self._run_encodes(batch, stats)

    def _admit(self, lane: str, pending: List[Request]) -> List[Request]:
        if lane == "encode":
            return list(pending)
        budget = self.out_budget if lane == "bulk" else self.small_budget
        batch, used = [], 0
        for r in pending:
            b = (self.out_bytes(r) + 255) // 256 * 256
            if b > budget:
                raise ValueError("a request's output exceeds its lane's output budget")
            if batch and used + b > budget:
                break
            batch.append(r)
            used += b
        return batch

    # -- lanes: persistent host threads (each keeps its library context warm)
    _ROUTE = {THUMB: "small", BRICK: "small", DECODE: "bulk", ENCODE: "encode"}

    def _lanes(self):
        if getattr(self, "_lane_q", None) is None:
            self._lane_q = {k: queue.Queue() for k in ("small", "bulk", "encode")}
            self._lane_th = []
            for k, q in self._lane_q.items():
                for w in range(self.workers[k]):
                    out = self.outs[k][w] if k in self.outs else None
                    th = threading.Thread(target=self._lane_loop, args=(k, q, out), daemon=True)
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

    def _lane_loop(self, name: str, q: "queue.Queue", out: Optional["dev.DeviceBuffer"]):
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
                stats = run["stats"]
                try:
                    batch = self._admit(name, pending)
                except BaseException as e:
                    run["errors"].append(e)
                    batch = list(pending)
                    pending = []
                else:
                    pending = pending[len(batch):]
                    try:
                        tp = time.perf_counter()
                        self._run_pass(name, out, batch, stats)
                        t = time.perf_counter() - run["t0"]
                        with stats.lock:
                            stats.pass_ms.append((time.perf_counter() - tp) * 1e3)
                            stats.passes += 1
                            stats.lane_passes[name] = stats.lane_passes.get(name, 0) + 1
                            for r in batch:
                                r.done = t
                                stats.latencies_ms.setdefault(r.kind, []).append((t - r.arrival) * 1e3)
                                stats.bytes_out += self.out_bytes(r)
                        if run["verify"] is not None and name != "encode":
                            run["verify"](self, batch)
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
        drained.  verify(server, batch) runs on the lane's thread after each
        decode pass, before the lane reuses its output memory."""
        lanes = self._lanes()
        stats = Stats()
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
