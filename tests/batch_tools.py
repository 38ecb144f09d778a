"""Device batch helpers for the GPU tests: run many codestreams or chunks
through one batch call and read every item's own outcome (its status, or
its decoded samples), as a caller that isolates per-item failures does."""

from __future__ import annotations

import ctypes as C
import importlib
from typing import List, Optional, Sequence, Tuple

import numpy as np

NDIC_UNSUPPORTED = 3


def reduced(n: int, levels: int) -> int:
    for _ in range(levels):
        n = (n + 1) // 2
    return n


def _arena(items: Sequence[bytes], align: int = 16) -> Tuple[bytes, List[int]]:
    offs, parts, o = [], [], 0
    for b in items:
        offs.append(o)
        pad = (-len(b)) % align
        parts.append(bytes(b) + bytes(pad))
        o += len(b) + pad
    return b"".join(parts) or b"\0", offs


def plane_batch(ndic, streams: Sequence[bytes], width: int, height: int, levels: int, max_res: int = -1,
                prefix: bool = False, irreversible: bool = False) -> List[Tuple[int, Optional[np.ndarray]]]:
    """Decodes every codestream (or prefix) with one ndic_htj2k_decode_plane_batch
    call (..._irreversible: 9/7 codestreams too); returns (status, int32
    samples or None) per stream."""
    serve = importlib.import_module("nd-image-codecs_amd.serve")
    dev = importlib.import_module("nd-image-codecs_amd.device")
    lib = serve._lib()
    blob, offs = _arena(streams)
    d_in = dev.DeviceBuffer.from_bytes(blob)
    each = (width * height * 4 + 255) // 256 * 256
    d_out = dev.DeviceBuffer(max(1, each * len(streams)), zero=True)
    arr = (serve.PlaneDesc * len(streams))()
    for i, s in enumerate(streams):
        d = arr[i]
        d.in_offset, d.in_len, d.out_offset = offs[i], len(s), i * each
        d.width, d.height, d.levels = width, height, levels
        d.max_res, d.prefix = max_res, int(prefix)
        d.status = -1
    err = ndic.Error()
    fn = lib.ndic_htj2k_decode_plane_batch_irreversible if irreversible else lib.ndic_htj2k_decode_plane_batch
    rc = fn(C.c_void_p(d_in.ptr), C.c_void_p(d_out.ptr), arr, len(streams), None, C.byref(err))
    if rc != 0 and all(arr[i].status in (0, -1) for i in range(len(streams))):
        ndic._check(rc, err)  # a call-level failure, not a per-item one
    out = d_out.download()
    res = []
    for i in range(len(streams)):
        d = arr[i]
        if d.status != 0:
            res.append((int(d.status), None))
            continue
        n = d.out_width * d.out_height
        a = np.frombuffer(out[i * each:i * each + 4 * n], np.int32).reshape(d.out_height, d.out_width).copy()
        res.append((0, a))
    return res


def chunk_batch(ndic, chunks: Sequence[bytes], shape, dtype: str, steps, levels: int
                ) -> List[Tuple[int, Optional[bytes]]]:
    """Decodes every nd-lift-ht chunk with one ndic_lift_ht_decode_batch call;
    returns (status, decoded bytes or None) per chunk."""
    dev = importlib.import_module("nd-image-codecs_amd.device")
    blob, offs = _arena(chunks)
    d_in = dev.DeviceBuffer.from_bytes(blob)
    nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
    each = (nbytes + 255) // 256 * 256
    d_out = dev.DeviceBuffer(max(1, each * len(chunks)), zero=True)
    arr = (ndic.ChunkDesc * len(chunks))(*[
        ndic.make_desc(shape, dtype, steps, levels, offs[i], len(c), i * each) for i, c in enumerate(chunks)])
    for d in arr:
        d.status = -1
    err = ndic.Error()
    rc = ndic.lib().ndic_lift_ht_decode_batch(C.c_void_p(d_in.ptr), C.c_void_p(d_out.ptr), arr, len(arr), 0,
                                              None, C.byref(err))
    if rc != 0 and all(d.status in (0, -1) for d in arr):
        ndic._check(rc, err)
    out = d_out.download()
    return [(int(d.status), out[i * each:i * each + nbytes] if d.status == 0 else None) for i, d in enumerate(arr)]
