"""MI355X-native nd-lift-ht / nd-zfp codec path (host-side Python mirror).

This module is the Python face of ``libndic_hip.so``: the same call shapes as
the reference's PyO3 binding (``bindings/python/nd-image-codecs/src/lib.rs:42-177``:
``htj2k_encode``/``htj2k_decode``) and its pure-NumPy ``nd_lift`` module
(``python/nd_image_codecs/_lift.py:325-359``: ``forward``/``inverse``), with
errors raised as ``ValueError`` like the binding's ``PyValueError``.  Every
call runs on the GPU; there is no CPU fallback, and importing the module
fails loudly when the HIP library is missing.

ZFP follows the binding's ``nd_zfp_encode``/``nd_zfp_decode``
(``lib.rs:117-166``) and ``ndic_zfp::decode_chunk_brick``.  The device-resident
batch entry points take device pointers (integers); ``.device`` provides
device buffers on the library's own HIP runtime.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Any, Sequence

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libndic_hip.so"

DTYPE_CODES = {"uint8": 0, "int8": 1, "uint16": 2, "int16": 3, "uint32": 4, "int32": 5, "uint64": 6, "int64": 7,
               "float32": 8, "float64": 9}
_ALIASES = {"|u1": "uint8", "u1": "uint8", "|i1": "int8", "i1": "int8", "<u2": "uint16", "u2": "uint16",
            "<i2": "int16", "i2": "int16", "<u4": "uint32", "u4": "uint32", "<i4": "int32", "i4": "int32",
            "<u8": "uint64", "u8": "uint64", "<i8": "int64", "i8": "int64", "<f4": "float32", "f4": "float32",
            "<f8": "float64", "f8": "float64"}
NP_OF = {0: np.uint8, 1: np.int8, 2: np.uint16, 3: np.int16, 4: np.uint32, 5: np.int32, 6: np.uint64, 7: np.int64,
         8: np.float32, 9: np.float64}
KINDS = {"delta": 0, "haar": 1, "lift53": 2}
PROGRESSIONS = {"LRCP": 0, "RLCP": 1, "RPCL": 2, "PCRL": 3, "CPRL": 4}
STATUS = {0: "Ok", 1: "InvalidArgument", 2: "Codestream", 3: "Unsupported", 4: "Io"}
MAX_NDIM = 8
MAX_STEPS = 4


class NdicError(ValueError):
    """A failed codec call; ``status`` is the ndic_core::Error variant."""

    def __init__(self, status: int, offset: int, message: str):
        super().__init__(message)
        self.status = status
        self.kind = STATUS.get(status, str(status))
        self.offset = offset


class Error(C.Structure):
    _fields_ = [("status", C.c_int32), ("offset", C.c_uint64), ("message", C.c_char * 256)]


class Buffer(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_uint8)), ("len", C.c_size_t)]


class LiftStep(C.Structure):
    _fields_ = [("dimension", C.c_uint32), ("kind", C.c_uint32), ("levels", C.c_uint32), ("group", C.c_uint32)]


class Htj2kConfig(C.Structure):
    _fields_ = [("xy_levels", C.c_uint32), ("reversible", C.c_int32), ("progression", C.c_int32),
                ("index", C.c_int32)]


class ZfpConfig(C.Structure):
    _fields_ = [("mode", C.c_int32), ("precision", C.c_uint32), ("rate", C.c_double), ("tolerance", C.c_double),
                ("dims", C.c_int32)]


class ChunkDesc(C.Structure):
    _fields_ = [("in_offset", C.c_uint64), ("in_len", C.c_uint64), ("out_offset", C.c_uint64),
                ("out_len", C.c_uint64), ("ndim", C.c_uint32), ("shape", C.c_uint32 * MAX_NDIM),
                ("dtype", C.c_int32), ("xy_levels", C.c_uint32), ("nsteps", C.c_uint32),
                ("steps", LiftStep * MAX_STEPS), ("status", C.c_int32)]


class ZfpDesc(C.Structure):
    _fields_ = [("in_offset", C.c_uint64), ("in_len", C.c_uint64), ("out_offset", C.c_uint64),
                ("out_len", C.c_uint64), ("ndim", C.c_uint32), ("shape", C.c_uint32 * MAX_NDIM),
                ("dtype", C.c_int32), ("config", ZfpConfig), ("status", C.c_int32),
                ("index_offset", C.c_uint64), ("index_len", C.c_uint64)]


class RangePlanC(C.Structure):
    _fields_ = [("target", C.c_int32), ("max_res", C.c_uint32), ("ndims", C.c_uint32),
                ("decoded_size", C.c_uint64 * 34), ("total_bytes", C.c_uint64), ("header_len", C.c_uint64),
                ("planes", Buffer), ("ranges", Buffer)]


class SelectDesc(C.Structure):
    _fields_ = [("in_offset", C.c_uint64), ("in_len", C.c_uint64), ("out_offset", C.c_uint64),
                ("ndim", C.c_uint32), ("shape", C.c_uint32 * MAX_NDIM), ("dtype", C.c_int32),
                ("xy_levels", C.c_uint32), ("target", C.c_int32), ("start", C.c_uint32 * MAX_NDIM),
                ("count", C.c_uint32 * MAX_NDIM), ("max_px", C.c_uint32), ("nsteps", C.c_uint32),
                ("steps", LiftStep * MAX_STEPS), ("max_res", C.c_uint32), ("out_ndim", C.c_uint32),
                ("out_shape", C.c_uint64 * MAX_NDIM), ("out_bytes", C.c_uint64), ("status", C.c_int32)]


ZFP_MODES = {"fixed_rate": 0, "fixed_accuracy": 1, "fixed_precision": 2, "reversible": 3}


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() (make -C nd-image-codecs_amd)")
    lib = C.CDLL(str(LIB_PATH))
    P, S, E, B = C.c_void_p, C.c_size_t, C.POINTER(Error), C.POINTER(Buffer)
    sigs = {
        "ndic_buffer_free": (None, [B]),
        "ndic_version": (C.c_char_p, []),
        "ndic_lift_forward_chunk": (C.c_int, [C.c_char_p, S, P, S, C.c_int, C.c_char_p, P, S, B, E]),
        "ndic_lift_inverse_chunk": (C.c_int, [C.c_char_p, S, P, S, C.c_int, C.c_char_p, P, S, B, E]),
        "ndic_htj2k_encode_chunk": (C.c_int, [C.c_char_p, S, P, S, C.c_int, C.POINTER(Htj2kConfig), B, E]),
        "ndic_htj2k_decode_chunk": (C.c_int, [C.c_char_p, S, P, S, C.c_int, B, E]),
        "ndic_htj2k_decode_plane": (C.c_int, [C.c_char_p, S, C.c_int32, C.c_int32, B, C.POINTER(C.c_size_t),
                                              C.POINTER(C.c_size_t), E]),
        "ndic_lift_ht_decode_batch": (C.c_int, [P, P, P, S, C.c_uint32, P, E]),
        "ndic_lift_ht_encode_batch": (C.c_int, [P, P, S, P, S, C.c_uint32, P, E]),
        "ndic_zfp_encode_chunk": (C.c_int, [C.c_char_p, S, P, S, C.c_int, C.POINTER(ZfpConfig), B, E]),
        "ndic_zfp_decode_chunk": (C.c_int, [C.c_char_p, S, P, S, C.c_int, C.POINTER(ZfpConfig), B, E]),
        "ndic_zfp_decode_brick": (C.c_int, [C.c_char_p, S, P, S, C.c_int, C.POINTER(ZfpConfig), P, S, B,
                                            P, E]),
        "ndic_zfp_encode_batch": (C.c_int, [P, P, S, P, S, P, E]),
        "ndic_zfp_decode_batch": (C.c_int, [P, P, P, S, P, E]),
        "ndic_zfp_encode_batch_indexed": (C.c_int, [P, P, S, P, S, P, S, P, E]),
        "ndic_zfp_decode_batch_indexed": (C.c_int, [P, P, P, P, S, P, E]),
        "ndic_range_plan_chunk": (C.c_int, [C.c_char_p, S, C.c_int32, C.c_uint64, P, S, C.POINTER(RangePlanC), E]),
        "ndic_htj2k_decode_select_batch": (C.c_int, [P, P, P, S, P, E]),
        "ndic_device_memory": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), E]),
        "ndic_zfp_decode_bricks": (C.c_int, [P, P, P, S, P, E]),
        "ndic_zfp_decode_bricks_noalloc": (C.c_int, [P, P, P, S, P, E]),
        "ndic_tune_set": (C.c_int32, [C.c_char_p, C.c_int64]),
        "ndic_tune_get": (C.c_int64, [C.c_char_p]),
        "ndic_tune_reset": (C.c_int32, [C.c_char_p]),
        "ndic_tune_name": (C.c_char_p, [C.c_int32]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = _load()


def lib() -> C.CDLL:
    return _lib


def _dtype_code(dtype: Any) -> int:
    if isinstance(dtype, (int, np.integer)):
        return int(dtype)
    name = np.dtype(dtype).name if not isinstance(dtype, str) else _ALIASES.get(dtype, dtype)
    if name not in DTYPE_CODES:
        raise NdicError(1, 0, f"unsupported dtype {dtype!r}")
    return DTYPE_CODES[name]


def _check(rc: int, err: Error) -> None:
    if rc != 0:
        raise NdicError(err.status, err.offset, err.message.decode(errors="replace"))


def _take(buf: Buffer) -> bytes:
    out = C.string_at(buf.data, buf.len) if buf.len else b""
    _lib.ndic_buffer_free(C.byref(buf))
    return out


def _shape(shape: Sequence[int]):
    return (C.c_size_t * max(1, len(shape)))(*[int(d) for d in shape]), len(shape)


def _steps(transforms: Sequence[dict]):
    arr = (LiftStep * max(1, len(transforms)))()
    for i, t in enumerate(transforms):
        arr[i].dimension = int(t["dimension"])
        arr[i].kind = KINDS[t["kind"]] if isinstance(t["kind"], str) else int(t["kind"])
        arr[i].levels = int(t.get("levels", 0))
        arr[i].group = int(t.get("group", 0))
    return arr, len(transforms)


def _as_bytes(chunk: Any) -> bytes:
    if isinstance(chunk, (bytes, bytearray, memoryview)):
        return bytes(chunk)
    return np.ascontiguousarray(chunk).tobytes()


# ---------------------------------------------------------------------------
# htj2k (bindings/python/nd-image-codecs/src/lib.rs:42-80)
# ---------------------------------------------------------------------------
def htj2k_encode(chunk: Any, shape: Sequence[int], dtype: Any, *, xy_levels: int = 5, reversible: bool = True,
                 progression: str = "RPCL", index: bool = True) -> bytes:
    if progression not in PROGRESSIONS:
        raise NdicError(3, 0, f"htj2k: unknown progression order {progression!r}")
    data = _as_bytes(chunk)
    cfg = Htj2kConfig(int(xy_levels), int(bool(reversible)), PROGRESSIONS[progression], int(bool(index)))
    sh, nd = _shape(shape)
    out, err = Buffer(), Error()
    rc = _lib.ndic_htj2k_encode_chunk(data, len(data), sh, nd, _dtype_code(dtype), C.byref(cfg), C.byref(out),
                                      C.byref(err))
    _check(rc, err)
    return _take(out)


def htj2k_decode(chunk: Any, shape: Sequence[int], dtype: Any) -> bytes:
    data = _as_bytes(chunk)
    sh, nd = _shape(shape)
    out, err = Buffer(), Error()
    rc = _lib.ndic_htj2k_decode_chunk(data, len(data), sh, nd, _dtype_code(dtype), C.byref(out), C.byref(err))
    _check(rc, err)
    return _take(out)


def htj2k_decode_plane(codestream: bytes, max_res: int = -1, prefix: bool = False) -> np.ndarray:
    """Decode resolutions 0..=max_res of one plane codestream (or a byte-range prefix of it)."""
    out, err = Buffer(), Error()
    w, h = C.c_size_t(), C.c_size_t()
    rc = _lib.ndic_htj2k_decode_plane(codestream, len(codestream), int(max_res), int(bool(prefix)), C.byref(out),
                                      C.byref(w), C.byref(h), C.byref(err))
    _check(rc, err)
    return np.frombuffer(_take(out), np.int32).reshape(h.value, w.value).copy()


# ---------------------------------------------------------------------------
# nd_lift (python/nd_image_codecs/_lift.py:325-359, crates/ndic-zarr/src/lift.rs:323-346)
# ---------------------------------------------------------------------------
def lift_forward_bytes(chunk: Any, shape: Sequence[int], dtype: Any, transforms: Sequence[dict],
                       version: str = "0.1") -> bytes:
    data = _as_bytes(chunk)
    sh, nd = _shape(shape)
    st, ns = _steps(transforms)
    out, err = Buffer(), Error()
    rc = _lib.ndic_lift_forward_chunk(data, len(data), sh, nd, _dtype_code(dtype), version.encode(), st, ns,
                                      C.byref(out), C.byref(err))
    _check(rc, err)
    return _take(out)


def lift_inverse_bytes(coeffs: Any, shape: Sequence[int], dtype: Any, transforms: Sequence[dict],
                       version: str = "0.1") -> bytes:
    data = _as_bytes(coeffs)
    sh, nd = _shape(shape)
    st, ns = _steps(transforms)
    out, err = Buffer(), Error()
    rc = _lib.ndic_lift_inverse_chunk(data, len(data), sh, nd, _dtype_code(dtype), version.encode(), st, ns,
                                      C.byref(out), C.byref(err))
    _check(rc, err)
    return _take(out)


def plane_dtype(dtype: Any) -> np.dtype:
    code = _dtype_code(dtype)
    return np.dtype(np.int64) if code in (6, 7) else np.dtype(np.int32)


def forward(chunk: np.ndarray, transforms: Sequence[dict], version: str = "0.1") -> np.ndarray:
    """nd_lift forward of an array: returns the int32 (int64) coefficient plane."""
    raw = lift_forward_bytes(chunk, chunk.shape, chunk.dtype, transforms, version)
    return np.frombuffer(raw, plane_dtype(chunk.dtype)).reshape(chunk.shape).copy()


def inverse(coeffs: np.ndarray, transforms: Sequence[dict], dtype: Any, version: str = "0.1") -> np.ndarray:
    """nd_lift inverse of a coefficient plane back to ``dtype``."""
    code = _dtype_code(dtype)
    raw = lift_inverse_bytes(coeffs, coeffs.shape, code, transforms, version)
    return np.frombuffer(raw, NP_OF[code]).reshape(coeffs.shape).copy()


# ---------------------------------------------------------------------------
# Device-resident batches of the fused nd-lift-ht chain
# ---------------------------------------------------------------------------
def make_desc(shape: Sequence[int], dtype: Any, transforms: Sequence[dict], xy_levels: int = 5,
              in_offset: int = 0, in_len: int = 0, out_offset: int = 0) -> ChunkDesc:
    if len(shape) > MAX_NDIM or len(transforms) > MAX_STEPS:
        raise NdicError(1, 0, "batch descriptors hold at most 8 dimensions and 4 lift steps")
    d = ChunkDesc()
    d.in_offset, d.in_len, d.out_offset = int(in_offset), int(in_len), int(out_offset)
    d.ndim = len(shape)
    for i, s in enumerate(shape):
        d.shape[i] = int(s)
    d.dtype = _dtype_code(dtype)
    d.xy_levels = int(xy_levels)
    d.nsteps = len(transforms)
    st, _ = _steps(transforms)
    for i in range(len(transforms)):
        d.steps[i] = st[i]
    return d


def decode_batch(d_in_ptr: int, d_out_ptr: int, descs, plain: bool = False, stream: int = 0) -> None:
    """Decode chunks resident in device memory (pointers as integers)."""
    arr = descs if isinstance(descs, C.Array) else (ChunkDesc * len(descs))(*descs)
    err = Error()
    rc = _lib.ndic_lift_ht_decode_batch(C.c_void_p(d_in_ptr), C.c_void_p(d_out_ptr), arr, len(arr),
                                        1 if plain else 0, C.c_void_p(stream), C.byref(err))
    _check(rc, err)


def encode_batch(d_in_ptr: int, d_out_ptr: int, capacity: int, descs, plain: bool = False, stream: int = 0) -> None:
    """Encode chunks resident in device memory; descs[i].out_offset/out_len report the codestreams."""
    arr = descs if isinstance(descs, C.Array) else (ChunkDesc * len(descs))(*descs)
    err = Error()
    rc = _lib.ndic_lift_ht_encode_batch(C.c_void_p(d_in_ptr), C.c_void_p(d_out_ptr), int(capacity), arr, len(arr),
                                        1 if plain else 0, C.c_void_p(stream), C.byref(err))
    _check(rc, err)


# ---------------------------------------------------------------------------
# byte-range plans and plane-selective decodes (crates/ndic-codestream/src/
# range.rs:141-369, crates/ndic-zarr/src/htj2k_codec.rs:319-411)
# ---------------------------------------------------------------------------
PLAN_TARGETS = {"thumbnail": 0, "plane": 1, "thumbnail_3d": 2}


def range_plan(prefix: bytes, target: str, arg: int, transforms: Sequence[dict] = ()) -> dict:
    """RangeIndex::from_chunk_prefix(prefix) then thumbnail(arg = max_px),
    plane(arg = z) or thumbnail_3d(arg = max_px, transforms): the Plan as a
    dict (target, decoded_size, max_res, planes, ranges as end-inclusive
    (start, end) pairs, total_bytes)."""
    data = _as_bytes(prefix)
    st, ns = _steps(transforms)
    out = RangePlanC()
    err = Error()
    rc = _lib.ndic_range_plan_chunk(data, len(data), PLAN_TARGETS[target], int(arg), st, ns, C.byref(out),
                                    C.byref(err))
    _check(rc, err)
    planes = np.frombuffer(_take(out.planes), np.uint64).tolist()
    flat = np.frombuffer(_take(out.ranges), np.uint64).tolist()
    return {"target": target, "decoded_size": [int(out.decoded_size[i]) for i in range(out.ndims)],
            "max_res": int(out.max_res), "planes": planes,
            "ranges": [(flat[2 * i], flat[2 * i + 1]) for i in range(len(flat) // 2)],
            "total_bytes": int(out.total_bytes)}


def make_select_desc(shape: Sequence[int], dtype: Any, xy_levels: int, in_offset: int, in_len: int,
                     out_offset: int = 0, *, subset: Sequence[tuple] = None, max_px: int = None,
                     transforms: Sequence[dict] = ()) -> SelectDesc:
    """A plane-selective decode request: subset = [(start, count)] per axis
    (the partial decoder), or max_px (+ the chunk's nd_lift transforms) for
    the thumbnail_3d planes."""
    if len(shape) > MAX_NDIM or len(transforms) > MAX_STEPS:
        raise NdicError(1, 0, "select descriptors hold at most 8 dimensions and 4 lift steps")
    d = SelectDesc()
    d.in_offset, d.in_len, d.out_offset = int(in_offset), int(in_len), int(out_offset)
    d.ndim = len(shape)
    for i, x in enumerate(shape):
        d.shape[i] = int(x)
    d.dtype = _dtype_code(dtype)
    d.xy_levels = int(xy_levels)
    if subset is not None:
        d.target = 0
        for i, (a, n) in enumerate(subset):
            d.start[i], d.count[i] = int(a), int(n)
    else:
        d.target = 1
        d.max_px = int(max_px)
        st, _ = _steps(transforms)
        d.nsteps = len(transforms)
        for i in range(len(transforms)):
            d.steps[i] = st[i]
    return d


def decode_select_batch(d_in_ptr: int, d_out_ptr: int, descs, stream: int = 0):
    """Plane-selective decodes of device-resident chunks; with d_out_ptr 0 the
    descriptors only receive their output shapes and sizes."""
    arr = descs if isinstance(descs, C.Array) else (SelectDesc * len(descs))(*descs)
    err = Error()
    rc = _lib.ndic_htj2k_decode_select_batch(C.c_void_p(d_in_ptr), C.c_void_p(d_out_ptr), arr, len(arr),
                                             C.c_void_p(stream), C.byref(err))
    _check(rc, err)
    return arr


# ---------------------------------------------------------------------------
# nd_zfp / zfp (bindings/python/nd-image-codecs/src/lib.rs:117-166,
# crates/ndic-zfp/src/chunk.rs:101-233)
# ---------------------------------------------------------------------------
def zfp_config(mode: str = "reversible", rate=None, tolerance=None, precision=None, dims=None) -> ZfpConfig:
    """NdZfpConfig with NdZfpConfig::zfp_mode's parameter rules (chunk.rs:162-215):
    exactly the mode's own parameter may be present."""
    def forbid(present, what):
        if present:
            raise NdicError(1, 0, f'zfp: "{mode}" mode does not take "{what}"')

    if dims is not None and not 1 <= int(dims) <= 4:
        raise NdicError(1, 0, f"zfp: dims must be 1..=4, got {dims}")
    c = ZfpConfig()
    c.dims = 0 if dims is None else int(dims)
    if mode == "reversible":
        forbid(rate is not None, "rate")
        forbid(tolerance is not None, "tolerance")
        forbid(precision is not None, "precision")
    elif mode == "fixed_rate":
        forbid(tolerance is not None, "tolerance")
        forbid(precision is not None, "precision")
        if rate is None:
            raise NdicError(1, 0, 'zfp: fixed_rate mode needs a "rate"')
        c.rate = float(rate)
    elif mode == "fixed_accuracy":
        forbid(rate is not None, "rate")
        forbid(precision is not None, "precision")
        if tolerance is None:
            raise NdicError(1, 0, 'zfp: fixed_accuracy mode needs a "tolerance"')
        c.tolerance = float(tolerance)
    elif mode == "fixed_precision":
        forbid(rate is not None, "rate")
        forbid(tolerance is not None, "tolerance")
        if precision is None:
            raise NdicError(1, 0, 'zfp: fixed_precision mode needs a "precision"')
        if int(precision) < 0:
            raise NdicError(1, 0, f"nd_zfp: fixed-precision mode needs 1..=64 bit planes, got {precision}")
        c.precision = int(precision)
    else:
        raise NdicError(3, 0, f"zfp: unknown mode {mode!r}")
    c.mode = ZFP_MODES[mode]
    return c


_ZFP_DTYPES = {"uint8": "uint8", "|u1": "uint8", "u1": "uint8", "int8": "int8", "|i1": "int8", "i1": "int8",
               "uint16": "uint16", "<u2": "uint16", "u2": "uint16", "int16": "int16", "<i2": "int16", "i2": "int16",
               "int32": "int32", "<i4": "int32", "i4": "int32", "int64": "int64", "<i8": "int64", "i8": "int64",
               "float32": "float32", "<f4": "float32", "f4": "float32", "float64": "float64", "<f8": "float64",
               "f8": "float64"}


def _zfp_dtype(dtype: Any) -> int:
    """ZfpDtype::from_zarr_name (crates/ndic-zfp/src/chunk.rs:53-65)."""
    name = dtype if isinstance(dtype, str) else np.dtype(dtype).name
    if name not in _ZFP_DTYPES:
        raise NdicError(1, 0, f"zfp has no path for dtype {name!r}")
    return DTYPE_CODES[_ZFP_DTYPES[name]]


def nd_zfp_encode(chunk: Any, shape: Sequence[int], dtype: Any, *, mode: str = "reversible", rate=None,
                  tolerance=None, precision=None, dims=None) -> bytes:
    cfg = zfp_config(mode, rate, tolerance, precision, dims)
    data = _as_bytes(chunk)
    sh, nd = _shape(shape)
    out, err = Buffer(), Error()
    rc = _lib.ndic_zfp_encode_chunk(data, len(data), sh, nd, _zfp_dtype(dtype), C.byref(cfg), C.byref(out),
                                    C.byref(err))
    _check(rc, err)
    return _take(out)


def nd_zfp_decode(chunk: Any, shape: Sequence[int], dtype: Any, *, mode: str = "reversible", rate=None,
                  tolerance=None, precision=None, dims=None) -> bytes:
    cfg = zfp_config(mode, rate, tolerance, precision, dims)
    data = _as_bytes(chunk)
    sh, nd = _shape(shape)
    out, err = Buffer(), Error()
    rc = _lib.ndic_zfp_decode_chunk(data, len(data), sh, nd, _zfp_dtype(dtype), C.byref(cfg), C.byref(out),
                                    C.byref(err))
    _check(rc, err)
    return _take(out)


def nd_zfp_decode_brick(chunk: Any, shape: Sequence[int], dtype: Any, brick: Sequence[int], *, rate,
                        dims=None, mode: str = "fixed_rate"):
    """One 4^d brick of a fixed-rate chunk (ndic_zfp::decode_chunk_brick):
    returns (bytes, brick_shape)."""
    cfg = zfp_config(mode, rate=rate if mode == "fixed_rate" else None, dims=dims)
    data = _as_bytes(chunk)
    sh, nd = _shape(shape)
    br = (C.c_size_t * max(1, len(brick)))(*[int(b) for b in brick])
    bs = (C.c_size_t * 4)()
    out, err = Buffer(), Error()
    rc = _lib.ndic_zfp_decode_brick(data, len(data), sh, nd, _zfp_dtype(dtype), C.byref(cfg), br, len(brick),
                                    C.byref(out), bs, C.byref(err))
    _check(rc, err)
    ndim_eff = len(shape) if dims is None else int(dims)
    return _take(out), [int(bs[i]) for i in range(ndim_eff)]


def make_zfp_desc(shape: Sequence[int], dtype: Any, config: ZfpConfig, in_offset: int = 0, in_len: int = 0,
                  out_offset: int = 0, index_offset: int = 0, index_len: int = 0) -> ZfpDesc:
    if len(shape) > MAX_NDIM:
        raise NdicError(1, 0, "batch descriptors hold at most 8 dimensions")
    d = ZfpDesc()
    d.in_offset, d.in_len, d.out_offset = int(in_offset), int(in_len), int(out_offset)
    d.ndim = len(shape)
    for i, s in enumerate(shape):
        d.shape[i] = int(s)
    d.dtype = _zfp_dtype(dtype)
    d.config = config
    d.index_offset, d.index_len = int(index_offset), int(index_len)
    return d


def zfp_encode_batch(d_in_ptr: int, d_out_ptr: int, capacity: int, descs, stream: int = 0, d_index_ptr: int = 0,
                     index_capacity: int = 0):
    """Encode device-resident chunks; descs[i].out_offset/out_len report the
    streams.  With an index arena (d_index_ptr) each variable-size stream's
    block index is written there too (descs[i].index_offset/index_len)."""
    arr = descs if isinstance(descs, C.Array) else (ZfpDesc * len(descs))(*descs)
    err = Error()
    if d_index_ptr:
        rc = _lib.ndic_zfp_encode_batch_indexed(C.c_void_p(d_in_ptr), C.c_void_p(d_out_ptr), int(capacity),
                                                C.c_void_p(d_index_ptr), int(index_capacity), arr, len(arr),
                                                C.c_void_p(stream), C.byref(err))
    else:
        rc = _lib.ndic_zfp_encode_batch(C.c_void_p(d_in_ptr), C.c_void_p(d_out_ptr), int(capacity), arr, len(arr),
                                        C.c_void_p(stream), C.byref(err))
    _check(rc, err)
    return arr


def zfp_decode_batch(d_in_ptr: int, d_out_ptr: int, descs, stream: int = 0, d_index_ptr: int = 0):
    """Decode device-resident streams; with an index arena, streams whose
    descriptor carries a block index decode without walking the stream."""
    arr = descs if isinstance(descs, C.Array) else (ZfpDesc * len(descs))(*descs)
    err = Error()
    if d_index_ptr:
        rc = _lib.ndic_zfp_decode_batch_indexed(C.c_void_p(d_in_ptr), C.c_void_p(d_index_ptr), C.c_void_p(d_out_ptr),
                                                arr, len(arr), C.c_void_p(stream), C.byref(err))
    else:
        rc = _lib.ndic_zfp_decode_batch(C.c_void_p(d_in_ptr), C.c_void_p(d_out_ptr), arr, len(arr),
                                        C.c_void_p(stream), C.byref(err))
    _check(rc, err)
    return arr


def device_memory() -> tuple:
    """(free, total) bytes of device memory, measured now."""
    f, t, err = C.c_size_t(), C.c_size_t(), Error()
    _check(_lib.ndic_device_memory(C.byref(f), C.byref(t), C.byref(err)), err)
    return int(f.value), int(t.value)


def version() -> str:
    return _lib.ndic_version().decode()


def tune_set(key: str, value: int) -> None:
    """Sets a launch-parameter knob of the library (see ndic_tune_set)."""
    rc = _lib.ndic_tune_set(key.encode(), int(value))
    if rc == -1:
        raise KeyError(key)
    if rc != 0:
        raise ValueError(f"{key}: value {value} out of range")


def tune_get(key: str) -> int:
    v = _lib.ndic_tune_get(key.encode())
    if v == -1 and key not in tune_keys():
        raise KeyError(key)
    return int(v)


def tune_reset(key: str) -> None:
    if _lib.ndic_tune_reset(key.encode()) != 0:
        raise KeyError(key)


def tune_keys() -> list:
    out, i = [], 0
    while True:
        n = _lib.ndic_tune_name(i)
        if not n:
            return out
        out.append(n.decode())
        i += 1


__all__ = ["htj2k_encode", "htj2k_decode", "htj2k_decode_plane", "forward", "inverse", "lift_forward_bytes",
           "lift_inverse_bytes", "decode_batch", "encode_batch", "make_desc", "nd_zfp_encode", "nd_zfp_decode",
           "nd_zfp_decode_brick", "zfp_config", "make_zfp_desc", "zfp_encode_batch", "zfp_decode_batch",
           "NdicError", "lib", "version", "LIB_PATH", "tune_set", "tune_get", "tune_reset", "tune_keys"]

if os.environ.get("NDIC_VERBOSE"):
    print(f"nd-image-codecs_amd: loaded {LIB_PATH}")
