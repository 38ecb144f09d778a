"""ctypes bindings to the CPU oracle: TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
this module, as the checker.  The product (nd-image-codecs_amd) never imports
it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libndic_oracle.so"

DTYPES = {"u8": 0, "i8": 1, "u16": 2, "i16": 3, "u32": 4, "i32": 5, "u64": 6, "i64": 7, "f32": 8, "f64": 9}
NP_DTYPES = {0: np.uint8, 1: np.int8, 2: np.uint16, 3: np.int16, 4: np.uint32, 5: np.int32, 6: np.uint64,
             7: np.int64, 8: np.float32, 9: np.float64}
KINDS = {"delta": 0, "haar": 1, "lift53": 2}


class OracleError(Exception):
    def __init__(self, kind: int, offset: int, message: str):
        super().__init__(message)
        self.kind = kind
        self.offset = offset
        self.message = message


class _Err(C.Structure):
    _fields_ = [("kind", C.c_int), ("offset", C.c_size_t), ("message", C.c_char * 512)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


def _declare(L) -> None:
    P, S, I, U = C.c_void_p, C.c_size_t, C.c_int, C.c_uint
    E = C.POINTER(_Err)
    sig = {
        "orc_free": [P],
        "orc_lift_chunk": [I, C.c_char_p, S, P, S, I, C.c_char_p, P, S, P, P, E],
        "orc_lift_plane_i64": [I, P, P, S, P, S, E],
        "orc_dwt53": [I, P, S, S, U, E],
        "orc_ht_tables": [P] * 6,
        "orc_ht_encode_block": [P, S, S, C.c_uint32, P, P, E],
        "orc_ht_decode_block": [C.c_char_p, S, P, S, S, C.c_uint32, C.c_uint32, S, S, I, E],
        "orc_encode_image": [P, S, S, U, I, U, P, P, E],
        "orc_decode_codestream": [C.c_char_p, S, I, I, I, P, P, P, E],
        "orc_dwt97": [I, P, S, S, U, E],
        "orc_encode_image_irreversible": [P, S, S, U, I, U, U, U, U, P, P, E],
        "orc_htj2k_encode_chunk": [C.c_char_p, S, P, S, I, U, I, I, I, P, P, E],
        "orc_htj2k_decode_chunk": [C.c_char_p, S, P, S, I, P, P, E],
        "orc_zfp_encode": [I, C.c_char_p, S, P, S, I, I, C.c_double, P, P, P, E],
        "orc_zfp_decode": [I, C.c_char_p, S, P, S, I, I, C.c_double, P, P, E],
        "orc_zfp_perm": [I, P],
        "orc_zfp_brick_index": [P, S, I, C.c_double, P, P, P, P, E],
        "orc_zfp_block_starts": [C.c_char_p, S, P, S, I, I, C.c_double, P, P, E],
        "orc_zfp_decode_brick": [I, C.c_char_p, S, P, S, I, C.c_double, P, P, P, P, E],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = None if name in ("orc_free", "orc_ht_tables", "orc_zfp_perm") else C.c_int


def _check(rc: int, err: _Err) -> None:
    if rc != 0:
        raise OracleError(err.kind, err.offset, err.message.decode(errors="replace"))


def _take(ptr, n: int) -> bytes:
    out = C.string_at(ptr, n) if n else b""
    lib().orc_free(ptr)
    return out


def _shape(shape):
    return (C.c_size_t * len(shape))(*shape), len(shape)


def _steps(steps):
    flat = []
    for s in steps:
        flat += [int(s["dimension"]), KINDS[s["kind"]], int(s.get("levels", 0)), int(s.get("group", 0))]
    return (C.c_int * max(1, len(flat)))(*flat), len(steps)


def lift_chunk(forward: bool, data: bytes, shape, dtype: str, steps, version: str = "0.1") -> bytes:
    L = lib()
    out = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    err = _Err()
    sh, nd = _shape(shape)
    st, ns = _steps(steps)
    rc = L.orc_lift_chunk(int(forward), data, len(data), sh, nd, DTYPES[dtype], version.encode(), st, ns,
                          C.byref(out), C.byref(n), C.byref(err))
    _check(rc, err)
    return _take(out, n.value)


def lift_plane_i64(forward: bool, values: np.ndarray, shape, steps) -> np.ndarray:
    v = np.ascontiguousarray(values, dtype=np.int64).copy()
    err = _Err()
    sh, nd = _shape(shape)
    st, ns = _steps(steps)
    rc = lib().orc_lift_plane_i64(int(forward), v.ctypes.data_as(C.POINTER(C.c_int64)), sh, nd, st, ns, C.byref(err))
    _check(rc, err)
    return v


def dwt53(forward: bool, plane: np.ndarray, levels: int) -> np.ndarray:
    p = np.ascontiguousarray(plane, dtype=np.int32).copy()
    h, w = p.shape
    err = _Err()
    rc = lib().orc_dwt53(int(forward), p.ctypes.data_as(C.POINTER(C.c_int32)), C.c_size_t(w), C.c_size_t(h),
                         C.c_uint(levels), C.byref(err))
    _check(rc, err)
    return p


def dwt97(forward: bool, plane: np.ndarray, levels: int) -> np.ndarray:
    """The reference's unscaled 9/7 lifting (dwt/mod.rs:103-325) on an f32 plane."""
    p = np.ascontiguousarray(plane, dtype=np.float32).copy()
    h, w = p.shape
    err = _Err()
    rc = lib().orc_dwt97(int(forward), p.ctypes.data_as(C.POINTER(C.c_float)), C.c_size_t(w), C.c_size_t(h),
                         C.c_uint(levels), C.byref(err))
    _check(rc, err)
    return p


def encode_image_irreversible(samples: np.ndarray, depth: int, signed: bool, levels: int, eps: int, mu: int = 0,
                              guard: int = 2) -> bytes:
    """A lossy 9/7 codestream (QCD style 2, every band at eps / mu) for the
    tests of the decode-only irreversible path."""
    s = np.ascontiguousarray(samples, dtype=np.int32)
    h, w = s.shape
    out = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    err = _Err()
    rc = lib().orc_encode_image_irreversible(s.ctypes.data_as(C.POINTER(C.c_int32)), C.c_size_t(w), C.c_size_t(h),
                                             C.c_uint(depth), C.c_int(int(signed)), C.c_uint(levels), C.c_uint(eps),
                                             C.c_uint(mu), C.c_uint(guard), C.byref(out), C.byref(n), C.byref(err))
    _check(rc, err)
    return _take(out, n.value)


def ht_tables() -> dict:
    arrs = {k: np.zeros(n, np.uint16) for k, n in (("DEC_VLC_TBL0", 1024), ("DEC_VLC_TBL1", 1024),
                                                   ("ENC_VLC_TBL0", 2048), ("ENC_VLC_TBL1", 2048),
                                                   ("UVLC_TBL0", 320), ("UVLC_TBL1", 256))}
    lib().orc_ht_tables(*[a.ctypes.data_as(C.POINTER(C.c_uint16)) for a in arrs.values()])
    return arrs


def ht_encode_block(sm: np.ndarray, missing_msbs: int) -> bytes:
    b = np.ascontiguousarray(sm, dtype=np.uint32)
    h, w = b.shape
    out = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    err = _Err()
    rc = lib().orc_ht_encode_block(b.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_size_t(w), C.c_size_t(h),
                                   C.c_uint32(missing_msbs), C.byref(out), C.byref(n), C.byref(err))
    _check(rc, err)
    return _take(out, n.value)


def ht_decode_block(coded: bytes, width: int, height: int, missing_msbs: int, num_passes: int = 1,
                    len1: int | None = None, len2: int = 0, stripe_causal: bool = False) -> np.ndarray:
    out = np.zeros((height, width), np.uint32)
    err = _Err()
    rc = lib().orc_ht_decode_block(coded, C.c_size_t(len(coded)), out.ctypes.data_as(C.POINTER(C.c_uint32)),
                                   C.c_size_t(width), C.c_size_t(height), C.c_uint32(missing_msbs),
                                   C.c_uint32(num_passes), C.c_size_t(len(coded) if len1 is None else len1),
                                   C.c_size_t(len2), C.c_int(int(stripe_causal)), C.byref(err))
    _check(rc, err)
    return out


def encode_image(samples: np.ndarray, depth: int, signed: bool, levels: int) -> bytes:
    s = np.ascontiguousarray(samples, dtype=np.int32)
    h, w = s.shape
    out = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    err = _Err()
    rc = lib().orc_encode_image(s.ctypes.data_as(C.POINTER(C.c_int32)), C.c_size_t(w), C.c_size_t(h),
                                C.c_uint(depth), C.c_int(int(signed)), C.c_uint(levels), C.byref(out),
                                C.byref(n), C.byref(err))
    _check(rc, err)
    return _take(out, n.value)


def decode_codestream(data: bytes, max_res: int = -1, prefix: bool = False, irreversible: bool = False) -> np.ndarray:
    out = C.POINTER(C.c_int32)()
    w = C.c_size_t()
    h = C.c_size_t()
    err = _Err()
    rc = lib().orc_decode_codestream(data, C.c_size_t(len(data)), C.c_int(max_res), C.c_int(int(prefix)),
                                     C.c_int(int(irreversible)), C.byref(out), C.byref(w), C.byref(h),
                                     C.byref(err))
    _check(rc, err)
    arr = np.frombuffer(C.string_at(out, 4 * w.value * h.value), np.int32).reshape(h.value, w.value).copy()
    lib().orc_free(out)
    return arr


def htj2k_encode_chunk(data: bytes, shape, dtype: str, xy_levels: int = 5, reversible: bool = True,
                       progression: int = 2, index: bool = True) -> bytes:
    out = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    err = _Err()
    sh, nd = _shape(shape)
    rc = lib().orc_htj2k_encode_chunk(data, C.c_size_t(len(data)), sh, nd, DTYPES[dtype], C.c_uint(xy_levels),
                                      C.c_int(int(reversible)), C.c_int(progression), C.c_int(int(index)),
                                      C.byref(out), C.byref(n), C.byref(err))
    _check(rc, err)
    return _take(out, n.value)


def htj2k_decode_chunk(data: bytes, shape, dtype: str) -> bytes:
    out = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    err = _Err()
    sh, nd = _shape(shape)
    rc = lib().orc_htj2k_decode_chunk(data, C.c_size_t(len(data)), sh, nd, DTYPES[dtype], C.byref(out),
                                      C.byref(n), C.byref(err))
    _check(rc, err)
    return _take(out, n.value)


def lift_ht_encode(samples: np.ndarray, dtype: str, steps, xy_levels: int = 5) -> bytes:
    """nd_lift forward then htj2k over the int32 coefficient plane (the codec chain)."""
    raw = np.ascontiguousarray(samples).tobytes()
    coeffs = lift_chunk(True, raw, samples.shape, dtype, steps)
    return htj2k_encode_chunk(coeffs, samples.shape, "i32", xy_levels=xy_levels)


def lift_ht_decode(chunk: bytes, shape, dtype: str, steps) -> np.ndarray:
    coeffs = htj2k_decode_chunk(chunk, shape, "i32")
    raw = lift_chunk(False, coeffs, shape, dtype, steps)
    return np.frombuffer(raw, NP_DTYPES[DTYPES[dtype]]).reshape(shape).copy()


if os.environ.get("NDIC_ORACLE_AUTOBUILD") == "1":
    build()


# ---------------------------------------------------------------------------
# ZFP
# ---------------------------------------------------------------------------
ZFP_MODES = {"fixed_rate": 0, "fixed_precision": 1, "fixed_accuracy": 2, "reversible": 3}
ZFP_TYPES = {"i32": 1, "i64": 2, "f32": 3, "f64": 4}


def zfp_mode(config: dict):
    """(kind, value) for a zfp configuration object (mode + its parameter)."""
    mode = config["mode"]
    value = {"fixed_rate": config.get("rate"), "fixed_precision": config.get("precision"),
             "fixed_accuracy": config.get("tolerance"), "reversible": 0.0}[mode]
    return ZFP_MODES[mode], float(value)


def zfp_effective_shape(shape, dims=None):
    """Field shape of a chunk: the chunk shape itself, or (legacy ``dims``)
    the non-singleton axes left-padded with 1s up to ``dims``."""
    if dims is None:
        return list(shape)
    squeezed = [int(d) for d in shape if d > 1]
    if len(squeezed) > dims:
        raise OracleError(1, 0, "zfp: too many non-singleton dimensions")
    return [1] * (dims - len(squeezed)) + squeezed


def zfp_encode_chunk(data: bytes, shape, dtype: str, config: dict) -> bytes:
    L = lib()
    sh, nd = _shape(zfp_effective_shape(shape, config.get("dims")))
    kind, value = zfp_mode(config)
    out, n, err = C.c_void_p(), C.c_size_t(), _Err()
    _check(L.orc_zfp_encode(1, data, len(data), sh, nd, DTYPES[dtype], kind, value, C.byref(out), C.byref(n), None,
                            C.byref(err)), err)
    return _take(out, n.value)


def zfp_decode_chunk(stream: bytes, shape, dtype: str, config: dict) -> bytes:
    L = lib()
    sh, nd = _shape(zfp_effective_shape(shape, config.get("dims")))
    kind, value = zfp_mode(config)
    out, n, err = C.c_void_p(), C.c_size_t(), _Err()
    _check(L.orc_zfp_decode(1, stream, len(stream), sh, nd, DTYPES[dtype], kind, value, C.byref(out), C.byref(n),
                            C.byref(err)), err)
    return _take(out, n.value)


def zfp_compress(values: np.ndarray, config: dict):
    """Raw ZFP stream of an int32/int64/float32/float64 array (row-major
    shape, x = last axis); returns (stream, header_bits)."""
    L = lib()
    t = {np.dtype(np.int32): 1, np.dtype(np.int64): 2, np.dtype(np.float32): 3, np.dtype(np.float64): 4}[values.dtype]
    values = np.ascontiguousarray(values)
    sh, nd = _shape(values.shape)
    kind, value = zfp_mode(config)
    out, n, hb, err = C.c_void_p(), C.c_size_t(), C.c_uint64(), _Err()
    _check(L.orc_zfp_encode(0, values.tobytes(), values.nbytes, sh, nd, t, kind, value, C.byref(out), C.byref(n),
                            C.byref(hb), C.byref(err)), err)
    return _take(out, n.value), hb.value


def zfp_decompress(stream: bytes, shape, np_dtype, config: dict) -> np.ndarray:
    L = lib()
    t = {np.dtype(np.int32): 1, np.dtype(np.int64): 2, np.dtype(np.float32): 3, np.dtype(np.float64): 4}[
        np.dtype(np_dtype)]
    sh, nd = _shape(shape)
    kind, value = zfp_mode(config)
    out, n, err = C.c_void_p(), C.c_size_t(), _Err()
    _check(L.orc_zfp_decode(0, stream, len(stream), sh, nd, t, kind, value, C.byref(out), C.byref(n), C.byref(err)),
           err)
    return np.frombuffer(_take(out, n.value), dtype=np_dtype).reshape(shape)


def zfp_perm(dims: int) -> list:
    buf = (C.c_uint8 * (1 << (2 * dims)))()
    lib().orc_zfp_perm(dims, buf)
    return list(buf)


def zfp_brick_index(shape, np_dtype, rate: float) -> dict:
    """BrickIndex::fixed_rate (lib.rs:398-500) over a row-major shape."""
    t = {np.dtype(np.int32): 1, np.dtype(np.int64): 2, np.dtype(np.float32): 3, np.dtype(np.float64): 4}[
        np.dtype(np_dtype)]
    sh, nd = _shape(shape)
    hb, bpb, sl, err = C.c_uint64(), C.c_uint64(), C.c_uint64(), _Err()
    grid = (C.c_uint64 * 4)()
    _check(lib().orc_zfp_brick_index(sh, nd, t, float(rate), C.byref(hb), C.byref(bpb), C.byref(sl), grid,
                                     C.byref(err)), err)
    return {"header_bits": hb.value, "bits_per_brick": bpb.value, "stream_len": sl.value,
            "grid": [int(grid[a]) for a in range(len(shape))]}


def zfp_byte_range(index: dict, k: int):
    """BrickIndex::byte_range (lib.rs:486-500): (offset, length) in bytes."""
    bit = index["header_bits"] + k * index["bits_per_brick"]
    start = bit // 8
    end = -(-(bit + index["bits_per_brick"]) // 8)
    return start, end - start


def zfp_decode_brick(stream: bytes, shape, dtype: str, rate: float, brick):
    """decode_chunk_brick (chunk.rs:532): (brick bytes, clipped brick shape);
    shape is the effective shape, dtype an oracle dtype name (u8 ... f64)."""
    sh, nd = _shape(shape)
    br = (C.c_size_t * len(shape))(*brick)
    bs = (C.c_size_t * len(shape))()
    out, n, err = C.c_void_p(), C.c_size_t(), _Err()
    _check(lib().orc_zfp_decode_brick(1, stream, len(stream), sh, nd, DTYPES[dtype], float(rate), br, C.byref(out),
                                      C.byref(n), bs, C.byref(err)), err)
    return _take(out, n.value), [int(bs[a]) for a in range(len(shape))]


def zfp_block_starts(stream: bytes, shape, np_dtype, config: dict) -> np.ndarray:
    """Absolute bit position of each block of a raw ZFP stream, then the end
    of the last block (uint64)."""
    t = {np.dtype(np.int32): 1, np.dtype(np.int64): 2, np.dtype(np.float32): 3, np.dtype(np.float64): 4}[
        np.dtype(np_dtype)]
    sh, nd = _shape(shape)
    kind, value = zfp_mode(config)
    out, n, err = C.c_void_p(), C.c_size_t(), _Err()
    _check(lib().orc_zfp_block_starts(stream, len(stream), sh, nd, t, kind, value, C.byref(out), C.byref(n),
                                      C.byref(err)), err)
    return np.frombuffer(_take(out, n.value), np.uint64).copy()
