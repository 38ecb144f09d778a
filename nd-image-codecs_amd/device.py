"""Device buffers on the library's own HIP runtime (include/ndic_hip_dev.h).

The batch entry points take device pointers.  A caller that already holds
device memory passes its pointers; this module gives the rest (the tests and
the benchmark) device memory without a second GPU framework in the process:
two HIP runtimes in one process do not share a device context.
"""

from __future__ import annotations

import ctypes as C
import importlib
from typing import Optional

import numpy as np

_pkg = importlib.import_module("nd-image-codecs_amd")
_bound = False


class SynthBlobs(C.Structure):
    _fields_ = [("count", C.c_int32), ("cz", C.c_float * 8), ("cy", C.c_float * 8), ("cx", C.c_float * 8),
                ("amp", C.c_float * 8), ("sigma", C.c_float * 8)]


SYNTH_ZERO, SYNTH_RANDOM, SYNTH_BLOBS = 0, 1, 2


def _lib():
    global _bound
    lib = _pkg.lib()
    if not _bound:
        P, S, E = C.c_void_p, C.c_size_t, C.POINTER(_pkg.Error)
        for name, res, args in [
            ("ndic_dev_malloc", C.c_int, [S, C.POINTER(C.c_void_p), E]),
            ("ndic_dev_free", None, [P]),
            ("ndic_dev_memcpy", C.c_int, [P, P, S, C.c_int32, E]),
            ("ndic_dev_memset", C.c_int, [P, C.c_int32, S, E]),
            ("ndic_dev_synchronize", C.c_int, [E]),
            ("ndic_dev_synth_chunk", C.c_int, [P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, C.c_int32,
                                               C.POINTER(SynthBlobs), C.c_uint64, E]),
            ("ndic_dev_count_mismatch", C.c_int, [P, P, S, C.POINTER(C.c_uint64), E]),
            ("ndic_dev_count_mismatch_async", C.c_int, [P, P, S, P, E]),
            ("ndic_dev_copy", C.c_int, [P, P, S, E]),
            ("ndic_dev_hash_async", C.c_int, [P, S, P, E]),
            ("ndic_dev_profile", None, [C.c_int32]),
            ("ndic_dev_stage_times", C.c_size_t, [C.POINTER(C.c_double), C.POINTER(C.c_double), S]),
            ("ndic_dev_zfp_times", C.c_size_t, [C.POINTER(C.c_double), C.POINTER(C.c_uint64), S]),
        ]:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return lib


def _call(fn, *args) -> None:
    err = _pkg.Error()
    _pkg._check(fn(*args, C.byref(err)), err)


class DeviceBuffer:
    """`nbytes` of device memory, freed with the object (or by free())."""

    def __init__(self, nbytes: int, zero: bool = False):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _call(_lib().ndic_dev_malloc, max(self.nbytes, 1), C.byref(p))
        self.ptr = p.value or 0
        if zero:
            self.memset(0)

    @classmethod
    def from_bytes(cls, data, pad: int = 0) -> "DeviceBuffer":
        data = bytes(data)
        b = cls(len(data) + pad, zero=pad > 0)
        b.upload(data)
        return b

    def upload(self, data, offset: int = 0) -> None:
        data = bytes(data)
        if offset + len(data) > self.nbytes:
            raise ValueError("upload past the end of the device buffer")
        if data:
            _call(_lib().ndic_dev_memcpy, self.ptr + offset, C.c_char_p(data), len(data), 0)

    def download(self, offset: int = 0, n: Optional[int] = None) -> bytes:
        n = self.nbytes - offset if n is None else int(n)
        if offset + n > self.nbytes:
            raise ValueError("download past the end of the device buffer")
        out = C.create_string_buffer(max(n, 1))
        if n:
            _call(_lib().ndic_dev_memcpy, C.addressof(out), self.ptr + offset, n, 1)
        return out.raw[:n]

    def memset(self, value: int, offset: int = 0, n: Optional[int] = None) -> None:
        n = self.nbytes - offset if n is None else int(n)
        _call(_lib().ndic_dev_memset, self.ptr + offset, int(value), n)

    def free(self) -> None:
        if self.ptr:
            _lib().ndic_dev_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def memcpy_d2d(dst: int, src: int, n: int) -> None:
    """Device-to-device copy of n bytes (the copy engine; returns when done)."""
    if n:
        _call(_lib().ndic_dev_memcpy, dst, src, int(n), 2)


def synchronize() -> None:
    _call(_lib().ndic_dev_synchronize)


def synth_chunk(ptr: int, shape, dtype: str, content: int, blobs=(), seed: int = 0) -> None:
    """Writes a synthetic z,y,x volume (u8/u16/i16) at device address ptr;
    blobs is a sequence of (cz, cy, cx, amplitude, sigma)."""
    z, y, x = (int(v) for v in shape)
    b = SynthBlobs()
    b.count = len(blobs)
    for i, (cz, cy, cx, a, s) in enumerate(blobs):
        b.cz[i], b.cy[i], b.cx[i], b.amp[i], b.sigma[i] = cz, cy, cx, a, s
    code = _pkg._dtype_code(dtype)
    _call(_lib().ndic_dev_synth_chunk, ptr, z, y, x, code, int(content), C.byref(b), int(seed))


def copy(dst: int, src: int, n: int) -> None:
    """Device copy through a plain kernel (not the copy engine)."""
    _call(_lib().ndic_dev_copy, dst, src, int(n))


def count_mismatch(a: int, b: int, n: int) -> int:
    out = C.c_uint64()
    _call(_lib().ndic_dev_count_mismatch, a, b, int(n), C.byref(out))
    return out.value


def count_mismatch_async(a: int, b: int, n: int, d_count: int) -> None:
    """Adds the mismatch count of a and b to the device counter at d_count
    on this thread's stream, without waiting."""
    _call(_lib().ndic_dev_count_mismatch_async, a, b, int(n), d_count)


def hash_async(ptr: int, n: int, d_slot: int) -> None:
    """Adds the content hash of n device bytes at ptr to the uint64 at d_slot
    (on this thread's stream, without waiting); see host_hash."""
    _call(_lib().ndic_dev_hash_async, ptr, int(n), d_slot)


_K = np.uint64(0x9E3779B97F4A7C15)


def _mix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = x + _K
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def host_hash(data) -> int:
    """The content hash ndic_dev_hash_async computes, on the host."""
    b = bytes(data)
    n = len(b)
    w = np.frombuffer(b + bytes((-n) % 8), np.uint64)
    with np.errstate(over="ignore"):
        idx = np.arange(len(w), dtype=np.uint64) * _K
        total = int(_mix64(w ^ idx).sum(dtype=np.uint64)) if len(w) else 0
        total += int(_mix64(np.array([n], np.uint64) ^ _K)[0])
    return total & 0xFFFFFFFFFFFFFFFF


def profile(enable: bool) -> None:
    _lib().ndic_dev_profile(1 if enable else 0)


def stage_times(n: int = 5):
    ms = (C.c_double * n)()
    by = (C.c_double * n)()
    windows = _lib().ndic_dev_stage_times(ms, by, n)
    return list(ms), list(by), int(windows)


ZFP_GROUPS = ("encode", "place", "decode", "walk")


def zfp_times() -> dict:
    """Device milliseconds of the ZFP kernel groups since profile(True):
    {group: (ms, spans)} for encode, place, decode and walk."""
    ms = (C.c_double * 4)()
    sp = (C.c_uint64 * 4)()
    _lib().ndic_dev_zfp_times(ms, sp, 4)
    return {g: (ms[i], int(sp[i])) for i, g in enumerate(ZFP_GROUPS)}


def as_array(buf: DeviceBuffer, dtype, offset: int = 0, count: Optional[int] = None) -> np.ndarray:
    dt = np.dtype(dtype)
    n = (buf.nbytes - offset) // dt.itemsize if count is None else int(count)
    return np.frombuffer(buf.download(offset, n * dt.itemsize), dtype=dt)
