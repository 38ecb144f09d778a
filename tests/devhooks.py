"""ctypes bindings of the device-kernel test hooks (include/ndic_hip_dev.h)."""

from __future__ import annotations

import ctypes as C

import numpy as np


def bind(ndic):
    lib = ndic.lib()
    E, B = C.POINTER(ndic.Error), C.POINTER(ndic.Buffer)
    lib.ndic_dev_dwt53.argtypes = [C.c_int32, C.c_void_p, C.c_size_t, C.c_size_t, C.c_uint32, E]
    lib.ndic_dev_dwt53.restype = C.c_int
    lib.ndic_dev_ht_encode_block.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_uint32, B, E]
    lib.ndic_dev_ht_encode_block.restype = C.c_int
    lib.ndic_dev_ht_decode_block.argtypes = [C.c_char_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t,
                                             C.c_uint32, C.c_uint32, C.c_size_t, C.c_size_t, C.c_int, E]
    lib.ndic_dev_ht_decode_block.restype = C.c_int
    return lib


def dwt53(ndic, forward: bool, plane: np.ndarray, levels: int) -> np.ndarray:
    lib = bind(ndic)
    p = np.ascontiguousarray(plane, np.int32).copy()
    err = ndic.Error()
    rc = lib.ndic_dev_dwt53(int(forward), p.ctypes.data, p.shape[1], p.shape[0], levels, C.byref(err))
    ndic._check(rc, err)
    return p


def ht_encode_block(ndic, sm: np.ndarray, missing_msbs: int) -> bytes:
    lib = bind(ndic)
    b = np.ascontiguousarray(sm, np.uint32)
    out, err = ndic.Buffer(), ndic.Error()
    rc = lib.ndic_dev_ht_encode_block(b.ctypes.data, b.shape[1], b.shape[0], missing_msbs, C.byref(out), C.byref(err))
    ndic._check(rc, err)
    return ndic._take(out)


def ht_decode_block(ndic, coded: bytes, w: int, h: int, missing_msbs: int, num_passes: int = 1,
                    len1: int | None = None, len2: int = 0, stripe_causal: bool = False) -> np.ndarray:
    lib = bind(ndic)
    out = np.zeros((h, w), np.uint32)
    err = ndic.Error()
    rc = lib.ndic_dev_ht_decode_block(coded, len(coded), out.ctypes.data, w, h, missing_msbs, num_passes,
                                      len(coded) if len1 is None else len1, len2, int(stripe_causal),
                                      C.byref(err))
    ndic._check(rc, err)
    return out
