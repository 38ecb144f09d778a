#!/usr/bin/env python3
"""Write ``zfp_libzfp_vectors.json``: stream and decode hashes from the LLNL
zfp C library 0.5.5 (``/opt/conda/lib/libzfp.so`` in the build image, run
through ctypes) for the extra ZFP cases of ``tests/zfp_fixtures.py``, and the
4D coefficient order that library uses.

The reference delegates ZFP to zfp-rs 0.1.1 (un-vendored), which states
bit-for-bit stream identity with this C library and checks it in the
reference's own differential test against imagecodecs
(bindings/python/nd-image-codecs/tests/test_nd_zfp_roundtrip.py:132-160).
The reference's 104 checksums (fixtures/zfp/checksums.json) are reproduced
by this library as well, which the test suite re-checks through the oracle.

The 4D order is measured, not read: each of the 256 coefficients is made the
only non-zero one of a reversible int32 block (inverse of the reversible
transform applied to a unit impulse), the block is compressed, and the run
length of the single one-bit in the last bit plane gives its rank.

Usage: make_zfp_vectors.py   (needs the C library; the tests only read the JSON)
"""

from __future__ import annotations

import ctypes as C
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
from zfp_fixtures import ZFP_EXTRA_TYPES, case_values, fnv1a64, zfp_extra_cases  # noqa: E402

LIB = "/opt/conda/lib/libzfp.so"
TYPES = {np.dtype(np.int32): 1, np.dtype(np.int64): 2, np.dtype(np.float32): 3, np.dtype(np.float64): 4}


def load():
    L = C.CDLL(LIB)
    P = C.c_void_p
    for name, res, args in [
        ("zfp_field_1d", P, [P, C.c_int, C.c_uint]),
        ("zfp_field_2d", P, [P, C.c_int, C.c_uint, C.c_uint]),
        ("zfp_field_3d", P, [P, C.c_int, C.c_uint, C.c_uint, C.c_uint]),
        ("zfp_field_4d", P, [P, C.c_int, C.c_uint, C.c_uint, C.c_uint, C.c_uint]),
        ("zfp_stream_open", P, [P]),
        ("zfp_stream_set_rate", C.c_double, [P, C.c_double, C.c_int, C.c_uint, C.c_int]),
        ("zfp_stream_set_precision", C.c_uint, [P, C.c_uint]),
        ("zfp_stream_set_accuracy", C.c_double, [P, C.c_double]),
        ("zfp_stream_set_reversible", None, [P]),
        ("zfp_stream_maximum_size", C.c_size_t, [P, P]),
        ("stream_open", P, [P, C.c_size_t]),
        ("stream_close", None, [P]),
        ("zfp_stream_set_bit_stream", None, [P, P]),
        ("zfp_stream_rewind", None, [P]),
        ("zfp_write_header", C.c_size_t, [P, P, C.c_uint]),
        ("zfp_read_header", C.c_size_t, [P, P, C.c_uint]),
        ("zfp_compress", C.c_size_t, [P, P]),
        ("zfp_decompress", C.c_size_t, [P, P]),
        ("zfp_field_free", None, [P]),
        ("zfp_stream_close", None, [P]),
    ]:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def field(L, arr):
    n = list(arr.shape[::-1])
    fn = [L.zfp_field_1d, L.zfp_field_2d, L.zfp_field_3d, L.zfp_field_4d][len(n) - 1]
    return fn(arr.ctypes.data_as(C.c_void_p), TYPES[arr.dtype], *n)


def configure(L, z, mode, t, dims):
    m = mode["mode"]
    if m == "fixed_rate":
        L.zfp_stream_set_rate(z, mode["rate"], t, dims, 0)
    elif m == "fixed_precision":
        L.zfp_stream_set_precision(z, mode["precision"])
    elif m == "fixed_accuracy":
        L.zfp_stream_set_accuracy(z, mode["tolerance"])
    else:
        L.zfp_stream_set_reversible(z)


def compress(L, arr, mode):
    arr = np.ascontiguousarray(arr)
    f = field(L, arr)
    z = L.zfp_stream_open(None)
    configure(L, z, mode, TYPES[arr.dtype], arr.ndim)
    cap = L.zfp_stream_maximum_size(z, f)
    buf = C.create_string_buffer(cap)
    s = L.stream_open(buf, cap)
    L.zfp_stream_set_bit_stream(z, s)
    L.zfp_stream_rewind(z)
    L.zfp_write_header(z, f, 7)
    n = L.zfp_compress(z, f)
    out = buf.raw[:n]
    L.zfp_field_free(f)
    L.zfp_stream_close(z)
    L.stream_close(s)
    return out


def decompress(L, stream, shape, dtype, mode):
    out = np.zeros(shape, dtype=dtype)
    f = field(L, out)
    z = L.zfp_stream_open(None)
    configure(L, z, mode, TYPES[out.dtype], out.ndim)
    buf = C.create_string_buffer(stream, len(stream))
    s = L.stream_open(buf, len(stream))
    L.zfp_stream_set_bit_stream(z, s)
    L.zfp_stream_rewind(z)
    L.zfp_read_header(z, f, 7)
    L.zfp_decompress(z, f)
    L.zfp_field_free(f)
    L.zfp_stream_close(z)
    L.stream_close(s)
    return out


def rev_inv_lift(a, axis):
    a = np.moveaxis(a, axis, -1)
    x, y, z, w = (a[..., i].copy() for i in range(4))
    w += z
    z += y
    y += x
    w += z
    z += y
    w += z
    return np.moveaxis(np.stack([x, y, z, w], axis=-1), -1, axis)


def measure_perm4(L):
    perm = [None] * 256
    for c in range(256):
        b = np.zeros(256, np.int64)
        b[c] = 1
        b = b.reshape(4, 4, 4, 4)
        for ax in range(4):
            b = rev_inv_lift(b, ax)
        s = compress(L, b.astype(np.int32), {"mode": "reversible"})
        bits = np.unpackbits(np.frombuffer(s, np.uint8), bitorder="little")
        pos = 96 + 5 + 31  # header, precision (32), 31 empty bit planes
        assert bits[pos] == 1 and not bits[96 + 5:pos].any()
        pos += 1
        rank = 0
        while rank < 255 and bits[pos] == 0:
            rank += 1
            pos += 1
        perm[rank] = c
    assert sorted(perm) == list(range(256))
    return perm


def main() -> None:
    L = load()
    cases = []
    for name, kind, tname, shape, mode, seed in zfp_extra_cases():
        dt = ZFP_EXTRA_TYPES[tname]
        arr = case_values(kind, dt, shape, seed)
        s = compress(L, arr, mode)
        back = decompress(L, s, shape, dt, mode)
        cases.append({"name": name, "kind": kind, "type": tname, "shape": shape, "mode": mode, "seed": seed,
                      "bytes": len(s), "stream_fnv1a64": "0x%016x" % fnv1a64(s),
                      "decoded_fnv1a64": "0x%016x" % fnv1a64(back.tobytes())})
    doc = {"$comment": "LLNL zfp 0.5.5 C library outputs for tests/zfp_fixtures.py zfp_extra_cases(); "
                       "written by tests/golden/make_zfp_vectors.py",
           "perm4": measure_perm4(L), "cases": cases}
    text = "{\n \"$comment\": %s,\n \"perm4\": %s,\n \"cases\": [\n%s\n ]\n}\n" % (
        json.dumps(doc["$comment"]), json.dumps(doc["perm4"]),
        ",\n".join("  " + json.dumps(c, sort_keys=True) for c in cases))
    (HERE / "zfp_libzfp_vectors.json").write_text(text)
    print(len(cases), "cases")


if __name__ == "__main__":
    main()
