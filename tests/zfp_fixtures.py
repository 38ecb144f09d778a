"""The checksum-matrix input generator and hash of the ZFP fixtures.

Restates crates/ndic-zfp/tests/common/mod.rs:17-46 (ramp_chunk, fnv1a64),
:52-56 (tiny_chunk_f32) and :60-101 (matrix_cases order) so the committed
fixtures/zfp/checksums.json hashes can be reproduced.
"""

from __future__ import annotations

import numpy as np

NP = {"uint8": np.uint8, "int8": np.int8, "uint16": np.uint16, "int16": np.int16, "int32": np.int32,
      "int64": np.int64, "float32": np.float32, "float64": np.float64}
SHORT = {"uint8": "u8", "int8": "i8", "uint16": "u16", "int16": "i16", "int32": "i32", "int64": "i64",
         "float32": "f32", "float64": "f64"}


def _trunc_mod(v: np.ndarray, m: int) -> np.ndarray:
    return np.sign(v) * (np.abs(v) % m)


def ramp_values(n: int) -> np.ndarray:
    """v_i = (i % 97) * 3 - 140 + noise_i % 5, noise a 13/17/5 xorshift32
    seeded 0x9e3779b9 and stepped before each element."""
    noise = np.empty(n, dtype=np.int64)
    x = 0x9E3779B9
    for i in range(n):
        x ^= (x << 13) & 0xFFFFFFFF
        x ^= x >> 17
        x ^= (x << 5) & 0xFFFFFFFF
        noise[i] = x % 5
    i = np.arange(n, dtype=np.int64)
    return (i % 97) * 3 - 140 + noise


def ramp_chunk(n: int, dtype: str) -> bytes:
    v = ramp_values(n)
    if dtype == "uint8":
        a = (v % 256).astype(np.uint8)
    elif dtype == "int8":
        a = _trunc_mod(v, 127).astype(np.int8)
    elif dtype == "uint16":
        a = (v % 65536).astype(np.uint16)
    elif dtype == "int16":
        a = _trunc_mod(v, 32000).astype(np.int16)
    elif dtype == "int32":
        a = (v * 100001).astype(np.int32)
    elif dtype == "int64":
        a = v * 100000007
    elif dtype == "float32":
        a = v.astype(np.float32) / np.float32(3.0)
    else:
        a = v.astype(np.float64) / 3.0
    return a.astype(NP[dtype]).tobytes()


def fnv1a64(data: bytes) -> int:
    h = 0xCBF29CE484222325
    for b in data:
        h ^= b
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def tiny_chunk_f32() -> bytes:
    i = np.arange(4 * 8 * 8)
    return (((i * 7) % 4096).astype(np.float32) / np.float32(3.0)).astype(np.float32).tobytes()


# ---------------------------------------------------------------------------
# Extra cases cross-checked against the LLNL zfp C library 0.5.5
# (tests/golden/make_zfp_vectors.py writes their hashes).  Inputs are made
# from integer arithmetic only, so they are identical on every host.
# ---------------------------------------------------------------------------
def _xorshift64(seed: int, n: int) -> list:
    x = (seed * 0x9E3779B97F4A7C15 + 1) & 0xFFFFFFFFFFFFFFFF
    out = []
    for _ in range(n):
        x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 7
        x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
        out.append(x)
    return out


def case_values(kind: str, np_dtype, shape, seed: int) -> np.ndarray:
    """kind 'smooth': a ramp with small noise; 'wild': random signs and
    magnitudes over 2^-60..2^60 for floats (the whole 30-bit range for
    int32, 62-bit for int64); 'sparse': mostly zero blocks, signed zeros
    and isolated values."""
    n = int(np.prod(shape))
    r = _xorshift64(seed, n)
    dt = np.dtype(np_dtype)
    i = np.arange(n, dtype=np.int64)
    if kind == "smooth":
        noise = np.array([v % 9 for v in r], dtype=np.int64) - 4
        base = (i * 37) % 1000 - 500 + noise
        if dt.kind == "f":
            return (base.astype(np.float64) / 7.0).astype(dt).reshape(shape)
        scale = 1000003 if dt == np.int64 else 1009
        return (base * scale).astype(dt).reshape(shape)
    if kind == "wild":
        if dt.kind == "f":
            mant = np.array([(v >> 11) & ((1 << 52) - 1) for v in r], dtype=np.float64) / float(1 << 52) + 1.0
            expo = np.array([(v % 121) - 60 for v in r], dtype=np.int64)
            sign = np.array([1.0 if (v >> 63) else -1.0 for v in r])
            return (sign * np.ldexp(mant, expo)).astype(dt).reshape(shape)
        bits = 62 if dt == np.int64 else 30
        vals = [((v >> 1) & ((1 << bits) - 1)) - (1 << (bits - 1)) for v in r]
        return np.array(vals, dtype=dt).reshape(shape)
    # sparse
    out = np.zeros(n, dtype=np.float64)
    for k, v in enumerate(r):
        if v % 11 == 0:
            out[k] = ((v >> 8) % 2001 - 1000) / 8.0
    a = out.astype(dt)
    if dt.kind == "f":
        a[1::13] = -0.0 * (a[1::13] == 0) + a[1::13]
    return a.reshape(shape)


ZFP_EXTRA_SHAPES = {1: [13], 2: [7, 9], 3: [5, 6, 7], 4: [3, 5, 4, 6]}
ZFP_EXTRA_MODES = [
    {"mode": "reversible"},
    {"mode": "fixed_rate", "rate": 1.5},
    {"mode": "fixed_rate", "rate": 4.0},
    {"mode": "fixed_rate", "rate": 8.0},
    {"mode": "fixed_rate", "rate": 16.0},
    {"mode": "fixed_rate", "rate": 33.0},
    {"mode": "fixed_precision", "precision": 1},
    {"mode": "fixed_precision", "precision": 7},
    {"mode": "fixed_precision", "precision": 24},
    {"mode": "fixed_precision", "precision": 64},
    {"mode": "fixed_accuracy", "tolerance": 0.0},
    {"mode": "fixed_accuracy", "tolerance": 1e-4},
    {"mode": "fixed_accuracy", "tolerance": 0.5},
    {"mode": "fixed_accuracy", "tolerance": 1e3},
]
ZFP_EXTRA_TYPES = {"int32": np.int32, "int64": np.int64, "float32": np.float32, "float64": np.float64}


def zfp_extra_cases():
    """(name, kind, type, shape, mode, seed) in a fixed order."""
    out = []
    seed = 1
    for d, shape in ZFP_EXTRA_SHAPES.items():
        for tname in ZFP_EXTRA_TYPES:
            for kind in ("smooth", "wild", "sparse"):
                modes = ZFP_EXTRA_MODES if kind == "smooth" else ZFP_EXTRA_MODES[::3]
                for m in modes:
                    out.append(("%dd-%s-%s-%d" % (d, tname, kind, seed), kind, tname, shape, m, seed))
                    seed += 1
    return out
