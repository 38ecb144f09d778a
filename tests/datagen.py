"""Seeded synthetic data generators for the parity tests and the benchmark.

Content recipes follow the configurations the benchmark is quoted on:
Gaussian blobs plus Poisson noise, clipped to the dtype range.  NumPy only;
nothing here runs on the GPU.
"""

from __future__ import annotations

import numpy as np


def blobs_volume(shape, n_blobs=8, amp=(200.0, 2500.0), sigma=(4.0, 20.0), base=1000.0, lo=0, hi=4095,
                 seed=0, dtype=np.uint16, poisson=True):
    """clip(round(base + sum of 3D Gaussian blobs + Poisson noise), lo, hi) over a (z, y, x) volume."""
    rng = np.random.default_rng(seed)
    z, y, x = shape
    zz = np.arange(z, dtype=np.float32)[:, None, None]
    yy = np.arange(y, dtype=np.float32)[None, :, None]
    xx = np.arange(x, dtype=np.float32)[None, None, :]
    vol = np.full(shape, base, np.float32)
    for _ in range(n_blobs):
        c = rng.uniform(0, 1, 3) * np.array([z, y, x])
        a = rng.uniform(*amp)
        s = rng.uniform(*sigma)
        vol += a * np.exp(-((zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2) / (2 * s * s))
    if poisson:
        vol = rng.poisson(np.maximum(vol, 0)).astype(np.float32)
    return np.clip(np.round(vol), lo, hi).astype(dtype)


def c0_volume(shape=(32, 256, 256), seed=0):
    """Config-0 content: 12-bit blobs + Poisson noise in uint16."""
    return blobs_volume(shape, seed=seed)


def random_plane_ints(rng, shape, lo, hi, dtype):
    return rng.integers(lo, hi + 1, size=shape, dtype=np.int64).astype(dtype)


def sm_block(rng, w, h, kmax, dens):
    """A random HT code-block in sign-magnitude form: magnitudes below 2**kmax
    in the top kmax bits under the sign, a fraction dens of them non-zero."""
    mag = rng.integers(0, 2 ** kmax, size=(h, w)).astype(np.int64)
    mag[rng.uniform(size=(h, w)) > dens] = 0
    if mag.max() == 0:
        mag[0, 0] = 1
    sign = rng.integers(0, 2, size=(h, w)).astype(np.uint32) << 31
    sm = (sign | (mag.astype(np.uint32) << (31 - kmax))).astype(np.uint32)
    sm[mag == 0] = 0
    return sm
