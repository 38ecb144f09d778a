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


def series_5d(shape=(8, 1, 32, 512, 512), n_blobs=24, sigma=(3.0, 30.0), amp=(2000.0, 40000.0), base=1000.0,
              fade=0.02, saturated=0.01, seed=1):
    """The 5D series recipe of the single-chunk 8x1x32x512x512 configuration:
    uint16 t,c,z,y,x over the full 16-bit range -- n_blobs 3D Gaussian blobs
    (sigma and amplitude uniform in the given ranges) that move 1 px per t step
    along a random direction in the (y, x) plane and fade by `fade` per t step,
    Poisson noise around the blob intensity, and a `saturated` fraction of
    pixels at 65535."""
    rng = np.random.default_rng(seed)
    T, C, Z, Y, X = shape
    lam = np.full(shape, base, np.float32)
    cz = rng.uniform(0, Z, n_blobs)
    cy = rng.uniform(0, Y, n_blobs)
    cx = rng.uniform(0, X, n_blobs)
    s = rng.uniform(*sigma, n_blobs)
    a = rng.uniform(*amp, n_blobs)
    ang = rng.uniform(0, 2 * np.pi, n_blobs)
    zz = np.arange(Z, dtype=np.float32)
    for t in range(T):
        for b in range(n_blobs):
            y0c, x0c = cy[b] + t * np.sin(ang[b]), cx[b] + t * np.cos(ang[b])
            r = 3.0 * s[b]
            ya, yb = max(0, int(y0c - r)), min(Y, int(y0c + r) + 1)
            xa, xb = max(0, int(x0c - r)), min(X, int(x0c + r) + 1)
            if ya >= yb or xa >= xb:
                continue
            k = -1.0 / (2 * s[b] * s[b])
            gz = np.exp(k * (zz - cz[b]) ** 2)
            gy = np.exp(k * (np.arange(ya, yb, dtype=np.float32) - y0c) ** 2)
            gx = np.exp(k * (np.arange(xa, xb, dtype=np.float32) - x0c) ** 2)
            w = np.float32(a[b] * (1.0 - fade) ** t)
            for c in range(C):
                lam[t, c, :, ya:yb, xa:xb] += w * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    vol = rng.poisson(lam).astype(np.float64)
    np.clip(vol, 0, 65535, out=vol)
    out = vol.astype(np.uint16)
    flat = out.reshape(-1)
    flat[rng.choice(flat.size, size=int(flat.size * saturated), replace=False)] = 65535
    return out


def zfp_bench_fixture(shape=(32, 64, 64)) -> np.ndarray:
    """The f32 chunk the reference's ZFP chunk benchmarks encode and decode
    (bench/rs/ndic-bench-cli/src/workloads/zfp.rs:58-81): at (z, y, x),
    (4 z (nz - z) + y (ny - y) + x (nx - x) + 500 + n) / 3 in f32, n the
    next value of a 32-bit xorshift (13, 17, 5; seed 0x9e3779b9) mod 5,
    drawn in z, y, x order."""
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = (4 * z * (nz - z) + y * (ny - y) + x * (nx - x)).reshape(-1)
    noise = np.empty(f.size, np.int64)
    s = 0x9E3779B9
    for i in range(f.size):
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        noise[i] = s % 5
    return ((f + 500 + noise).astype(np.float32) / np.float32(3.0)).reshape(shape)
