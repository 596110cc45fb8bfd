"""Seeded synthetic inputs for the LES hot path (grids, velocity fields, masks).

The same host arrays feed the HIP library, the oracle and the reference
build, so every generator here is deterministic in (seed, stream, cell id)
and independent of thread count or device.

Grid convention: an n^3 grid on the unit cube with dx = dy = dz = 1/n, the
convention of the reference driver (src/main.cpp:53-55, 100 cells of 0.01).
Layout: flat, idx = i*ny*nz + j*nz + k (src/field.h:40).
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) applied elementwise."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def counter_bits(seed: int, stream: int, n: int) -> np.ndarray:
    """64 random bits per cell id 0..n-1 for (seed, stream)."""
    ids = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = splitmix64(np.array([seed], dtype=np.uint64) * np.uint64(0xD1B54A32D192ED03)
                          + np.uint64(stream) * np.uint64(0x8CB92BA72F3D8DD7))
        return splitmix64(ids ^ base[0])


def uniform(seed: int, stream: int, n: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    """U(lo, hi) doubles: the top 53 bits scaled to [0, 1)."""
    u = (counter_bits(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return lo + (hi - lo) * u


def mesh(n: int, ny: int | None = None, nz: int | None = None):
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    return dict(nx=n, ny=ny, nz=nz, dx=1.0 / n, dy=1.0 / ny, dz=1.0 / nz)


def box_mask(nx, ny, nz, boxes) -> np.ndarray:
    """uint8 mask (1 = solid) from half-open boxes ((i0,i1),(j0,j1),(k0,k1))."""
    m = np.zeros((nx, ny, nz), dtype=np.uint8)
    for (i0, i1), (j0, j1), (k0, k1) in boxes:
        m[i0:i1, j0:j1, k0:k1] = 1
    return m.reshape(-1)


def centred_box_mask(n: int, side: int) -> np.ndarray:
    lo = (n - side) // 2
    return box_mask(n, n, n, [((lo, lo + side),) * 3])


def random_boxes_mask(n: int, count: int, seed: int, smin: int = 10, smax: int = 20) -> np.ndarray:
    """`count` boxes with sides uniform in [smin, smax], placed inside
    [1, n-2] so no box touches the inlet plane i = 0."""
    rng = np.random.default_rng(seed)
    boxes = []
    for _ in range(count):
        box = []
        for _ax in range(3):
            s = int(rng.integers(smin, smax + 1))
            s = min(s, n - 3)
            lo = int(rng.integers(1, n - 1 - s + 1))
            box.append((lo, lo + s))
        boxes.append(tuple(box))
    return box_mask(n, n, n, boxes)


def bernoulli_mask(n: int, frac: float, seed: int, stream: int = 7) -> np.ndarray:
    u = uniform(seed, stream, n * n * n, 0.0, 1.0)
    return (u < frac).astype(np.uint8)


def spectral_velocity(n: int, seed: int):
    """Three velocity components with E(k) ~ k^(-5/3): complex Gaussian
    Fourier coefficients scaled by |k|^(-11/6) for 1 <= |k| <= n/3, inverse
    FFT, each component normalised to rms 1."""
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(n, 1.0 / n)
    kz = np.fft.rfftfreq(n, 1.0 / n)
    kk = np.sqrt(kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2)
    amp = np.zeros_like(kk)
    band = (kk >= 1.0) & (kk <= n / 3.0)
    amp[band] = kk[band] ** (-11.0 / 6.0)
    out = []
    for _c in range(3):
        coef = (rng.standard_normal(kk.shape) + 1j * rng.standard_normal(kk.shape)) * amp
        f = np.fft.irfftn(coef, s=(n, n, n), axes=(0, 1, 2))
        f /= np.sqrt(np.mean(f * f))
        out.append(np.ascontiguousarray(f.reshape(-1)))
    return out


def config_c1():
    """32^3, u,v,w ~ U(-1,1) from seed 0, centred 8^3 obstacle, p = 0."""
    n = 32
    N = n ** 3
    u, v, w = (uniform(0, s, N) for s in range(3))
    return mesh(n), (u, v, w), centred_box_mask(n, 8)
