"""Seeded test cases shared by the CPU and GPU tests."""
from __future__ import annotations

import numpy as np

import les_workloads as wl


def random_case(nx, ny, nz, seed, frac=0.08, anisotropic=True):
    """Grid with unequal spacings, a Bernoulli solid mask (1-cell gaps, solids
    on every face) and eight U(-1,1) fields."""
    if anisotropic:
        m = dict(nx=nx, ny=ny, nz=nz, dx=1.0 / nx, dy=0.7 / ny, dz=1.3 / nz)
    else:
        m = dict(nx=nx, ny=ny, nz=nz, dx=1.0 / nx, dy=1.0 / ny, dz=1.0 / nz)
    N = nx * ny * nz
    solid = (wl.uniform(seed, 9, N, 0, 1) < frac).astype(np.uint8)
    f = [wl.uniform(seed, s, N) for s in range(8)]
    return m, solid, f


def box_case(n, seed, nboxes=2, inlet_solids=False):
    """Cubic grid with box obstacles (convergent PCG) and U(-1,1) fields.
    With inlet_solids, a few cells of the inlet plane are solid too."""
    m = wl.mesh(n)
    N = n ** 3
    rng = np.random.default_rng(seed)
    boxes = []
    for _ in range(nboxes):
        box = []
        for _ax in range(3):
            s = int(rng.integers(2, max(3, n // 3)))
            lo = int(rng.integers(1, n - 1 - s))
            box.append((lo, lo + s))
        boxes.append(tuple(box))
    solid = wl.box_mask(n, n, n, boxes)
    if inlet_solids:
        sj = n // 2
        solid[sj * n + 2: sj * n + 5] = 1  # i = 0, j = n/2, k = 2..4
    f = [wl.uniform(seed, s, N) for s in range(8)]
    return m, solid, f


def bits_equal(a, b):
    """Bitwise equality, except that any two NaNs count as equal (x86 and the
    GPU produce different NaN payloads)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    ok = ~na
    return np.array_equal(a[ok].view(np.uint64), b[ok].view(np.uint64))


def rel_err(a, b):
    d = np.linalg.norm(np.asarray(a) - np.asarray(b))
    s = max(np.linalg.norm(np.asarray(b)), 1e-300)
    return d / s
