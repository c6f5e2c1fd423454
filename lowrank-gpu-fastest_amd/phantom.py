"""Synthetic 4D (X, Y, Z, T) float32 volumes for the denoiser.

Two generators:

* ``make_lowrank_volume`` -- rank-r spatio-temporal signal plus i.i.d. noise,
  the workload definition of the benchmarks. Specification (RNG draws in
  this order from ``numpy.random.default_rng(seed)``):

  1. maps: for each of r components, an N(0, 1) field of shape (X, Y, Z),
     smoothed by a separable Gaussian (sigma voxels, kernel truncated at
     4 sigma, mirror-reflect boundary ``d c b a | a b c d | d c b a``) and
     min-max scaled to [0, 1];
  2. curves: for each component, ``period ~ U(4, 16)``, ``phase ~ U(0, 2 pi)``,
     ``tau ~ U(4, 20)`` frames; curve ``sin(2 pi t / period + phase) * exp(-t / tau)``
     for t = 0 .. T-1;
  3. ``clean = maps @ curves`` (float64), ``noisy = clean + N(0, noise^2)``,
     both returned as float32.

* ``make_phantom`` -- restatement of the reference's own phantom (common.py:47-77):
  products of cosine fields, a ~1000 baseline, amplitude-20 sinusoidal
  courses and sigma = 4 noise. Draw order and arithmetic follow the reference
  so that the same seed gives the same arrays.
"""

from __future__ import annotations

import numpy as np


def _gaussian_kernel(sigma: float) -> np.ndarray:
    radius = int(4.0 * sigma + 0.5)
    x = np.arange(-radius, radius + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum()


def gaussian_smooth(field: np.ndarray, sigma: float) -> np.ndarray:
    """Separable Gaussian filter, truncate 4 sigma, mirror-reflect boundary."""
    out = np.asarray(field, dtype=np.float64)
    if sigma <= 0:
        return out.copy()
    k = _gaussian_kernel(sigma)
    r = (len(k) - 1) // 2
    for ax in range(out.ndim):
        n = out.shape[ax]
        # reflect index map (edge sample repeated), valid for any radius
        idx = np.arange(-r, n + r)
        period = 2 * n
        idx = np.mod(idx, period)
        idx = np.where(idx >= n, period - 1 - idx, idx)
        padded = np.take(out, idx, axis=ax)
        moved = np.moveaxis(padded, ax, -1)
        acc = np.zeros(moved.shape[:-1] + (n,), dtype=np.float64)
        for j, kj in enumerate(k):
            acc += kj * moved[..., j : j + n]
        out = np.moveaxis(acc, -1, ax)
    return out


def make_lowrank_volume(shape, rank, smooth_sigma, noise_sigma, seed):
    """Rank-``rank`` signal plus N(0, noise_sigma^2) noise: ``(noisy, clean)`` float32."""
    x, y, z, t = (int(s) for s in shape)
    rng = np.random.default_rng(seed)
    maps = np.empty((x, y, z, rank), dtype=np.float64)
    for r in range(rank):
        f = gaussian_smooth(rng.standard_normal((x, y, z)), smooth_sigma)
        lo, hi = f.min(), f.max()
        maps[..., r] = (f - lo) / (hi - lo) if hi > lo else 0.0
    tt = np.arange(t, dtype=np.float64)
    curves = np.empty((rank, t), dtype=np.float64)
    for r in range(rank):
        period = rng.uniform(4.0, 16.0)
        phase = rng.uniform(0.0, 2.0 * np.pi)
        tau = rng.uniform(4.0, 20.0)
        curves[r] = np.sin(2.0 * np.pi * tt / period + phase) * np.exp(-tt / tau)
    clean = maps @ curves
    noisy = clean + rng.normal(0.0, noise_sigma, size=clean.shape)
    return noisy.astype(np.float32), clean.astype(np.float32)


def make_phantom(shape=(96, 96, 48, 48), rank=5, sigma=4.0, seed=0):
    """The reference's phantom (common.py:47-77): ``(noisy, clean)`` float32.

    Cosine-product spatial maps, sinusoidal time courses of amplitude 20, a
    ``1000 * (1.2 + field)`` baseline and N(0, sigma^2) noise.
    """
    rng = np.random.default_rng(seed)
    spatial, t = tuple(shape[:-1]), int(shape[-1])
    axes = [np.linspace(0.0, 1.0, s) for s in spatial]
    grids = np.meshgrid(*axes, indexing="ij")

    def cosine_field():
        freq = rng.uniform(0.5, 3.0, size=3)
        phase = rng.uniform(0.0, 2.0 * np.pi, size=3)
        terms = [np.cos(2.0 * np.pi * f * g + ph) for f, g, ph in zip(freq, grids, phase)]
        return np.prod(terms, axis=0)

    maps = np.stack([cosine_field() for _ in range(rank)], axis=-1)
    time_axis = np.linspace(0.0, 1.0, t)
    rows = []
    for _ in range(rank):
        f = rng.uniform(1.0, 6.0)
        ph = rng.uniform(0.0, 6.0)
        rows.append(np.sin(2.0 * np.pi * f * time_axis + ph))
    courses = np.stack(rows)
    baseline = 1000.0 * (1.2 + cosine_field())[..., None]
    clean = baseline + 20.0 * maps @ courses
    noisy = clean + rng.normal(scale=sigma, size=tuple(shape))
    return noisy.astype(np.float32), clean.astype(np.float32)


# Named workloads: (shape, patch, stride, rank, smooth_sigma, noise_sigma, seed)
WORKLOADS = {
    "small": ((32, 32, 24, 16), (5, 5, 5), 2, 4, 2.0, 0.05, 0),
    "default": ((96, 96, 48, 48), (11, 11, 11), 3, 6, 4.0, 0.03, 1),
    "t96": ((96, 96, 48, 96), (11, 11, 11), 3, 8, 4.0, 0.03, 1),
}


def workload_volume(name: str):
    """``(noisy, clean, patch, stride)`` of a named workload."""
    shape, patch, stride, rank, ss, ns, seed = WORKLOADS[name]
    noisy, clean = make_lowrank_volume(shape, rank, ss, ns, seed)
    return noisy, clean, patch, stride


# The six geometries of the mixed-shape and serving workloads:
# (shape (X, Y, Z, T), patch, stride).
MIXED_GEOMETRIES = (
    ((64, 64, 40, 30), (7, 7, 7), 2),
    ((96, 96, 48, 48), (11, 11, 11), 3),
    ((128, 128, 60, 100), (9, 9, 9), 3),
    ((80, 80, 32, 64), (7, 7, 7), 2),
    ((112, 112, 56, 24), (11, 11, 11), 4),
    ((48, 48, 48, 128), (5, 5, 5), 2),
)


def mixed_volume(k: int, seed: int):
    """``(noisy, patch, stride)`` of request ``seed`` on mixed geometry ``k``:
    the same generator, rank drawn from 4..8, noise sigma from U(0.02, 0.08),
    maps smoothed at sigma 4 voxels (all drawn from ``seed``)."""
    shape, patch, stride = MIXED_GEOMETRIES[k]
    rng = np.random.default_rng(seed)
    rank = int(rng.integers(4, 9))
    noise = float(rng.uniform(0.02, 0.08))
    noisy, _ = make_lowrank_volume(shape, rank, 4.0, noise, seed)
    return noisy, patch, stride
