"""Patch geometry of the denoiser (host side, integer work only).

The corner rule is the reference's (common.py:37-44): on each axis the
strided corners ``0, s, 2s, ...`` that fit, plus one last corner flush with
the border when the stride does not land on it; patches are ordered x outer,
z inner (``itertools.product``), so patch ``b = (ix*ncy + iy)*ncz + iz``.
Here the per-axis corner lists are closed-form, which is also how the kernels
compute a patch's corner from its index without any uploaded table.
"""

from __future__ import annotations

import numpy as np


def axis_corners(size: int, p: int, stride: int) -> np.ndarray:
    """Corners along one axis: ``i*stride`` for ``i*stride <= size-p``, plus ``size-p``."""
    if not (1 <= p <= size):
        raise ValueError(f"patch extent {p} must be in [1, {size}]")
    if stride < 1:
        raise ValueError(f"stride {stride} must be >= 1")
    last = size - p
    n_strided = last // stride + 1
    c = np.arange(n_strided, dtype=np.int64) * stride
    if last % stride:
        c = np.append(c, last)
    return c


def patch_corners(vol_shape, patch, stride) -> np.ndarray:
    """All patch corners ``(P, 3)`` int64 in the reference's order."""
    ax = [axis_corners(int(s), int(p), int(stride)) for s, p in zip(vol_shape[:3], patch)]
    g = np.meshgrid(*ax, indexing="ij")
    return np.stack([a.reshape(-1) for a in g], axis=-1)


def num_patches(vol_shape, patch, stride) -> int:
    n = 1
    for s, p in zip(vol_shape[:3], patch):
        n *= len(axis_corners(int(s), int(p), int(stride)))
    return n


def check_geometry(vol_shape, patch, stride):
    """Validate a (X, Y, Z, T) shape against a patch and stride; raise ValueError."""
    if len(vol_shape) != 4:
        raise ValueError(f"volume must be 4D (X, Y, Z, T), got shape {tuple(vol_shape)}")
    if len(patch) != 3:
        raise ValueError(f"patch must have 3 extents, got {patch}")
    x, y, z, t = (int(s) for s in vol_shape)
    if min(x, y, z) < 1:
        raise ValueError(f"empty volume {tuple(vol_shape)}")
    if t < 2:
        raise ValueError(f"need T >= 2 frames, got {t}")
    if t > 128:
        raise ValueError(f"T = {t} frames exceeds the supported 128")
    for s, p in zip((x, y, z), patch):
        if not (1 <= int(p) <= s):
            raise ValueError(f"patch {tuple(patch)} does not fit volume {tuple(vol_shape)}")
    if int(stride) < 1:
        raise ValueError(f"stride must be >= 1, got {stride}")
    if int(np.prod(patch)) > 4096:
        raise ValueError(f"patch of {int(np.prod(patch))} voxels exceeds 4096")
