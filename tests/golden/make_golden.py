"""Write golden vectors by running the reference's own step-0 code on small inputs.

The reference (paquiteau/lowrank-gpu-fastest) ships no tests and no fixtures;
its only correctness oracle is the NumPy loop in step0_numpy_loop.py, whose MP
rule constants and count semantics no closed form pins. This script imports
that code read-only from a checkout of the reference and records
inputs/outputs of:

  * common.patch_corners            (common.py:37-44)
  * common.make_phantom             (common.py:47-77)
  * step0_numpy_loop.mppca_patch    (step0_numpy_loop.py:8-25)
  * step0_numpy_loop.denoise        (step0_numpy_loop.py:28-39)

into ``golden.npz`` next to this file. Only data is written; nothing of the
reference's source is stored.

Run from the repository root:  python tests/golden/make_golden.py [REF_DIR]
(REF_DIR defaults to /root/reference). Importing common.py would create a
``results/`` directory inside the reference checkout (common.py:33-34); the
import below suppresses that one mkdir, and bytecode caching, so the checkout
is never written.
"""

from __future__ import annotations

import hashlib
import importlib
import os
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent


def _import_reference(ref_dir: str):
    os.environ.setdefault("VOL", "96,96,48,48")
    os.environ.setdefault("STEP", "3")
    sys.dont_write_bytecode = True  # no __pycache__ inside the checkout
    sys.path.insert(0, ref_dir)
    real_mkdir = pathlib.Path.mkdir
    pathlib.Path.mkdir = lambda self, *a, **k: None  # keep the checkout untouched
    try:
        common = importlib.import_module("common")
        step0 = importlib.import_module("step0_numpy_loop")
    finally:
        pathlib.Path.mkdir = real_mkdir
        sys.path.remove(ref_dir)
    return common, step0


def _lowrank(shape, rank, noise, seed, dtype=np.float32, offset=0.0):
    """Small low-rank-plus-noise input (own generator, independent of the package)."""
    rng = np.random.default_rng(seed)
    *sp, t = shape
    maps = rng.random(tuple(sp) + (rank,))
    curves = rng.standard_normal((rank, t))
    clean = offset + maps @ curves
    return (clean + noise * rng.standard_normal(shape)).astype(dtype)


def main(ref_dir="/root/reference"):
    common, step0 = _import_reference(ref_dir)
    out = {}

    # ---- corners -------------------------------------------------------
    geoms = [
        ((32, 32, 24), (5, 5, 5), 2),
        ((16, 16, 12), (5, 5, 5), 2),
        ((14, 12, 10), (7, 7, 7), 3),
        ((64, 64, 40), (7, 7, 7), 2),
        ((112, 112, 56), (11, 11, 11), 4),
        ((48, 48, 48), (5, 5, 5), 2),
        ((11, 11, 11), (11, 11, 11), 3),
    ]
    geoms.append(((96, 96, 48), (11, 11, 11), 3))
    for gi, (shape, patch, stride) in enumerate(geoms):
        c = np.ascontiguousarray(common.patch_corners(shape, patch, stride), dtype=np.int64)
        out[f"corners{gi}_geom"] = np.array([*shape, *patch, stride], dtype=np.int64)
        out[f"corners{gi}_count"] = np.array(len(c), dtype=np.int64)
        out[f"corners{gi}_sha256"] = np.frombuffer(
            hashlib.sha256(c.tobytes()).digest(), dtype=np.uint8
        )
        if len(c) <= 3000:
            out[f"corners{gi}"] = c
        else:  # large sets: the hash pins every row, head/tail show them
            out[f"corners{gi}_head"] = c[:64]
            out[f"corners{gi}_tail"] = c[-64:]

    # ---- make_phantom ----------------------------------------------------
    for pi, (shape, rank, sigma, seed) in enumerate(
        [((12, 10, 8, 6), 3, 4.0, 0), ((8, 8, 6, 16), 5, 4.0, 1)]
    ):
        noisy, clean = common.make_phantom(shape, rank=rank, sigma=sigma, seed=seed)
        out[f"phantom{pi}_args"] = np.array([*shape, rank, sigma, seed], dtype=np.float64)
        out[f"phantom{pi}_noisy"] = noisy
        out[f"phantom{pi}_clean"] = clean

    # ---- mppca_patch -----------------------------------------------------
    mats = [
        _lowrank((125, 16), 4, 0.05, 100),
        _lowrank((343, 30), 6, 0.05, 101),
        _lowrank((1331, 48), 6, 0.03, 102),
        _lowrank((125, 16), 4, 0.05, 103, dtype=np.float64),
        _lowrank((343, 24), 5, 0.08, 104, dtype=np.float64),
        np.full((64, 12), 3.25, dtype=np.float32),
        _lowrank((343, 20), 3, 4.0, 105, offset=1000.0),
        (0.05 * np.random.default_rng(106).standard_normal((216, 20))).astype(np.float32),
        _lowrank((64, 64), 3, 0.02, 107),  # N = T: smallest legal N for step 0
    ]
    for mi, x in enumerate(mats):
        xd, p = step0.mppca_patch(x)
        out[f"patch{mi}_x"] = x
        out[f"patch{mi}_xd"] = xd
        out[f"patch{mi}_p"] = np.array(p, dtype=np.int64)

    # ---- denoise (step-0 loop) ---------------------------------------------
    vols = [
        (_lowrank((16, 16, 12, 10), 3, 0.05, 200), (5, 5, 5), 2),
        (_lowrank((14, 12, 10, 24), 4, 0.05, 201), (7, 7, 7), 3),
        (_lowrank((12, 12, 10, 8), 2, 0.05, 202, dtype=np.float64), (5, 5, 5), 2),
        (_lowrank((16, 16, 14, 12), 3, 0.05, 203), (11, 11, 11), 3),
    ]
    saved = (step0.PATCH, step0.patch_corners)
    try:
        for vi, (vol, patch, stride) in enumerate(vols):
            step0.PATCH = patch
            step0.patch_corners = (
                lambda shape=vol.shape[:3], patch=patch, stride=stride:
                common.patch_corners(shape, patch, stride)
            )
            res = step0.denoise(vol)
            out[f"vol{vi}_x"] = vol
            out[f"vol{vi}_geom"] = np.array([*patch, stride], dtype=np.int64)
            out[f"vol{vi}_out"] = res
    finally:
        step0.PATCH, step0.patch_corners = saved

    out["numpy_version"] = np.array(np.__version__)
    np.savez_compressed(HERE / "golden.npz", **out)
    print(f"wrote {HERE / 'golden.npz'} ({len(out)} arrays)")


if __name__ == "__main__":
    main(*sys.argv[1:])
