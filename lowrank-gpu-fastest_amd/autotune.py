"""Autotuner of the tile schedule: which kernel each stage runs for a geometry.

The reference picks its launch configuration by a sweep written into the
source (BLOCK_N / num_warps constants, step8_tuned.py:22-23, found by
micro/tune_recon.py:47-71). Here the choices are made per (N, T) at run time,
by timing the whole device path on the GPU it runs on:

* the FP64 Gram's kernel: int8 sliding window, FP64 sliding window or
  per-patch sums (``MPPCA_GRAM_*``; a kernel that cannot take the geometry is
  skipped);
* the eigensolver's layout: register lane grid or matrix in LDS;
* the fused reconstruction's x-slabs per wavefront: 1, 2 or 4.

The search is coordinate-wise (the Gram, then the eigensolver, then the
reconstruction tile, each with the best so far for the others) over one
batch of synthetic volumes of the geometry. :func:`install` writes the
winners into the library's autotune table, which is read-only once set: every
entry point then uses them for those (N, T). Every choice computes the same
arithmetic up to rounding (each Gram kernel within its stated bound of the
exact Gram), so a schedule changes speed, not results beyond rounding.
"""

from __future__ import annotations

import json
import os
import time

import numpy as np

from . import _lib
from .batch import BatchDenoiser

GRAMS = ("limb", "slide", "patch")
EIGS = ("grid", "lds")
TILES = (1, 2, 4)


def _time(bd, n, tuning, reps):
    """Median ms of ``reps`` batched passes with ``tuning`` (None if the
    schedule cannot run this geometry)."""
    torch = bd.torch
    bd.tuning = tuning
    try:
        bd.run(n)
        torch.cuda.synchronize()
    except _lib.MppcaError as e:
        if e.code == -2:  # MPPCA_EUNSUPPORTED: this kernel does not take the geometry
            return None
        raise
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        bd.run(n)
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(ts))


def tune(shape, patch=(11, 11, 11), stride=3, batch=1, reps=5, form="plain", vol=None):
    """Tile schedule of one geometry: ``(tuning, timings)`` where ``tuning`` is
    an ``mppca_tuning`` for (N, T) and ``timings`` maps each schedule tried to
    its median ms per batch of ``batch`` volumes."""
    import torch

    shape = tuple(int(s) for s in shape)
    patch = tuple(int(p) for p in patch)
    N, T = int(np.prod(patch)), shape[3]
    bd = BatchDenoiser(shape, patch, stride, capacity=batch, form=form)
    if vol is None:
        from .phantom import make_lowrank_volume

        vol, _ = make_lowrank_volume(shape, 6, 2.0, 0.04, 7)
    for k in range(batch):
        bd.vol_d[k].copy_(torch.from_numpy(np.ascontiguousarray(vol)))
    best = {"gram": 0, "eig": 0, "recon_sx": 0}
    timings = {}

    def trial(**kw):
        cfg = dict(best, **kw)
        t = _time(bd, batch, _lib.Tuning(N, T, cfg["gram"], cfg["eig"], cfg["recon_sx"]), reps)
        timings[f"gram={cfg['gram']} eig={cfg['eig']} recon_sx={cfg['recon_sx']}"] = t
        return t

    try:
        for field, values in (("gram", [_lib.GRAM_KERNELS[g] for g in GRAMS]),
                              ("eig", [_lib.EIG_LAYOUTS[e] for e in EIGS]),
                              ("recon_sx", list(TILES) if form != "fp32" else [])):
            if form != "plain" and field != "recon_sx":
                continue  # the fp32 and mixed forms have one Gram kernel and one eigensolver
            results = {v: trial(**{field: v}) for v in values}
            ok = {v: t for v, t in results.items() if t is not None}
            if ok:
                best[field] = min(ok, key=ok.get)
    finally:
        bd.close()
    return _lib.Tuning(N, T, best["gram"], best["eig"], best["recon_sx"]), timings


def install(tunings):
    """Install the autotune table (once per process; the library keeps it
    read-only afterwards). ``tunings``: a list of ``mppca_tuning``."""
    lib = _lib.load()
    arr = (_lib.Tuning * len(tunings))(*tunings)
    _lib.check(lib.mppca_tune_table_set(arr, len(tunings)), "mppca_tune_table_set")


def to_dict(t):
    return {"N": t.N, "T": t.T, "gram": t.gram, "eig": t.eig, "recon_sx": t.recon_sx}


def from_dict(d):
    return _lib.Tuning(int(d["N"]), int(d["T"]), int(d["gram"]), int(d["eig"]),
                       int(d["recon_sx"]))


def tune_table(geometries, path=None, **kw):
    """Tune every (shape, patch, stride) of ``geometries`` and return the
    table (one entry per (N, T)); with ``path``, load it from that JSON file
    when it was made for this GPU and library, else tune and save it."""
    import torch

    key = {"device": torch.cuda.get_device_name(0), "abi": _lib.ABI_VERSION}
    if path and os.path.exists(path):
        with open(path) as f:
            saved = json.load(f)
        if saved.get("key") == key:
            return [from_dict(d) for d in saved["entries"]]
    entries = {}
    for shape, patch, stride in geometries:
        t, _ = tune(shape, patch, stride, **kw)
        entries[(t.N, t.T)] = t
    table = list(entries.values())
    if path:
        with open(path, "w") as f:
            json.dump({"key": key, "entries": [to_dict(t) for t in table]}, f, indent=1)
    return table

