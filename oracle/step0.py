"""NumPy restatement of the reference's step-0 MP-PCA loop (test infrastructure).

Reference: paquiteau/lowrank-gpu-fastest
  * corner rule ............ common.py:37-44
  * per-patch MP-PCA ....... step0_numpy_loop.py:8-25
  * weighted recombination . step0_numpy_loop.py:28-39
  * Gram-route equivalent .. step4_gram_eigh.py:40-67

Numerics follow step 0 operation by operation (float32 SVD on float32 data,
float32 accumulators, NumPy 2 promotion of the MP inequality to float64), so
that the golden vectors written by the reference compare bit-for-bit or to
the last ulp. Geometry (``patch``, ``stride``) is an argument here, where the
reference binds it at import time (common.py:27-29).
"""

from __future__ import annotations

import itertools
import os

import numpy as np


# --------------------------------------------------------------------------
# geometry
# --------------------------------------------------------------------------
def patch_corners(vol_shape, patch, stride) -> np.ndarray:
    """Patch corners ``(P, 3)`` int64 in lexicographic (x outer, z inner) order.

    common.py:37-44: on each axis the strided positions ``0, s, 2s, ...`` that
    fit, plus one shifted last patch flush with the border when the stride
    does not land on it.
    """
    per_axis = []
    for size, p in zip(vol_shape[:3], patch):
        if p > size:
            raise ValueError(f"patch {p} larger than volume axis {size}")
        last = size - p
        per_axis.append(sorted(set(range(0, last + 1, stride)) | {last}))
    return np.array(list(itertools.product(*per_axis)), dtype=np.int64)


# --------------------------------------------------------------------------
# Marchenko-Pastur rank
# --------------------------------------------------------------------------
def mp_rank(eigs: np.ndarray, n: int, m: int | None = None):
    """MP-PCA rank from descending noise-normalised eigenvalues.

    step0_numpy_loop.py:18-24 (and step4_gram_eigh.py:49-56): with
    ``rcum[i] = sum_{j>=i} eigs[j]``, index ``i`` is signal when

        (eigs[i] - eigs[m-1]) * (m - i) * (n - i) > 4 * rcum[i] * sqrt(m * n)

    and ``p`` is the COUNT of signal indices (not the first crossing).

    ``eigs`` has length ``m`` (= T-1 when N >= T-1). The arithmetic keeps
    step 0's dtypes: ``eigs`` and ``rcum`` in the input dtype, the two sides
    of the inequality promoted to float64 by the int64 index arrays.

    Returns ``(p, margin)`` where ``margin[i] = (lhs - rhs) / (|lhs| + |rhs|)``
    is the signed relative margin of each inequality (0 where both sides are 0).
    """
    eigs = np.asarray(eigs)
    if m is None:
        m = eigs.shape[0]
    rcum = np.cumsum(eigs[::-1])[::-1]
    idx = np.arange(m)
    lhs = (eigs - eigs[-1]) * (m - idx) * (n - idx)
    rhs = 4 * rcum * np.sqrt(m * n)
    signal = lhs > rhs
    den = np.abs(lhs) + np.abs(rhs)
    margin = np.where(den > 0, (lhs - rhs) / np.where(den > 0, den, 1.0), 0.0)
    return int(signal.sum()), margin


def _mp_dims(n_vox: int, t: int):
    """(M, n) of the MP rule: M = min(N, T-1) kept values, n = max(N, T-1).

    For N >= T-1 this is step 0's (T-1, N) (step0_numpy_loop.py:10,15-17).
    For N < T-1 step 0 raises (its T-1 index range broadcasts against only N
    singular values); the transposed-problem convention used here is the
    build's own definition.
    """
    return min(n_vox, t - 1), max(n_vox, t - 1)


# --------------------------------------------------------------------------
# one patch
# --------------------------------------------------------------------------
def mppca_patch(x: np.ndarray, force_rank: int | None = None, details: bool = False):
    """MP-PCA of one ``(N, T)`` Casorati matrix: ``(denoised, p)``.

    step0_numpy_loop.py:8-25. Centre each voxel row over time, SVD, drop the
    null (constant time course) direction, MP rank, rebuild
    ``U_p S_p V_p^T + m`` in the input dtype.

    ``force_rank`` replaces the MP rank by a given ``p`` (used to compare
    against a GPU run whose rank differs on a patch whose MP margin is below
    the eigenvalue error bound). ``details`` adds the eigenvalues and margins.
    """
    n_vox, t = x.shape
    mean = x.mean(axis=1, keepdims=True)
    u, s, vt = np.linalg.svd(x - mean, full_matrices=False)
    m, n_eff = _mp_dims(n_vox, t)
    u, s, vt = u[:, :m], s[:m], vt[:m]
    eigs = s**2 / n_eff
    p, margin = mp_rank(eigs, n_eff, m)
    if force_rank is not None:
        p = int(force_rank)
    xd = (u[:, :p] * s[:p]) @ vt[:p] + mean
    if details:
        return xd, p, {"eigs": eigs, "margin": margin}
    return xd, p


def gram_route_patch(x: np.ndarray):
    """Gram-route MP-PCA of one patch in the input dtype: ``(denoised, p)``.

    step4_gram_eigh.py:40-67: ascending ``eigh(Xc^T Xc)``; drop the smallest
    (null) eigenvalue, clamp the rest at 0, divide by n, MP rank, then
    ``Xc V_p V_p^T + m`` -- U is never formed.
    """
    n_vox, t = x.shape
    mean = x.mean(axis=1, keepdims=True)
    xc = x - mean
    lam, v = np.linalg.eigh(xc.T @ xc)  # ascending, V in columns
    m, n_eff = _mp_dims(n_vox, t)
    desc = lam[::-1][:m]
    eigs = np.maximum(desc, 0) / n_eff
    p, _ = mp_rank(eigs, n_eff, m)
    vp = v[:, ::-1][:, :p]
    return xc @ (vp @ vp.T) + mean, p


# --------------------------------------------------------------------------
# whole volume
# --------------------------------------------------------------------------
def _accumulate(vol, patch, corners, results, out, wsum):
    px, py, pz = patch
    for (i, j, k), (xd, p) in zip(corners, results):
        sl = np.s_[i : i + px, j : j + py, k : k + pz]
        w = 1.0 / (2.0 + p)
        out[sl] += w * xd.reshape(px, py, pz, -1)
        wsum[sl] += w


def denoise(vol, patch=(11, 11, 11), stride=3, force_ranks=None, details=False):
    """Step-0 denoiser: ``(X, Y, Z, T)`` volume -> denoised volume, same dtype.

    step0_numpy_loop.py:28-39: every patch in corner order, ``w = 1/(2+p)``,
    accumulate ``w * xd`` and ``w``, divide by ``max(wsum, 1e-12)``. Float32
    input gives the FP32 oracle, float64 input the FP64 oracle (F9b: the
    accumulators follow the input dtype).

    ``force_ranks`` (P,) overrides the MP rank patch by patch.
    With ``details`` returns ``(out, ranks, min_abs_margin)``.
    """
    vol = np.asarray(vol)
    corners = patch_corners(vol.shape, patch, stride)
    out = np.zeros_like(vol)
    wsum = np.zeros(vol.shape[:3] + (1,), dtype=vol.dtype)
    px, py, pz = patch
    ranks = np.zeros(len(corners), dtype=np.int64)
    margins = np.zeros(len(corners), dtype=np.float64)
    for b, (i, j, k) in enumerate(corners):
        x = vol[i : i + px, j : j + py, k : k + pz].reshape(-1, vol.shape[-1])
        fr = None if force_ranks is None else force_ranks[b]
        xd, p, d = mppca_patch(x, force_rank=fr, details=True)
        ranks[b] = p
        margins[b] = np.min(np.abs(d["margin"])) if d["margin"].size else 1.0
        _accumulate(vol, patch, [(i, j, k)], [(xd, p)], out, wsum)
    res = out / np.maximum(wsum, 1e-12)
    if details:
        return res, ranks, margins
    return res


# ---- process-parallel form (same accumulation order, bit-identical) --------
_SHARED = {}


def _init_worker(vol, patch):
    try:
        from threadpoolctl import threadpool_limits

        threadpool_limits(1)
    except Exception:  # pragma: no cover - threadpoolctl is optional
        pass
    _SHARED["vol"] = vol
    _SHARED["patch"] = patch


def _work(args):
    corners, force = args
    vol, (px, py, pz) = _SHARED["vol"], _SHARED["patch"]
    res = []
    for b, (i, j, k) in enumerate(corners):
        x = vol[i : i + px, j : j + py, k : k + pz].reshape(-1, vol.shape[-1])
        fr = None if force is None else force[b]
        xd, p, d = mppca_patch(x, force_rank=fr, details=True)
        mg = float(np.min(np.abs(d["margin"]))) if d["margin"].size else 1.0
        res.append((xd, p, mg))
    return res


def denoise_parallel(vol, patch=(11, 11, 11), stride=3, workers=None,
                     force_ranks=None, chunk=64):
    """``denoise(..., details=True)`` spread over worker processes.

    Workers compute per-patch ``(xd, p)``; the parent accumulates them in
    corner order, so the result is bit-identical to the serial loop.
    Returns ``(out, ranks, min_abs_margin)``.
    """
    import multiprocessing as mp

    vol = np.asarray(vol)
    corners = patch_corners(vol.shape, patch, stride)
    if workers is None:
        workers = min(8, os.cpu_count() or 1)
    out = np.zeros_like(vol)
    wsum = np.zeros(vol.shape[:3] + (1,), dtype=vol.dtype)
    ranks = np.zeros(len(corners), dtype=np.int64)
    margins = np.zeros(len(corners), dtype=np.float64)
    jobs = []
    for c0 in range(0, len(corners), chunk):
        fr = None if force_ranks is None else np.asarray(force_ranks[c0 : c0 + chunk])
        jobs.append((corners[c0 : c0 + chunk], fr))
    # spawn, not fork: a forked child would inherit the parent's GPU device
    # handles when the parent has already initialised the GPU runtime
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers, initializer=_init_worker, initargs=(vol, patch)) as pool:
        b = 0
        for (cs, _), res in zip(jobs, pool.imap(_work, jobs)):
            _accumulate(vol, patch, cs, [(xd, p) for xd, p, _ in res], out, wsum)
            for _, p, mg in res:
                ranks[b] = p
                margins[b] = mg
                b += 1
    return out / np.maximum(wsum, 1e-12), ranks, margins
