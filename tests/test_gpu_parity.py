"""End-to-end parity of the HIP denoiser with the step-0 oracle.

Bar (plain FP32 form): relative L2 error <= 5e-7 against step 0, with the
rank of every patch identical to step 0's except where the MP decision is
within the stated eigenvalue error bound (GPU margin < 1, counted and
reported). Where such patches exist, the 5e-7 comparison is made against
step 0 run with those patches' ranks forced to the GPU's.
"""

import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

REL_TOL = 5e-7


def _parity(out, info, vol, patch, stride, ref=None, ranks=None, parallel=False, tol=REL_TOL):
    if ref is None:
        if parallel:
            ref, ranks, _ = oracle.denoise_parallel(vol, patch, stride,
                                                    workers=min(8, os.cpu_count() or 1))
        else:
            ref, ranks, _ = oracle.denoise(vol, patch, stride, details=True)
    bad = np.flatnonzero(info["rank"] != ranks)
    assert np.all(info["margin"][bad] < 1.0), (
        f"rank mismatch outside the error bound at patches {bad[info['margin'][bad] >= 1][:10]}")
    raw = float(np.linalg.norm(out - ref) / np.linalg.norm(ref))
    if bad.size:
        if parallel:
            ref, _, _ = oracle.denoise_parallel(vol, patch, stride, force_ranks=info["rank"],
                                                workers=min(8, os.cpu_count() or 1))
        else:
            ref = oracle.denoise(vol, patch, stride, force_ranks=info["rank"])
    rel = float(np.linalg.norm(out - ref) / np.linalg.norm(ref))
    print(f"\n  {vol.shape} patch {patch} stride {stride}: {len(ranks)} patches, "
          f"rank mismatches {bad.size} (all within bound), rel err {rel:.2e} "
          f"(raw {raw:.2e}), sweeps max {info['sweeps'].max()}")
    assert np.all(info["sweeps"] >= 0)
    assert rel <= tol
    return rel, bad.size


def test_golden_volumes(pkg, golden):
    """The reference's own step-0 outputs (tests/golden) for float32 volumes."""
    i = 0
    while f"vol{i}_x" in golden:
        x, g = golden[f"vol{i}_x"], golden[f"vol{i}_geom"]
        i += 1
        if x.dtype != np.float32:
            continue
        patch, stride = tuple(int(v) for v in g[:3]), int(g[3])
        out, info = pkg.denoise(x, patch, stride, return_info=True)
        ref, ranks, _ = oracle.denoise(x, patch, stride, details=True)
        np.testing.assert_array_equal(ref, golden[f"vol{i - 1}_out"])
        _parity(out, info, x, patch, stride, ref=ref, ranks=ranks)


@pytest.mark.parametrize("shape,patch,stride,rank,noise", [
    ((32, 32, 24, 16), (5, 5, 5), 2, 4, 0.05),      # small workload, 2475 patches
    ((20, 18, 16, 30), (7, 7, 7), 2, 5, 0.04),
    ((24, 20, 18, 48), (11, 11, 11), 3, 6, 0.03),
    # Z - 11 = 1 mod 3: the clamped last corner opens a fifth window beside
    # four strided ones (as at Z = 48), every slot of the Gram's register ring
    ((22, 20, 24, 48), (11, 11, 11), 3, 6, 0.03),
    ((14, 13, 12, 17), (3, 5, 4), 2, 3, 0.05),      # odd T, anisotropic patch
    ((12, 12, 12, 64), (7, 7, 7), 2, 6, 0.03),
    ((11, 11, 11, 20), (11, 11, 11), 3, 4, 0.05),   # one patch = the whole volume
    ((16, 16, 14, 56), (9, 9, 9), 3, 5, 0.04),      # T in 49..64, 81-voxel footprint
])
def test_configs(pkg, shape, patch, stride, rank, noise):
    vol, _ = pkg.make_lowrank_volume(shape, rank, 2.0, noise, 11)
    out, info = pkg.denoise(vol, patch, stride, return_info=True)
    _parity(out, info, vol, patch, stride)


@pytest.mark.parametrize("shape,patch,stride,rank,noise", [
    # T above 64: per-patch FP64 Gram, FP64 eigensolver on the four-wavefront
    # lane grid, rank-p factors and the fused reconstruction
    ((17, 17, 17, 96), (11, 11, 11), 3, 8, 0.03),     # the T = 96 variant's patch
    ((15, 15, 15, 100), (9, 9, 9), 3, 6, 0.05),       # mixed shape c: 9^3, stride 3
    ((12, 12, 12, 128), (5, 5, 5), 2, 6, 0.04),       # mixed shape f: N = 125 < T - 1
    ((13, 12, 11, 65), (5, 4, 6), 2, 5, 0.05),        # first T past 64 (80-frame tiles)
    # the other mixed geometries at reduced X, Y, Z (their patch, stride and T)
    ((16, 16, 14, 30), (7, 7, 7), 2, 5, 0.06),        # shape a: 7^3 / 2, T = 30
    ((14, 14, 12, 64), (7, 7, 7), 2, 6, 0.04),        # shape d: 7^3 / 2, T = 64
    ((23, 23, 19, 24), (11, 11, 11), 4, 4, 0.07),     # shape e: 11^3 / 4, T = 24
])
def test_mixed_geometries(pkg, shape, patch, stride, rank, noise):
    """Every patch / stride / T of the mixed-shape workload against step 0."""
    vol, _ = pkg.make_lowrank_volume(shape, rank, 2.0, noise, 12)
    out, info = pkg.denoise(vol, patch, stride, return_info=True)
    _parity(out, info, vol, patch, stride)


def test_reference_phantom_baseline_1000(pkg):
    """The reference's own phantom: ~1000 baseline, sigma 4 (common.py:47-77)."""
    vol, _ = pkg.make_phantom((24, 22, 20, 24), rank=5, seed=0)
    out, info = pkg.denoise(vol, (7, 7, 7), 3, return_info=True)
    _parity(out, info, vol, (7, 7, 7), 3)


def test_default_workload_full_size(pkg):
    """96x96x48x48, 11^3 patches, stride 3 (12,600 patches) against step 0."""
    vol, _, patch, stride = pkg.workload_volume("default")
    out, info = pkg.denoise(vol, patch, stride, return_info=True)
    _parity(out, info, vol, patch, stride, parallel=True)


def test_t96_workload_full_size(pkg):
    """96x96x48x96, 11^3 patches, stride 3 (the T = 96 variant, 12,600 patches)
    against step 0."""
    vol, _, patch, stride = pkg.workload_volume("t96")
    out, info = pkg.denoise(vol, patch, stride, return_info=True)
    _parity(out, info, vol, patch, stride, parallel=True)


def test_fp32_form_reference_phantom(pkg):
    """The all-FP32 form (the reference's step-8 arithmetic) meets 5e-7 on the
    reference's own phantom, whose signal-to-noise ratio keeps the Gram
    condition number small."""
    vol, _ = pkg.make_phantom((24, 22, 20, 24), rank=5, seed=0)
    out, info = pkg.denoise(vol, (7, 7, 7), 3, form="fp32", return_info=True)
    _parity(out, info, vol, (7, 7, 7), 3)


@pytest.mark.parametrize("shape,patch,stride", [((32, 32, 24, 16), (5, 5, 5), 2),
                                                ((24, 20, 18, 48), (11, 11, 11), 3)])
def test_fp32_form_conditioning_limit(pkg, shape, patch, stride):
    """On the rank-r generator (strong signal, weak noise) the FP32 Gram route
    loses ~u * lambda_max / gap on the projector: its error against step 0 is
    ~1e-6 (NumPy's float32 Gram + eigh shows the same, 7.6e-7 and 1.2e-6 on
    these two volumes). Bound stated here: 2e-5."""
    vol, _ = pkg.make_lowrank_volume(shape, 6 if shape[3] == 48 else 4, 2.0,
                                     0.03 if shape[3] == 48 else 0.05, 11)
    out, info = pkg.denoise(vol, patch, stride, form="fp32", return_info=True)
    _parity(out, info, vol, patch, stride, tol=2e-5)


def test_constant_volume(pkg):
    vol = np.full((10, 9, 8, 12), 3.5, dtype=np.float32)
    out, info = pkg.denoise(vol, (5, 5, 5), 2, return_info=True)
    assert np.all(info["rank"] == 0)
    np.testing.assert_allclose(out, vol, rtol=2e-7, atol=0)


def test_gaps_when_stride_exceeds_patch(pkg):
    """Voxels no patch covers come out 0, as step 0's out / max(wsum, 1e-12)."""
    vol, _ = pkg.make_lowrank_volume((12, 9, 8, 10), 3, 2.0, 0.05, 2)
    patch, stride = (4, 3, 5), 5
    out = pkg.denoise(vol, patch, stride)
    ref = oracle.denoise(vol, patch, stride)
    gaps = ref == 0
    assert gaps.any()
    assert np.all(out[gaps] == 0)
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) <= REL_TOL


def test_repeat_calls_agree(pkg):
    """Float atomics reorder sums; two runs agree to a few ulps, ranks exactly."""
    vol, _ = pkg.make_lowrank_volume((20, 20, 16, 24), 4, 2.0, 0.04, 3)
    a, ia = pkg.denoise(vol, (5, 5, 5), 2, return_info=True)
    b, ib = pkg.denoise(vol, (5, 5, 5), 2, return_info=True)
    np.testing.assert_array_equal(ia["rank"], ib["rank"])
    assert np.linalg.norm(a - b) / np.linalg.norm(a) <= 1e-6


def test_device_tensor_path(pkg):
    import torch

    vol, _ = pkg.make_lowrank_volume((16, 16, 12, 10), 3, 2.0, 0.05, 0)
    host = pkg.denoise(vol, (5, 5, 5), 2)
    dev = pkg.denoise(torch.from_numpy(vol).cuda(), (5, 5, 5), 2)
    assert dev.is_cuda and dev.dtype == torch.float32
    got = dev.cpu().numpy()
    assert np.linalg.norm(got - host) / np.linalg.norm(host) <= 1e-6


def test_bad_input_raises(pkg):
    with pytest.raises(TypeError):
        pkg.denoise(np.zeros((8, 8, 8, 6), dtype=np.float64), (3, 3, 3), 2)
    with pytest.raises(ValueError):
        pkg.denoise(np.zeros((8, 8, 8, 6), dtype=np.float32), (9, 3, 3), 2)
    with pytest.raises(ValueError):
        pkg.denoise(np.zeros((8, 8, 8, 200), dtype=np.float32), (3, 3, 3), 2)


def test_nonfinite_input_status(pkg):
    """A NaN in the volume flags its patches (status -2): the host checks the
    status after the call and raises by default (step 0's SVD raises on such
    input), warns or passes on request; the device-tensor path checks too."""
    import torch

    vol, _ = pkg.make_lowrank_volume((14, 13, 12, 16), 3, 2.0, 0.05, 4)
    vol[6, 6, 6, 5] = np.nan
    with pytest.raises(pkg.SolverStatusError) as e:
        pkg.denoise(vol, (5, 5, 5), 2)
    assert (e.value.status == -2).any() and (e.value.status >= -2).all()
    with pytest.warns(RuntimeWarning):
        out = pkg.denoise(vol, (5, 5, 5), 2, on_bad="warn")
    assert np.isnan(out[6, 6, 6]).all()
    out, info = pkg.denoise(vol, (5, 5, 5), 2, on_bad="ignore", return_info=True)
    assert np.isfinite(out[0, 0, 0]).all() and (info["sweeps"] == -2).any()
    with pytest.raises(pkg.SolverStatusError):
        pkg.denoise(torch.from_numpy(vol).cuda(), (5, 5, 5), 2)


def test_device_path_permuted_input(pkg):
    """A non-contiguous CUDA input gets a contiguous output of its shape."""
    import torch

    vol, _ = pkg.make_lowrank_volume((12, 12, 10, 12), 3, 2.0, 0.05, 6)
    t = torch.from_numpy(np.ascontiguousarray(vol.transpose(1, 0, 2, 3))).cuda().permute(1, 0, 2, 3)
    assert not t.is_contiguous()
    dev = pkg.denoise(t, (5, 5, 5), 2)
    assert dev.is_contiguous() and tuple(dev.shape) == vol.shape
    host = pkg.denoise(vol, (5, 5, 5), 2)
    assert np.linalg.norm(dev.cpu().numpy() - host) / np.linalg.norm(host) <= 1e-6
