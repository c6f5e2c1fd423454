"""CPU tests: the oracle against the reference's golden vectors and the
mathematics, and the package's host-side geometry and generators."""

import hashlib

import numpy as np
import pytest

import oracle
from oracle.step0 import _mp_dims


# ---------------------------------------------------------------- geometry
def _corner_sets(golden):
    i = 0
    while f"corners{i}_geom" in golden:
        g = golden[f"corners{i}_geom"]
        yield i, tuple(g[:3]), tuple(g[3:6]), int(g[6])
        i += 1


def _check_corners(c, golden, i):
    c = np.ascontiguousarray(c, dtype=np.int64)
    assert len(c) == int(golden[f"corners{i}_count"])
    assert hashlib.sha256(c.tobytes()).digest() == golden[f"corners{i}_sha256"].tobytes()
    if f"corners{i}" in golden:
        np.testing.assert_array_equal(c, golden[f"corners{i}"])
    else:
        np.testing.assert_array_equal(c[:64], golden[f"corners{i}_head"])
        np.testing.assert_array_equal(c[-64:], golden[f"corners{i}_tail"])


def test_corners_match_reference(golden, pkg):
    n = 0
    for i, shape, patch, stride in _corner_sets(golden):
        _check_corners(oracle.patch_corners(shape, patch, stride), golden, i)
        _check_corners(pkg.patch_corners(shape, patch, stride), golden, i)
        assert pkg.num_patches(shape, patch, stride) == int(golden[f"corners{i}_count"])
        n += 1
    assert n >= 8


def test_default_patch_count(pkg):
    # README.md:29-30 -- 12,600 patches at 96x96x48, 11^3, stride 3
    assert pkg.num_patches((96, 96, 48, 48), (11, 11, 11), 3) == 12600
    # mixed shapes (own arithmetic of the corner rule)
    assert pkg.num_patches((64, 64, 40), (7, 7, 7), 2) == 16200
    assert pkg.num_patches((128, 128, 60), (9, 9, 9), 3) == 30258
    assert pkg.num_patches((80, 80, 32), (7, 7, 7), 2) == 20216
    assert pkg.num_patches((112, 112, 56), (11, 11, 11), 4) == 9477
    assert pkg.num_patches((48, 48, 48), (5, 5, 5), 2) == 12167


@pytest.mark.parametrize("shape,patch,stride", [
    ((9, 7, 6), (3, 4, 2), 2), ((10, 10, 10), (5, 5, 5), 3), ((7, 5, 4), (7, 5, 4), 1),
    ((12, 9, 8), (4, 3, 5), 5)])
def test_voxel_coverage(pkg, shape, patch, stride):
    """Every voxel is covered when stride <= patch on each axis; a stride larger
    than the patch leaves gaps (the reference's rule does not fill them)."""
    cov = np.zeros(shape, dtype=np.int64)
    for i, j, k in pkg.patch_corners(shape, patch, stride):
        cov[i:i + patch[0], j:j + patch[1], k:k + patch[2]] += 1
    if stride <= min(patch):
        assert cov.min() >= 1
    else:
        assert cov.min() == 0


def test_geometry_rejects_bad_input(pkg):
    with pytest.raises(ValueError):
        pkg.patch_corners((4, 4, 4), (5, 5, 5), 1)
    with pytest.raises(ValueError):
        pkg.axis_corners(10, 3, 0)


# ---------------------------------------------------------------- MP-PCA
def _patches(golden):
    i = 0
    while f"patch{i}_x" in golden:
        yield i, golden[f"patch{i}_x"], golden[f"patch{i}_xd"], int(golden[f"patch{i}_p"])
        i += 1


def test_mppca_patch_matches_reference(golden):
    n = 0
    for i, x, xd_ref, p_ref in _patches(golden):
        xd, p = oracle.mppca_patch(x)
        assert p == p_ref, f"patch{i}: rank {p} != {p_ref}"
        assert xd.dtype == xd_ref.dtype
        np.testing.assert_array_equal(xd, xd_ref, err_msg=f"patch{i}")
        n += 1
    assert n >= 8


def test_golden_ranks_are_informative(golden):
    ranks = [p for _, _, _, p in _patches(golden)]
    assert 0 in ranks and max(ranks) >= 4  # both the null and the signal branches


def test_denoise_matches_reference(golden):
    i = 0
    while f"vol{i}_x" in golden:
        x, g = golden[f"vol{i}_x"], golden[f"vol{i}_geom"]
        out = oracle.denoise(x, tuple(g[:3]), int(g[3]))
        assert out.dtype == x.dtype
        np.testing.assert_array_equal(out, golden[f"vol{i}_out"], err_msg=f"vol{i}")
        i += 1
    assert i >= 4


def test_parallel_oracle_is_bit_identical(golden):
    x, g = golden["vol1_x"], golden["vol1_geom"]
    a, ra, ma = oracle.denoise(x, tuple(g[:3]), int(g[3]), details=True)
    b, rb, mb = oracle.denoise_parallel(x, tuple(g[:3]), int(g[3]), workers=2, chunk=7)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(ma, mb)


def test_constant_patch_and_volume():
    x = np.full((64, 12), 7.5, dtype=np.float32)
    xd, p = oracle.mppca_patch(x)
    assert p == 0
    np.testing.assert_array_equal(xd, x)
    vol = np.full((8, 7, 6, 5), -2.25, dtype=np.float32)
    np.testing.assert_array_equal(oracle.denoise(vol, (3, 3, 3), 2), vol)


def test_mp_rank_flat_spectrum_is_zero():
    for m, n in [(15, 125), (47, 1331), (63, 64)]:
        assert oracle.mp_rank(np.full(m, 0.37), n, m)[0] == 0


def test_mp_rank_counts_dominant_spikes():
    rng = np.random.default_rng(3)
    n, t = 1331, 48
    noise = rng.standard_normal((n, t - 1)) * 0.1
    s = np.linalg.svd(noise, compute_uv=False)
    for r in (1, 3, 6):
        eigs = np.sort(np.concatenate([np.full(r, 1e3), s[r:] ** 2]))[::-1] / n
        assert oracle.mp_rank(eigs, n, t - 1)[0] == r


def test_mp_rank_scale_invariant():
    rng = np.random.default_rng(4)
    for _ in range(20):
        m, n = 23, 343
        eigs = np.sort(rng.gamma(1.0, 1.0, m) * rng.choice([1, 1, 50], m))[::-1]
        p = oracle.mp_rank(eigs, n, m)[0]
        for c in (1e-3, 4.0, 1e5):  # powers of 2 are exact; others within margin
            if c == 4.0:
                assert oracle.mp_rank(eigs * c, n, m)[0] == p


def test_gram_route_equals_svd_route_fp64():
    """U_p S_p V_p^T + m == Xc V_p V_p^T + m (step4_gram_eigh.py:1-17)."""
    rng = np.random.default_rng(5)
    for n, t, r in [(125, 16, 3), (343, 30, 5), (1331, 48, 6)]:
        x = rng.random((n, r)) @ rng.standard_normal((r, t)) + 0.05 * rng.standard_normal((n, t))
        a, pa = oracle.mppca_patch(x)
        b, pb = oracle.gram_route_patch(x)
        assert pa == pb
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-10)


def test_wide_patch_convention_fp64():
    """N < T-1 (mixed shape f): M = min(N, T-1), n = max(N, T-1); SVD == Gram route."""
    rng = np.random.default_rng(6)
    n, t = 40, 64
    x = rng.random((n, 3)) @ rng.standard_normal((3, t)) + 0.05 * rng.standard_normal((n, t))
    assert _mp_dims(n, t) == (40, 63)
    a, pa = oracle.mppca_patch(x)
    b, pb = oracle.gram_route_patch(x)
    assert pa == pb and pa >= 3
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-10)


def test_force_ranks_reproduce_mp_ranks(golden):
    x, g = golden["vol0_x"], golden["vol0_geom"]
    out, ranks, _ = oracle.denoise(x, tuple(g[:3]), int(g[3]), details=True)
    forced = oracle.denoise(x, tuple(g[:3]), int(g[3]), force_ranks=ranks)
    np.testing.assert_array_equal(out, forced)


# ---------------------------------------------------------------- generators
def test_make_phantom_matches_reference(pkg, golden):
    i = 0
    while f"phantom{i}_args" in golden:
        a = golden[f"phantom{i}_args"]
        shape, rank, sigma, seed = tuple(int(v) for v in a[:4]), int(a[4]), float(a[5]), int(a[6])
        noisy, clean = pkg.make_phantom(shape, rank=rank, sigma=sigma, seed=seed)
        np.testing.assert_array_equal(noisy, golden[f"phantom{i}_noisy"])
        np.testing.assert_array_equal(clean, golden[f"phantom{i}_clean"])
        i += 1
    assert i >= 2


def test_gaussian_smooth_matches_scipy(pkg):
    ndimage = pytest.importorskip("scipy.ndimage")
    rng = np.random.default_rng(7)
    f = rng.standard_normal((13, 9, 11))
    for s in (0.7, 2.0, 4.0):
        ref = ndimage.gaussian_filter(f, s, mode="reflect", truncate=4.0)
        np.testing.assert_allclose(pkg.gaussian_smooth(f, s), ref, rtol=0, atol=1e-12)


def test_lowrank_volume_spec(pkg):
    noisy, clean = pkg.make_lowrank_volume((12, 10, 8, 16), 4, 2.0, 0.05, 0)
    assert noisy.dtype == np.float32 and noisy.shape == (12, 10, 8, 16)
    again, _ = pkg.make_lowrank_volume((12, 10, 8, 16), 4, 2.0, 0.05, 0)
    np.testing.assert_array_equal(noisy, again)
    flat = clean.reshape(-1, 16).astype(np.float64)
    assert np.linalg.matrix_rank(flat, tol=1e-4 * np.abs(flat).max()) == 4
    assert abs(float((noisy - clean).std()) - 0.05) < 0.005
