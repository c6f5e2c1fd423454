"""GPU tests of each kernel through the C ABI, against float64 NumPy and
closed forms. Tolerances are stated per test (FP32 arithmetic)."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)


def _torch():
    import torch

    return torch


def _stream():
    return _torch().cuda.current_stream().cuda_stream


def _dev(a, dtype=None):
    torch = _torch()
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def _sync():
    _torch().cuda.synchronize()


def _gather(vol, patch, stride):
    """(P, N, T) patches of a volume in the reference's voxel order."""
    import oracle as o

    px, py, pz = patch
    c = o.patch_corners(vol.shape, patch, stride)
    return np.stack([vol[i:i + px, j:j + py, k:k + pz].reshape(-1, vol.shape[-1])
                     for i, j, k in c])


def test_centre(pkg, lib):
    torch = _torch()
    rng = np.random.default_rng(0)
    for t in (2, 17, 48, 128):
        vol = (rng.standard_normal((7, 6, 5, t)) * 3 + 100).astype(np.float32)
        d = _dev(vol)
        xc = torch.empty_like(d)
        mean = torch.empty(vol.shape[:3], dtype=torch.float32, device="cuda")
        pkg._lib.check(lib.mppca_centre(d.data_ptr(), xc.data_ptr(), mean.data_ptr(),
                                        7 * 6 * 5, t, _stream()), "centre")
        _sync()
        m64 = vol.astype(np.float64).mean(-1)
        # float32 sum of T terms, then one division: |err| <= T*eps*max|x|
        tol = t * EPS32 * np.abs(vol).max()
        assert np.abs(mean.cpu().numpy() - m64).max() <= tol
        np.testing.assert_allclose(xc.cpu().numpy(), vol - mean.cpu().numpy()[..., None],
                                   rtol=0, atol=0)


@pytest.mark.parametrize("shape,patch,stride", [
    ((12, 11, 10, 10), (5, 5, 5), 2),
    ((14, 12, 12, 16), (3, 4, 5), 3),
    ((16, 16, 14, 30), (7, 7, 7), 3),
    ((13, 13, 13, 48), (11, 11, 11), 2),
    ((10, 9, 8, 64), (5, 5, 5), 2),
    ((9, 8, 8, 100), (4, 4, 4), 2),
    ((8, 8, 8, 128), (5, 5, 5), 3),
    ((6, 6, 6, 2), (3, 3, 3), 1),
])
def test_gram_vs_fp64(pkg, lib, shape, patch, stride):
    torch = _torch()
    rng = np.random.default_rng(1)
    vol = (rng.random(shape[:3] + (3,)) @ rng.standard_normal((3, shape[3]))
           + 0.05 * rng.standard_normal(shape)).astype(np.float32)
    xc = (vol - vol.mean(-1, keepdims=True, dtype=np.float64)).astype(np.float32)
    P = pkg.num_patches(shape, patch, stride)
    t = shape[3]
    g = torch.empty((P, t, t), dtype=torch.float32, device="cuda")
    # two launches over a split patch range exercise b0 / nb
    half = P // 2
    dx = _dev(xc)
    for b0, nb in ((0, half), (half, P - half)):
        gv = g[b0:]
        pkg._lib.check(lib.mppca_gram(dx.data_ptr(), gv.data_ptr(), *shape, *patch, stride,
                                      b0, nb, _stream()), "gram")
    _sync()
    X = _gather(xc, patch, stride).astype(np.float64)
    G64 = np.einsum("bnt,bns->bts", X, X)
    A = np.abs(X)
    bound = np.einsum("bnt,bns->bts", A, A) * (np.sqrt(X.shape[1]) + 2) * EPS32
    Gg = g.cpu().numpy().astype(np.float64)
    assert np.all(np.abs(Gg - G64) <= bound + 1e-30)
    np.testing.assert_array_equal(Gg, np.swapaxes(Gg, 1, 2))  # exactly symmetric


def _random_spd(rng, b, t, cond=1e4, null=True):
    q, _ = np.linalg.qr(rng.standard_normal((b, t, t)))
    lam = np.exp(rng.uniform(0, np.log(cond), (b, t)))
    if null:
        lam[:, 0] = 0.0
    return np.einsum("bij,bj,bkj->bik", q, lam, q), lam


@pytest.mark.parametrize("t", [2, 3, 16, 17, 48, 64, 96, 128])
def test_eigh_jacobi(pkg, lib, t):
    torch = _torch()
    rng = np.random.default_rng(2)
    b = 24
    G64, _ = _random_spd(rng, b, t)
    G = G64.astype(np.float32)
    dg = _dev(G)
    lam = torch.empty((b, t), dtype=torch.float32, device="cuda")
    vec = torch.empty((b, t, t), dtype=torch.float32, device="cuda")
    info = torch.empty(b, dtype=torch.int32, device="cuda")
    pkg._lib.check(lib.mppca_eigh_jacobi(dg.data_ptr(), lam.data_ptr(), vec.data_ptr(),
                                         info.data_ptr(), b, t, 30, _stream()), "eigh")
    _sync()
    L, V, inf = (lam.cpu().numpy().astype(np.float64), vec.cpu().numpy().astype(np.float64),
                 info.cpu().numpy())
    assert np.all(inf >= 0), f"no convergence: {inf}"
    ref = np.linalg.eigvalsh(G.astype(np.float64))
    nrm = np.linalg.norm(G.astype(np.float64), axis=(1, 2), ord=2)[:, None]
    assert np.all(np.diff(L, axis=1) >= 0)  # ascending
    assert np.all(np.abs(L - ref).max(axis=1) <= 8 * t * EPS32 * nrm[:, 0])
    eye = np.eye(t)
    assert np.abs(np.swapaxes(V, 1, 2) @ V - eye).max() <= 8 * t * EPS32
    resid = np.linalg.norm(G.astype(np.float64) @ V - V * L[:, None, :], axis=(1, 2))
    assert np.all(resid <= 8 * t * EPS32 * nrm[:, 0])


def test_eigh_closed_forms(pkg, lib):
    torch = _torch()
    t = 6
    mats = np.zeros((4, t, t), dtype=np.float32)
    mats[1] = np.diag([5, -1, 3, 0, 2, 7]).astype(np.float32)  # diagonal: no rotation
    a, b, c = 2.0, 3.0, -1.0
    mats[2, 0, 0], mats[2, 1, 1], mats[2, 0, 1], mats[2, 1, 0] = a, c, b, b  # 2x2 block
    mats[3] = np.full((t, t), 1.5, dtype=np.float32)  # rank one: eigenvalues 0 x5, 9
    dg = _dev(mats)
    lam = torch.empty((4, t), dtype=torch.float32, device="cuda")
    vec = torch.empty((4, t, t), dtype=torch.float32, device="cuda")
    info = torch.empty(4, dtype=torch.int32, device="cuda")
    pkg._lib.check(lib.mppca_eigh_jacobi(dg.data_ptr(), lam.data_ptr(), vec.data_ptr(),
                                         info.data_ptr(), 4, t, 30, _stream()), "eigh")
    _sync()
    L = lam.cpu().numpy()
    inf = info.cpu().numpy()
    np.testing.assert_array_equal(L[0], np.zeros(t))
    assert inf[0] == 0 and inf[1] == 0
    np.testing.assert_array_equal(L[1], np.sort([5, -1, 3, 0, 2, 7]).astype(np.float32))
    disc = np.sqrt(((a - c) / 2) ** 2 + b * b)
    two = np.sort([(a + c) / 2 - disc, (a + c) / 2 + disc, 0, 0, 0, 0])
    np.testing.assert_allclose(L[2], two, rtol=0, atol=4 * EPS32 * 4)
    np.testing.assert_allclose(L[3], [0, 0, 0, 0, 0, 9.0], rtol=0, atol=16 * EPS32 * 9)


def _mp_reference(lam_asc, n_vox, t):
    """Oracle MP rank in step 0's dtypes from ascending eigenvalues."""
    m, n = min(n_vox, t - 1), max(n_vox, t - 1)
    desc = lam_asc[::-1][:m]
    eigs = np.maximum(desc, np.float32(0)) / np.float32(n)
    assert eigs.dtype == np.float32
    return oracle.mp_rank(eigs, n, m)[0]


@pytest.mark.parametrize("n_vox,t", [(1331, 48), (125, 16), (343, 30), (729, 100),
                                     (125, 128), (64, 64), (27, 2)])
def test_mp_threshold_bit_exact(pkg, lib, n_vox, t):
    """Ranks equal the oracle's exactly on the same eigenvalues (integer result)."""
    torch = _torch()
    rng = np.random.default_rng(3)
    b = 400
    m = min(n_vox, t - 1)
    lam = np.zeros((b, t), dtype=np.float32)
    for i in range(b):
        r = rng.integers(0, min(m, 10) + 1)
        noise = np.sort(rng.gamma(4.0, 0.25, m))[::-1]
        spikes = rng.uniform(1.5, 30.0, r) * noise.max()
        vals = np.sort(np.concatenate([spikes, noise[r:]]))[::-1] * n_vox * 1e-3
        full = np.zeros(t)
        full[:m] = vals
        full[m:] = rng.uniform(-1e-6, 1e-6, t - m)  # null / rank-deficient tail
        lam[i] = np.sort(full).astype(np.float32)
    dl = _dev(lam)
    rank = torch.empty(b, dtype=torch.int32, device="cuda")
    w = torch.empty(b, dtype=torch.float32, device="cuda")
    mg = torch.empty(b, dtype=torch.float32, device="cuda")
    pkg._lib.check(lib.mppca_mp_threshold(dl.data_ptr(), b, t, n_vox, rank.data_ptr(),
                                          w.data_ptr(), mg.data_ptr(), _stream()), "mp")
    _sync()
    got = rank.cpu().numpy()
    want = np.array([_mp_reference(lam[i], n_vox, t) for i in range(b)])
    np.testing.assert_array_equal(got, want)
    if m >= 8:
        assert len(np.unique(want)) > 3
    np.testing.assert_array_equal(w.cpu().numpy(),
                                  (1.0 / (2.0 + want.astype(np.float64))).astype(np.float32))


@pytest.mark.parametrize("shape,patch,stride", [
    ((13, 12, 11, 24), (5, 4, 6), 2),     # NTB = 2
    ((13, 12, 11, 80), (5, 4, 6), 2),     # NTB = 5
    ((13, 12, 11, 100), (5, 4, 6), 2),    # NTB = 7
    ((13, 12, 11, 128), (5, 4, 6), 2),    # NTB = 8: over 64 KiB of LDS
    ((17, 16, 16, 100), (16, 16, 15), 1),  # NTB = 7 with N = 3840: over 64 KiB
])
def test_recon_and_normalize_vs_fp64(pkg, lib, shape, patch, stride):
    torch = _torch()
    rng = np.random.default_rng(4)
    t = shape[3]
    vol = rng.standard_normal(shape).astype(np.float32)
    mean = vol.mean(-1, dtype=np.float64).astype(np.float32)
    xc = vol - mean[..., None]
    P = pkg.num_patches(shape, patch, stride)
    q = rng.standard_normal((P, t, 3))
    proj = (q @ np.swapaxes(q, 1, 2)).astype(np.float32)
    wts = (1.0 / (2.0 + rng.integers(0, 8, P))).astype(np.float32)
    acc = torch.zeros(shape, dtype=torch.float32, device="cuda")
    dxc, dp, dw, dm = _dev(xc), _dev(proj), _dev(wts), _dev(mean)
    pkg._lib.check(lib.mppca_recon_scatter(dxc.data_ptr(), dp.data_ptr(), dw.data_ptr(),
                                           acc.data_ptr(), *shape, *patch, stride, 0, P,
                                           _stream()), "recon")
    out = torch.empty_like(acc)
    pkg._lib.check(lib.mppca_normalize(acc.data_ptr(), dm.data_ptr(), dw.data_ptr(),
                                       out.data_ptr(), *shape, *patch, stride, _stream()),
                   "normalize")
    _sync()
    ref_acc = np.zeros(shape)
    wsum = np.zeros(shape[:3])
    px, py, pz = patch
    for b, (i, j, k) in enumerate(pkg.patch_corners(shape, patch, stride)):
        sl = np.s_[i:i + px, j:j + py, k:k + pz]
        ref_acc[sl] += wts[b] * (xc[sl].astype(np.float64) @ proj[b].astype(np.float64))
        wsum[sl] += wts[b]
    scale = np.abs(xc).max() * np.abs(proj).max() * t * 40
    np.testing.assert_allclose(acc.cpu().numpy(), ref_acc, rtol=0, atol=scale * 4 * EPS32)
    ref_out = mean[..., None] + ref_acc / wsum[..., None]
    np.testing.assert_allclose(out.cpu().numpy(), ref_out, rtol=0,
                               atol=4 * EPS32 * (np.abs(ref_out).max() + scale))
    # fused factor form on P_b = q q^T: same result, no zeroed accumulator needed
    fused = torch.full_like(acc, float("nan"))
    fac = _dev(_factor_blocks([q[b].astype(np.float32) for b in range(P)], wts, t))
    rk = _dev(np.full(P, 3, np.int32))
    pkg._lib.check(lib.mppca_recon_factors(dxc.data_ptr(), fac.data_ptr(), rk.data_ptr(),
                                           dw.data_ptr(), dm.data_ptr(), fused.data_ptr(), *shape,
                                           *patch, stride, _stream()), "recon_factors")
    _sync()
    np.testing.assert_allclose(fused.cpu().numpy(), ref_out, rtol=0,
                               atol=4 * EPS32 * (np.abs(ref_out).max() + scale))


def _factor_blocks(vecs, wts, t):
    """Factor blocks in the layout mppca_eig64_factors writes (include/mppca.h):
    TF = 16*ceil(T/16); Ut[nc][c][t] at nc*16*TF + c*TF + t, Uw[nc][t][c] at
    TF*TF + nc*16*TF + t*16 + c; vector g of a tile of q vectors in slot
    4*(g // S2) + g % S2 with S2 = ceil(q/2) for rank p <= 8, else ceil(q/4)."""
    tf = 16 * ((t + 15) // 16)
    F = np.zeros((len(vecs), 2 * tf * tf), np.float32)
    for b, V in enumerate(vecs):
        p = V.shape[1]
        for g in range(p):
            nc, g16 = divmod(g, 16)
            q = min(16, p - 16 * nc)
            s2 = (q + 1) // 2 if p <= 8 else (q + 3) // 4
            c = 4 * (g16 // s2) + g16 % s2
            F[b, nc * 16 * tf + c * tf + np.arange(t)] = V[:, g]
            F[b, tf * tf + nc * 16 * tf + np.arange(t) * 16 + c] = (
                np.float64(wts[b]) * V[:, g].astype(np.float64)).astype(np.float32)
    return F


@pytest.mark.parametrize("shape,patch,stride", [
    ((21, 19, 37, 48), (11, 11, 11), 3),   # tiles cut by the volume edge on every axis
    ((9, 10, 13, 16), (3, 4, 5), 2),
    ((12, 9, 8, 10), (4, 3, 5), 5),        # gaps: uncovered voxels must come out 0
    ((7, 7, 7, 64), (7, 7, 7), 1),         # one patch
    ((11, 9, 10, 33), (5, 4, 6), 2),       # T not a multiple of 4: scalar loads
    ((10, 12, 9, 40), (6, 6, 6), 2),       # ranks up to 20: two slot tiles
])
def test_recon_factors_shapes(pkg, lib, shape, patch, stride):
    """Fused factor reconstruction + normalisation against float64 NumPy, with
    ranks 0 .. min(T, 20) (empty patches, partial slot groups, two slot tiles)."""
    torch = _torch()
    rng = np.random.default_rng(12)
    t = shape[3]
    xc = rng.standard_normal(shape).astype(np.float32)
    mean = rng.standard_normal(shape[:3]).astype(np.float32)
    P = pkg.num_patches(shape, patch, stride)
    ranks = rng.integers(0, min(t, 20) + 1, P).astype(np.int32)
    vecs = [np.linalg.qr(rng.standard_normal((t, t)))[0][:, :r].astype(np.float32)
            for r in ranks]
    wts = (1.0 / (2.0 + ranks)).astype(np.float32)
    out = torch.full(shape, float("nan"), dtype=torch.float32, device="cuda")
    dxc, dw, dm = _dev(xc), _dev(wts), _dev(mean)
    fac, rk = _dev(_factor_blocks(vecs, wts, t)), _dev(ranks)
    assert lib.mppca_factor_floats(t) == fac.shape[1]
    pkg._lib.check(lib.mppca_recon_factors(dxc.data_ptr(), fac.data_ptr(), rk.data_ptr(),
                                           dw.data_ptr(), dm.data_ptr(), out.data_ptr(), *shape,
                                           *patch, stride, _stream()), "recon_factors")
    _sync()
    acc = np.zeros(shape)
    wsum = np.zeros(shape[:3])
    px, py, pz = patch
    for b, (i, j, k) in enumerate(pkg.patch_corners(shape, patch, stride)):
        sl = np.s_[i:i + px, j:j + py, k:k + pz]
        V = vecs[b].astype(np.float64)
        acc[sl] += wts[b] * (xc[sl].astype(np.float64) @ V @ V.T)
        wsum[sl] += wts[b]
    with np.errstate(invalid="ignore", divide="ignore"):
        ref = np.where(wsum[..., None] > 0, mean[..., None] + acc / wsum[..., None], 0.0)
    got = out.cpu().numpy()
    assert np.isfinite(got).all()
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * (np.abs(ref).max() + 1))


def test_recon_factors_rejects_large_t(pkg, lib):
    torch = _torch()
    shape = (6, 6, 6, 130)
    d = torch.zeros(shape, dtype=torch.float32, device="cuda")
    m = torch.zeros(shape[:3], dtype=torch.float32, device="cuda")
    assert lib.mppca_factor_floats(130) < 0
    rc = lib.mppca_recon_factors(d.data_ptr(), d.data_ptr(), d.data_ptr(), d.data_ptr(),
                                 m.data_ptr(), m.data_ptr(), *shape, 3, 3, 3, 2, _stream())
    assert rc == -2  # MPPCA_EUNSUPPORTED: T above 128


def _slide_bound(xc, patch, stride):
    """Rounding bound of the sliding-window Gram (T <= 64): G = P(cz + pz) - P(cz)
    with P(z) the FP64 prefix over the planes z' < z of the patch's x-y
    footprint. Products of float32 values are exact; each prefix carries at
    most (cz + pz) * px * py rounded additions of terms bounded by
    Q = sum_{z' < cz + pz} |x| |x|^T, and the difference one more rounding."""
    import oracle as o

    px, py, pz = patch
    A = np.abs(xc.astype(np.float64))
    out = []
    for i, j, k in o.patch_corners(xc.shape, patch, stride):
        F = A[i:i + px, j:j + py, :k + pz].reshape(-1, xc.shape[-1])
        out.append(2.0 * (k + pz) * px * py * 2.0 ** -52 * (F.T @ F))
    return np.stack(out)


def _limb_bound(xc, patch, stride):
    """Quantisation bound of the int8-limb sliding Gram: each plane of a
    column's footprint is rounded to integers at 2^-30 of its largest |x|
    (M), and limb pairs worth at most 2^-34.9 M^2 per voxel are left out, so
    entry (t, s) of a window is off by at most
    sum_z 2^-30 M_z (S_z[t] + S_z[s]) + F 2^-34 M_z^2, S_z = sum of |x| over
    the plane's voxels (the square of the rounding error included)."""
    import oracle as o

    px, py, pz = patch
    A = np.abs(xc.astype(np.float64))
    out = []
    for i, j, k in o.patch_corners(xc.shape, patch, stride):
        F = A[i:i + px, j:j + py, k:k + pz]          # px, py, pz, T
        M = F.max(axis=(0, 1, 3))                    # per plane
        S = F.sum(axis=(0, 1))                       # pz, T
        b = 2.0 ** -30 * (M[:, None, None] * (S[:, :, None] + S[:, None, :])).sum(0)
        out.append(b + px * py * 2.0 ** -34 * (M ** 2).sum())
    return np.stack(out)


def test_gram64_exact(pkg, lib):
    """FP64 Gram. A launch over whole x-y columns slides a window along each
    column (on int8 limbs of each plane rounded at 2^-30 of its largest value
    up to 128 frames with a footprint of at most 128 voxels, or on FP64 MFMA
    up to 64 frames) and is held to the prefix-difference bound plus the
    quantisation bound; a launch over part of the patches (not whole columns)
    takes the per-patch kernel (products of float32 values exact in FP64,
    sums of 53 bits), held to the patch-sum bound."""
    torch = _torch()
    rng = np.random.default_rng(8)
    for shape, patch, stride in [((13, 13, 13, 48), (11, 11, 11), 2), ((9, 8, 8, 100), (4, 4, 4), 2),
                                 ((10, 9, 8, 17), (5, 3, 4), 2), ((8, 8, 8, 128), (5, 5, 5), 3),
                                 ((9, 7, 30, 33), (3, 4, 5), 3), ((8, 9, 23, 64), (4, 3, 3), 5),
                                 ((6, 5, 40, 2), (2, 2, 7), 1),
                                 # T = 24, 32 and 20 with 81-, 121- and 100-voxel footprints:
                                 # the 48-frame int8 kernel (wavefronts of blocks past T
                                 # stage and convert but issue no MFMA)
                                 ((12, 12, 10, 24), (9, 9, 3), 2), ((13, 13, 9, 32), (11, 11, 3), 2),
                                 ((12, 12, 8, 20), (10, 10, 4), 3),
                                 # T = 56 with a 156-voxel footprint: the FP64 sliding kernel
                                 ((15, 14, 10, 56), (13, 12, 3), 2),
                                 # 65..128 frames: the wide int8 kernel (one and two
                                 # 64-voxel halves per tile row, 5..8 frame blocks)
                                 ((13, 13, 12, 96), (11, 11, 5), 3), ((11, 10, 9, 100), (9, 9, 3), 2),
                                 ((10, 10, 8, 80), (4, 4, 4), 2), ((10, 10, 7, 128), (9, 8, 3), 2),
                                 ((9, 9, 8, 72), (3, 5, 4), 3)]:
        print("gram64", shape, patch, stride, flush=True)
        vol = (rng.random(shape[:3] + (3,)) @ rng.standard_normal((3, shape[3]))
               + 0.05 * rng.standard_normal(shape)).astype(np.float32)
        xc = (vol - vol.mean(-1, keepdims=True, dtype=np.float64)).astype(np.float32)
        P, t = pkg.num_patches(shape, patch, stride), shape[3]
        X = _gather(xc, patch, stride).astype(np.float64)
        G64 = np.einsum("bnt,bns->bts", X, X)
        A = np.abs(X)
        exact = np.einsum("bnt,bns->bts", A, A) * X.shape[1] * 2.0 ** -52
        # whole columns: a sliding kernel (int8 limbs up to 128 frames with a
        # footprint of at most 128 voxels, else FP64 MFMA up to 64 frames)
        sliding = t <= 64 or patch[0] * patch[1] <= 128
        slide = (np.maximum(exact, _slide_bound(xc, patch, stride) + _limb_bound(xc, patch, stride)
                            + np.abs(G64) * 2.0 ** -52) if sliding else exact)
        dx = _dev(xc)
        for b0, nb in ((0, P), (1, P - 2)):  # whole volume; a range of partial columns
            bound = slide if b0 == 0 else exact
            g = torch.full((nb, t, t), float("nan"), dtype=torch.float64, device="cuda")
            pkg._lib.check(lib.mppca_gram64(dx.data_ptr(), g.data_ptr(), *shape, *patch, stride,
                                            b0, nb, _stream()), "gram64")
            _sync()
            assert np.all(np.abs(g.cpu().numpy() - G64[b0:b0 + nb]) <= bound[b0:b0 + nb])


def test_gram64_nonfinite(pkg, lib):
    """A NaN or an infinity in the volume makes the Gram of every patch that
    holds it non-finite (the eigensolver flags such matrices). Patches of
    other x-y columns, and windows of the same column that end before its
    plane, keep their bounds."""
    torch = _torch()
    rng = np.random.default_rng(11)
    for shape, patch, stride, bad in [((13, 13, 16, 48), (11, 11, 5), 2, np.nan),
                                      ((12, 12, 10, 96), (11, 11, 3), 2, np.nan),
                                      ((10, 9, 12, 17), (5, 3, 4), 2, np.inf),
                                      ((8, 8, 10, 64), (4, 3, 3), 3, np.nan)]:
        print("gram64 nonfinite", shape, patch, stride, bad, flush=True)
        xc = (rng.standard_normal(shape) * rng.uniform(0.5, 2.0)).astype(np.float32)
        v = (shape[0] // 2, shape[1] // 2, shape[2] // 2)
        xc[v + (shape[3] // 3,)] = bad
        P, t = pkg.num_patches(shape, patch, stride), shape[3]
        import oracle as o
        corners = o.patch_corners(shape, patch, stride)
        ok = xc.copy()
        ok[v + (shape[3] // 3,)] = 0.0
        X = _gather(ok, patch, stride).astype(np.float64)
        G64 = np.einsum("bnt,bns->bts", X, X)
        bound = (_slide_bound(ok, patch, stride) + _limb_bound(ok, patch, stride)
                 + np.abs(G64) * 2.0 ** -52)
        g = torch.full((P, t, t), 0.0, dtype=torch.float64, device="cuda")
        pkg._lib.check(lib.mppca_gram64(_dev(xc).data_ptr(), g.data_ptr(), *shape, *patch, stride,
                                        0, P, _stream()), "gram64")
        _sync()
        got = g.cpu().numpy()
        px, py, pz = patch
        for b, (i, j, k) in enumerate(corners):
            inx = i <= v[0] < i + px and j <= v[1] < j + py
            if inx and k <= v[2] < k + pz:
                assert not np.isfinite(got[b]).all(), (b, (i, j, k))
            elif not inx or k + pz <= v[2]:
                assert np.all(np.abs(got[b] - G64[b]) <= bound[b]), (b, (i, j, k))


def _eig64(pkg, lib, G, n_vox):
    torch = _torch()
    b, t, _ = G.shape
    dg = _dev(G.astype(np.float64))
    proj = torch.empty((b, t, t), dtype=torch.float32, device="cuda")
    rank = torch.empty(b, dtype=torch.int32, device="cuda")
    w = torch.empty(b, dtype=torch.float32, device="cuda")
    mg = torch.empty(b, dtype=torch.float32, device="cuda")
    lam = torch.empty((b, t), dtype=torch.float32, device="cuda")
    info = torch.empty(b, dtype=torch.int32, device="cuda")
    pkg._lib.check(lib.mppca_eig64_mp(dg.data_ptr(), b, t, n_vox, proj.data_ptr(),
                                      rank.data_ptr(), w.data_ptr(), mg.data_ptr(),
                                      lam.data_ptr(), info.data_ptr(), _stream()), "eig64_mp")
    _sync()
    return (proj.cpu().numpy(), rank.cpu().numpy(), w.cpu().numpy(), mg.cpu().numpy(),
            lam.cpu().numpy(), info.cpu().numpy())


@pytest.mark.parametrize("n_vox,t", [(1331, 48), (125, 16), (343, 30), (729, 96), (1331, 128),
                                     (64, 2), (125, 3), (60, 17), (40, 64), (343, 64), (200, 49),
                                     (729, 57)])
def test_eig64_mp_vs_numpy(pkg, lib, n_vox, t):
    """Eigenvalues, MP rank and rank-p projector of FP64 Gram matrices of
    low-rank + noise patches against numpy.linalg.eigh in float64."""
    rng = np.random.default_rng(9)
    b = 32
    X = np.empty((b, n_vox, t))
    for i in range(b):
        r = int(rng.integers(0, min(8, t - 1) + 1))
        sig = rng.random((n_vox, r)) @ rng.standard_normal((r, t)) * rng.uniform(0.2, 3.0)
        X[i] = sig + rng.uniform(0.01, 0.1) * rng.standard_normal((n_vox, t))
    X -= X.mean(axis=2, keepdims=True)
    G = np.einsum("bnt,bns->bts", X, X)
    proj, rank, w, mg, lam, info = _eig64(pkg, lib, G, n_vox)
    assert np.all(info == 0)
    ref_l, ref_v = np.linalg.eigh(G)
    ref_l, ref_v = ref_l[:, ::-1], ref_v[:, :, ::-1]
    nrm = ref_l[:, :1]
    m, n = min(n_vox, t - 1), max(n_vox, t - 1)
    tol = 4 * EPS32 * np.abs(ref_l) + 1e-12 * nrm
    # N < T - 1: the T - N trailing eigenvalues are an exact cluster at 0 (the
    # Gram's rank is N); the lane-grid solver resolves a cluster to 2^-33 of
    # the norm (isolated eigenvalues to full precision). The MP rule uses
    # only the first M = min(N, T - 1).
    tol[:, m:] += 2.0 ** -33 * nrm
    assert np.all(np.abs(lam - ref_l) <= tol)
    want = np.array([oracle.mp_rank(np.maximum(lam[i, :m], np.float32(0)) / np.float32(n), n, m)[0]
                     for i in range(b)])
    np.testing.assert_array_equal(rank, want)
    assert len(np.unique(rank)) > 1 or t <= 3
    np.testing.assert_array_equal(w, (1.0 / (2.0 + rank)).astype(np.float32))
    for i in range(b):
        vp = ref_v[i, :, :rank[i]]
        np.testing.assert_allclose(proj[i], vp @ vp.T, rtol=0, atol=1e-6)


def test_eig64_degenerate(pkg, lib):
    """Zero matrix, a repeated signal eigenvalue and a non-finite entry."""
    t, n_vox = 12, 200
    G = np.zeros((3, t, t))
    q, _ = np.linalg.qr(np.random.default_rng(10).standard_normal((t, t)))
    lam = np.full(t, 0.01)
    lam[:3] = 50.0  # triple eigenvalue, clearly signal
    lam[-1] = 0.0
    G[1] = (q * lam) @ q.T
    G[2] = G[1]
    G[2, 3, 4] = G[2, 4, 3] = np.nan
    proj, rank, w, mg, lamo, info = _eig64(pkg, lib, G, n_vox)
    assert rank[0] == 0 and info[0] == 0 and np.all(proj[0] == 0)
    assert rank[1] == 3 and info[1] == 0
    np.testing.assert_allclose(proj[1], q[:, :3] @ q[:, :3].T, rtol=0, atol=1e-6)
    assert info[2] == -2 and rank[2] == 0


@pytest.mark.parametrize("t", [48, 17, 33, 64, 50])
def test_eig64_structured(pkg, lib, t):
    """Matrices that stay split after the Householder reduction: zero couplings
    (splits, exact zero minors in the Sturm counts), exactly repeated
    eigenvalues (clusters that never isolate), exact zero eigenvalues and a
    multiple of the identity. Eigenvalues against numpy in float64, ranks
    against the MP rule on them."""
    rng = np.random.default_rng(33)
    n_vox = 1331
    mats = []
    d = np.sort(rng.uniform(0.5, 2.0, t))[::-1]
    d[:3] = [400.0, 150.0, 60.0]
    mats.append(np.diag(d))                                   # every coupling zero
    rep = np.full(t, 1.25)
    rep[:4] = [300.0, 300.0, 80.0, 80.0]                      # repeated signal and noise
    mats.append(np.diag(rep))
    blk = np.zeros((t, t))
    h = t // 2
    for a, b in ((0, h), (h, t)):                             # two blocks, one split
        x = rng.standard_normal((n_vox, b - a)) * 0.05
        x[:, 0] += rng.standard_normal(n_vox) * (3.0 if a == 0 else 1.5)
        blk[a:b, a:b] = x.T @ x
    mats.append(blk)
    zer = np.diag(np.concatenate([[500.0, 90.0], rng.uniform(0.8, 1.2, t - 5), [0.0, 0.0, 0.0]]))
    mats.append(zer)                                          # exact zero eigenvalues
    mats.append(np.eye(t) * 2.5)                              # one cluster of t copies
    G = np.stack(mats)
    proj, rank, w, mg, lam, info = _eig64(pkg, lib, G, n_vox)
    assert np.all(info == 0)
    ref_l = np.linalg.eigvalsh(G)[:, ::-1]
    nrm = np.abs(ref_l).max(axis=1, keepdims=True)
    # copies of a repeated eigenvalue never isolate: their bracket is resolved
    # to 2^-33 of the norm (isolated ones to full precision)
    assert np.all(np.abs(lam - ref_l) <= 4 * EPS32 * np.abs(ref_l) + 2.0**-33 * nrm)
    want = np.array([_mp_reference(np.sort(lam[i]).astype(np.float32), n_vox, t)
                     for i in range(len(mats))])
    np.testing.assert_array_equal(rank, want)
    np.testing.assert_array_equal(w, (1.0 / (2.0 + rank)).astype(np.float32))
    # projectors of the well-separated signal subspaces
    for i in (0, 2, 3):
        vl, vv = np.linalg.eigh(G[i])
        vp = vv[:, ::-1][:, :rank[i]]
        np.testing.assert_allclose(proj[i], vp @ vp.T, rtol=0, atol=1e-6)


@pytest.mark.parametrize("n_vox,t", [(1331, 48), (125, 16), (343, 30), (60, 17), (40, 64),
                                     (200, 40), (343, 64), (150, 52), (1331, 96), (343, 80),
                                     (125, 128), (729, 100)])
def test_eig64_factors_match_projector(pkg, lib, n_vox, t):
    """mppca_eig64_factors: same rank / weight as mppca_eig64_mp, sum of the
    slot outer products equals the projector, Uw = w * Ut^T, empty slots and
    padding frames exactly zero. Ranks up to 24 cover two slot tiles."""
    torch = _torch()
    rng = np.random.default_rng(21)
    b = 24
    X = np.empty((b, n_vox, t))
    for i in range(b):
        r = int(rng.integers(0, min(24, t - 2) + 1))
        sig = rng.random((n_vox, r)) @ rng.standard_normal((r, t)) * rng.uniform(0.5, 3.0)
        X[i] = sig + 0.01 * rng.standard_normal((n_vox, t))
    X -= X.mean(axis=2, keepdims=True)
    G = np.einsum("bnt,bns->bts", X, X)
    proj, rank, w, _, _, info = _eig64(pkg, lib, G, n_vox)
    dg = _dev(G)
    nf = lib.mppca_factor_floats(t)
    fac = torch.full((b, nf), float("nan"), dtype=torch.float32, device="cuda")
    rank2 = torch.empty(b, dtype=torch.int32, device="cuda")
    w2 = torch.empty(b, dtype=torch.float32, device="cuda")
    pkg._lib.check(lib.mppca_eig64_factors(dg.data_ptr(), b, t, n_vox, fac.data_ptr(),
                                           rank2.data_ptr(), w2.data_ptr(), None, None, None,
                                           _stream()), "eig64_factors")
    _sync()
    np.testing.assert_array_equal(rank2.cpu().numpy(), rank)
    np.testing.assert_array_equal(w2.cpu().numpy(), w)
    if t >= 30:
        assert rank.max() > 16
    F = fac.cpu().numpy()
    tf = 16 * ((t + 15) // 16)
    for i in range(b):
        p = int(rank[i])
        nct = (p + 15) // 16
        ut = F[i, :nct * 16 * tf].reshape(nct, 16, tf)
        uw = F[i, tf * tf:tf * tf + nct * 16 * tf].reshape(nct, tf, 16)
        assert np.isfinite(ut).all() and np.isfinite(uw).all()
        assert np.all(ut[:, :, t:] == 0) and np.all(uw[:, t:, :] == 0)
        U = ut.reshape(nct * 16, tf)[:, :t].astype(np.float64)
        np.testing.assert_allclose(U.T @ U, proj[i], rtol=0, atol=1e-6)
        filled = np.abs(U).sum(1) > 0
        assert filled.sum() == p
        if p <= 8:  # compact layout: slots 0..7 of the first tile only
            assert not filled[8:].any()
        if p == 0:  # the first tile is still written, as zeros
            assert np.all(F[i, :16 * tf] == 0) and np.all(F[i, tf * tf:tf * tf + 16 * tf] == 0)
        np.testing.assert_allclose(np.swapaxes(uw, 1, 2).reshape(nct * 16, tf)[:, :t],
                                   np.float64(w[i]) * U, rtol=2 * EPS32, atol=0)
