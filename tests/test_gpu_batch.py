"""Batched form (mppca_denoise_device_batch, lowrank-gpu-fastest_amd/batch.py):
volumes of one geometry denoised by shared launches give each volume the
result it gets alone; HIP-graph replays give the eager result; volumes of
the six mixed geometries (the batched mixed-shape workload) match step 0."""

import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _single(pkg, vol, patch, stride, form):
    import torch

    d = pkg.Denoiser(vol.shape, patch, stride, form=form)
    v = torch.from_numpy(vol).cuda()
    o = torch.empty_like(v)
    d.run_device(v, o)
    torch.cuda.synchronize()
    return o.cpu().numpy(), d.info()


@pytest.mark.parametrize("shape,patch,stride,form", [
    ((20, 18, 16, 30), (7, 7, 7), 2, "plain"),      # int8 sliding Gram, factor route
    ((24, 20, 18, 48), (11, 11, 11), 3, "plain"),
    ((14, 13, 12, 17), (3, 5, 4), 2, "plain"),      # odd T, anisotropic patch
    ((15, 14, 10, 56), (13, 12, 3), 2, "plain"),    # FP64 sliding Gram
    ((12, 12, 12, 128), (5, 5, 5), 2, "plain"),     # T = 128, N < T - 1
    ((13, 12, 11, 80), (5, 4, 6), 2, "plain"),
    ((16, 16, 14, 24), (5, 5, 5), 2, "fp32"),       # all-FP32 form
])
def test_batch_matches_single(pkg, shape, patch, stride, form):
    vols = [pkg.make_lowrank_volume(shape, 4, 2.0, 0.05, 30 + k)[0] for k in range(3)]
    bd = pkg.BatchDenoiser(shape, patch, stride, capacity=3, form=form)
    got = bd(vols)
    exact = bd.route == "factors"  # no atomics: bit for bit
    for k, v in enumerate(vols):
        print("batch", shape, patch, stride, form, "volume", k, flush=True)
        want, winfo = _single(pkg, v, patch, stride, form)
        info = bd.info(k)
        np.testing.assert_array_equal(info["rank"], winfo["rank"])
        np.testing.assert_array_equal(info["sweeps"], winfo["sweeps"])
        if exact:
            np.testing.assert_array_equal(got[k], want)
        else:
            np.testing.assert_allclose(got[k], want, rtol=0, atol=1e-6 * np.abs(want).max())
    # a smaller batch in the same buffers
    got2 = bd(vols[1:2])
    if exact:
        np.testing.assert_array_equal(got2[0], got[1])
    bd.close()


@pytest.mark.parametrize("shape,patch,stride", [((20, 18, 16, 30), (7, 7, 7), 2),
                                                ((12, 12, 12, 100), (5, 5, 5), 2)])
def test_batch_graph_replays(pkg, shape, patch, stride):
    """A captured batch replayed several times, with eager work and new input
    data between the replays, gives the eager batch's result every time."""
    import torch

    bd = pkg.BatchDenoiser(shape, patch, stride, capacity=2)
    exact = bd.route == "factors"
    for rep in range(3):
        vols = [pkg.make_lowrank_volume(shape, 4, 2.0, 0.05, 50 + 2 * rep + k)[0]
                for k in range(2)]
        for k in range(2):
            bd.vol_d[k].copy_(torch.from_numpy(vols[k]))
        bd.run(2)
        eager = bd.out_d[:2].cpu().numpy()
        eager_rank = bd.rank.cpu().numpy().copy()
        bd.out_d.fill_(float("nan"))
        bd.replay(2)
        torch.cuda.synchronize()
        graphed = bd.out_d[:2].cpu().numpy()
        print("graph replay", rep, shape, flush=True)
        np.testing.assert_array_equal(bd.rank.cpu().numpy(), eager_rank)
        if exact:
            np.testing.assert_array_equal(graphed, eager)
        else:
            np.testing.assert_allclose(graphed, eager, rtol=0, atol=1e-6 * np.abs(eager).max())
        # eager single-volume work on the same GPU between replays
        pkg.denoise(vols[0], patch, stride)
    bd.close()


def test_denoise_many_mixed_small(pkg):
    """denoise_many over the six mixed geometries at reduced X, Y, Z (their
    patch, stride and T), two volumes of some: each equals its lone denoise."""
    geoms = [((16, 16, 14, 30), (7, 7, 7), 2), ((24, 23, 19, 48), (11, 11, 11), 3),
             ((15, 15, 15, 100), (9, 9, 9), 3), ((14, 14, 12, 64), (7, 7, 7), 2),
             ((23, 23, 19, 24), (11, 11, 11), 4), ((12, 12, 12, 128), (5, 5, 5), 2)]
    vols, patches, strides = [], [], []
    for k, (shape, patch, stride) in enumerate(geoms):
        for j in range(2 if k % 2 == 0 else 1):
            vols.append(pkg.make_lowrank_volume(shape, 5, 2.0, 0.04, 70 + 10 * k + j)[0])
            patches.append(patch)
            strides.append(stride)
    out, infos = pkg.denoise_many(vols, patches, strides, return_info=True)
    for v, p, s, o, inf in zip(vols, patches, strides, out, infos):
        want, winfo = _single(pkg, v, p, s, "plain")
        np.testing.assert_array_equal(inf["rank"], winfo["rank"])
        np.testing.assert_array_equal(o, want)  # the factor route at every T


def test_mixed_shapes_workload_full_size(pkg):
    """The batched mixed-shape workload at full size: the six geometries,
    seeds 10-15 (rank 4-8, noise U(0.02, 0.08)), batched and HIP-graph
    replayed, each against step 0 (relative L2 <= 5e-7, ranks equal except
    within the eigenvalue error bound)."""
    vols, patches, strides = [], [], []
    for k in range(len(pkg.MIXED_GEOMETRIES)):
        v, p, s = pkg.mixed_volume(k, 10 + k)
        vols.append(v)
        patches.append(p)
        strides.append(s)
    out, infos = pkg.denoise_many(vols, patches, strides, graph=True, return_info=True)
    workers = min(8, os.cpu_count() or 1)
    for v, p, s, o, inf in zip(vols, patches, strides, out, infos):
        ref, ranks, _ = oracle.denoise_parallel(v, p, s, workers=workers)
        bad = np.flatnonzero(inf["rank"] != ranks)
        assert np.all(inf["margin"][bad] < 1.0)
        if bad.size:
            ref, _, _ = oracle.denoise_parallel(v, p, s, workers=workers,
                                                force_ranks=inf["rank"])
        rel = float(np.linalg.norm(o - ref) / np.linalg.norm(ref))
        print(f"\n  {v.shape} patch {p} stride {s}: rel err {rel:.2e}, "
              f"rank mismatches {bad.size}", flush=True)
        assert rel <= 5e-7
