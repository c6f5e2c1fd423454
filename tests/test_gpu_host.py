"""Host -> host pipeline (mppca_denoise_host): the device-resident path's
result for every slab count and route, through the C ABI.

The pipeline's x-slab Gram launches have fewer columns than a whole-volume
launch, so they split each column's z walk into runs (the prefix restarts at a
run's first window): the Gram matrices agree up to FP64 rounding of the prefix
differences, the ranks are the same and the volumes agree far below the
parity tolerance (asserted here at 1e-9 of the largest value)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _volume(shape, seed, rank=4):
    rng = np.random.default_rng(seed)
    maps = rng.random(shape[:3] + (rank,))
    curves = rng.standard_normal((rank, shape[3]))
    return (maps @ curves + 0.05 * rng.standard_normal(shape)).astype(np.float32)


def _assert_same(got, want, route):
    # factor route: rounding of the Gram's prefix differences only; projector
    # route: float atomics in the scatter as well
    tol = 1e-9 if route == "factors" else 1e-6
    np.testing.assert_allclose(got, want, rtol=0, atol=tol * np.abs(want).max())


def _device_result(d, vol):
    torch = _torch()
    v = torch.from_numpy(vol).cuda()
    o = torch.empty_like(v)
    d.run_device(v, o)
    torch.cuda.synchronize()
    return o.cpu().numpy(), d.info()["rank"].copy()


@pytest.mark.parametrize("shape,patch,stride,form", [
    ((40, 22, 18, 24), (5, 5, 5), 2, "plain"),
    ((33, 20, 16, 16), (7, 6, 5), 3, "plain"),   # odd X, shifted last corners
    ((24, 12, 12, 12), (3, 3, 3), 5, "plain"),   # stride > patch: uncovered planes
    ((20, 12, 12, 80), (4, 4, 4), 2, "plain"),   # T > 64: per-patch FP64 Gram on slabs
    ((24, 16, 16, 16), (5, 5, 5), 2, "fp32"),    # projector route: whole-volume
    ((40, 22, 18, 24), (5, 5, 5), 2, "mixed"),   # precision form, per-slab FP64 re-solve
    ((20, 12, 12, 96), (4, 4, 4), 2, "mixed"),
])
def test_host_pipeline_matches_device_path(pkg, shape, patch, stride, form):
    vol = _volume(shape, 5)
    d = pkg.Denoiser(shape, patch, stride, form=form)
    want, want_rank = _device_result(d, vol)
    for nslabs in (1, 3, 6, shape[0] + 5):
        d.nslabs = nslabs
        print("host pipeline", shape, patch, stride, form, "nslabs", nslabs, flush=True)
        got = d(vol)
        if form == "mixed":  # per-patch Grams and per-plane scales: the same bits
            np.testing.assert_array_equal(got, want)
        else:
            _assert_same(got, want, d.route)
        np.testing.assert_array_equal(d.info()["rank"], want_rank)
        # the same call twice: bit for bit (same launches, no atomics on the factor route)
        if d.route == "factors":
            np.testing.assert_array_equal(d(vol), got)


def test_host_pipeline_pageable_output_through_staging(pkg, lib):
    torch = _torch()
    shape, patch, stride = (36, 20, 16, 24), (5, 5, 5), 3
    vol = _volume(shape, 7)
    d = pkg.Denoiser(shape, patch, stride)
    want, _ = _device_result(d, vol)
    stage_in = torch.empty(shape, dtype=torch.float32, pin_memory=True)
    stage_out = torch.empty(shape, dtype=torch.float32, pin_memory=True)
    out = np.full(shape, np.nan, np.float32)  # pageable
    streams = [torch.cuda.Stream() for _ in range(3)]
    torch.cuda.synchronize()
    rc = lib.mppca_denoise_host(
        vol.ctypes.data, out.ctypes.data, *shape, *patch, stride, 0, d._ws_ptr, d._ws_bytes,
        d.vol_d.data_ptr(), d.out_d.data_ptr(), stage_in.data_ptr(), stage_out.data_ptr(),
        4, 3, *(s.cuda_stream for s in streams), None, None, None, None)
    pkg._lib.check(rc, "mppca_denoise_host")
    _assert_same(out, want, "factors")


def test_host_pipeline_one_stream(pkg, lib):
    """All three streams the same (here the null stream): serial, same result."""
    torch = _torch()
    shape, patch, stride = (30, 14, 14, 16), (5, 5, 5), 2
    vol = _volume(shape, 9)
    d = pkg.Denoiser(shape, patch, stride)
    want, _ = _device_result(d, vol)
    pin_in = torch.empty(shape, dtype=torch.float32, pin_memory=True)
    out = torch.empty(shape, dtype=torch.float32, pin_memory=True)
    rc = lib.mppca_denoise_host(
        vol.ctypes.data, out.data_ptr(), *shape, *patch, stride, 0, d._ws_ptr, d._ws_bytes,
        d.vol_d.data_ptr(), d.out_d.data_ptr(), pin_in.data_ptr(), None, 5, 2,
        None, None, None, None, None, None, None)
    pkg._lib.check(rc, "mppca_denoise_host")
    _assert_same(out.numpy(), want, "factors")


def test_host_pipeline_alternating_volumes(pkg):
    """Calls that alternate between volumes, reusing the kept threads, events
    and signals: each result is its own volume's (no data of the previous
    call survives in the device buffers or in the caches the uploads bypass);
    the status the call reports agrees with the per-patch array."""
    shape, patch, stride = (40, 22, 18, 24), (5, 5, 5), 2
    vols = [_volume(shape, s) for s in (21, 22, 23)]
    d = pkg.Denoiser(shape, patch, stride)
    wants = [_device_result(d, v)[0] for v in vols]
    for k in range(9):
        print("alternating call", k, flush=True)
        got = d(vols[k % 3])
        _assert_same(got, wants[k % 3], "factors")
        assert int(d.info()["sweeps"].min()) >= 0


def test_host_pipeline_slabs_without_corners(pkg):
    """Slabs narrower than the stride complete no corner and only centre their
    planes; the next set's Gram reads those planes, whichever compute stream
    either set ran on. Alternating volumes (the workspace holds the other
    volume's centred planes when a call starts) each give their own
    device-path result and ranks, for several slab counts."""
    shape, patch, stride = (23, 23, 19, 24), (11, 11, 11), 4
    vols = [_volume(shape, s) for s in (31, 32)]
    d = pkg.Denoiser(shape, patch, stride)
    wants = [_device_result(d, v) for v in vols]
    for nslabs in (9, 12, 23):
        d.nslabs = nslabs
        for k in range(6):
            print("slabs without corners: nslabs", nslabs, "call", k, flush=True)
            got = d(vols[k % 2])
            _assert_same(got, wants[k % 2][0], "factors")
            np.testing.assert_array_equal(d.info()["rank"], wants[k % 2][1])


def test_result_buffers_outlive_later_calls(pkg):
    """Results are pinned buffers that a later call reuses only once the
    caller has dropped the array (or every view of it)."""
    import gc
    shape, patch, stride = (30, 14, 14, 16), (5, 5, 5), 2
    v1, v2 = _volume(shape, 41), _volume(shape, 42)
    d = pkg.Denoiser(shape, patch, stride)
    a1 = d(v1)
    c1 = a1.copy()
    view = a1[3:]
    del a1
    a2 = d(v2)
    c2 = a2.copy()
    np.testing.assert_array_equal(view, c1[3:])  # still v1's result
    assert not np.array_equal(c1, c2)
    del view, a2
    gc.collect()
    for _ in range(4):
        np.testing.assert_array_equal(d(v1), c1)  # each result dropped at once
    assert len(d._outs) == 2
