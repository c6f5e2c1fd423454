"""mppca_denoise_device_noalloc (workspace in the library's static device
scratch, no allocation) against mppca_denoise_device (caller workspace) on the
same inputs, through the C ABI: two calls in a row, the second on a smaller
volume, so the second call runs over the first one's leftovers in the shared
scratch. The plain form's factor route is deterministic, so the outputs,
ranks, margins and status must be identical bit for bit; the fp32 form's
projector route sums with float atomics and is held to 1e-6 of the largest
value."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PLAIN, FP32 = 0, 1


def _run(pkg, lib, vol, patch, stride, form, noalloc):
    import torch

    shape = vol.shape
    P = pkg.num_patches(shape, patch, stride)
    v = torch.from_numpy(vol).cuda()
    out = torch.full(shape, float("nan"), dtype=torch.float32, device="cuda")
    rank = torch.full((P,), -7, dtype=torch.int32, device="cuda")
    margin = torch.full((P,), float("nan"), dtype=torch.float32, device="cuda")
    info = torch.full((P,), -7, dtype=torch.int32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    geom = (*shape, *patch, stride, form)
    if noalloc:
        rc = lib.mppca_denoise_device_noalloc(v.data_ptr(), out.data_ptr(), *geom,
                                              rank.data_ptr(), margin.data_ptr(),
                                              info.data_ptr(), st)
        pkg._lib.check(rc, "mppca_denoise_device_noalloc")
    else:
        ws_bytes = lib.mppca_workspace_bytes(*geom)
        ws = torch.empty(ws_bytes + 256, dtype=torch.uint8, device="cuda")
        ptr = (ws.data_ptr() + 255) // 256 * 256
        rc = lib.mppca_denoise_device(v.data_ptr(), out.data_ptr(), *geom, ptr, ws_bytes,
                                      rank.data_ptr(), margin.data_ptr(), info.data_ptr(), st)
        pkg._lib.check(rc, "mppca_denoise_device")
    torch.cuda.synchronize()
    return (out.cpu().numpy(), rank.cpu().numpy(), margin.cpu().numpy(), info.cpu().numpy())


@pytest.mark.parametrize("cases,form,exact", [
    # factor route: a larger volume, then a smaller one
    ([((32, 32, 24, 16), (5, 5, 5), 2, 0), ((16, 16, 12, 10), (5, 5, 5), 2, 1)], PLAIN, True),
    ([((24, 20, 18, 48), (11, 11, 11), 3, 2), ((14, 13, 12, 17), (3, 5, 4), 2, 3)], PLAIN, True),
    # T = 128 (N < T - 1), then T = 24 on a smaller volume
    ([((12, 12, 12, 128), (5, 5, 5), 2, 4), ((13, 12, 11, 24), (5, 4, 6), 2, 5)], PLAIN, True),
    # the all-FP32 form
    ([((20, 18, 16, 30), (7, 7, 7), 2, 6), ((12, 12, 10, 12), (5, 5, 5), 2, 7)], FP32, False),
])
def test_noalloc_matches_device_path(pkg, lib, cases, form, exact):
    for shape, patch, stride, seed in cases:
        print("noalloc", shape, patch, stride, "form", form, flush=True)
        vol, _ = pkg.make_lowrank_volume(shape, 4, 2.0, 0.05, seed)
        want = _run(pkg, lib, vol, patch, stride, form, noalloc=False)
        got = _run(pkg, lib, vol, patch, stride, form, noalloc=True)
        assert np.isfinite(got[0]).all()
        if exact:
            np.testing.assert_array_equal(got[0], want[0])
        else:
            np.testing.assert_allclose(got[0], want[0], rtol=0,
                                       atol=1e-6 * np.abs(want[0]).max())
        np.testing.assert_array_equal(got[1], want[1])
        np.testing.assert_array_equal(got[3], want[3])
        if exact:
            np.testing.assert_array_equal(got[2], want[2])


def test_noalloc_rejects_oversize(pkg, lib):
    import torch

    shape = (96, 96, 48, 48)
    v = torch.zeros(shape, dtype=torch.float32, device="cuda")
    o = torch.zeros_like(v)
    rc = lib.mppca_denoise_device_noalloc(v.data_ptr(), o.data_ptr(), *shape, 11, 11, 11, 3,
                                          PLAIN, None, None, None,
                                          torch.cuda.current_stream().cuda_stream)
    assert rc == -3
    torch.cuda.synchronize()
    assert float(o.abs().max()) == 0.0  # nothing was enqueued
