"""Autotuner (lowrank-gpu-fastest_amd/autotune.py): every schedule it tries
gives the same denoised volume up to rounding, it picks a schedule that runs,
and the installed table is used and stays read-only."""

import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,patch,stride", [((24, 20, 18, 48), (11, 11, 11), 3),
                                                ((20, 18, 16, 30), (7, 7, 7), 2),
                                                ((12, 12, 12, 100), (5, 5, 5), 2)])
def test_every_schedule_same_result(pkg, shape, patch, stride):
    import torch

    vol, _ = pkg.make_lowrank_volume(shape, 5, 2.0, 0.04, 3)
    tuning, timings = pkg.autotune.tune(shape, patch, stride, reps=2, vol=vol)
    print(shape, {k: (round(v, 3) if v else v) for k, v in timings.items()}, flush=True)
    lib = pkg._lib
    assert tuning.N == int(np.prod(patch)) and tuning.T == shape[3]
    assert tuning.gram in lib.GRAM_KERNELS.values() and tuning.eig in lib.EIG_LAYOUTS.values()
    ran = [k for k, v in timings.items() if v is not None]
    assert ran
    bd = pkg.BatchDenoiser(shape, patch, stride, capacity=1)
    bd.vol_d[0].copy_(torch.from_numpy(vol))
    bd.run(1)
    want = bd.out_d[0].cpu().numpy()
    want_rank = bd.rank.cpu().numpy().copy()
    for gram in (1, 2, 3):
        for eig in (1, 2):
            for sx in (1, 2, 4):
                bd.tuning = lib.Tuning(tuning.N, tuning.T, gram, eig, sx)
                try:
                    bd.run(1)
                except lib.MppcaError as e:
                    assert e.code == -2  # this kernel does not take the geometry
                    continue
                got = bd.out_d[0].cpu().numpy()
                info = bd.info(0)
                bad = np.flatnonzero(info["rank"] != want_rank)
                assert np.all(info["margin"][bad] < 1.0)
                if bad.size == 0:
                    np.testing.assert_allclose(got, want, rtol=0,
                                               atol=1e-6 * np.abs(want).max())
    bd.close()


INSTALL = r'''
import sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import __graft_entry__
pkg = __graft_entry__.load_package()
lib = pkg._lib.load()
geom = ((20, 18, 16, 30), (7, 7, 7), 2)
vol, _ = pkg.make_lowrank_volume(geom[0], 5, 2.0, 0.04, 4)
before = pkg.denoise(vol, *geom[1:])
table = pkg.autotune.tune_table([geom], reps=2)
assert lib.mppca_tune_table_size() == 0
pkg.autotune.install(table)
assert lib.mppca_tune_table_size() == len(table) == 1
try:
    pkg.autotune.install(table)
    raise SystemExit("a second install must fail")
except pkg._lib.MppcaError:
    pass
after = pkg.get_denoiser(geom[0], *geom[1:])(vol)
assert np.abs(after - before).max() <= 1e-6 * np.abs(before).max()
print("installed", pkg.autotune.to_dict(table[0]))
'''


def test_install_table_once():
    """In a fresh process: tune, install, use; a second install is refused."""
    r = subprocess.run([sys.executable, "-c", INSTALL, str(ROOT)], capture_output=True,
                       text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "installed" in r.stdout
