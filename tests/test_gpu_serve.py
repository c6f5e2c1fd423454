"""Serving loop: results are the same alone and in a concurrent mix, admission
respects the HBM budget, Poisson driving reports latencies."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _volume(shape, seed, rank=4):
    rng = np.random.default_rng(seed)
    maps = rng.random(shape[:3] + (rank,))
    curves = rng.standard_normal((rank, shape[3]))
    return (maps @ curves + 0.05 * rng.standard_normal(shape)).astype(np.float32)


# (shape, patch, stride): factor route (T <= 64) and one projector route (T > 64)
GEOMS = [((32, 24, 20, 16), (5, 5, 5), 2),
         ((28, 28, 16, 30), (7, 7, 7), 2),
         ((24, 20, 20, 48), (5, 5, 5), 3),
         ((16, 16, 16, 80), (4, 4, 4), 2)]


def _alone(pkg, vol, patch, stride):
    return pkg.Denoiser(vol.shape, patch, stride)(vol)


def test_results_same_alone_and_in_mix(pkg):
    vols = [(_volume(s, 20 + i), p, st) for i, (s, p, st) in enumerate(GEOMS)]
    alone = [_alone(pkg, v, p, st) for v, p, st in vols]
    order = [0, 1, 2, 3, 2, 0, 1, 3, 0, 2, 1, 0]
    with pkg.Server(workers=3) as srv:
        futs = [srv.submit(*vols[k]) for k in order]  # a burst: several in flight
        outs = [f.result(timeout=120) for f in futs]
        assert srv.admitted >= len(GEOMS)
    for k, out in zip(order, outs):
        print("request on geometry", GEOMS[k], flush=True)
        if GEOMS[k][0][3] <= 64:  # factor route: no atomics, bit for bit
            np.testing.assert_array_equal(out, alone[k])
        else:  # projector route: float atomics, same up to summation order
            np.testing.assert_allclose(out, alone[k], rtol=0, atol=1e-6 * np.abs(alone[k]).max())


def test_admission_within_budget(pkg):
    a = (_volume(GEOMS[0][0], 1),) + GEOMS[0][1:]
    b = (_volume(GEOMS[1][0], 2),) + GEOMS[1][1:]
    need = max(pkg.device_bytes(*GEOMS[0]), pkg.device_bytes(*GEOMS[1]))
    want_a, want_b = _alone(pkg, *a), _alone(pkg, *b)
    with pkg.Server(workers=2, hbm_budget=need) as srv:  # room for one buffer set only
        futs = [srv.submit(*x) for x in (a, b, a, b, a)]
        outs = [f.result(timeout=120) for f in futs]
        assert srv.used_bytes <= need
        assert srv.evictions >= 1
    for out, want in zip(outs, (want_a, want_b, want_a, want_b, want_a)):
        np.testing.assert_array_equal(out, want)


def test_request_larger_than_budget_fails_cleanly(pkg):
    vol = _volume(GEOMS[0][0], 3)
    with pkg.Server(workers=1, hbm_budget=1 << 20) as srv:
        with pytest.raises(pkg.AdmissionError):
            srv.submit(vol, *GEOMS[0][1:]).result(timeout=60)
        assert srv.used_bytes == 0


def test_poisson_run_reports_latency(pkg):
    vols = [(_volume(s, 40 + i), p, st) for i, (s, p, st) in enumerate(GEOMS[:3])]
    with pkg.Server(workers=2) as srv:
        res = pkg.poisson_run(srv, vols, rate_hz=200.0, n_requests=24, seed=1, keep=True)
    lat = res["latency_ms"]
    assert 0 < lat["p50"] <= lat["p99"] <= lat["max"]
    assert res["volumes_per_s"] > 0
    for f, k in zip(res["futures"], res["picks"]):
        assert f.result().shape == vols[k][0].shape
