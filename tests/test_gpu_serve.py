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


# (shape, patch, stride): T up to 64 and one above (all on the factor route)
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
        np.testing.assert_array_equal(out, alone[k])  # factor route: no atomics


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


# ---- continuous batching (batch_serve.BatchServer) ---------------------------------

def test_batch_server_same_alone_and_in_mix(pkg):
    """A burst over four geometries (batches form as requests wait): every
    result equals the volume denoised alone."""
    vols = [(_volume(s, 60 + i), p, st) for i, (s, p, st) in enumerate(GEOMS)]
    alone = [_alone(pkg, v, p, st) for v, p, st in vols]
    order = [0, 1, 2, 3, 2, 0, 1, 3, 0, 2, 1, 0, 0, 0, 2, 2]
    with pkg.BatchServer(max_batch=3, batch_patches=10 ** 9) as srv:
        futs = [srv.submit(*vols[k]) for k in order]
        outs = [f.result(timeout=120) for f in futs]
        assert srv.batched_volumes == len(order)
        print(f"{srv.batches} batches for {len(order)} requests", flush=True)
    for k, out in zip(order, outs):
        np.testing.assert_array_equal(out, alone[k])


def test_batch_server_mixed_shapes_small(pkg):
    """The six mixed geometries (their patch, stride and T at reduced X, Y,
    Z), per-request seeds, rank and noise, in one stream: alone == in the mix."""
    geoms = [((16, 16, 14, 30), (7, 7, 7), 2), ((24, 23, 19, 48), (11, 11, 11), 3),
             ((15, 15, 15, 100), (9, 9, 9), 3), ((14, 14, 12, 64), (7, 7, 7), 2),
             ((23, 23, 19, 24), (11, 11, 11), 4), ((12, 12, 12, 128), (5, 5, 5), 2)]
    rng = np.random.default_rng(5)
    reqs = []
    for j in range(14):
        k = int(rng.integers(0, len(geoms)))
        shape, patch, stride = geoms[k]
        seed = 100 + j
        r = np.random.default_rng(seed)
        vol, _ = pkg.make_lowrank_volume(shape, int(r.integers(4, 9)), 2.0,
                                         float(r.uniform(0.02, 0.08)), seed)
        reqs.append((vol, patch, stride))
    alone = [_alone(pkg, *r) for r in reqs]
    with pkg.BatchServer(max_batch=2, graphs=True, batch_patches=10 ** 9) as srv:
        futs = [srv.submit(*r) for r in reqs]
        outs = [f.result(timeout=120) for f in futs]
    for (vol, patch, stride), out, want in zip(reqs, outs, alone):
        np.testing.assert_array_equal(out, want)


def test_batch_server_budget_and_eviction(pkg):
    a = (_volume(GEOMS[0][0], 1),) + GEOMS[0][1:]
    b = (_volume(GEOMS[1][0], 2),) + GEOMS[1][1:]
    need = max(pkg.lane_bytes(*GEOMS[0], 1, nbuf=2), pkg.lane_bytes(*GEOMS[1], 1, nbuf=2))
    want_a, want_b = _alone(pkg, *a), _alone(pkg, *b)
    with pkg.BatchServer(max_batch=4, hbm_budget=need, nbuf=2) as srv:  # one lane, 1 volume
        outs = [srv.submit(*x).result(timeout=120) for x in (a, b, a, b)]
        assert srv.used_bytes <= need and srv.evictions >= 1
    for out, want in zip(outs, (want_a, want_b, want_a, want_b)):
        np.testing.assert_array_equal(out, want)
    with pkg.BatchServer(hbm_budget=1 << 20) as srv:
        with pytest.raises(pkg.AdmissionError):
            srv.submit(*a).result(timeout=60)


def test_batch_server_status_and_poisson(pkg):
    bad = _volume(GEOMS[0][0], 7)
    bad[3, 3, 3, 2] = np.nan
    good = (_volume(GEOMS[2][0], 8),) + GEOMS[2][1:]
    with pkg.BatchServer(max_batch=2) as srv:
        with pytest.raises(pkg.SolverStatusError):
            srv.submit(bad, *GEOMS[0][1:]).result(timeout=60)
        res = pkg.poisson_run(srv, [good], rate_hz=300.0, n_requests=20, seed=2, keep=True)
    assert 0 < res["latency_ms"]["p50"] <= res["latency_ms"]["p99"]
    want = _alone(pkg, *good)
    for f in res["futures"]:
        np.testing.assert_array_equal(f.result(), want)


def test_batch_server_precision_form(pkg):
    """The precision form through the continuous-batching server: every
    result equals the volume denoised alone in that form (the scale is per
    volume and every patch, re-solved or not, is solved on its own)."""
    vols = [(_volume(s, 80 + i), p, st) for i, (s, p, st) in enumerate(GEOMS)]
    alone = [pkg.denoise(v, p, st, form="mixed") for v, p, st in vols]
    order = [0, 1, 2, 3, 1, 0, 2, 0, 0, 3]
    with pkg.BatchServer(max_batch=3, form="mixed", batch_patches=10 ** 9) as srv:
        futs = [srv.submit(*vols[k]) for k in order]
        outs = [f.result(timeout=120) for f in futs]
    for k, out in zip(order, outs):
        np.testing.assert_array_equal(out, alone[k])
