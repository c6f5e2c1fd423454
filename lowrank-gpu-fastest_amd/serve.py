"""Serving loop: volumes arrive as requests, are admitted against an HBM budget
and denoised concurrently; latency is measured from arrival to result.

The reference only times one call at a time (common.py:129-143 reports the
best of N); this module is the request-driven form around the same
``denoise`` entry point.

* ``Server(workers, hbm_budget)`` runs ``workers`` host threads. A request
  checks out a per-geometry set of device buffers and three HIP streams (a
  ``Denoiser``) from a shared pool and runs the native host->host pipeline
  (``mppca_denoise_host``: staging copies, uploads, kernels and downloads
  overlapped slab by slab). The ctypes call releases the GIL, so requests
  held by different workers overlap on the device: one request's transfers
  run under another's kernels.
* Admission: a request starts only when its buffers (workspace and two
  volumes) are idle in the pool or fit in the HBM budget beside every buffer
  set in the pool; idle sets of other geometries are dropped first (least
  recently used), and otherwise the request waits for a set in use to return.
* ``poisson_run`` drives a Server with Poisson arrivals at a fixed rate and
  returns per-request latencies (arrival to result), p50/p99 and volumes/s.
* ``freeze_gc`` (default on): Python's cyclic garbage collector stops every
  thread while it scans all tracked objects, and with torch loaded a full
  pass takes tens of milliseconds: at 100 requests/s it set the p99 latency
  (40 ms against a 5 ms median). The Server collects once and freezes the
  objects that exist when it starts (``gc.freeze``), so later passes scan only
  what the requests create. Frozen objects are still freed by reference
  counting; only cycles among them are never collected.

Every request's output is the same as denoising it alone: the kernels are
deterministic per volume (bit for bit on the factor route; up to float atomic
summation order on the projector route, T > 64 or the fp32 form).
"""

from __future__ import annotations

import collections
import gc
import queue
import threading
import time
from concurrent.futures import Future

import numpy as np

from .denoise import Denoiser


class AdmissionError(RuntimeError):
    """A request whose buffers alone exceed the HBM budget."""


def device_bytes(shape, patch, stride, form="plain") -> int:
    """Device memory one request holds: workspace plus input and output volumes."""
    from . import _lib

    lib = _lib.load()
    x, y, z, t = shape
    ws = int(lib.mppca_workspace_bytes(x, y, z, t, *patch, stride, _lib.FORM_CODES[form]))
    if ws < 0:
        _lib.check(ws, "mppca_workspace_bytes")
    return ws + 2 * 4 * x * y * z * t + 256


class Server:
    """Concurrent host->host denoising of submitted volumes under an HBM budget."""

    def __init__(self, workers: int = 2, hbm_budget: int | None = None, form: str = "plain",
                 nslabs: int = 4, nthreads: int = 4, freeze_gc: bool = True):
        import torch

        if workers < 1:
            raise ValueError("workers must be >= 1")
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(0.5 * free)
        self.form = form
        self.nslabs = nslabs
        self.nthreads = nthreads
        self.hbm_budget = int(hbm_budget)
        self._used = 0
        self._cv = threading.Condition()
        self._q: queue.Queue = queue.Queue()
        self._idle: collections.OrderedDict = collections.OrderedDict()  # key -> [(d, bytes)]
        self._closed = False
        self.admitted = 0
        self.evictions = 0
        if freeze_gc:
            gc.collect()
            gc.freeze()
        self._threads = [threading.Thread(target=self._work, args=(i,), daemon=True)
                         for i in range(workers)]
        for th in self._threads:
            th.start()

    # -- public ------------------------------------------------------------------
    def submit(self, vol: np.ndarray, patch=(11, 11, 11), stride: int = 3,
               keep: bool = True) -> Future:
        """Queue one float32 (X, Y, Z, T) volume; the Future yields the denoised
        volume (and carries ``t_submit`` / ``t_done`` perf_counter stamps).
        ``keep=False`` (load generators) resolves the Future to None once the
        result is in host memory, releasing its pinned buffer at once."""
        if self._closed:
            raise RuntimeError("server closed")
        fut: Future = Future()
        fut.t_submit = time.perf_counter()
        self._q.put((vol, tuple(int(p) for p in patch), int(stride), bool(keep), fut))
        return fut

    def close(self):
        """Finish queued requests, stop the workers and release their buffers."""
        if self._closed:
            return
        self._closed = True
        for _ in self._threads:
            self._q.put(None)
        for th in self._threads:
            th.join()
        with self._cv:
            self._idle.clear()
            self._used = 0

    @property
    def used_bytes(self) -> int:
        return self._used

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- workers -----------------------------------------------------------------
    def _checkout(self, shape, patch, stride):
        key = (shape, patch, stride)
        need = device_bytes(shape, patch, stride, self.form)
        if need > self.hbm_budget:
            raise AdmissionError(f"request needs {need} B of HBM, budget {self.hbm_budget} B")
        with self._cv:
            while True:
                sets = self._idle.get(key)
                if sets:
                    d, nbytes = sets.pop()
                    if not sets:
                        del self._idle[key]
                    return key, d, nbytes
                if self._used + need <= self.hbm_budget:
                    self._used += need
                    self.admitted += 1
                    break
                if self._idle:  # drop the least recently used idle set
                    k = next(iter(self._idle))
                    _, nb = self._idle[k].pop(0)
                    if not self._idle[k]:
                        del self._idle[k]
                    self._used -= nb
                    self.evictions += 1
                    continue
                self._cv.wait()  # every set is in use: wait for one to return
        try:
            d = Denoiser(shape, patch, stride, form=self.form)
        except BaseException:
            with self._cv:
                self._used -= need
                self._cv.notify_all()
            raise
        d.nslabs, d.nthreads = self.nslabs, self.nthreads
        return key, d, need

    def _checkin(self, key, d, nbytes):
        with self._cv:
            self._idle.setdefault(key, []).append((d, nbytes))
            self._idle.move_to_end(key)
            self._cv.notify_all()

    def _work(self, i: int):
        import torch

        torch.cuda.set_device(0)
        while True:
            item = self._q.get()
            if item is None:
                return
            vol, patch, stride, keep, fut = item
            held = None
            try:
                vol = np.ascontiguousarray(vol, dtype=np.float32)
                held = self._checkout(tuple(vol.shape), patch, stride)
                out = held[1](vol)
                fut.t_done = time.perf_counter()
                fut.set_result(out if keep else None)
                del out
            except BaseException as e:  # noqa: BLE001 -- handed to the caller
                fut.t_done = time.perf_counter()
                fut.set_exception(e)
            finally:
                if held is not None:
                    self._checkin(*held)


def poisson_run(server: Server, volumes, rate_hz: float, n_requests: int, seed: int = 0,
                timeout_s: float = 120.0, keep: bool = False) -> dict:
    """Submit ``n_requests`` at Poisson arrival times of rate ``rate_hz``, each a
    volume drawn uniformly from ``volumes`` (a list of ``(vol, patch, stride)``);
    latency = time the result is in host memory - arrival time. Results are
    dropped on arrival unless ``keep``."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_hz, n_requests)
    picks = rng.integers(0, len(volumes), n_requests)
    futs = []
    t0 = time.perf_counter()
    t_next = t0
    for k in range(n_requests):
        t_next += gaps[k]
        while True:
            now = time.perf_counter()
            if now >= t_next:
                break
            time.sleep(min(t_next - now, 0.002))
        vol, patch, stride = volumes[picks[k]]
        f = server.submit(vol, patch, stride, keep=keep)
        f.t_submit = t_next  # the arrival time, not the (later) submission instant
        futs.append(f)
    for f in futs:
        f.result(timeout=timeout_s)
    t_end = max(f.t_done for f in futs)
    lat = np.array([(f.t_done - f.t_submit) * 1e3 for f in futs])
    return {
        "requests": n_requests,
        "rate_hz": rate_hz,
        "volumes_per_s": n_requests / (t_end - t0),
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "mean": float(lat.mean()), "max": float(lat.max())},
        "picks": picks.tolist(),
        "futures": futs,
    }
