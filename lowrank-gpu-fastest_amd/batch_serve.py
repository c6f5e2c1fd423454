"""Continuous-batching serving loop: requests of any of several geometries
arrive at any time and share kernel launches on the GPU.

The reference times one call at a time (common.py:129-143) and has no serving
path. :class:`Server` (serve.py) runs one whole host -> host pipeline per
request; this server instead splits serving into three stages that run
concurrently, each on its own host thread, so that the PCIe transfers of some
requests overlap the kernels of others and requests that wait are batched:

* upload: a request is admitted into a free slot of its geometry's *lane*
  (device buffers for up to ``max_batch`` volumes -- fewer for volumes that
  fill the GPU alone -- allocated under the HBM
  budget, least recently used lanes dropped first), staged into pinned memory
  by host threads in pieces and uploaded piece by piece (DMA);
* compute: whenever the GPU has finished the previous batch, every request
  uploaded (or uploading) into a lane's filling buffer is denoised by ONE
  batched pass (``mppca_denoise_device_batch``: each stage one launch over the
  whole batch), replayed from a captured HIP graph when ``graphs``; each lane
  has ``nbuf`` (3) buffers, so that one fills while another computes and a
  third downloads;
* download: each result is copied into pinned host memory on a DMA engine
  (``mppca_copy_d2h``; not a copy kernel on the CUs), its solver status is
  checked and its Future resolved.

Each result equals the volume denoised alone (bit for bit on the factor route,
plain and mixed forms; up to float atomic summation order in the fp32 form): the
batched kernels compute every volume of a batch exactly as alone.
"""

from __future__ import annotations

import collections
import ctypes
import gc
import threading
import time
from concurrent.futures import Future

import numpy as np

from . import _lib
from .batch import BatchDenoiser
from .denoise import SolverStatusError
from .serve import AdmissionError


def lane_bytes(shape, patch, stride, max_batch, form="plain", nbuf=3) -> int:
    """Device memory of one lane: ``nbuf`` buffers of ``max_batch`` volumes."""
    lib = _lib.load()
    x, y, z, t = shape
    ws = int(lib.mppca_batch_workspace_bytes(max_batch, x, y, z, t, *patch, stride,
                                             _lib.FORM_CODES[form]))
    if ws < 0:
        _lib.check(ws, "mppca_batch_workspace_bytes")
    return nbuf * (ws + 2 * max_batch * 4 * x * y * z * t + 256)


class _Buffer:
    """One batch buffer of a lane: device state plus pinned staging."""

    FILLING, COMPUTING, FREE = 0, 1, 2

    def __init__(self, torch, shape, patch, stride, cap, form, graphs):
        self.bd = BatchDenoiser(shape, patch, stride, capacity=cap, form=form)
        self.stage = torch.empty((cap,) + tuple(shape), dtype=torch.float32, pin_memory=True)
        self.state = _Buffer.FREE
        self.reqs = []  # (request, slot, upload event)
        self.graphs = graphs
        self.done_event = None


class _Lane:
    def __init__(self, key, nbytes, bufs):
        self.key = key
        self.nbytes = nbytes
        self.bufs = bufs
        self.filling = None  # the buffer uploads go to (FILLING), or None
        self.inflight = 0    # requests admitted and not yet downloaded


class _Request:
    __slots__ = ("vol", "key", "keep", "fut", "lane", "buf", "slot")

    def __init__(self, vol, key, keep, fut):
        self.vol, self.key, self.keep, self.fut = vol, key, keep, fut
        self.lane = self.buf = self.slot = None


class BatchServer:
    """Continuous-batching host -> host denoising under an HBM budget."""

    def __init__(self, max_batch=8, hbm_budget=None, form="plain", graphs=True,
                 stage_threads=8, piece_mb=16, nbuf=3, batch_patches=12000, freeze_gc=True,
                 on_bad="raise"):
        import torch

        self.torch = torch
        if max_batch < 1:
            raise ValueError("max_batch must be >= 1")
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(0.5 * free)
        self.lib = _lib.load()
        self.form = form
        self.max_batch = int(max_batch)
        self.hbm_budget = int(hbm_budget)
        self.graphs = bool(graphs)
        if nbuf < 2:
            raise ValueError("nbuf must be >= 2")
        self.nbuf = int(nbuf)
        # a lane batches up to max(1, batch_patches // P) volumes (at most
        # max_batch): a volume of the default size fills the GPU alone (batching
        # it only makes requests wait for each other), small ones do not
        self.batch_patches = int(batch_patches)
        self.stage_threads = int(stage_threads)
        self.piece_bytes = int(piece_mb) << 20
        self.on_bad = on_bad
        self._cv = threading.Condition()
        self._lanes: collections.OrderedDict = collections.OrderedDict()
        self._used = 0
        self._q: collections.deque = collections.deque()
        self._dq: collections.deque = collections.deque()  # computed buffers to download
        self._closed = False
        self.admitted = 0
        self.evictions = 0
        self.batches = 0
        self.batched_volumes = 0
        self._s_up = torch.cuda.Stream()
        self._s_cmp = torch.cuda.Stream()
        if freeze_gc:  # see serve.Server: full GC passes stall every thread
            gc.collect()
            gc.freeze()
        self._threads = [threading.Thread(target=f, daemon=True)
                         for f in (self._upload_loop, self._compute_loop, self._download_loop)]
        for th in self._threads:
            th.start()

    # -- public -------------------------------------------------------------------
    def submit(self, vol, patch=(11, 11, 11), stride=3, keep=True) -> Future:
        """Queue one float32 (X, Y, Z, T) volume; the Future yields the denoised
        volume (``t_submit`` / ``t_done`` perf_counter stamps on the Future)."""
        if self._closed:
            raise RuntimeError("server closed")
        a = np.asarray(vol)
        if a.dtype != np.float32 or a.ndim != 4:
            raise ValueError(f"expected a float32 (X, Y, Z, T) volume, got {a.dtype} {a.shape}")
        fut: Future = Future()
        fut.t_submit = time.perf_counter()
        key = (tuple(a.shape), tuple(int(p) for p in patch), int(stride))
        with self._cv:
            self._q.append(_Request(np.ascontiguousarray(a), key, bool(keep), fut))
            self._cv.notify_all()
        return fut

    def close(self):
        """Finish queued requests, stop the threads, release the lanes."""
        if self._closed:
            return
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        for th in self._threads:
            th.join()
        with self._cv:
            for lane in self._lanes.values():
                for b in lane.bufs:
                    b.bd.close()
            self._lanes.clear()
            self._used = 0

    @property
    def used_bytes(self):
        return self._used

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- admission ----------------------------------------------------------------
    def _lane_for(self, key):
        """The lane of a geometry, created under the budget (caller holds _cv)."""
        lane = self._lanes.get(key)
        if lane is not None:
            self._lanes.move_to_end(key)
            return lane
        shape, patch, stride = key
        from .geometry import num_patches

        cap = max(1, min(self.max_batch, self.batch_patches // num_patches(shape, patch, stride)))
        need = lane_bytes(shape, patch, stride, cap, self.form, self.nbuf)
        while cap > 1 and need > self.hbm_budget:
            cap -= 1
            need = lane_bytes(shape, patch, stride, cap, self.form, self.nbuf)
        if need > self.hbm_budget:
            raise AdmissionError(f"request needs {need} B of HBM, budget {self.hbm_budget} B")
        while self._used + need > self.hbm_budget:
            idle = [k for k, ln in self._lanes.items() if ln.inflight == 0]
            if not idle:
                return None  # wait for a lane to drain
            old = self._lanes.pop(idle[0])
            for b in old.bufs:
                b.bd.close()
            self._used -= old.nbytes
            self.evictions += 1
        bufs = [_Buffer(self.torch, shape, patch, stride, cap, self.form, self.graphs)
                for _ in range(self.nbuf)]
        lane = _Lane(key, need, bufs)
        self._lanes[key] = lane
        self._used += need
        self.admitted += 1
        return lane

    def _slot_for(self, req):
        """(lane, buffer, slot) for a request, or None to wait (caller holds _cv)."""
        lane = self._lane_for(req.key)
        if lane is None:
            return None
        buf = lane.filling
        if buf is None:
            free = [b for b in lane.bufs if b.state == _Buffer.FREE]
            if not free:
                return None
            buf = free[0]
            buf.state = _Buffer.FILLING
            buf.reqs = []
            lane.filling = buf
        if len(buf.reqs) >= buf.bd.capacity:
            return None
        return lane, buf, len(buf.reqs)

    # -- stage 1: staging + upload ---------------------------------------------------
    def _upload_loop(self):
        torch = self.torch
        torch.cuda.set_device(0)
        while True:
            with self._cv:
                while not self._q and not self._closed:
                    self._cv.wait()
                if not self._q:
                    return
                req = self._q[0]
                try:
                    got = self._slot_for(req)
                except BaseException as e:  # noqa: BLE001 -- handed to the caller
                    self._q.popleft()
                    req.fut.t_done = time.perf_counter()
                    req.fut.set_exception(e)
                    continue
                if got is None:  # no slot: wait for a buffer or a lane to free up
                    self._cv.wait(timeout=0.05)
                    continue
                self._q.popleft()
                lane, buf, slot = got
                req.lane, req.buf, req.slot = lane, buf, slot
                lane.inflight += 1
                buf.reqs.append([req, slot, None])  # reserve the slot
            # stage and upload in pieces outside the lock
            src = req.vol.reshape(-1).view(np.uint8)
            dst = buf.stage[slot].view(-1)
            dev = buf.bd.vol_d[slot].view(-1)
            nbytes = src.nbytes
            step = max(self.piece_bytes // 4 * 4, 4)
            with torch.cuda.stream(self._s_up):
                for o in range(0, nbytes, step):
                    n = min(step, nbytes - o)
                    rc = self.lib.mppca_host_copy(dst.data_ptr() + o, src.ctypes.data + o, n,
                                                  self.stage_threads)
                    _lib.check(rc, "mppca_host_copy")
                    dev[o // 4:(o + n) // 4].copy_(dst[o // 4:(o + n) // 4], non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(self._s_up)
            with self._cv:
                for r in buf.reqs:
                    if r[0] is req:
                        r[2] = ev
                self._cv.notify_all()

    # -- stage 2: batched compute ----------------------------------------------------
    def _compute_loop(self):
        torch = self.torch
        torch.cuda.set_device(0)
        last = None
        rr = 0
        while True:
            if last is not None:
                last.synchronize()  # the GPU is done: batch what arrived meanwhile
            with self._cv:
                while True:
                    ready = [ln for ln in self._lanes.values()
                             if ln.filling is not None and ln.filling.reqs
                             and all(r[2] is not None for r in ln.filling.reqs)]
                    if ready or (self._closed and not self._q and not any(
                            ln.filling is not None and ln.filling.reqs
                            for ln in self._lanes.values())):
                        break
                    self._cv.wait(timeout=0.05)
                if not ready:
                    self._dq.append(None)
                    self._cv.notify_all()
                    return
                lane = ready[rr % len(ready)]
                rr += 1
                buf = lane.filling
                buf.state = _Buffer.COMPUTING
                lane.filling = None
                reqs = list(buf.reqs)
            n = len(reqs)
            s = self._s_cmp
            for _, _, ev in reqs:
                s.wait_event(ev)
            with torch.cuda.stream(s):
                if buf.graphs:
                    buf.bd.replay(n, stream=s)
                else:
                    buf.bd.run(n, stream=s)
                done = torch.cuda.Event()
                done.record(s)
            buf.done_event = done
            last = done
            self.batches += 1
            self.batched_volumes += n
            with self._cv:
                self._dq.append(buf)
                self._cv.notify_all()

    # -- stage 3: download -------------------------------------------------------------
    def _download_loop(self):
        torch = self.torch
        torch.cuda.set_device(0)
        while True:
            with self._cv:
                while not self._dq:
                    self._cv.wait()
                buf = self._dq.popleft()
            if buf is None:
                return
            bd = buf.bd
            P = bd.n_patches
            ev = ctypes.c_void_p(buf.done_event.cuda_event)
            status = np.empty(len(buf.reqs) * P, dtype=np.int32)
            rc = self.lib.mppca_copy_d2h(status.ctypes.data, bd.sweeps.data_ptr(),
                                         status.nbytes, ev)
            for k, (req, slot, _) in enumerate(buf.reqs):
                try:
                    _lib.check(rc, "mppca_copy_d2h")
                    host = torch.empty(bd.shape, dtype=torch.float32, pin_memory=True)
                    _lib.check(self.lib.mppca_copy_d2h(host.data_ptr(), bd.out_d[slot].data_ptr(),
                                                       host.numel() * 4, ev), "mppca_copy_d2h")
                    st = status[slot * P:(slot + 1) * P]
                    if self.on_bad != "ignore" and st.min() < 0:
                        msg = (f"{int((st == -2).sum())} patch(es) hold NaN or Inf, "
                               f"{int((st == -1).sum())} did not converge, "
                               f"{int((st == -3).sum())} past the re-solve capacity")
                        if self.on_bad == "raise":
                            raise SolverStatusError(msg, st.copy())
                    req.fut.t_done = time.perf_counter()
                    req.fut.set_result(host.numpy() if req.keep else None)
                except BaseException as e:  # noqa: BLE001 -- handed to the caller
                    req.fut.t_done = time.perf_counter()
                    req.fut.set_exception(e)
            with self._cv:
                lane = buf.reqs[0][0].lane if buf.reqs else None
                if lane is not None:
                    lane.inflight -= len(buf.reqs)
                buf.state = _Buffer.FREE
                buf.reqs = []
                self._cv.notify_all()
