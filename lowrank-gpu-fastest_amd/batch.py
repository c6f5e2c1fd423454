"""Batched form: several volumes denoised by the same kernel launches.

The reference denoises one volume per call (step8_tuned.py:116-146) and
batches only patches inside it (BATCH = 2048, step8_tuned.py:21). Here a
batch is ``n`` volumes of one geometry held back to back in HBM: every stage
of the path is one launch whose grid spans all ``n`` volumes
(``mppca_denoise_device_batch``), so small volumes fill the GPU together and
share the launch overheads. Volumes of different geometries form one batch
per geometry (:func:`denoise_many`).

* :class:`BatchDenoiser` holds the device buffers of up to ``capacity``
  volumes of one geometry and runs a batch eagerly or as a captured HIP graph
  (``mppca_graph_create``: one replay call per batch, for fixed buffers).
* Tile schedules come from the autotune table when one is installed
  (:mod:`.autotune`), or from an explicit :class:`~._lib.Tuning`.

Each volume's result equals ``denoise`` on it alone: bit for bit on the
factor route (plain and mixed forms), to float atomic summation order on
the projector route (fp32 form).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .denoise import FACTOR_FORMS, FORMS, FACTOR_MAX_T, SolverStatusError
from .geometry import check_geometry, num_patches


class BatchDenoiser:
    """Device state for batches of up to ``capacity`` volumes of one geometry."""

    def __init__(self, shape, patch=(11, 11, 11), stride=3, capacity=4, form="plain",
                 device=None, tuning=None):
        import torch

        if form not in FORMS:
            raise ValueError(f"unknown form {form!r}; available: {FORMS}")
        if capacity < 1:
            raise ValueError("capacity must be >= 1")
        self.torch = torch
        self.shape = tuple(int(s) for s in shape)
        self.patch = tuple(int(p) for p in patch)
        self.stride = int(stride)
        self.form = form
        self.capacity = int(capacity)
        check_geometry(self.shape, self.patch, self.stride)
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU visible: the HIP denoiser has no CPU fallback")
        self.device = torch.device(device if device is not None else "cuda")
        self.lib = _lib.load()
        self.form_code = _lib.FORM_CODES[form]
        self.n_patches = num_patches(self.shape, self.patch, self.stride)
        self.tuning = tuning
        ws = int(self.lib.mppca_batch_workspace_bytes(self.capacity, *self._geom_args(),
                                                      self.form_code))
        if ws < 0:
            _lib.check(ws, "mppca_batch_workspace_bytes")
        dev = self.device
        self._ws = torch.empty(ws + 256, dtype=torch.uint8, device=dev)
        self._ws_ptr = (self._ws.data_ptr() + 255) // 256 * 256
        self._ws_bytes = ws
        self.vol_d = torch.empty((self.capacity,) + self.shape, dtype=torch.float32, device=dev)
        self.out_d = torch.empty((self.capacity,) + self.shape, dtype=torch.float32, device=dev)
        P = self.capacity * self.n_patches
        self.rank = torch.empty(P, dtype=torch.int32, device=dev)
        self.margin = torch.empty(P, dtype=torch.float32, device=dev)
        self.sweeps = torch.empty(P, dtype=torch.int32, device=dev)
        self.route = ("factors" if form in FACTOR_FORMS and self.shape[3] <= FACTOR_MAX_T
                      else "projector")
        self._graphs = {}

    def _geom_args(self):
        x, y, z, t = self.shape
        px, py, pz = self.patch
        return (x, y, z, t, px, py, pz, self.stride)

    def _tuning_ptr(self):
        return None if self.tuning is None else ctypes.byref(self.tuning)

    def run(self, n=None, stream=None):
        """Enqueue the whole path over volumes ``vol_d[:n]`` into ``out_d[:n]``."""
        n = self.capacity if n is None else int(n)
        if not 1 <= n <= self.capacity:
            raise ValueError(f"batch of {n} volumes outside [1, {self.capacity}]")
        torch = self.torch
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        rc = self.lib.mppca_denoise_device_batch(
            self.vol_d.data_ptr(), self.out_d.data_ptr(), n, *self._geom_args(), self.form_code,
            self._tuning_ptr(), self._ws_ptr, self._ws_bytes, self.rank.data_ptr(),
            self.margin.data_ptr(), self.sweeps.data_ptr(), stream.cuda_stream)
        _lib.check(rc, "mppca_denoise_device_batch")

    # -- HIP graphs ---------------------------------------------------------
    def graph(self, n=None):
        """The executable graph of a batch of ``n`` volumes (captured once)."""
        n = self.capacity if n is None else int(n)
        if not 1 <= n <= self.capacity:
            raise ValueError(f"batch of {n} volumes outside [1, {self.capacity}]")
        g = self._graphs.get(n)
        if g is None:
            # one eager run first: the launchers' one-time set-up (LDS limits,
            # occupancy queries) happens outside the capture
            self.run(n)
            self.torch.cuda.current_stream(self.device).synchronize()
            h = ctypes.c_void_p()
            rc = self.lib.mppca_graph_create(
                ctypes.byref(h), self.vol_d.data_ptr(), self.out_d.data_ptr(), n,
                *self._geom_args(), self.form_code, self._tuning_ptr(), self._ws_ptr,
                self._ws_bytes, self.rank.data_ptr(), self.margin.data_ptr(),
                self.sweeps.data_ptr())
            _lib.check(rc, "mppca_graph_create")
            g = self._graphs[n] = h.value
        return g

    def replay(self, n=None, stream=None):
        """Replay the captured graph of a batch of ``n`` volumes on ``stream``."""
        g = self.graph(n)
        if stream is None:
            stream = self.torch.cuda.current_stream(self.device)
        _lib.check(self.lib.mppca_graph_launch(g, stream.cuda_stream), "mppca_graph_launch")

    def close(self):
        for g in self._graphs.values():
            self.lib.mppca_graph_destroy(g)
        self._graphs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 -- interpreter shutdown
            pass

    # -- host -> host ---------------------------------------------------------
    def __call__(self, vols, graph=False, on_bad="raise"):
        """Denoise a list of up to ``capacity`` float32 volumes of this
        geometry as one batch: upload, one batched pass, download."""
        torch = self.torch
        n = len(vols)
        if not 1 <= n <= self.capacity:
            raise ValueError(f"batch of {n} volumes outside [1, {self.capacity}]")
        for k, v in enumerate(vols):
            a = np.asarray(v)
            if a.dtype != np.float32 or a.shape != self.shape:
                raise ValueError(f"volume {k}: expected float32 {self.shape}, got "
                                 f"{a.dtype} {a.shape}")
            self.vol_d[k].copy_(torch.from_numpy(np.ascontiguousarray(a)))
        if graph:
            self.replay(n)
        else:
            self.run(n)
        self.check_status(n, on_bad)
        return [self.out_d[k].cpu().numpy() for k in range(n)]

    def check_status(self, n, on_bad="raise"):
        """Per-patch solver status of the last batch of ``n`` volumes, as
        :meth:`Denoiser.check_status`."""
        if on_bad == "ignore":
            return
        P = self.n_patches
        st = self.sweeps[: n * P]
        if int(st.min().item()) >= 0:
            return
        st = st.cpu().numpy()
        msg = (f"{int((st == -2).sum())} patch(es) hold NaN or Inf, "
               f"{int((st == -1).sum())} did not converge, {int((st == -3).sum())} past the "
               f"re-solve capacity, in a batch of {n} volumes")
        if on_bad == "raise":
            raise SolverStatusError(msg, st)
        import warnings

        warnings.warn(msg, RuntimeWarning, stacklevel=3)

    def info(self, k):
        """Per-patch rank / margin / status of volume ``k`` of the last batch."""
        self.torch.cuda.current_stream(self.device).synchronize()
        P = self.n_patches
        sl = slice(k * P, (k + 1) * P)
        return {"rank": self.rank[sl].cpu().numpy().astype(np.int64),
                "margin": self.margin[sl].cpu().numpy(),
                "sweeps": self.sweeps[sl].cpu().numpy()}


def denoise_many(vols, patches, strides, form="plain", graph=False, return_info=False):
    """Denoise volumes of possibly different geometries: one batch per
    geometry (shape, patch, stride), each batch one pass of batched launches.
    Returns the results in input order."""
    if isinstance(patches, tuple) and len(patches) == 3 and all(isinstance(p, int)
                                                                 for p in patches):
        patches = [patches] * len(vols)
    if isinstance(strides, int):
        strides = [strides] * len(vols)
    groups = {}
    for k, (v, p, s) in enumerate(zip(vols, patches, strides)):
        key = (tuple(np.shape(v)), tuple(int(x) for x in p), int(s))
        groups.setdefault(key, []).append(k)
    out = [None] * len(vols)
    infos = [None] * len(vols)
    for (shape, patch, stride), idx in groups.items():
        bd = BatchDenoiser(shape, patch, stride, capacity=len(idx), form=form)
        try:
            res = bd([vols[k] for k in idx], graph=graph)
            for j, k in enumerate(idx):
                out[k] = res[j]
                if return_info:
                    infos[k] = bd.info(j)
        finally:
            bd.close()
    return (out, infos) if return_info else out
