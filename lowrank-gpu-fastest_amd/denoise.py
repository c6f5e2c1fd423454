"""Host driver of the HIP denoiser: the reference's ``denoise`` entry point.

``denoise(vol, patch, stride)`` takes a float32 (X, Y, Z, T) volume and
returns the denoised volume, like the reference's ``stepN.denoise(vol_np)``
(step8_tuned.py:116-146, step0_numpy_loop.py:28-39) with the patch shape and
stride as arguments instead of module globals (common.py:27-29).

* NumPy (or CPU tensor) in -> NumPy out: host -> host, the timed form of the
  reference's harness (common.py:129-143).
* CUDA tensor in -> CUDA tensor out: device -> device, no host copies.

All device work is enqueued on torch's current stream through the C ABI of
``libmppca.so``; device buffers are torch tensors held by a per-geometry
:class:`Denoiser` and reused across calls.
"""

from __future__ import annotations

import ctypes
import os
import threading
import warnings
import weakref

import numpy as np

from . import _lib
from .geometry import check_geometry, num_patches

# "plain": FP64-accumulated MFMA Gram + FP64 eigensolver, FP32 reconstruction
#          (matches step 0 to ~1e-7 on any input); the default.
# "fp32":  the reference's all-FP32 arithmetic (FP32 MFMA Gram, FP32 Jacobi).
# "mixed": the precision form: FP16 hi/lo three-product Gram on MFMA (FP32
#          accumulation), FP32 Householder reduction, FP64 MP decision, and an
#          exact FP64 re-solve of every patch whose relative MP margin is below
#          1e-4; rank-p factors and FP32 reconstruction as "plain". Stated
#          against the FP64 oracle (step 0 on float64 data).
FORMS = ("plain", "fp32", "mixed")
FACTOR_FORMS = ("plain", "mixed")  # forms with the rank-p factor route
FACTOR_MAX_T = 128


class SolverStatusError(RuntimeError):
    """Some patch's eigensolve failed: NaN / Inf in the patch (status -2) or
    Jacobi sweeps that did not converge (status -1). ``status`` holds the
    per-patch codes."""

    def __init__(self, msg, status):
        super().__init__(msg)
        self.status = status


class Denoiser:
    """Device-resident state for one (shape, patch, stride): workspace, buffers.

    Not thread-safe: one Denoiser serves one host thread at a time (the
    module-level :func:`denoise` keeps one per thread and geometry).
    """

    def __init__(self, shape, patch=(11, 11, 11), stride=3, device=None, form="plain"):
        import torch

        if form not in FORMS:
            raise ValueError(f"unknown form {form!r}; available: {FORMS}")
        self.torch = torch
        self.shape = tuple(int(s) for s in shape)
        self.patch = tuple(int(p) for p in patch)
        self.stride = int(stride)
        self.form = form
        check_geometry(self.shape, self.patch, self.stride)
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU visible: the HIP denoiser has no CPU fallback")
        self.device = torch.device(device if device is not None else "cuda")
        self.lib = _lib.load()
        self.n_patches = num_patches(self.shape, self.patch, self.stride)
        self.form_code = _lib.FORM_CODES[form]
        ws = int(self.lib.mppca_workspace_bytes(*self._geom_args(), self.form_code))
        if ws < 0:
            _lib.check(ws, "mppca_workspace_bytes")
        dev = self.device
        self._ws = torch.empty(ws + 256, dtype=torch.uint8, device=dev)
        self._ws_ptr = (self._ws.data_ptr() + 255) // 256 * 256
        self._ws_bytes = ws
        self.vol_d = torch.empty(self.shape, dtype=torch.float32, device=dev)
        self.out_d = torch.empty(self.shape, dtype=torch.float32, device=dev)
        # plain and mixed forms: eigensolver -> rank-p factors -> fused recon +
        # normalise; fp32 form: projector -> atomic scatter -> normalise
        self.route = ("factors" if form in FACTOR_FORMS and self.shape[3] <= FACTOR_MAX_T
                      else "projector")
        self.rank = torch.empty(self.n_patches, dtype=torch.int32, device=dev)
        self.margin = torch.empty(self.n_patches, dtype=torch.float32, device=dev)
        self.sweeps = torch.empty(self.n_patches, dtype=torch.int32, device=dev)
        self._pin_in = None
        # pinned result buffers, each with a weak reference to the array last
        # returned on it: a buffer is reused once that array is gone
        self._outs = []
        # host -> host pipeline: x-slabs and host copy threads (a sweep of 5-20
        # slabs on the default volume: 9-11 slabs best, 2.73-2.75 ms against
        # 2.80 at 7 and 2.89 at 16; more slabs are slower because every launch
        # of a slab costs at least one pass of its longest workgroup, however
        # few workgroups it has; 6 threads: more take cores from the threads
        # that drive the GPU). The precision form re-solves the close calls
        # of every slab (a chain of small launches), so fewer slabs suit it
        # at 48 frames: 7 slabs 2.90 ms against 3.06 at 9 (5-9 swept); at 96
        # frames 9 stay best (7.09 ms against 7.25 at 7 and 7.61 at 11)
        self.nslabs = 7 if self.form == "mixed" and self.shape[3] <= 64 else 9
        self.nthreads = max(1, min(6, os.cpu_count() or 1))

    def _geom_args(self):
        x, y, z, t = self.shape
        px, py, pz = self.patch
        return (x, y, z, t, px, py, pz, self.stride)

    # -- device -> device ---------------------------------------------------
    def run_device(self, vol_d, out_d, stream=None):
        """Enqueue the whole path on ``stream`` (default: torch's current)."""
        torch = self.torch
        if vol_d.dtype != torch.float32 or tuple(vol_d.shape) != self.shape:
            raise ValueError(f"expected float32 {self.shape}, got {vol_d.dtype} {tuple(vol_d.shape)}")
        if not vol_d.is_contiguous() or not out_d.is_contiguous():
            raise ValueError("device volumes must be contiguous")
        if vol_d.data_ptr() == out_d.data_ptr():
            raise ValueError("input and output device buffers must be distinct")
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        rc = self.lib.mppca_denoise_device(
            vol_d.data_ptr(), out_d.data_ptr(), *self._geom_args(), self.form_code,
            self._ws_ptr, self._ws_bytes,
            self.rank.data_ptr(), self.margin.data_ptr(), self.sweeps.data_ptr(),
            stream.cuda_stream,
        )
        _lib.check(rc, "mppca_denoise_device")
        return out_d

    # -- device -> device, one C-ABI call per stage ---------------------------
    @property
    def STAGES(self):
        if self.route == "factors":
            return ("centre", "gram", "eigh_mp", "recon")
        return ("centre", "gram", "eigh_mp", "zero", "recon", "normalize")

    def run_staged(self, vol_d, out_d, mark=None, max_sweeps=20):
        """Same kernels as :meth:`run_device`, launched stage by stage (the
        plain and fp32 forms; the mixed form runs as one call).

        ``mark(name)`` (optional) is called before each stage and once more
        with ``None`` after the last, so a caller can record events between
        stages on the current stream. Intermediates stay available in
        ``self.stage_buffers`` for inspection.
        """
        if self.form == "mixed":
            raise NotImplementedError("the mixed form runs as one call (run_device)")
        torch = self.torch
        x, y, z, t = self.shape
        nvox = x * y * z
        P = self.n_patches
        if not hasattr(self, "stage_buffers"):
            dev = self.device
            gdt = torch.float64 if self.form == "plain" else torch.float32
            self.stage_buffers = {
                "xc": torch.empty(self.shape, dtype=torch.float32, device=dev),
                "mean": torch.empty(self.shape[:3], dtype=torch.float32, device=dev),
                "gram": torch.empty((P, t, t), dtype=gdt, device=dev),
                "weight": torch.empty(P, dtype=torch.float32, device=dev),
            }
            if self.route == "factors":
                nf = int(self.lib.mppca_factor_floats(t))
                self.stage_buffers["factors"] = torch.empty((P, nf), dtype=torch.float32,
                                                            device=dev)
            else:
                self.stage_buffers["proj"] = torch.empty((P, t, t), dtype=torch.float32,
                                                         device=dev)
        sb = self.stage_buffers
        st = torch.cuda.current_stream(self.device).cuda_stream
        L, chk, g = self.lib, _lib.check, self._geom_args()
        n_vox = int(np.prod(self.patch))
        mark = mark or (lambda name: None)
        mark("centre")
        chk(L.mppca_centre(vol_d.data_ptr(), sb["xc"].data_ptr(), sb["mean"].data_ptr(),
                           nvox, t, st), "mppca_centre")
        mark("gram")
        if self.form == "plain":
            chk(L.mppca_gram64(sb["xc"].data_ptr(), sb["gram"].data_ptr(), *g, 0, P, st),
                "mppca_gram64")
            mark("eigh_mp")
            if self.route == "factors":
                chk(L.mppca_eig64_factors(sb["gram"].data_ptr(), P, t, n_vox,
                                          sb["factors"].data_ptr(), self.rank.data_ptr(),
                                          sb["weight"].data_ptr(), self.margin.data_ptr(), None,
                                          self.sweeps.data_ptr(), st), "mppca_eig64_factors")
                mark("recon")
                chk(L.mppca_recon_factors(sb["xc"].data_ptr(), sb["factors"].data_ptr(),
                                          self.rank.data_ptr(), sb["weight"].data_ptr(),
                                          sb["mean"].data_ptr(), out_d.data_ptr(), *g, st),
                    "mppca_recon_factors")
                mark(None)
                return out_d
            chk(L.mppca_eig64_mp(sb["gram"].data_ptr(), P, t, n_vox, sb["proj"].data_ptr(),
                                 self.rank.data_ptr(), sb["weight"].data_ptr(),
                                 self.margin.data_ptr(), None, self.sweeps.data_ptr(), st),
                "mppca_eig64_mp")
        else:
            chk(L.mppca_gram(sb["xc"].data_ptr(), sb["gram"].data_ptr(), *g, 0, P, st),
                "mppca_gram")
            mark("eigh_mp")
            chk(L.mppca_eigh_mp(sb["gram"].data_ptr(), P, t, n_vox, max_sweeps,
                                sb["proj"].data_ptr(), self.rank.data_ptr(),
                                sb["weight"].data_ptr(), self.margin.data_ptr(), None,
                                self.sweeps.data_ptr(), st), "mppca_eigh_mp")
        mark("zero")
        out_d.zero_()
        mark("recon")
        chk(L.mppca_recon_scatter(sb["xc"].data_ptr(), sb["proj"].data_ptr(),
                                  sb["weight"].data_ptr(), out_d.data_ptr(), *g, 0, P, st),
            "mppca_recon_scatter")
        mark("normalize")
        chk(L.mppca_normalize(out_d.data_ptr(), sb["mean"].data_ptr(), sb["weight"].data_ptr(),
                              out_d.data_ptr(), *g, st), "mppca_normalize")
        mark(None)
        return out_d

    # -- per-patch status --------------------------------------------------
    def check_status(self, on_bad="raise"):
        """Check the per-patch solver status of the last call (a LAPACK-style
        ``info``: -1 the Jacobi sweeps did not converge, -2 a non-finite patch,
        -3 a mixed-form patch whose MP margin asked for an FP64 re-solve past
        the re-solve list's capacity; 1 marks a patch that was re-solved).

        ``on_bad``: "raise" (:class:`SolverStatusError`), "warn" or "ignore".
        Waits for the device work of the last call. Step 0 raises on such
        input too (``numpy.linalg.svd`` does not converge on NaN or Inf).
        """
        if on_bad == "ignore":
            return
        if on_bad not in ("raise", "warn"):
            raise ValueError(f"on_bad must be 'raise', 'warn' or 'ignore', not {on_bad!r}")
        worst = int(self.sweeps.min().item())
        if worst >= 0:
            return
        st = self.sweeps.cpu().numpy()
        nan, nconv, over = int((st == -2).sum()), int((st == -1).sum()), int((st == -3).sum())
        msg = (f"{nan} patch(es) hold NaN or Inf, {nconv} did not converge, {over} past the "
               f"re-solve capacity (of {self.n_patches}; first at patch "
               f"{int(np.flatnonzero(st < 0)[0])})")
        if on_bad == "raise":
            raise SolverStatusError(msg, st)
        warnings.warn(msg, RuntimeWarning, stacklevel=3)

    # -- host -> host -------------------------------------------------------
    def __call__(self, vol, on_bad="raise") -> np.ndarray:
        torch = self.torch
        a = vol.numpy() if isinstance(vol, torch.Tensor) else np.asarray(vol)
        if a.dtype != np.float32:
            raise TypeError(f"volume must be float32, got {a.dtype}")
        if a.shape != self.shape:
            raise ValueError(f"volume shape {a.shape} != {self.shape}")
        a = np.ascontiguousarray(a)
        if self._pin_in is None:
            self._pin_in = torch.empty(self.shape, dtype=torch.float32, pin_memory=True)
            self._streams = [torch.cuda.Stream(self.device) for _ in range(3)]
            self._ev_prior = torch.cuda.Event()
        # the result lands in a pinned buffer, returned as is (as
        # step8_tuned.py:144-146 returns its pinned tensor); buffers whose
        # arrays the caller has dropped are reused (a new pinned tensor from
        # torch's host allocator cost ~0.1 ms a call)
        slot = next((k for k, (_, ref) in enumerate(self._outs) if ref() is None), None)
        if slot is None:
            self._outs.append([torch.empty(self.shape, dtype=torch.float32, pin_memory=True),
                               lambda: None])
            slot = len(self._outs) - 1
        host = self._outs[slot][0]
        sc, si, so = self._streams
        # earlier work on the buffers (the compute and output streams wait; the
        # input stream's sets wait for the compute stream's first centring)
        self._ev_prior.record(torch.cuda.current_stream(self.device))
        sc.wait_event(self._ev_prior)
        so.wait_event(self._ev_prior)
        worst = ctypes.c_int32(0)  # the lowest per-patch status, read by the call
        rc = self.lib.mppca_denoise_host(
            a.ctypes.data, host.data_ptr(), *self._geom_args(), self.form_code,
            self._ws_ptr, self._ws_bytes, self.vol_d.data_ptr(), self.out_d.data_ptr(),
            self._pin_in.data_ptr(), None, self.nslabs, self.nthreads,
            sc.cuda_stream, si.cuda_stream, so.cuda_stream,
            self.rank.data_ptr(), self.margin.data_ptr(), self.sweeps.data_ptr(),
            ctypes.byref(worst),
        )
        _lib.check(rc, "mppca_denoise_host")
        if worst.value < 0:  # some patch failed: the per-patch report
            self.check_status(on_bad)
        arr = host.numpy()
        self._outs[slot][1] = weakref.ref(arr)
        return arr

    def info(self) -> dict:
        """Per-patch rank, MP margin (units of the error bound) and solver status
        ("sweeps": Jacobi sweeps in the fp32 form, 0 in the plain form; -1 no
        convergence, -2 non-finite input)."""
        self.torch.cuda.current_stream(self.device).synchronize()
        return {
            "rank": self.rank.cpu().numpy().astype(np.int64),
            "margin": self.margin.cpu().numpy(),
            "sweeps": self.sweeps.cpu().numpy(),
        }


_CACHE = threading.local()


def get_denoiser(shape, patch=(11, 11, 11), stride=3, device=None, form="plain") -> Denoiser:
    """The calling thread's cached :class:`Denoiser` for a geometry."""
    if not hasattr(_CACHE, "d"):
        _CACHE.d = {}
    key = (tuple(int(s) for s in shape), tuple(int(p) for p in patch), int(stride), str(device), form)
    d = _CACHE.d.get(key)
    if d is None:
        d = Denoiser(shape, patch, stride, device=device, form=form)
        _CACHE.d[key] = d
    return d


def denoise(vol, patch=(11, 11, 11), stride=3, *, form="plain", return_info=False,
            on_bad="raise"):
    """MP-PCA denoise a float32 (X, Y, Z, T) volume on the GPU.

    NumPy in -> NumPy out (host -> host); a CUDA tensor in -> a new CUDA
    tensor out. ``form`` is one of :data:`FORMS`. ``return_info`` adds a dict
    with the per-patch rank, MP margin and solver status. Raises on invalid
    input; never falls back to a CPU path. A patch holding NaN or Inf gets
    rank 0 and status -2 (the NaN reaches the output, as in any float path);
    the per-patch status is checked after the call and ``on_bad`` ("raise",
    the default, "warn" or "ignore") says what a negative status does.
    """
    try:
        import torch
    except ImportError as e:  # pragma: no cover
        raise ImportError("PyTorch (ROCm) is required by the HIP denoiser") from e
    if isinstance(vol, torch.Tensor) and vol.is_cuda:
        d = get_denoiser(tuple(vol.shape), patch, stride, device=vol.device, form=form)
        # a fresh contiguous output whatever the input's strides
        out = torch.empty(vol.shape, dtype=vol.dtype, device=vol.device)
        d.run_device(vol.contiguous(), out)
        d.check_status(on_bad)
        res = out
    else:
        shape = tuple(vol.shape)
        d = get_denoiser(shape, patch, stride, form=form)
        res = d(vol, on_bad=on_bad)
    if return_info:
        return res, d.info()
    return res
