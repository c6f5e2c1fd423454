"""MI355X-native patch-wise MP-PCA denoiser (hand-written HIP kernels for gfx950).

The hot path of paquiteau/lowrank-gpu-fastest -- ``denoise(volume) -> volume``
(step8_tuned.py:116-146) -- re-implemented behind a C-ABI library
(``libmppca.so``, declared in include/mppca.h) driven from Python.

This directory's name is not a Python identifier; it is imported as
``lowrank_gpu_fastest_amd`` through ``__graft_entry__.load_package()``.
"""

from .geometry import axis_corners, num_patches, patch_corners  # noqa: F401
from .phantom import (  # noqa: F401
    WORKLOADS,
    gaussian_smooth,
    make_lowrank_volume,
    make_phantom,
    mixed_volume,
    workload_volume,
    MIXED_GEOMETRIES,
)
from .denoise import FORMS, Denoiser, SolverStatusError, denoise, get_denoiser  # noqa: F401
from .serve import AdmissionError, Server, device_bytes, poisson_run  # noqa: F401
from .batch import BatchDenoiser, denoise_many  # noqa: F401
from .batch_serve import BatchServer, lane_bytes  # noqa: F401
from . import autotune  # noqa: F401

__all__ = [
    "denoise",
    "Denoiser",
    "get_denoiser",
    "SolverStatusError",
    "FORMS",
    "patch_corners",
    "axis_corners",
    "num_patches",
    "make_lowrank_volume",
    "make_phantom",
    "workload_volume",
    "WORKLOADS",
    "gaussian_smooth",
    "mixed_volume",
    "MIXED_GEOMETRIES",
    "Server",
    "poisson_run",
    "device_bytes",
    "AdmissionError",
    "BatchDenoiser",
    "denoise_many",
    "BatchServer",
    "lane_bytes",
]
