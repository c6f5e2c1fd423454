"""ctypes binding of ``libmppca.so`` (declarations: include/mppca.h).

The library is loaded only after ``import torch``: it is linked against the
HIP runtime by SONAME and so binds to the runtime torch has already loaded.
There is no fallback: a missing or stale library raises ``ImportError``.
"""

from __future__ import annotations

import ctypes
import pathlib

PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG / "libmppca.so"
ABI_VERSION = 106
FORM_CODES = {"plain": 0, "fp32": 1, "mixed": 2}

_i32, _i64, _vp, _f = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_float
_cp = ctypes.c_char_p

# name -> (restype, argtypes); every symbol declared in include/mppca.h
SIGNATURES = {
    "mppca_abi_version": (_i32, []),
    "mppca_last_error": (_cp, []),
    "mppca_num_patches": (_i64, [_i32] * 7),
    "mppca_centre": (_i32, [_vp, _vp, _vp, _i64, _i32, _vp]),
    "mppca_gram": (_i32, [_vp, _vp] + [_i32] * 8 + [_i64, _i64, _vp]),
    "mppca_gram64": (_i32, [_vp, _vp] + [_i32] * 8 + [_i64, _i64, _vp]),
    "mppca_eigh_jacobi": (_i32, [_vp, _vp, _vp, _vp, _i64, _i32, _i32, _vp]),
    "mppca_mp_threshold": (_i32, [_vp, _i64, _i32, _i32, _vp, _vp, _vp, _vp]),
    "mppca_eigh_mp": (_i32, [_vp, _i64, _i32, _i32, _i32] + [_vp] * 7),
    "mppca_eig64_mp": (_i32, [_vp, _i64, _i32, _i32] + [_vp] * 7),
    "mppca_recon_scatter": (_i32, [_vp] * 4 + [_i32] * 8 + [_i64, _i64, _vp]),
    "mppca_normalize": (_i32, [_vp] * 4 + [_i32] * 8 + [_vp]),
    "mppca_factor_floats": (_i64, [_i32]),
    "mppca_eig64_factors": (_i32, [_vp, _i64, _i32, _i32] + [_vp] * 7),
    "mppca_mixed_eigenvalues": (_i32, [_vp] + [_i32] * 8 + [_vp, _i64, _vp, _vp]),
    "mppca_recon_factors": (_i32, [_vp] * 6 + [_i32] * 8 + [_vp]),
    "mppca_workspace_bytes": (_i64, [_i32] * 9),
    "mppca_denoise_device": (_i32, [_vp, _vp] + [_i32] * 9 + [_vp, _i64, _vp, _vp, _vp, _vp]),
    "mppca_denoise_host": (_i32, [_vp, _vp] + [_i32] * 9 + [_vp, _i64] + [_vp] * 4
                           + [_i32, _i32] + [_vp] * 7),
    "mppca_batch_workspace_bytes": (_i64, [_i32] * 10),
    "mppca_denoise_device_batch": (_i32, [_vp, _vp] + [_i32] * 10 + [_vp, _vp, _i64] + [_vp] * 4),
    "mppca_graph_create": (_i32, [_vp, _vp, _vp] + [_i32] * 10 + [_vp, _vp, _i64] + [_vp] * 3),
    "mppca_tune_table_set": (_i32, [_vp, _i32]),
    "mppca_tune_table_size": (_i32, []),
    "mppca_graph_launch": (_i32, [_vp, _vp]),
    "mppca_graph_destroy": (_i32, [_vp]),
    "mppca_copy_d2h": (_i32, [_vp, _vp, _i64, _vp]),
    "mppca_host_copy": (_i32, [_vp, _vp, _i64, _i32]),
    "mppca_noalloc_scratch_bytes": (_i64, []),
    "mppca_denoise_device_noalloc": (_i32, [_vp, _vp] + [_i32] * 9 + [_vp] * 4),
}


class Tuning(ctypes.Structure):
    """``mppca_tuning`` (include/mppca.h): the tile schedule of one geometry."""

    _fields_ = [("N", ctypes.c_int32), ("T", ctypes.c_int32), ("gram", ctypes.c_int32),
                ("eig", ctypes.c_int32), ("recon_sx", ctypes.c_int32)]


GRAM_KERNELS = {"auto": 0, "limb": 1, "slide": 2, "patch": 3}
EIG_LAYOUTS = {"auto": 0, "grid": 1, "lds": 2}


class MppcaError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, what: str, code: int, message: str):
        super().__init__(f"{what} failed with status {code}: {message}")
        self.code = code


_LIB = None


def load(path: pathlib.Path | None = None):
    """Load and type the library once per process (fails loudly if absent)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    import torch  # noqa: F401  (the HIP runtime must be torch's)

    p = pathlib.Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise ImportError(
            f"{p} is missing: build it with `python lowrank-gpu-fastest_amd/build.py` "
            "(or __graft_entry__.build()); there is no CPU fallback"
        )
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    got = lib.mppca_abi_version()
    if got != ABI_VERSION:
        raise ImportError(f"{p}: ABI version {got}, expected {ABI_VERSION}; rebuild it")
    if path is None:
        _LIB = lib
    return lib


def check(rc: int, what: str):
    if rc != 0:
        msg = load().mppca_last_error()
        raise MppcaError(what, rc, msg.decode() if msg else "")
    return rc
