"""Build ``libmppca.so`` (gfx950) in-tree with hipcc.

``python lowrank-gpu-fastest_amd/build.py`` or ``__graft_entry__.build()``.
The library is linked against the HIP runtime by SONAME (libamdhip64.so.7);
loaded after ``import torch`` it binds to the runtime torch already loaded,
so the process has one HIP runtime.
"""

from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
# HIP sources, plus host-only C++ (compiled by hipcc as plain C++)
SRCS = sorted((PKG / "csrc").glob("*.hip")) + sorted((PKG / "csrc").glob("*.cpp"))
OUT = PKG / "libmppca.so"
OBJ_DIR = PKG / "_build"  # per-source objects (kept out of git and of GPU runs)
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 VALU forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    # MFMA accumulators in VGPRs (gfx950 allows it): no AGPR<->VGPR copies of
    # loop-carried accumulators around every MFMA group
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build)")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr.strip():
        # hipcc's host pass reports the device-only target feature; keep the rest
        lines = [ln for ln in res.stderr.splitlines() if "packed-fp32-ops" not in ln]
        if any(ln.strip() for ln in lines):
            print("\n".join(lines), file=sys.stderr)


def build(verbose: bool = False, extra_flags=(), jobs: int | None = None) -> pathlib.Path:
    """Compile every source to an object (in parallel), then link the library."""
    tmp = OUT.with_suffix(".so.tmp")
    rocm = pathlib.Path(hipcc()).resolve().parent.parent
    OBJ_DIR.mkdir(exist_ok=True)
    objs = [OBJ_DIR / (s.name + ".o") for s in SRCS]
    jobs = jobs or max(1, min(len(SRCS), 8, os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda so: _run([hipcc(), *HIPCC_FLAGS, *extra_flags, f"-I{ROOT / 'include'}",
                                     "-c", str(so[0]), "-o", str(so[1])], verbose),
                    zip(SRCS, objs)))
    # the host pipeline's DMA copies call the HSA runtime (linked by SONAME,
    # libhsa-runtime64.so.1, so the process keeps the one torch's HIP runtime loaded)
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *[str(o) for o in objs],
          f"-L{rocm / 'lib'}", "-lhsa-runtime64", "-o", str(tmp)], verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
