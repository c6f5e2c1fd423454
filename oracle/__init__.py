"""CPU oracle for the patch-wise MP-PCA denoiser -- TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy, the reference's step-0 algorithm
(paquiteau/lowrank-gpu-fastest: step0_numpy_loop.py:8-39, common.py:37-44,
step4_gram_eigh.py:40-57). It is the checker that the HIP path is compared
against, and the CPU baseline that bench.py times beside it.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it. The product path (the package under
``lowrank-gpu-fastest_amd/``) never imports, calls or falls back to it.

Pinning status of each function (how a transcription mistake would be caught):

* ``patch_corners``     -- golden vectors written by the reference's own
                           ``common.patch_corners`` (tests/golden); closed-form
                           count 12,600 at 96x96x48 / 11^3 / stride 3
                           (README.md:29-30); brute-force voxel coverage.
* ``mppca_patch``       -- golden vectors written by the reference's own
                           ``step0_numpy_loop.mppca_patch`` (float32 and float64
                           matrices); closed forms: constant matrix -> p = 0 and
                           output = input; r dominant spikes -> p = r.
* ``mp_rank``           -- golden ranks of the reference's ``mppca_patch``;
                           flat spectrum -> 0; brute force on tiny spectra.
* ``denoise``           -- golden volumes written by the reference's step-0
                           loop (float32 and float64); constant volume -> itself.
* ``gram_route_patch``  -- equals ``mppca_patch`` to rounding in float64 (the
                           identity U_p S_p V_p^T = Xc V_p V_p^T).

The N < T-1 convention (``mp_rank`` with M = min(N, T-1) and
n = max(N, T-1)) has no reference answer: step 0 raises there. It is pinned
only by brute-force FP64 SVD agreement with the Gram route -- parity
unpinned against the reference for that case.
"""

from .step0 import (  # noqa: F401
    denoise,
    denoise_parallel,
    gram_route_patch,
    mp_rank,
    mppca_patch,
    patch_corners,
)
