"""CPU tests of the C ABI: the library loads, exports every declared symbol,
and validates arguments on the host (no GPU work is launched here)."""

import re

import pytest

from conftest import ROOT

HEADER = ROOT / "include" / "mppca.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mppca_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_stages():
    names = declared_functions()
    for stage in ("mppca_centre", "mppca_gram", "mppca_gram64", "mppca_eig64_mp",
                  "mppca_eigh_jacobi", "mppca_mp_threshold",
                  "mppca_eigh_mp", "mppca_recon_scatter", "mppca_normalize",
                  "mppca_denoise_device", "mppca_workspace_bytes", "mppca_last_error"):
        assert stage in names


def test_library_exports_every_declared_symbol(lib, pkg):
    names = declared_functions()
    for n in names:
        assert hasattr(lib, n), f"{n} declared in mppca.h but not exported"
    assert set(names) == set(pkg._lib.SIGNATURES), "ctypes signatures out of sync with mppca.h"


def test_abi_version(lib, pkg):
    assert lib.mppca_abi_version() == pkg._lib.ABI_VERSION


def test_num_patches_and_errors(lib):
    assert lib.mppca_num_patches(96, 96, 48, 11, 11, 11, 3) == 12600
    assert lib.mppca_num_patches(32, 32, 24, 5, 5, 5, 2) == 2475
    assert lib.mppca_num_patches(5, 5, 5, 11, 11, 11, 3) < 0
    assert b"larger than volume" in lib.mppca_last_error()
    assert lib.mppca_num_patches(16, 16, 16, 4, 4, 4, 0) < 0


PLAIN, FP32 = 0, 1


def test_workspace_bytes(lib):
    x, y, z, t = 96, 96, 48, 48
    nvox, p = x * y * z, 12600
    ws = lib.mppca_workspace_bytes(x, y, z, t, 11, 11, 11, 3, FP32)
    assert 4 * (nvox * t + nvox + p * t * t + 4 * p) <= ws
    assert ws < 4 * (nvox * t + nvox + p * t * t + 4 * p) + 256 * 16
    # plain, T <= 64: FP64 Gram + rank-p factor blocks (2 * TF^2 floats, TF = 48 here)
    ws = lib.mppca_workspace_bytes(x, y, z, t, 11, 11, 11, 3, PLAIN)
    assert lib.mppca_factor_floats(t) == 2 * 48 * 48
    assert 4 * (nvox * t + nvox + 2 * p * t * t + 2 * p * 48 * 48 + 4 * p) <= ws
    assert ws < 4 * (nvox * t + nvox + 2 * p * t * t + 2 * p * 48 * 48 + 4 * p) + 256 * 16
    # plain, T > 64: FP64 Gram + rank-p factor blocks (TF = 96)
    nv, pp, tt = 16 * 16 * 16, 6 ** 3, 96
    ws = lib.mppca_workspace_bytes(16, 16, 16, tt, 6, 6, 6, 2, PLAIN)
    assert 4 * (nv * tt + nv + 4 * pp * tt * tt + 4 * pp) <= ws
    assert ws < 4 * (nv * tt + nv + 4 * pp * tt * tt + 4 * pp) + 256 * 16
    assert lib.mppca_workspace_bytes(x, y, z, t, 11, 11, 11, 3, 7) < 0  # unknown form
    assert lib.mppca_workspace_bytes(x, y, z, 129, 11, 11, 11, 3, 0) < 0  # T > 128
    assert lib.mppca_workspace_bytes(x, y, z, 1, 11, 11, 11, 3, 0) < 0   # T < 2
    assert lib.mppca_workspace_bytes(40, 40, 40, 48, 17, 17, 17, 3, 0) < 0  # N > 4096


def test_factor_entry_points_validate(lib):
    assert lib.mppca_factor_floats(17) == 2 * 32 * 32
    assert lib.mppca_factor_floats(64) == 2 * 64 * 64
    assert lib.mppca_factor_floats(65) == 2 * 80 * 80
    assert lib.mppca_factor_floats(128) == 2 * 128 * 128
    assert lib.mppca_factor_floats(129) < 0 and lib.mppca_factor_floats(1) < 0
    assert lib.mppca_eig64_factors(None, 4, 48, 100, None, None, None, None, None, None,
                                   None) < 0
    assert lib.mppca_eig64_factors(16, 4, 130, 100, 16, 16, 16, None, None, None, None) == -2
    assert lib.mppca_eig64_factors(16, 4, 48, 100, 18, 16, 16, None, None, None, None) == -1
    assert b"aligned" in lib.mppca_last_error()
    assert lib.mppca_recon_factors(None, None, None, None, None, None, 16, 16, 16, 48,
                                   5, 5, 5, 2, None) < 0
    assert lib.mppca_recon_factors(16, 16, 16, 16, 16, 32, 16, 16, 16, 130,
                                   5, 5, 5, 2, None) == -2
    assert lib.mppca_recon_factors(16, 16, 16, 16, 16, 16, 16, 16, 16, 48,
                                   5, 5, 5, 2, None) == -1  # out aliases xc


def test_denoise_host_validates_before_any_call(lib):
    args = [16, 16, 16, 8, 5, 5, 5, 2]
    # null host / device pointers, bad slab and thread counts, unknown form
    assert lib.mppca_denoise_host(None, 1, *args, 0, 256, 1 << 30, 1, 2, None, None, 4, 2,
                                  None, None, None, None, None, None, None) == -1
    assert lib.mppca_denoise_host(1, 1, *args, 0, 256, 1 << 30, 1, 1, None, None, 4, 2,
                                  None, None, None, None, None, None, None) == -1  # d_vol == d_out
    assert lib.mppca_denoise_host(1, 1, *args, 0, 256, 1 << 30, 1, 2, None, None, 0, 2,
                                  None, None, None, None, None, None, None) == -1
    assert lib.mppca_denoise_host(1, 1, *args, 0, 256, 1 << 30, 1, 2, None, None, 4, 0,
                                  None, None, None, None, None, None, None) == -1
    assert lib.mppca_denoise_host(1, 1, *args, 9, 256, 1 << 30, 1, 2, None, None, 4, 2,
                                  None, None, None, None, None, None, None) == -1
    assert lib.mppca_denoise_host(1, 1, *args, 0, 256, 16, 1, 2, None, None, 4, 2,
                                  None, None, None, None, None, None, None) == -3  # workspace


def test_host_side_validation_launches_nothing(lib):
    # null pointers and bad ranges are rejected before any launch
    assert lib.mppca_gram(None, None, 16, 16, 16, 8, 5, 5, 5, 2, 0, 1, None) < 0
    assert lib.mppca_gram(1, 1, 16, 16, 16, 8, 5, 5, 5, 2, 0, 10**6, None) < 0
    assert b"outside" in lib.mppca_last_error()
    assert lib.mppca_eigh_jacobi(None, None, None, None, 4, 8, 10, None) < 0
    assert lib.mppca_eigh_mp(1, 4, 200, 100, 10, 1, 1, 1, None, None, None, None) < 0
    assert lib.mppca_eig64_mp(None, 4, 8, 100, None, None, None, None, None, None, None) < 0
    assert lib.mppca_gram64(1, 1, 16, 16, 16, 8, 5, 5, 5, 2, -1, 1, None) < 0
    dd = lib.mppca_denoise_device
    assert dd(1, 1, 16, 16, 16, 8, 5, 5, 5, 2, 0, 1, 10, None, None, None, None) < 0
    assert dd(256, 512, 16, 16, 16, 8, 5, 5, 5, 2, 0, 256, 10, None, None, None,
              None) == -3  # workspace too small
    assert dd(256, 256, 16, 16, 16, 8, 5, 5, 5, 2, 0, 256, 10, None, None, None,
              None) == -1  # in place is not allowed
    assert dd(256, 512, 16, 16, 16, 8, 5, 5, 5, 2, 5, 256, 1 << 30, None, None, None,
              None) == -1  # unknown form


def test_check_raises(pkg, lib):
    lib.mppca_num_patches(5, 5, 5, 11, 11, 11, 3)
    with pytest.raises(pkg._lib.MppcaError):
        pkg._lib.check(-1, "probe")


def test_noalloc_limits_on_host(lib):
    """The no-allocation entry point states its scratch size and rejects a
    geometry whose workspace exceeds it before any GPU work (dummy pointers)."""
    assert lib.mppca_noalloc_scratch_bytes() == 64 << 20
    big = lib.mppca_workspace_bytes(96, 96, 48, 48, 11, 11, 11, 3, PLAIN)
    assert big > lib.mppca_noalloc_scratch_bytes()
    rc = lib.mppca_denoise_device_noalloc(16, 32, 96, 96, 48, 48, 11, 11, 11, 3, PLAIN,
                                          None, None, None, None)
    assert rc == -3  # MPPCA_ENOSPACE
    assert b"static scratch" in lib.mppca_last_error()
    assert lib.mppca_denoise_device_noalloc(16, 16, 8, 8, 8, 6, 3, 3, 3, 2, PLAIN,
                                            None, None, None, None) == -1  # vol == out
