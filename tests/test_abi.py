"""The C ABI library loads and exports every entry point include/mslr.h declares (no GPU needed)."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def header_symbols():
    txt = (ROOT / "include" / "mslr.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mslr_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = header_symbols()
    for s in ["mslr_ctx_create", "mslr_sense_create", "mslr_sense_forward", "mslr_sense_adjoint",
              "mslr_img2patches", "mslr_patches2img", "mslr_svst_batched", "mslr_pogm_run",
              "mslr_sense_poweriter", "mslr_last_error"]:
        assert s in syms


def test_library_exports_every_symbol():
    from mslr_hip import _lib
    lib = _lib.load()
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding declares a signature for every header symbol
    assert set(header_symbols()) <= set(_lib.SIGNATURES)


def test_version_and_error_string():
    from mslr_hip import _lib
    lib = _lib.load()
    assert lib.mslr_version() >= 100
    assert isinstance(lib.mslr_last_error(), bytes)


def test_patch_geometry_is_host_only():
    """mslr_patch_geometry needs no device; matches src/recon.jl:58-63."""
    from mslr_hip import recon
    assert recon.patch_geometry((90, 90, 60), (10, 10, 10), (10, 10, 10)) == ((10, 10, 10), (8, 8, 5), 486)
    assert recon.patch_geometry((90, 90, 60), (6, 6, 6), (3, 3, 3))[2] == 15979
    assert recon.patch_geometry((10, 9, 7), (20, 4, 3), (5, 6, 4)) == ((10, 4, 3), (0, 1, 1), 4)


def test_invalid_geometry_is_argument_error():
    from mslr_hip import recon
    with pytest.raises(ValueError):
        recon.patch_geometry((10, 10, 10), (0, 4, 4), (1, 1, 1))


def test_null_context_rejected():
    from mslr_hip import _lib
    lib = _lib.load()
    code = lib.mslr_ctx_sync(None)
    assert code == _lib.MSLR_ERR_ARG
    assert b"null context" in lib.mslr_last_error()


def test_no_device_fails_loudly():
    """Without a GPU the library reports a device error instead of computing on the CPU."""
    import os
    from mslr_hip import _lib
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.mslr_ctx_create(0, C.byref(h)) == _lib.MSLR_ERR_DEVICE
