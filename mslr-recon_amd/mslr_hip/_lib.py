"""ctypes binding of libmslr_hip.so (the C ABI declared in include/mslr.h).

Every entry point returns a status code; `check` turns a non-zero code into the
Python exception that matches the reference's error kind (ArgumentError ->
ValueError, AssertionError -> AssertionError, ...).  The library must be
built (mslr_hip.build) -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmslr_hip.so"

MSLR_OK = 0
MSLR_ERR_ARG = 1
MSLR_ERR_SHAPE = 2
MSLR_ERR_DEVICE = 3
MSLR_ERR_NOMEM = 4
MSLR_ERR_UNSUPPORTED = 5
MSLR_ERR_NUMERIC = 6

FORM_PLAIN = 0
FORM_BATCHED = 1
FORM_MIXED = 2
PRODUCTS_FP32 = 0
PRODUCTS_SPLIT_BF16 = 1
MOM_POGM, MOM_FPGM, MOM_PGM = 0, 1, 2
RESTART_GR, RESTART_NONE = 0, 1


class MSLRError(RuntimeError):
    """Base class of library errors."""


class MSLRDeviceError(MSLRError):
    """No usable GPU / HIP runtime failure (the reference's error("use_gpu=true requires ..."))."""


class MSLRUnsupported(MSLRError):
    """Size outside what the kernels support."""


class MSLRNumericError(MSLRError, ArithmeticError):
    """Non-finite values or eigensolver failure."""


def _exc(code: int, msg: str) -> Exception:
    if code == MSLR_ERR_ARG:
        return ValueError(msg)
    if code == MSLR_ERR_SHAPE:
        return AssertionError(msg)
    if code == MSLR_ERR_DEVICE:
        return MSLRDeviceError(msg)
    if code == MSLR_ERR_NOMEM:
        return MemoryError(msg)
    if code == MSLR_ERR_UNSUPPORTED:
        return MSLRUnsupported(msg)
    if code == MSLR_ERR_NUMERIC:
        return MSLRNumericError(msg)
    return MSLRError(f"status {code}: {msg}")


class mslr_pogm_desc(C.Structure):
    _fields_ = [
        ("N", C.c_int64 * 3),
        ("Nt", C.c_int64),
        ("nscales", C.c_int64),
        ("patch_sizes", C.POINTER(C.c_int64)),
        ("strides", C.POINTER(C.c_int64)),
        ("lambdas", C.POINTER(C.c_double)),
        ("L", C.c_double),
        ("niter", C.c_int),
        ("mom", C.c_int),
        ("restart", C.c_int),
        ("restart_cutoff", C.c_double),
        ("bsig", C.c_double),
        ("form", C.c_int),
        ("log_costs", C.c_int),
        ("checkpoint", C.c_int),
    ]


class mslr_tuning(C.Structure):
    _fields_ = [
        ("gram_chunk", C.c_int),
        ("eig_lanes_limit", C.c_int),
        ("svst_apply", C.c_int),
        ("sense_chunks", C.c_int),
        ("yz_threads_fwd", C.c_int),
        ("yz_threads_adj", C.c_int),
        ("update_lines", C.c_int),
        ("sb2st_schedule", C.c_int),
        ("apply_fp32", C.c_int),
        ("reserved", C.c_int * 3),
    ]


_i64 = C.c_int64
_p = C.c_void_p
_pi64 = C.POINTER(C.c_int64)
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)

# name -> (restype, argtypes)
SIGNATURES: dict[str, tuple] = {
    "mslr_last_error": (C.c_char_p, []),
    "mslr_version": (C.c_int, []),
    "mslr_prof_enable": (C.c_int, [_p, C.c_int]),
    "mslr_prof_report": (C.c_int, [_p, C.c_char_p, C.c_size_t]),
    "mslr_ctx_create": (C.c_int, [C.c_int, C.POINTER(_p)]),
    "mslr_ctx_destroy": (C.c_int, [_p]),
    "mslr_ctx_sync": (C.c_int, [_p]),
    "mslr_ctx_set_products": (C.c_int, [_p, C.c_int]),
    "mslr_ctx_set_tuning": (C.c_int, [_p, C.POINTER(mslr_tuning)]),
    "mslr_ctx_get_tuning": (C.c_int, [_p, C.POINTER(mslr_tuning)]),
    "mslr_ctx_info": (C.c_int, [_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "mslr_malloc": (C.c_int, [_p, C.c_size_t, C.POINTER(_p)]),
    "mslr_free": (C.c_int, [_p, _p]),
    "mslr_upload": (C.c_int, [_p, _p, _p, C.c_size_t]),
    "mslr_download": (C.c_int, [_p, _p, _p, C.c_size_t]),
    "mslr_copy_d2d": (C.c_int, [_p, _p, _p, C.c_size_t]),
    "mslr_zero": (C.c_int, [_p, _p, C.c_size_t]),
    "mslr_sense_create": (C.c_int, [_p, _pi64, _i64, _i64, _p, _p, C.c_int, C.c_int, C.POINTER(_p)]),
    "mslr_sense_destroy": (C.c_int, [_p]),
    "mslr_sense_nsamples": (_i64, [_p]),
    "mslr_sense_forward": (C.c_int, [_p, _p, _p, _i64, _i64]),
    "mslr_sense_forward_noalloc": (C.c_int, [_p, _p, _p, _i64, _i64]),
    "mslr_sense_adjoint": (C.c_int, [_p, _p, _p, _i64, _i64]),
    "mslr_sense_forward_f64": (C.c_int, [_p, _p, _p, _i64, _i64]),
    "mslr_sense_adjoint_f64": (C.c_int, [_p, _p, _p, _i64, _i64]),
    "mslr_sense_grad": (C.c_int, [_p, _p, _i64, _p, _p, _pd]),
    "mslr_sense_residual": (C.c_int, [_p, _p, _i64, _p, _p, _pd]),
    "mslr_sense_poweriter": (C.c_int, [_p, C.c_int, C.c_double, _pd, _pi]),
    "mslr_patch_geometry": (C.c_int, [_pi64, _pi64, _pi64, _pi64, _pi64, _pi64]),
    "mslr_img2patches": (C.c_int, [_p, _p, _pi64, _i64, _pi64, _pi64, _p]),
    "mslr_patches2img": (C.c_int, [_p, _p, _pi64, _i64, _pi64, _pi64, _p]),
    "mslr_svst_batched": (C.c_int, [_p, _p, _i64, _i64, _i64, _pd, C.c_int]),
    "mslr_patch_nucnorm": (C.c_int, [_p, _p, _i64, _i64, _i64, _pd]),
    "mslr_patch_nucnorm_weighted": (C.c_int, [_p, _p, _i64, _i64, _i64, _pd, _pd]),
    "mslr_svst_vjp": (C.c_int, [_p, _p, _p, _p, _i64, _i64, _i64, _pd, C.c_int, _pd]),
    "mslr_tsnr": (C.c_int, [_p, _p, _i64, _i64, _p]),
    "mslr_normalise_smaps": (C.c_int, [_p, _p, C.POINTER(C.c_int64), _i64]),
    "mslr_kspace_flatten": (C.c_int, [_p, _p, C.POINTER(C.c_int64), _i64, _i64, _p, _p, _i64,
                                      C.POINTER(C.c_int64)]),
    "mslr_sense_scale_factor": (C.c_int, [_p, _p, C.c_double, C.POINTER(C.c_double)]),
    "mslr_scale_kspace": (C.c_int, [_p, _p, _i64, C.c_double]),
    "mslr_detrend": (C.c_int, [_p, _p, _i64, _i64]),
    "mslr_patch_svst": (C.c_int, [_p, _p, _p, _pi64, _i64, _pi64, _pi64, C.c_double]),
    "mslr_prox_scales": (C.c_int, [_p, _p, _pi64, _i64, _i64, _pi64, _pi64, _pd]),
    "mslr_pogm_create": (C.c_int, [_p, C.POINTER(mslr_pogm_desc), C.POINTER(_p)]),
    "mslr_pogm_destroy": (C.c_int, [_p]),
    "mslr_pogm_init": (C.c_int, [_p, _p, _p]),
    "mslr_pogm_run": (C.c_int, [_p, C.c_int]),
    "mslr_pogm_x": (_p, [_p]),
    "mslr_pogm_log": (C.c_int, [_p, _pi, _pi]),
    "mslr_pogm_thresholds": (C.c_int, [_p, _pd]),
    "mslr_pogm_costs": (C.c_int, [_p, _pi, _pd, _pd]),
    "mslr_pogm_lambda_vjp": (C.c_int, [_p, _p, _pd]),
}

_lib: C.CDLL | None = None


def load(path: Path | str | None = None) -> C.CDLL:
    """Load the shared library (once) and declare every signature."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise MSLRDeviceError(
            f"{p} is not built; run `python -m mslr_hip.build` (there is no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != MSLR_OK:
        msg = load().mslr_last_error().decode(errors="replace")
        raise _exc(code, msg)


def i64x3(v) -> C.Array:
    a = (C.c_int64 * 3)()
    for i in range(3):
        a[i] = int(v[i])
    return a
