"""Device context and column-major device arrays over the C ABI.

Host arrays follow the reference's Julia layout: numpy arrays with the Julia
shape (e.g. (Nx, Ny, Nz, Nt)) stored in Fortran (column-major) order.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _lib

_DTYPES = {np.dtype(np.complex64), np.dtype(np.complex128), np.dtype(np.float32),
           np.dtype(np.float64), np.dtype(np.uint8), np.dtype(np.int32)}


class Context:
    """One HIP device + stream (mslr_ctx)."""

    def __init__(self, device: int = 0):
        lib = _lib.load()
        h = C.c_void_p()
        _lib.check(lib.mslr_ctx_create(int(device), C.byref(h)))
        self.handle = h
        self.device = device
        self._lib = lib

    def sync(self) -> None:
        _lib.check(self._lib.mslr_ctx_sync(self.handle))

    def info(self) -> dict:
        name = C.create_string_buffer(256)
        tot = C.c_size_t()
        fr = C.c_size_t()
        _lib.check(self._lib.mslr_ctx_info(self.handle, name, 256, C.byref(tot), C.byref(fr)))
        return {"name": name.value.decode(), "total_bytes": tot.value, "free_bytes": fr.value}

    def set_products(self, products: str) -> None:
        """SVST Gram / reconstruction products: "fp32" (FP64 Gram products of the
        FP32 data, the default) or "split_bf16" (hi + lo BF16 on the matrix cores,
        the mixed-precision form)."""
        code = {"fp32": _lib.PRODUCTS_FP32, "split_bf16": _lib.PRODUCTS_SPLIT_BF16}[products]
        _lib.check(self._lib.mslr_ctx_set_products(self.handle, code))

    TUNING_FIELDS = ("gram_chunk", "eig_lanes_limit", "svst_apply", "sense_chunks", "yz_threads_fwd",
                     "yz_threads_adj", "update_lines", "sb2st_schedule", "apply_fp32")

    def set_tuning(self, **kw) -> None:
        """Launch shapes (mslr_ctx_set_tuning; an omitted or 0 field = built-in):
        gram_chunk (contraction rows per Gram workgroup, a multiple of 256),
        eig_lanes_limit (patches x m limit for 16 lanes per eigenvector, -1: always 4),
        svst_apply (1: dense apply, 2: low-rank products never fused), sense_chunks (frame groups of the gradient, at most
        64), yz_threads_fwd / yz_threads_adj (512, 768 or 1024 threads of the compiled
        y-z FFT passes of the 90 x 60 plane), update_lines (yz-lines per workgroup of the fused transpose + update, a
        multiple of 8 in [8, 64]), sb2st_schedule (0-3: the
        global-band tridiagonalisation's wave schedule), apply_fp32 (split-BF16 form:
        1 + a mask of the SVST reconstruction products run in FP32 instead, bit 0 the
        dense X W, bit 1 X U_r, bit 2 T f U_r^H)."""
        bad = set(kw) - set(self.TUNING_FIELDS)
        if bad:
            raise TypeError(f"unknown tuning fields {sorted(bad)}")
        t = _lib.mslr_tuning(*(int(kw.get(f, 0)) for f in self.TUNING_FIELDS))
        _lib.check(self._lib.mslr_ctx_set_tuning(self.handle, C.byref(t)))

    def tuning(self) -> dict:
        t = _lib.mslr_tuning()
        _lib.check(self._lib.mslr_ctx_get_tuning(self.handle, C.byref(t)))
        return {f: int(getattr(t, f)) for f in self.TUNING_FIELDS}

    def prof_enable(self, on: bool = True) -> None:
        """Time every kernel launch with HIP events (clears earlier records)."""
        _lib.check(self._lib.mslr_prof_enable(self.handle, int(bool(on))))

    def prof_report(self) -> dict:
        """{kernel: (launches, total_ms)} of the launches recorded since prof_enable."""
        buf = C.create_string_buffer(1 << 16)
        _lib.check(self._lib.mslr_prof_report(self.handle, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, n, ms = line.split()
            out[name] = (int(n), float(ms))
        return out

    def close(self) -> None:
        if getattr(self, "handle", None):
            self._lib.mslr_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


_default: Context | None = None
_default_lock = threading.Lock()


def default_context() -> Context:
    global _default
    with _default_lock:
        if _default is None:
            _default = Context(0)
        return _default


class DeviceArray:
    """A column-major device buffer with a Julia-style shape."""

    def __init__(self, shape, dtype, ctx: Context | None = None, zero: bool = False):
        self.ctx = ctx or default_context()
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        if self.dtype not in _DTYPES:
            raise TypeError(f"unsupported dtype {self.dtype}")
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        _lib.check(self.ctx._lib.mslr_malloc(self.ctx.handle, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p
        self._owner = True
        if zero:
            self.zero()

    @classmethod
    def view(cls, ptr: int, shape, dtype, ctx: Context | None = None, keepalive=None):
        """Non-owning view of device memory (e.g. the POGM state)."""
        obj = cls.__new__(cls)
        obj.ctx = ctx or default_context()
        obj.shape = tuple(int(s) for s in shape)
        obj.dtype = np.dtype(dtype)
        obj.nbytes = int(np.prod(obj.shape, dtype=np.int64)) * obj.dtype.itemsize
        obj.ptr = C.c_void_p(ptr)
        obj._owner = False
        obj._keep = keepalive
        return obj

    @classmethod
    def from_host(cls, a: np.ndarray, ctx: Context | None = None, dtype=None) -> "DeviceArray":
        a = np.asarray(a)
        if dtype is not None:
            a = a.astype(dtype, copy=False)
        a = np.asfortranarray(a)
        d = cls(a.shape, a.dtype, ctx)
        _lib.check(d.ctx._lib.mslr_upload(d.ctx.handle, d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.asfortranarray(np.asarray(a, dtype=self.dtype))
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _lib.check(self.ctx._lib.mslr_upload(self.ctx.handle, self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        _lib.check(self.ctx._lib.mslr_download(self.ctx.handle, out.ctypes.data_as(C.c_void_p),
                                               self.ptr, self.nbytes))
        return out

    def zero(self) -> None:
        _lib.check(self.ctx._lib.mslr_zero(self.ctx.handle, self.ptr, self.nbytes))

    def copy(self) -> "DeviceArray":
        d = DeviceArray(self.shape, self.dtype, self.ctx)
        _lib.check(self.ctx._lib.mslr_copy_d2d(self.ctx.handle, d.ptr, self.ptr, self.nbytes))
        return d

    def reshape(self, *shape) -> "DeviceArray":
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
            shape = tuple(shape[0])
        if int(np.prod(shape, dtype=np.int64)) * self.dtype.itemsize != self.nbytes:
            raise ValueError("reshape changes the size")
        return DeviceArray.view(self.ptr.value, shape, self.dtype, self.ctx, keepalive=self)

    def offset_view(self, elem_offset: int, shape) -> "DeviceArray":
        base = self.ptr.value + int(elem_offset) * self.dtype.itemsize
        return DeviceArray.view(base, shape, self.dtype, self.ctx, keepalive=self)

    def __del__(self):  # pragma: no cover
        try:
            if getattr(self, "_owner", False) and self.ptr and self.ctx.handle:
                self.ctx._lib.mslr_free(self.ctx.handle, self.ptr)
                self.ptr = None
        except Exception:
            pass

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"


def as_device(x, dtype=np.complex64, ctx: Context | None = None) -> DeviceArray:
    if isinstance(x, DeviceArray):
        if x.dtype != np.dtype(dtype):
            raise TypeError(f"expected {np.dtype(dtype)} device array, got {x.dtype}")
        return x
    return DeviceArray.from_host(np.asarray(x), ctx, dtype=dtype)
