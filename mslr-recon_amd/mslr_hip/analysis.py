"""Post-reconstruction analysis on the GPU (src/analysis.jl).

`tSNR(img)` and `detrend_(img)` (in place, the reference's `detrend!`) act on a 4D
image series (Nx, Ny, Nz, Nt), a host array or a DeviceArray.  `plotOpt` (plotting)
is not served here.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .device import Context, DeviceArray, as_device, default_context


def _sizes(d: DeviceArray):
    if len(d.shape) < 2:
        raise AssertionError("img must have a time axis last")
    V = int(np.prod(d.shape[:-1]))
    return V, int(d.shape[-1])


def tSNR(img, ctx: Context | None = None) -> DeviceArray:
    """mean(|img|) / (std(|img|) + eps(Float32)) along time (src/analysis.jl:39-44);
    returns a Float32 DeviceArray of the spatial shape."""
    ctx = ctx or default_context()
    d = as_device(img, np.complex64, ctx)
    V, Nt = _sizes(d)
    out = DeviceArray(d.shape[:-1], np.float32, ctx)
    _lib.check(ctx._lib.mslr_tsnr(ctx.handle, d.ptr, V, Nt, out.ptr))
    return out


def detrend_(img, ctx: Context | None = None) -> DeviceArray:
    """detrend!(img) (src/analysis.jl:122-132): subtracts t * slope from each voxel's time
    course (t = 1..Nt), in place on a DeviceArray (a host array is copied to the device)."""
    ctx = ctx or default_context()
    d = img if isinstance(img, DeviceArray) else as_device(img, np.complex64, ctx)
    if d.dtype != np.complex64:
        raise TypeError("img must be ComplexF32")
    V, Nt = _sizes(d)
    _lib.check(ctx._lib.mslr_detrend(ctx.handle, d.ptr, V, Nt))
    return d
