"""Patch operators and SVST on the GPU, with the signatures of src/recon.jl.

    img2patches(img, patch_size, stride_size) -> P              (src/recon.jl:53-78)
    patches2img(P, patch_size, stride_size, og_size) -> img     (src/recon.jl:96-126)
    SVST(X, beta) -> X_lr                                       (src/recon.jl:183-198)
    patchSVST(img, beta | lambdas, patch_size, stride_size)     (src/recon.jl:251-269)
    patch_nucnorm(P[, lambdas]) -> cost                         (src/recon.jl:138-166)

Arguments may be host numpy arrays (Julia shapes, any memory order) or
DeviceArray; results are DeviceArray (call .to_host()).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .device import Context, DeviceArray, as_device, default_context


def patch_geometry(N, patch_size, stride_size):
    """(clamped patch size, Nsteps, Np) of src/recon.jl:58-63."""
    ps = (C.c_int64 * 3)()
    ns = (C.c_int64 * 3)()
    Np = C.c_int64()
    _lib.check(_lib.load().mslr_patch_geometry(_lib.i64x3(N), _lib.i64x3(patch_size),
                                               _lib.i64x3(stride_size), ps, ns, C.byref(Np)))
    return tuple(ps), tuple(ns), Np.value


def img2patches(img, patch_size, stride_size, ctx: Context | None = None) -> DeviceArray:
    ctx = ctx or default_context()
    d = as_device(img, np.complex64, ctx)
    if len(d.shape) != 4:
        raise ValueError("img must be (Nx, Ny, Nz, Nt)")
    N, Nt = d.shape[:3], d.shape[3]
    ps, _, Np = patch_geometry(N, patch_size, stride_size)
    P = DeviceArray((int(np.prod(ps)), Nt, Np), np.complex64, ctx)
    _lib.check(ctx._lib.mslr_img2patches(ctx.handle, d.ptr, _lib.i64x3(N), Nt, _lib.i64x3(patch_size),
                                         _lib.i64x3(stride_size), P.ptr))
    return P


def patches2img(P, patch_size, stride_size, og_size, ctx: Context | None = None) -> DeviceArray:
    ctx = ctx or default_context()
    d = as_device(P, np.complex64, ctx)
    if len(d.shape) != 3:
        raise ValueError("P must be (p, Nt, Np)")
    Nt = d.shape[1]
    ps, _, Np = patch_geometry(og_size, patch_size, stride_size)
    if d.shape[0] != int(np.prod(ps)) or d.shape[2] != Np:
        raise AssertionError("P does not match the patch geometry")
    img = DeviceArray((*og_size, Nt), np.complex64, ctx)
    _lib.check(ctx._lib.mslr_patches2img(ctx.handle, d.ptr, _lib.i64x3(og_size), Nt,
                                         _lib.i64x3(patch_size), _lib.i64x3(stride_size), img.ptr))
    return img


def svst_batched(P, beta, ctx: Context | None = None, inplace: bool = False) -> DeviceArray:
    """SVST of every P[:, :, i] with threshold beta (scalar) or beta[i] (vector)."""
    ctx = ctx or default_context()
    if isinstance(P, DeviceArray):
        d = P if inplace else P.copy()
    else:
        d = DeviceArray.from_host(np.asarray(P), ctx, dtype=np.complex64)
    if len(d.shape) == 2:
        p, Nt, Np = d.shape[0], d.shape[1], 1
    else:
        p, Nt, Np = d.shape
    b = np.atleast_1d(np.asarray(beta, dtype=np.float64))
    per = int(b.size > 1 or (np.ndim(beta) > 0 and Np > 1))
    if per and b.size != Np:
        raise AssertionError("need one threshold per patch")
    b = np.ascontiguousarray(b)
    _lib.check(ctx._lib.mslr_svst_batched(ctx.handle, d.ptr, p, Nt, Np,
                                          b.ctypes.data_as(C.POINTER(C.c_double)), per))
    return d


def g_prox(X, c, lambdas, patch_sizes, strides, ctx: Context | None = None) -> DeviceArray:
    """run_recon's in-place prox over all scales (scripts/reconstruct.jl:196-202):
    X[..., k] <- patchSVST(X[..., k], c * lambdas[k], patch_sizes[k], strides[k]) for the
    5D array X (Nx, Ny, Nz, Nt, S); a DeviceArray is updated in place and returned."""
    ctx = ctx or default_context()
    d = X if isinstance(X, DeviceArray) else as_device(X, np.complex64, ctx)
    if len(d.shape) != 5:
        raise AssertionError("X must be (Nx, Ny, Nz, Nt, S)")
    S = d.shape[4]
    if len(patch_sizes) != S or len(strides) != S or len(lambdas) != S:
        raise AssertionError("one patch size, stride and lambda per scale")
    ps = np.ascontiguousarray(np.asarray(patch_sizes, dtype=np.int64).reshape(S, 3))
    ss = np.ascontiguousarray(np.asarray(strides, dtype=np.int64).reshape(S, 3))
    b = np.ascontiguousarray([float(c) * float(l) for l in lambdas], dtype=np.float64)
    N = (C.c_int64 * 3)(*d.shape[:3])
    _lib.check(ctx._lib.mslr_prox_scales(ctx.handle, d.ptr, N, d.shape[3], S,
                                         ps.ctypes.data_as(C.POINTER(C.c_int64)),
                                         ss.ctypes.data_as(C.POINTER(C.c_int64)),
                                         b.ctypes.data_as(C.POINTER(C.c_double))))
    return d


def svst_vjp(Z, Ybar, beta, ctx: Context | None = None):
    """Pullback of svst_batched: (Zbar DeviceArray, betabar) for Y = SVST(Z, beta) and the
    cotangent Ybar (same shape as Z: (p, Nt) or (p, Nt, Np), any p and Nt).  betabar has one
    value per patch when beta does, else it is the scalar d<Ybar, Y>/d beta."""
    ctx = ctx or default_context()
    z = as_device(Z, np.complex64, ctx)
    yb = as_device(Ybar, np.complex64, ctx)
    if z.shape != yb.shape:
        raise AssertionError("Z and Ybar differ in shape")
    p, Nt = z.shape[0], z.shape[1]
    Np = 1 if len(z.shape) == 2 else z.shape[2]
    b = np.atleast_1d(np.asarray(beta, dtype=np.float64))
    per = int(b.size > 1 or (np.ndim(beta) > 0 and Np > 1))
    if per and b.size != Np:
        raise AssertionError("need one threshold per patch")
    b = np.ascontiguousarray(b)
    out = DeviceArray(z.shape, np.complex64, ctx)
    bb = np.zeros(Np if per else 1)
    _lib.check(ctx._lib.mslr_svst_vjp(ctx.handle, z.ptr, yb.ptr, out.ptr, p, Nt, Np,
                                      b.ctypes.data_as(C.POINTER(C.c_double)), per,
                                      bb.ctypes.data_as(C.POINTER(C.c_double))))
    return out, (bb if per else float(bb[0]))


def SVST(X, beta, ctx: Context | None = None) -> DeviceArray:
    """Singular value soft-thresholding of one matrix X (p x Nt)."""
    d = as_device(X, np.complex64, ctx)
    if len(d.shape) != 2:
        raise ValueError("X must be a matrix")
    if float(beta) < 0:
        raise ValueError("beta must be >= 0")
    return svst_batched(d, float(beta), ctx)


def patchSVST(img, beta, patch_size, stride_size, ctx: Context | None = None) -> DeviceArray:
    """Patch-wise SVST with a global threshold (scalar beta) or per-patch thresholds (vector)."""
    ctx = ctx or default_context()
    d = as_device(img, np.complex64, ctx)
    N, Nt = d.shape[:3], d.shape[3]
    if np.ndim(beta) == 0:
        out = DeviceArray(d.shape, np.complex64, ctx)
        _lib.check(ctx._lib.mslr_patch_svst(ctx.handle, d.ptr, out.ptr, _lib.i64x3(N), Nt,
                                            _lib.i64x3(patch_size), _lib.i64x3(stride_size),
                                            float(beta)))
        return out
    P = img2patches(d, patch_size, stride_size, ctx)
    P = svst_batched(P, np.asarray(beta, dtype=np.float64), ctx, inplace=True)
    return patches2img(P, patch_size, stride_size, N, ctx)


def patch_nucnorm(P, lambdas=None, ctx: Context | None = None) -> float:
    """Sum over patches of the nuclear norm; with per-patch weights: mean of
    lambda_i * ||P_i||_* over patches whose top singular value is > 0 (src/recon.jl:157-166)."""
    ctx = ctx or default_context()
    d = as_device(P, np.complex64, ctx)
    if len(d.shape) != 3:
        raise AssertionError("P must be (space x time x patches)")
    p, Nt, Np = d.shape
    if lambdas is None:
        out = C.c_double()
        _lib.check(ctx._lib.mslr_patch_nucnorm(ctx.handle, d.ptr, p, Nt, Np, C.byref(out)))
        return out.value
    lam = np.ascontiguousarray(lambdas, dtype=np.float64)
    if lam.size != Np:
        raise AssertionError("need one weight per patch")
    out = C.c_double()
    _lib.check(ctx._lib.mslr_patch_nucnorm_weighted(ctx.handle, d.ptr, p, Nt, Np,
                                                    lam.ctypes.data_as(C.POINTER(C.c_double)),
                                                    C.byref(out)))
    return out.value
