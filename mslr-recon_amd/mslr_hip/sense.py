"""SENSE encoding operator on the GPU.

`Asense_hip(samp, smaps; fft_forward, unitary)` mirrors `Asense_gpu`
(src/sense_gpu.jl:56-108): a per-frame operator with idim = N and
odim = (K, Nc).  `block_diag` stacks per-frame operators over time like
`LinearMapsAA.block_diag` does in scripts/reconstruct.jl:142; the result is a
single all-frames plan (`SenseOp`) with idim (Nx,Ny,Nz,Nt) and odim (K,Nc,Nt).

The adjoint is the exact conjugate transpose of the forward.  (The reference
GPU adjoint applies a normalised ifft, src/sense_gpu.jl:93, i.e. 1/prod(N)
times this operator.)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .device import Context, DeviceArray, as_device, default_context


class SenseOp:
    """All-frames SENSE operator A: (Nx,Ny,Nz,Nt) -> (K,Nc,Nt)."""

    def __init__(self, mask, smaps, *, fft_forward: bool = True, unitary: bool = True,
                 ctx: Context | None = None):
        self.ctx = ctx or default_context()
        mask = np.asarray(mask)
        if mask.ndim == 3:
            mask = mask[..., None]
        smaps = np.asarray(smaps)
        if smaps.ndim != 4:
            raise ValueError("smaps must be (Nx, Ny, Nz, Nc)")
        if mask.shape[:3] != smaps.shape[:3]:
            raise AssertionError("mask and smaps spatial sizes differ")
        self.N = tuple(int(s) for s in smaps.shape[:3])
        self.Nc = int(smaps.shape[3])
        self.Nt = int(mask.shape[3])
        m8 = np.asfortranarray(mask != 0, dtype=np.uint8)
        s64 = np.asfortranarray(smaps, dtype=np.complex64)
        h = C.c_void_p()
        lib = self.ctx._lib
        _lib.check(lib.mslr_sense_create(self.ctx.handle, _lib.i64x3(self.N), self.Nc, self.Nt,
                                         s64.ctypes.data_as(C.c_void_p),
                                         m8.ctypes.data_as(C.c_void_p),
                                         int(bool(fft_forward)), int(bool(unitary)), C.byref(h)))
        self.handle = h
        self.K = int(lib.mslr_sense_nsamples(h))
        self.fft_forward = fft_forward
        self.unitary = unitary

    # LinearMapAA-style metadata
    @property
    def idim(self):
        return (*self.N, self.Nt)

    @property
    def odim(self):
        return (self.K, self.Nc, self.Nt)

    @property
    def shape(self):
        return (self.K * self.Nc * self.Nt, int(np.prod(self.N)) * self.Nt)

    def _frames(self, x: DeviceArray, ncomp: int):
        nf = x.nbytes // (ncomp * x.dtype.itemsize)
        return nf

    def forward(self, x, out: DeviceArray | None = None, t0: int = 0, f64: bool = False):
        """y = A x for frames t0.. (x: (Nx,Ny,Nz,nframes) device or host)."""
        dt = np.complex128 if f64 else np.complex64
        xd = as_device(x, dt, self.ctx)
        V = int(np.prod(self.N))
        nf = xd.nbytes // (V * xd.dtype.itemsize)
        if nf * V * xd.dtype.itemsize != xd.nbytes:
            raise AssertionError("x size is not a whole number of frames")
        y = out or DeviceArray((self.K, self.Nc, nf), dt, self.ctx)
        fn = self.ctx._lib.mslr_sense_forward_f64 if f64 else self.ctx._lib.mslr_sense_forward
        _lib.check(fn(self.handle, xd.ptr, y.ptr, int(t0), int(nf)))
        return y

    def forward_noalloc(self, x: DeviceArray, out: DeviceArray, t0: int = 0):
        """y = A x for frames t0.. into `out` without any device allocation
        (mslr_sense_forward_noalloc): x and out must be device arrays."""
        if not isinstance(x, DeviceArray) or not isinstance(out, DeviceArray):
            raise TypeError("forward_noalloc takes device arrays")
        if x.dtype != np.complex64 or out.dtype != np.complex64:
            raise TypeError("forward_noalloc is ComplexF32")
        V = int(np.prod(self.N))
        nf = x.nbytes // (V * 8)
        if nf * V * 8 != x.nbytes or out.nbytes != self.K * self.Nc * nf * 8:
            raise AssertionError("x and out sizes do not match a whole number of frames")
        _lib.check(self.ctx._lib.mslr_sense_forward_noalloc(self.handle, x.ptr, out.ptr, int(t0), int(nf)))
        return out

    def adjoint(self, y, out: DeviceArray | None = None, t0: int = 0, f64: bool = False):
        """x = A' y for frames t0.. (y: (K,Nc,nframes))."""
        dt = np.complex128 if f64 else np.complex64
        yd = as_device(y, dt, self.ctx)
        per = self.K * self.Nc * yd.dtype.itemsize
        nf = yd.nbytes // per if per else self.Nt
        x = out or DeviceArray((*self.N, nf), dt, self.ctx)
        fn = self.ctx._lib.mslr_sense_adjoint_f64 if f64 else self.ctx._lib.mslr_sense_adjoint
        _lib.check(fn(self.handle, yd.ptr, x.ptr, int(t0), int(nf)))
        return x

    def __matmul__(self, x):
        return self.forward(x)

    @property
    def H(self):
        return _Adjoint(self)

    # dc_cost_grad (scripts/reconstruct.jl:180-183) on the scale sum
    def grad(self, X, y, out: DeviceArray | None = None, want_cost: bool = False):
        Xd = as_device(X, np.complex64, self.ctx)
        yd = as_device(y, np.complex64, self.ctx)
        V = int(np.prod(self.N))
        S = Xd.nbytes // (V * self.Nt * 8)
        g = out or DeviceArray((*self.N, self.Nt), np.complex64, self.ctx)
        cost = C.c_double(0.0)
        _lib.check(self.ctx._lib.mslr_sense_grad(self.handle, Xd.ptr, int(S), yd.ptr, g.ptr,
                                                 C.byref(cost) if want_cost else None))
        return (g, cost.value) if want_cost else g

    def residual(self, X, y, want_residual: bool = False):
        """||A sum_k X_k - y||^2 (the data term of dc_cost, scripts/reconstruct.jl:176-178),
        with the residual (K, Nc, Nt) as a DeviceArray when asked."""
        Xd = as_device(X, np.complex64, self.ctx)
        yd = as_device(y, np.complex64, self.ctx)
        V = int(np.prod(self.N))
        S = Xd.nbytes // (V * self.Nt * 8)
        r = DeviceArray((self.K, self.Nc, self.Nt), np.complex64, self.ctx) if want_residual else None
        n2 = C.c_double(0.0)
        _lib.check(self.ctx._lib.mslr_sense_residual(self.handle, Xd.ptr, int(S), yd.ptr,
                                                     r.ptr if r is not None else None, C.byref(n2)))
        return (n2.value, r) if want_residual else n2.value

    def poweriter(self, niter: int = 200, tol: float = 1e-6):
        """(sigma1, iterations) by FP64 power iteration from x0 = ones."""
        s = C.c_double()
        it = C.c_int()
        _lib.check(self.ctx._lib.mslr_sense_poweriter(self.handle, int(niter), float(tol),
                                                      C.byref(s), C.byref(it)))
        return s.value, it.value

    def close(self):
        if getattr(self, "handle", None):
            self.ctx._lib.mslr_sense_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class _Adjoint:
    def __init__(self, op: SenseOp):
        self.op = op

    def __matmul__(self, y):
        return self.op.adjoint(y)


class Asense_hip(SenseOp):
    """Per-frame operator with the signature of Asense_gpu(samp, smaps; fft_forward, unitary)."""

    def __init__(self, samp, smaps, *, fft_forward: bool = True, unitary: bool = True,
                 ctx: Context | None = None):
        samp = np.asarray(samp)
        if samp.ndim != 3:
            raise ValueError("samp must be (Nx, Ny, Nz)")
        self.samp = samp.astype(bool)
        self._smaps = smaps
        super().__init__(samp[..., None], smaps, fft_forward=fft_forward, unitary=unitary, ctx=ctx)

    @property
    def idim(self):
        return self.N

    @property
    def odim(self):
        return (self.K, self.Nc)

    @property
    def shape(self):
        return (self.K * self.Nc, int(np.prod(self.N)))


def block_diag(ops) -> SenseOp:
    """Stack per-frame Asense_hip operators that share smaps into one all-frames plan."""
    ops = list(ops)
    if not ops:
        raise ValueError("no operators")
    first = ops[0]
    for o in ops:
        if not isinstance(o, Asense_hip):
            raise TypeError("block_diag expects Asense_hip operators")
        if o._smaps is not first._smaps and not np.array_equal(np.asarray(o._smaps), np.asarray(first._smaps)):
            raise AssertionError("block_diag: operators use different smaps")
        if o.fft_forward != first.fft_forward or o.unitary != first.unitary:
            raise AssertionError("block_diag: operators use different conventions")
    mask = np.stack([o.samp for o in ops], axis=-1)
    return SenseOp(mask, first._smaps, fft_forward=first.fft_forward, unitary=first.unitary,
                   ctx=first.ctx)
