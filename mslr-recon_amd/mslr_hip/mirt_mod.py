"""POGM and power iteration on the GPU (src/mirt_mod.jl).

`pogm_mslr` is the device form of
    pogm_mod(X0, comp_cost, dc_cost_grad, L; mom=:pogm, niter, g_prox, fun)
as called by run_recon (scripts/reconstruct.jl:220-226) with the in-place
patch-SVST prox (scripts/reconstruct.jl:196-202): the whole iteration runs in
the library with the scalar recurrence on the device.

`poweriter_mod(A; niter, tol)` mirrors src/mirt_mod.jl:215-238 on the GPU in FP64.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceArray, as_device
from .sense import SenseOp

_MOM = {"pogm": _lib.MOM_POGM, "fpgm": _lib.MOM_FPGM, "pgm": _lib.MOM_PGM}
_RESTART = {"gr": _lib.RESTART_GR, "none": _lib.RESTART_NONE}
_FORM = {"plain": _lib.FORM_PLAIN, "batched": _lib.FORM_BATCHED, "mixed": _lib.FORM_MIXED}


def _sym(v) -> str:
    return str(v).lstrip(":")


class POGM:
    """Device POGM state for the MSLR problem (mslr_pogm)."""

    def __init__(self, A: SenseOp, *, patch_sizes, strides, lambdas, L: float, niter: int,
                 mom="pogm", restart="gr", restart_cutoff: float = 0.0, bsig: float = 1.0,
                 f_mu: float = 0.0, form="plain", log_costs: bool = False,
                 checkpoint: bool = False):
        mom, restart = _sym(mom), _sym(restart)
        if mom not in _MOM:
            raise ValueError(f"mom={mom}")
        if restart not in ("none", "gr", "fr"):
            raise ValueError(f"restart={restart}")
        if L < 0:
            raise ValueError(f"f_L={L} < 0")
        if f_mu < 0:
            raise ValueError(f"f_mu={f_mu} < 0")
        if bsig < 0:
            raise ValueError(f"bsig={bsig} < 0")
        if not abs(restart_cutoff) < 1:
            raise ValueError(f"restart_cutoff={restart_cutoff}")
        if restart == "fr":
            raise _lib.MSLRUnsupported("restart=:fr needs Fcost; the device loop implements :gr and :none")
        if f_mu != 0:
            raise _lib.MSLRUnsupported("the device loop implements f_mu = 0")
        self.A = A
        self.ctx = A.ctx
        S = len(patch_sizes)
        if len(strides) != S or len(lambdas) != S:
            raise AssertionError("patch_sizes, strides and lambdas differ in length")
        self.S = S
        self.niter = int(niter)
        ps = np.ascontiguousarray(np.asarray(patch_sizes, dtype=np.int64).reshape(S, 3))
        ss = np.ascontiguousarray(np.asarray(strides, dtype=np.int64).reshape(S, 3))
        lam = np.ascontiguousarray(np.asarray(lambdas, dtype=np.float64))
        self._keep = (ps, ss, lam)
        d = _lib.mslr_pogm_desc()
        for i in range(3):
            d.N[i] = A.N[i]
        d.Nt = A.Nt
        d.nscales = S
        d.patch_sizes = ps.ctypes.data_as(C.POINTER(C.c_int64))
        d.strides = ss.ctypes.data_as(C.POINTER(C.c_int64))
        d.lambdas = lam.ctypes.data_as(C.POINTER(C.c_double))
        d.L = float(L)
        d.niter = self.niter
        d.mom = _MOM[mom]
        d.restart = _RESTART[restart]
        d.restart_cutoff = float(restart_cutoff)
        d.bsig = float(bsig)
        d.form = _FORM[form] if isinstance(form, str) else int(form)
        d.log_costs = 1 if log_costs else 0
        d.checkpoint = 1 if checkpoint else 0
        h = C.c_void_p()
        _lib.check(self.ctx._lib.mslr_pogm_create(A.handle, C.byref(d), C.byref(h)))
        self.handle = h

    def init(self, y, X0=None) -> None:
        yd = as_device(y, np.complex64, self.ctx)
        self._y = yd
        x0p = None
        if X0 is not None:
            self._x0 = as_device(X0, np.complex64, self.ctx)
            x0p = self._x0.ptr
        _lib.check(self.ctx._lib.mslr_pogm_init(self.handle, yd.ptr, x0p))

    def run(self, n: int) -> None:
        _lib.check(self.ctx._lib.mslr_pogm_run(self.handle, int(n)))

    @property
    def X(self) -> DeviceArray:
        p = self.ctx._lib.mslr_pogm_x(self.handle)
        return DeviceArray.view(p, (*self.A.N, self.A.Nt, self.S), np.complex64, self.ctx, keepalive=self)

    def log(self):
        it = C.c_int()
        buf = (C.c_int * max(1, self.niter))()
        _lib.check(self.ctx._lib.mslr_pogm_log(self.handle, C.byref(it), buf))
        return it.value, [bool(buf[i]) for i in range(it.value)]

    def costs(self):
        """The logger's (dc_costs, reg_costs) of x_0 .. x_k after k iterations
        (scripts/reconstruct.jl:176-191,205,228-229); needs log_costs=True."""
        n = C.c_int()
        dc = (C.c_double * (self.niter + 1))()
        reg = (C.c_double * (self.niter + 1))()
        _lib.check(self.ctx._lib.mslr_pogm_costs(self.handle, C.byref(n), dc, reg))
        return [dc[i] for i in range(n.value)], [reg[i] for i in range(n.value)]

    def logger_out(self):
        """pogm_mod's `out` for run_recon's logger: [(dc_cost, reg_cost, is_restart)]
        for iterations 0..k (src/mirt_mod.jl:106,181)."""
        dc, reg = self.costs()
        _, rs = self.log()
        flags = [False] + rs
        return [(dc[i], reg[i], flags[i]) for i in range(len(dc))]

    def lambda_vjp(self, xbar):
        """d<xbar, sum_s X_s>/d lambda_s through the iterations run so far (the pullback
        of the unrolled POGM with respect to the per-scale weights); needs checkpoint=True.
        xbar: cotangent of the final image sum, (Nx, Ny, Nz, Nt)."""
        xb = as_device(xbar, np.complex64, self.ctx)
        if tuple(xb.shape) != (*self.A.N, self.A.Nt):
            raise AssertionError("xbar must be (Nx, Ny, Nz, Nt)")
        out = (C.c_double * self.S)()
        _lib.check(self.ctx._lib.mslr_pogm_lambda_vjp(self.handle, xb.ptr, out))
        return [out[i] for i in range(self.S)]

    def thresholds(self):
        b = (C.c_double * self.S)()
        _lib.check(self.ctx._lib.mslr_pogm_thresholds(self.handle, b))
        return list(b)

    def close(self):
        if getattr(self, "handle", None):
            self.ctx._lib.mslr_pogm_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def pogm_mslr(A: SenseOp, y, *, patch_sizes, strides, lambdas, L: float, niter: int,
              X0=None, form="plain", **kw):
    """Run the whole POGM loop on the device; returns (X DeviceArray, restarts list, state)."""
    st = POGM(A, patch_sizes=patch_sizes, strides=strides, lambdas=lambdas, L=L, niter=niter,
              form=form, **kw)
    st.init(y, X0)
    st.run(niter)
    _, restarts = st.log()
    return st.X, restarts, st


def pogm_lambda_grad(A: SenseOp, y, x_ref, *, patch_sizes, strides, lambdas, L: float, niter: int,
                     form="plain"):
    """The differentiation form: loss = mean |sum_s X_s - x_ref|^2 after `niter` unrolled
    POGM iterations and d loss / d lambda_s (per-iteration checkpoints, backward pass on
    the device, restart decisions held fixed).  Returns (loss, grads, X DeviceArray, state)."""
    st = POGM(A, patch_sizes=patch_sizes, strides=strides, lambdas=lambdas, L=L, niter=niter,
              form=form, checkpoint=True)
    st.init(y)
    st.run(niter)
    xs = st.X.to_host().sum(axis=4)
    r = xs.astype(np.complex128) - np.asarray(x_ref, dtype=np.complex128)
    loss = float(np.vdot(r.ravel(), r.ravel()).real) / r.size
    xbar = np.asfortranarray((2.0 / r.size) * r).astype(np.complex64)
    return loss, st.lambda_vjp(xbar), st.X, st


def poweriter_mod(A: SenseOp, *, niter: int = 200, tol: float = 1e-6, x0=None, chat: bool = True):
    """sigma1(A) by power iteration from x0 = ones in FP64 on the GPU.

    Returns (None, sigma1): the dominant singular vector is not returned to the host.
    """
    if x0 is not None:
        raise _lib.MSLRUnsupported("the device power iteration starts from ones (the default x0)")
    s, it = A.poweriter(niter, tol)
    if chat and it < niter:
        print(f"Power iteration converged at iter {it}")
    return None, s
