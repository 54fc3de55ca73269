"""run_recon with the keyword API of scripts/reconstruct.jl:59-72, on the GPU library.

    run_recon(; fn_ksp, fn_smaps, fn_recon_base, N, Nvc, Nt, FOV, PATCH_SIZES,
                STRIDES, NITERS, σ1A_PRECOMPUTED, use_gpu=true)

`run_recon_arrays` takes the k-space and coil maps as arrays instead of file
names (the reference reads HDF5 k-space, which needs h5py -- not available in
this image) and returns the dict the reference writes to its .mat file
(scripts/reconstruct.jl:242-260).  Setup (smaps normalisation, mask inference and
validation, the k-space flatten, the 0.99-quantile scale factor of the SENSE
combination, scripts/reconstruct.jl:86-118,145-147) runs on the device.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import _lib
from .device import Context, DeviceArray, as_device, default_context
from .mirt_mod import POGM, poweriter_mod
from .sense import SenseOp

EPS32 = float(np.finfo(np.float32).eps)
EPS64 = float(np.finfo(np.float64).eps)


def lambdas_ong_lustig(N, Nt, patch_sizes):
    """Regularisation weights (scripts/reconstruct.jl:163-169)."""
    Nvox = int(np.prod(N))
    return [math.sqrt(int(np.prod(ps))) + math.sqrt(Nt) +
            math.sqrt(math.log(Nvox * Nt / max(int(np.prod(ps)), Nt))) for ps in patch_sizes]


def normalise_smaps(smaps_raw, ctx: Context | None = None) -> np.ndarray:
    """smaps ./ (sqrt.(sum(abs2.(smaps); dims=4)) .+ eps(Float32)) (scripts/reconstruct.jl:88-89)
    on the device (mslr_normalise_smaps); returns the ComplexF32 maps on the host, which
    the SENSE plan is built from."""
    ctx = ctx or default_context()
    s = np.asfortranarray(np.asarray(smaps_raw).astype(np.complex64))
    if s.ndim != 4:
        raise ValueError("smaps must be (Nx, Ny, Nz, Nc)")
    d = DeviceArray.from_host(s, ctx)
    _lib.check(ctx._lib.mslr_normalise_smaps(ctx.handle, d.ptr, _lib.i64x3(s.shape[:3]), s.shape[3]))
    return d.to_host()


def sense_comb_scale(A: SenseOp, y, q: float = 0.99) -> float:
    """quantile(vec(abs.(sense_comb(ksp0, smaps))), q) (scripts/reconstruct.jl:102-103) from the
    flattened k-space y (K, Nc, Nt) through the plan's exact adjoint, on the device
    (mslr_sense_scale_factor; any 2/3/5 image size)."""
    yd = as_device(y, np.complex64, A.ctx)
    sf = C.c_double(0.0)
    _lib.check(A.ctx._lib.mslr_sense_scale_factor(A.handle, yd.ptr, float(q), C.byref(sf)))
    return sf.value


def scale_kspace(yd: DeviceArray, sf: float) -> None:
    """y ./= max(sf, eps(Float32)) in place (scripts/reconstruct.jl:104; FP64 division)."""
    n = yd.nbytes // 8
    _lib.check(yd.ctx._lib.mslr_scale_kspace(yd.ctx.handle, yd.ptr, n, float(sf)))


def prepare_flat(smaps_raw, mask, y_raw, ctx: Context | None = None):
    """Setup on flattened k-space y (K, Nc, Nt) in reference sample order.

    Returns (A, smaps, y_device_normalised, scale_factor, R).
    """
    ctx = ctx or default_context()
    smaps = normalise_smaps(smaps_raw, ctx)
    A = SenseOp(mask, smaps, ctx=ctx)
    # a caller's device buffer is copied, not scaled in place
    yd = y_raw.copy() if isinstance(y_raw, DeviceArray) else as_device(y_raw, np.complex64, ctx)
    sf = sense_comb_scale(A, yd)
    scale_kspace(yd, sf)
    R = float(np.prod(A.N)) / float(A.K)
    return A, smaps, yd, sf, R


def flatten_kspace(ksp0, Nvc: int, Nt: int, ctx: Context | None = None):
    """Mask inference, validation and the (K, Nvc, Nt) flatten of the zero-filled k-space
    (scripts/reconstruct.jl:109-118,145-147) on the device (mslr_kspace_flatten).
    ksp0: (Nx, Ny, Nz, Nvc, Nt), host or device.  Returns (omega host bool (Nx,Ny,Nz,Nt),
    y DeviceArray (K, Nvc, Nt)); raises AssertionError with the reference's messages."""
    ctx = ctx or default_context()
    kd = as_device(ksp0, np.complex64, ctx)
    N = tuple(int(v) for v in kd.shape[:3])
    if tuple(kd.shape) != (*N, Nvc, Nt):
        raise AssertionError(f"Unexpected k-space shape: {kd.shape}")
    md = DeviceArray((*N, Nt), np.uint8, ctx)
    K = C.c_int64(0)
    _lib.check(ctx._lib.mslr_kspace_flatten(ctx.handle, kd.ptr, _lib.i64x3(N), Nvc, Nt, md.ptr, None, 0,
                                            C.byref(K)))
    yd = DeviceArray((K.value, Nvc, Nt), np.complex64, ctx)
    _lib.check(ctx._lib.mslr_kspace_flatten(ctx.handle, kd.ptr, _lib.i64x3(N), Nvc, Nt, md.ptr, yd.ptr,
                                            K.value * Nvc * Nt, C.byref(K)))
    return md.to_host().astype(bool), yd


def prepare_device(ksp0, smaps_raw, N, Nvc: int, Nt: int, ctx: Context | None = None):
    """The setup of run_recon (scripts/reconstruct.jl:86-118,145-147) on the device:
    normalised maps, the mask and its validation, the flattened k-space divided by the
    0.99-quantile scale factor.  Returns (A, smaps, y DeviceArray, sf, R, omega)."""
    ctx = ctx or default_context()
    N = tuple(int(n) for n in N)
    if tuple(np.shape(ksp0)) != (*N, Nvc, Nt):
        raise AssertionError(f"Unexpected k-space shape: {tuple(np.shape(ksp0))}")
    omega, yd = flatten_kspace(ksp0, Nvc, Nt, ctx)
    smaps = normalise_smaps(smaps_raw, ctx)
    A = SenseOp(omega, smaps, ctx=ctx)
    sf = sense_comb_scale(A, yd)
    scale_kspace(yd, sf)
    R = float(np.prod(N)) / float(A.K)
    return A, smaps, yd, sf, R, omega


def run_recon_arrays(ksp0, smaps_raw, *, N, Nvc, Nt, PATCH_SIZES, STRIDES, NITERS,
                     σ1A_PRECOMPUTED=None, form="plain", verbose=True, ctx: Context | None = None):
    """scripts/reconstruct.jl:83-238 on in-memory arrays; returns the output dict."""
    ctx = ctx or default_context()
    N = tuple(int(n) for n in N)
    ksp0 = np.asarray(ksp0)
    if ksp0.shape != (*N, Nvc, Nt):
        raise AssertionError(f"Unexpected k-space shape: {ksp0.shape}")
    S = len(PATCH_SIZES)
    A, smaps, yd, sf, R, omega = prepare_device(ksp0, smaps_raw, N, Nvc, Nt, ctx)
    if verbose:
        print(f"  Scale factor = {sf:.4f}\nAcceleration factor R ≈ {R:.2f}")
    if σ1A_PRECOMPUTED is None:
        _, s1 = poweriter_mod(A, chat=verbose)
    else:
        s1 = float(σ1A_PRECOMPUTED)
    L = S * s1 ** 2
    lams = lambdas_ong_lustig(N, Nt, PATCH_SIZES)
    st = POGM(A, patch_sizes=PATCH_SIZES, strides=STRIDES, lambdas=lams, L=L, niter=NITERS, form=form,
              log_costs=True)
    st.init(yd)
    t0 = time.time()
    st.run(NITERS)
    ctx.sync()
    if verbose:
        print(f"POGM: {NITERS} iterations in {time.time() - t0:.2f} s")
    costs = st.logger_out()
    X = st.X.to_host()
    return {
        "X": X, "X_recon": X.sum(axis=4), "omega": omega,
        "dc_costs": np.array([c[0] for c in costs]), "reg_costs": np.array([c[1] for c in costs]),
        "restarts": np.array([c[2] for c in costs]), "R": R,
        "sigma1A": s1, "L": L, "Nscales": S, "patch_sizes": PATCH_SIZES, "strides": STRIDES,
        "lambdas": lams, "Niters": NITERS, "scale_factor": sf, "used_gpu": True,
    }


def run_recon(*, fn_ksp: str, fn_smaps: str, fn_recon_base: str, N, Nvc: int, Nt: int, FOV,
              PATCH_SIZES, STRIDES, NITERS: int, σ1A_PRECOMPUTED=None, use_gpu: bool = True,
              form: str = "plain"):
    """Keyword API of scripts/reconstruct.jl:59-72.  use_gpu=false is not served here."""
    if not use_gpu:
        raise _lib.MSLRUnsupported("this package is the GPU path; the CPU path is the reference's")
    import scipy.io as sio
    try:
        smaps_raw = sio.loadmat(fn_smaps)["smaps"]
    except NotImplementedError as e:  # MAT v7.3 (HDF5)
        raise _lib.MSLRUnsupported(f"{fn_smaps}: HDF5-based .mat needs h5py") from e
    if str(fn_ksp).endswith((".h5", ".hdf5")):
        raise _lib.MSLRUnsupported("reading HDF5 k-space needs h5py; use run_recon_arrays")
    ksp0 = np.load(fn_ksp, allow_pickle=False)
    out = run_recon_arrays(ksp0, smaps_raw, N=N, Nvc=Nvc, Nt=Nt, PATCH_SIZES=PATCH_SIZES,
                           STRIDES=STRIDES, NITERS=NITERS, σ1A_PRECOMPUTED=σ1A_PRECOMPUTED, form=form)
    fn_out = f"{fn_recon_base}_{len(PATCH_SIZES)}scales.mat"
    sio.savemat(fn_out, mat_dict(out), do_compression=True, oned_as="column")
    print(f"\n✓ Saved → {fn_out}")
    return None


def mat_dict(out: dict) -> dict:
    """The saved variables of scripts/reconstruct.jl:241-260 in MAT-v5 form: vectors as
    columns (Julia's Vector), patch_sizes / strides (Vector of Vectors) as cell arrays."""
    d = {}
    for k, v in out.items():
        if v is None:
            continue
        if k in ("patch_sizes", "strides"):
            cell = np.empty(len(v), dtype=object)
            for i, row in enumerate(v):
                cell[i] = np.asarray(row, dtype=np.float64)
            v = cell
        elif k in ("lambdas", "dc_costs", "reg_costs"):
            v = np.asarray(v, dtype=np.float64)
        elif k == "restarts":
            v = np.asarray(v, dtype=bool)
        d[k] = v
    return d
