"""run_recon with the keyword API of scripts/reconstruct.jl:59-72, on the GPU library.

    run_recon(; fn_ksp, fn_smaps, fn_recon_base, N, Nvc, Nt, FOV, PATCH_SIZES,
                STRIDES, NITERS, σ1A_PRECOMPUTED, use_gpu=true)

`run_recon_arrays` takes the k-space and coil maps as arrays instead of file
names (the reference reads HDF5 k-space, which needs h5py -- not available in
this image) and returns the dict the reference writes to its .mat file
(scripts/reconstruct.jl:242-260).  Setup (smaps normalisation, the 0.99-quantile
k-space scale, mask inference and validation, scripts/reconstruct.jl:86-118)
runs on the host; its one transform (the SENSE combination) runs through the
GPU adjoint.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import _lib
from .device import Context, DeviceArray, default_context
from .mirt_mod import POGM, poweriter_mod
from .sense import SenseOp

EPS32 = float(np.finfo(np.float32).eps)
EPS64 = float(np.finfo(np.float64).eps)


def lambdas_ong_lustig(N, Nt, patch_sizes):
    """Regularisation weights (scripts/reconstruct.jl:163-169)."""
    Nvox = int(np.prod(N))
    return [math.sqrt(int(np.prod(ps))) + math.sqrt(Nt) +
            math.sqrt(math.log(Nvox * Nt / max(int(np.prod(ps)), Nt))) for ps in patch_sizes]


def normalise_smaps(smaps_raw) -> np.ndarray:
    """ComplexF32 maps divided by their root-sum-of-squares + eps(Float32) (scripts/reconstruct.jl:88-89)."""
    s = np.asarray(smaps_raw).astype(np.complex64)
    rss = np.sqrt(np.sum(np.abs(s) ** 2, axis=3, keepdims=True))
    return np.asfortranarray(s / (rss + np.float32(EPS32)))


def sense_comb_scale(A: SenseOp, y, smaps: np.ndarray) -> float:
    """0.99 quantile of |sense_comb(ksp)| (scripts/reconstruct.jl:102-103).

    For even sizes sense_comb(ksp) = A'y / (sqrt(prod(N)) * (sum|S|^2 + eps)), with
    A' the exact adjoint computed on the GPU.
    """
    if any(n % 2 for n in A.N):
        raise _lib.MSLRUnsupported("the scale factor uses the adjoint; odd image sizes shift differently")
    img = A.adjoint(y).to_host()
    V = float(np.prod(A.N))
    den = (np.sum(np.abs(smaps.astype(np.complex128)) ** 2, axis=3) + EPS64) * math.sqrt(V)
    mag = np.abs(img)
    mag /= den[..., None].astype(np.float32)
    return float(np.quantile(mag.ravel(order="F").astype(np.float64), 0.99, method="linear"))


def prepare_flat(smaps_raw, mask, y_raw, ctx: Context | None = None):
    """Setup on flattened k-space y (K, Nc, Nt) in reference sample order.

    Returns (A, smaps, y_device_normalised, scale_factor, R).
    """
    ctx = ctx or default_context()
    smaps = normalise_smaps(smaps_raw)
    A = SenseOp(mask, smaps, ctx=ctx)
    yd = DeviceArray.from_host(np.asarray(y_raw), ctx, dtype=np.complex64)
    sf = sense_comb_scale(A, yd, smaps)
    yh = np.asfortranarray(np.asarray(y_raw, dtype=np.complex64) / np.float32(max(sf, EPS32)))
    yd.upload(yh)
    R = float(np.prod(A.N)) / float(A.K)
    return A, smaps, yd, sf, R


def flatten_kspace(ksp0, Nvc, Nt):
    """Mask inference, validation and the (K, Nvc, Nt) flatten (scripts/reconstruct.jl:109-118,145-147)."""
    ksp0 = np.asarray(ksp0)
    omega = ksp0[:, :, :, 0, :] != 0
    for ic in range(1, Nvc):
        if not np.array_equal(omega, ksp0[:, :, :, ic, :] != 0):
            raise AssertionError(f"Coil {ic + 1} has a different sampling pattern")
    counts = omega.reshape(-1, Nt, order="F").sum(axis=0)
    for it in range(1, Nt):
        if counts[it] != counts[it - 1]:
            raise AssertionError(f"Frame {it + 1} has a different sample count")
    flat = ksp0.reshape(-1, Nvc, Nt, order="F")
    y = np.stack([flat[np.flatnonzero(omega[..., t].ravel(order="F")), :, t] for t in range(Nt)], axis=2)
    return omega, np.asfortranarray(y.astype(np.complex64))


def run_recon_arrays(ksp0, smaps_raw, *, N, Nvc, Nt, PATCH_SIZES, STRIDES, NITERS,
                     σ1A_PRECOMPUTED=None, form="plain", verbose=True, ctx: Context | None = None):
    """scripts/reconstruct.jl:83-238 on in-memory arrays; returns the output dict."""
    ctx = ctx or default_context()
    N = tuple(int(n) for n in N)
    ksp0 = np.asarray(ksp0)
    if ksp0.shape != (*N, Nvc, Nt):
        raise AssertionError(f"Unexpected k-space shape: {ksp0.shape}")
    S = len(PATCH_SIZES)
    omega, y_raw = flatten_kspace(ksp0.astype(np.complex64), Nvc, Nt)
    A, smaps, yd, sf, R = prepare_flat(smaps_raw, omega, y_raw, ctx)
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
