"""Prepared synthetic MSLR problems for the parity tests (generator + reference setup)."""
import functools

import numpy as np

import oracle as O
from mslr_hip import synth


SMALL = {
    "small": {},
    # overlapping 4^3 patches (stride 2) and the 1x1x1 voxel scale
    "small_overlap": dict(patch_sizes=([12, 12, 10], [4, 4, 4], [1, 1, 1]),
                          strides=([12, 12, 10], [2, 2, 2], [1, 1, 1])),
    # patches with fewer voxels than frames (2^3 = 8 < Nt = 10) and the 1x1x1 scale
    "small_vjp": dict(patch_sizes=([12, 12, 10], [2, 2, 2], [1, 1, 1]),
                      strides=([12, 12, 10], [2, 2, 2], [1, 1, 1])),
    # odd image sizes (15 = 3*5, 9 = 3^2): the shifts differ from the even-size sign flips
    "small_odd": dict(N=(15, 9, 10), Nt=6, patch_sizes=([15, 9, 10], [5, 3, 5]),
                      strides=([15, 9, 10], [5, 3, 5])),
}


@functools.lru_cache(maxsize=4)
def _problem(name: str, seed: int):
    cfg = synth.small_config(**SMALL[name]) if name in SMALL else synth.CONFIGS[name]
    if name not in SMALL:
        seed = synth.SEEDS[name]
    return cfg, synth.generate(cfg, seed=seed), seed


def truth(name: str = "small", seed: int = 7):
    """The synthetic image (Nx,Ny,Nz,Nt) the k-space was made from."""
    return _problem(name, seed)[1].x_true


@functools.lru_cache(maxsize=8)
def raw(name: str = "small", seed: int = 7):
    """(cfg, ksp0 (Nx,Ny,Nz,Nc,Nt) zero-filled, raw smaps): the run_recon inputs."""
    cfg, pb, seed = _problem(name, seed)
    N, Nt, Nc = tuple(cfg["N"]), cfg["Nt"], cfg["Nc"]
    y = O.A_forward(pb.x_true, pb.mask, pb.smaps)
    y = synth.add_noise(y, seed)
    ksp0 = np.zeros((*N, Nc, Nt), dtype=np.complex64, order="F")
    flat = ksp0.reshape(-1, Nc, Nt, order="F")
    for t in range(Nt):
        flat[O.sample_index(pb.mask[..., t]), :, t] = y[:, :, t]
    return cfg, ksp0, pb.smaps


@functools.lru_cache(maxsize=8)
def prepared(name: str = "small", seed: int = 7):
    """(cfg, smaps, ksp (K,Nc,Nt) normalised, omega, sf) following scripts/reconstruct.jl:86-147."""
    cfg, ksp0, smaps_raw = raw(name, seed)
    N, Nt, Nc = tuple(cfg["N"]), cfg["Nt"], cfg["Nc"]
    smaps, ksp, omega, R, sf = O.prepare(ksp0, smaps_raw, N, Nc, Nt)
    return cfg, smaps, ksp, omega, sf


def production_inputs(name: str, ctx=None):
    """The synthetic workload at its full size as tests/golden/make_recon_golden.py builds it:
    (cfg, problem, normalised smaps, y_raw (K, Nc, Nt) complex64) with y = A x_true in FP64
    (here the device's FP64 SENSE forward, there the oracle's) rounded to ComplexF32, plus
    the seeded noise.  Needs the GPU."""
    from mslr_hip import synth
    from mslr_hip.device import DeviceArray, default_context
    from mslr_hip.reconstruct import normalise_smaps
    from mslr_hip.sense import SenseOp
    ctx = ctx or default_context()
    cfg = synth.CONFIGS[name]
    pb = synth.generate(name)
    smaps = normalise_smaps(pb.smaps)
    A = SenseOp(pb.mask, smaps, ctx=ctx)
    xd = DeviceArray.from_host(pb.x_true, ctx, dtype=np.complex128)
    y = A.forward(xd, f64=True)
    del xd
    yh = y.to_host().astype(np.complex64)
    del y
    A.close()
    return cfg, pb, smaps, synth.add_noise(yh, synth.SEEDS[name])


def recon_golden(name: str):
    """(record dict, subset voxel indices, subset values (nvox, Nt) complex128) of
    tests/golden/recon_<name>.json / _subset.npz, or None when absent."""
    import json
    from pathlib import Path
    here = Path(__file__).resolve().parent / "golden"
    js, npz = here / f"recon_{name}.json", here / f"recon_{name}_subset.npz"
    if not (js.exists() and npz.exists()):
        return None
    rec = json.loads(js.read_text())
    with np.load(npz, allow_pickle=False) as z:
        vox, sub = z["voxels"], z["re"] + 1j * z["im"]
    return rec, vox, sub
