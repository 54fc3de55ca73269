"""Prepared synthetic MSLR problems for the parity tests (generator + reference setup)."""
import functools

import numpy as np

import oracle as O
from mslr_hip import synth


@functools.lru_cache(maxsize=8)
def prepared(name: str = "small", seed: int = 7):
    """(cfg, smaps, ksp (K,Nc,Nt) normalised, omega, sf) following scripts/reconstruct.jl:86-147."""
    cfg = synth.small_config() if name == "small" else synth.CONFIGS[name]
    if name != "small":
        seed = synth.SEEDS[name]
    pb = synth.generate(cfg, seed=seed)
    N, Nt, Nc = tuple(cfg["N"]), cfg["Nt"], cfg["Nc"]
    y = O.A_forward(pb.x_true, pb.mask, pb.smaps)
    y = synth.add_noise(y, seed)
    ksp0 = np.zeros((*N, Nc, Nt), dtype=np.complex64, order="F")
    flat = ksp0.reshape(-1, Nc, Nt, order="F")
    for t in range(Nt):
        flat[O.sample_index(pb.mask[..., t]), :, t] = y[:, :, t]
    smaps, ksp, omega, R, sf = O.prepare(ksp0, pb.smaps, N, Nc, Nt)
    return cfg, smaps, ksp, omega, sf
