"""Writes recon_<config>.json and recon_<config>_subset.npz: the FP64 oracle's POGM run
of run_recon (scripts/reconstruct.jl:159-231) on the synthetic workload at its full
size, for the GPU parity tests and bench.py's accuracy figure.

Inputs (the GPU side rebuilds them the same way): the synthetic problem of
mslr_hip.synth (seed of the config), coil maps normalised as
scripts/reconstruct.jl:89, k-space y = A x_true evaluated in FP64 and rounded to
ComplexF32, plus complex Gaussian noise at 1e-3 of its RMS (synth.add_noise); the
scale factor sf = the type-7 0.99 quantile of |sense_comb| of the zero-filled
k-space (scripts/reconstruct.jl:102-104, src/recon.jl:312-323); y / Float32(sf).

Recorded: sf, |y|^2, the restart flags, dc_cost and reg_cost of every iterate (the
logger, scripts/reconstruct.jl:176-191,205), the NRMSE of every iterate's image sum
against the synthetic truth (in the truth's units, sum_s X_s * sf) and the norm of
each scale, the CPU time; and the final image sum at a fixed set of voxels (all
frames, FP64) in the .npz.

    python tests/golden/make_recon_golden.py c1 12
"""
import json
import math
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parents[1] / "mslr-recon_amd"))

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from mslr_hip import synth  # noqa: E402

NVOX_SUBSET = 512


def subset_voxels(name: str, V: int) -> np.ndarray:
    """The fixed voxel set of the image-sum subset (0-based column-major indices)."""
    u = synth.uniform(synth.SEEDS[name], 777, V)
    return np.sort(np.argsort(u, kind="stable")[:NVOX_SUBSET])


def kspace_and_scale(name: str, log=print):
    """(cfg, pb, smaps, y_raw (K,Nc,Nt) complex64, sf) as described in the module docstring."""
    cfg = synth.CONFIGS[name]
    seed = synth.SEEDS[name]
    N, Nt, Nc = tuple(cfg["N"]), cfg["Nt"], cfg["Nc"]
    pb = synth.generate(name)
    smaps = O.normalise_smaps(pb.smaps)
    t = time.time()
    y = np.empty((int(pb.mask[..., 0].sum()), Nc, Nt), dtype=np.complex64, order="F")
    sm = smaps.astype(np.complex128)
    for f in range(Nt):
        y[:, :, f] = O.Asense_forward(pb.x_true[..., f].astype(np.complex128), pb.mask[..., f], sm)
    y = synth.add_noise(y, seed)
    log(f"k-space {y.shape} in {time.time() - t:.0f} s")
    t = time.time()
    V = int(np.prod(N))
    mag = np.empty((V, Nt), dtype=np.float32)
    for f in range(Nt):
        k0 = np.zeros((V, Nc, 1), dtype=np.complex64)
        k0[O.sample_index(pb.mask[..., f]), :, 0] = y[:, :, f]
        img = O.sense_comb(k0.reshape(*N, Nc, 1, order="F"), smaps)
        mag[:, f] = np.abs(img[..., 0]).ravel(order="F")
    sf = O.quantile_linear(mag.ravel(order="F"), 0.99)
    log(f"scale factor {sf!r} in {time.time() - t:.0f} s")
    return cfg, pb, smaps, y, sf


def main(name: str, niter: int) -> None:
    t0 = time.time()
    log = lambda m: print(f"[{time.time() - t0:7.0f} s] {m}", flush=True)  # noqa: E731
    cfg, pb, smaps, y, sf = kspace_and_scale(name, log)
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    V = int(np.prod(N))
    y_norm = (y / np.float32(max(sf, O.EPS32))).astype(np.complex64)
    xt = pb.x_true.astype(np.complex128)
    xt_norm = float(np.linalg.norm(xt))
    vox = subset_voxels(name, V)
    per_iter = []
    final = {}

    def cb(it, X, restart):
        s = X.sum(axis=4)
        nr = float(np.linalg.norm(s * sf - xt) / xt_norm)
        norms = [float(np.linalg.norm(X[..., k])) for k in range(X.shape[4])]
        per_iter.append({"iter": it, "nrmse_vs_truth": nr, "scale_norms": norms})
        if it == niter:
            final["subset"] = s.reshape(V, Nt, order="F")[vox, :].copy()
        log(f"iteration {it}: restart {restart} NRMSE {nr:.9f}")

    X, restarts, ex = O.run_recon_lean(y_norm, pb.mask, smaps, N=N, Nt=Nt, PATCH_SIZES=cfg["patch_sizes"],
                                       STRIDES=cfg["strides"], NITERS=niter, sigma1A=1.0,
                                       dtype=np.complex128, callback=cb)
    secs = time.time() - t0
    out = {"config": name, "niter": niter, "sigma1A": 1.0, "mode": "FP64 oracle (run_recon_lean)",
           "scale_factor": sf, "y_norm2": float(np.vdot(y.ravel(), y.ravel()).real),
           "lambdas": ex["lambdas"], "restarts": restarts,
           "dc_costs": ex["dc_costs"], "reg_costs": ex["reg_costs"],
           "nrmse_vs_truth": [r["nrmse_vs_truth"] for r in per_iter],
           "scale_norms": [r["scale_norms"] for r in per_iter],
           "subset_voxels": NVOX_SUBSET, "cpu_seconds": round(secs, 1),
           "cpu_threads": O._workers()}
    (HERE / f"recon_{name}.json").write_text(json.dumps(out, indent=1) + "\n")
    sub = final["subset"]
    np.savez(HERE / f"recon_{name}_subset.npz", voxels=vox.astype(np.int64), re=sub.real.copy(),
             im=sub.imag.copy())
    log(f"done: {json.dumps({k: out[k] for k in ('scale_factor', 'restarts', 'cpu_seconds')})}")


if __name__ == "__main__":
    main(sys.argv[1], int(sys.argv[2]))
