"""Writes lambda_grad_<config>.json: the FP64 forward-mode oracle's loss and
d loss / d lambda_k for the differentiation workload (MSE of the image sum after
NITERS unrolled POGM iterations to the synthetic truth in the k-space's units).

    python tests/golden/make_lambda_grad.py c4 10
"""
import json
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parents[1] / "mslr-recon_amd"))

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from problems import prepared, truth  # noqa: E402


def main(name: str, niter: int) -> None:
    cfg, smaps, ksp, omega, sf = prepared(name)
    x_ref = truth(name).astype(np.complex128) / sf
    t = time.time()
    loss, grads, rs, _ = O.pogm_lambda_jvp(ksp, omega, smaps, x_ref, N=tuple(cfg["N"]), Nt=cfg["Nt"],
                                           PATCH_SIZES=cfg["patch_sizes"], STRIDES=cfg["strides"],
                                           NITERS=niter,
                                           progress=lambda it: print(f"iteration {it}: {time.time() - t:.0f} s",
                                                                     flush=True))
    out = {"config": name, "niter": niter, "sigma1A": 1.0, "loss": loss, "grads": grads,
           "restarts": rs, "scale_factor": sf,
           "lambdas": O.lambdas_ong_lustig(tuple(cfg["N"]), cfg["Nt"], cfg["patch_sizes"]),
           "seconds": round(time.time() - t, 1)}
    (HERE / f"lambda_grad_{name}.json").write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out))


if __name__ == "__main__":
    main(sys.argv[1], int(sys.argv[2]))
