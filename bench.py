#!/usr/bin/env python3
"""POGM iterations/s of the MSLR hot path on one MI355X.

Workload (default): 90x90x60 image, Nt=300 frames, 18 coils, 8x ky-kz
undersampling, three scales [[90,90,60],[10,10,10],[1,1,1]] with
non-overlapping strides, sigma1(A)=1 (the production shape).  Synthetic data
of that shape from mslr_hip.synth (seed 1).  One step = one POGM iteration:
gradient (A and A' over all frames) + momentum/restart update + patch-SVST
prox over all scales.  Inputs are resident in HBM when timing starts.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "mslr-recon_amd"))
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# --------------------------------------------------------------------------- algorithmic work
def kernel_bytes_per_iter(cfg, K, lines_total, nchunk=1):
    """Algorithmic HBM bytes of each kernel per POGM iteration (inputs read once, outputs
    written once).  lines_total: sampled (ky, kz) lines summed over frames; nchunk: frame
    chunks of the SENSE passes (coil maps are read once per chunk)."""
    N, Nt, Nc = cfg["N"], cfg["Nt"], cfg["Nc"]
    V = int(np.prod(N))
    S = len(cfg["patch_sizes"])
    b = 8
    lines_b = Nc * N[0] * lines_total * b      # sampled-line intermediate, all frames
    out = {
        # S scale images in, plane-major sum out
        "sum_planes": S * V * Nt * b + V * Nt * b,
        # image planes in, coil maps in, sampled lines out
        "fwd_yz": V * Nt * b + nchunk * Nc * V * b + lines_b,
        # lines in, y in, residual out, sample table in
        "fwd_x": lines_b + 2 * K * Nc * Nt * b + K * Nt * 4,
        # residual in, sample table in, lines out
        "adj_x": K * Nc * Nt * b + K * Nt * 4 + lines_b,
        # lines in, coil maps in, image planes out
        "adj_yz": lines_b + nchunk * Nc * V * b + V * Nt * b,
        "planes_to_image": 2 * V * Nt * b,
        # X, U in and out (all scales), g in
        "pogm_update": 4 * S * V * Nt * b + V * Nt * b,
    }
    img2p = p2img = svst_rw = 0
    for ps, ss in zip(cfg["patch_sizes"], cfg["strides"]):
        psz = [min(ps[a], N[a]) for a in range(3)]
        ns = [math.ceil((N[a] - psz[a]) / ss[a]) + 1 for a in range(3)]
        Np = ns[0] * ns[1] * ns[2]
        p = psz[0] * psz[1] * psz[2]
        img2p += V * Nt * b + p * Nt * Np * b
        p2img += p * Nt * Np * b + V * Nt * b
        svst_rw += 2 * p * Nt * Np * b
    out["img2patches"] = img2p
    out["patches2img"] = p2img
    out["apply"] = svst_rw
    return out


def gram_flops_per_iter(cfg):
    """Complex Gram products (full-matrix convention 8*max(p,Nt)*m^2 per patch)."""
    N, Nt = cfg["N"], cfg["Nt"]
    tot = 0
    for ps, ss in zip(cfg["patch_sizes"], cfg["strides"]):
        psz = [min(ps[a], N[a]) for a in range(3)]
        ns = [math.ceil((N[a] - psz[a]) / ss[a]) + 1 for a in range(3)]
        Np = ns[0] * ns[1] * ns[2]
        p = psz[0] * psz[1] * psz[2]
        m = min(p, Nt)
        if m > 1:
            tot += 8 * max(p, Nt) * m * m * Np
    return tot


# --------------------------------------------------------------------------- CPU baseline
def cpu_baseline(cfg, pb, y_norm, smaps, budget_s=30.0):
    """The oracle (a restatement of the reference CPU path) timed on samples of one
    POGM iteration and scaled to a full iteration.  Returns (iter/s, cores, sample text)."""
    import oracle as O
    N, Nt, Nc = tuple(cfg["N"]), cfg["Nt"], cfg["Nc"]
    V = int(np.prod(N))
    cores = O._workers()
    parts = {}
    mask = pb.mask
    # SENSE A and A' for a few frames, all coils
    fs = 2
    x = pb.x_true[..., :fs].astype(np.complex64)
    t = time.perf_counter()
    r = O.A_forward(x, mask[..., :fs], smaps) - y_norm[:, :, :fs]
    O.A_adjoint(r, mask[..., :fs], smaps)
    parts["sense"] = (time.perf_counter() - t) * Nt / fs
    lams = O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    img = pb.x_true
    for k, (ps, ss) in enumerate(zip(cfg["patch_sizes"], cfg["strides"])):
        psc, nst, Np = O.patch_geometry(N, ps, ss)
        p = int(np.prod(psc))
        beta = 0.5 * lams[k]
        if Np == 1:
            # one global patch: gesdd cost is linear in the rows of a tall matrix
            rows = max(Nt, p // 10)
            X = img.reshape(V, Nt, order="F")[:rows]
            t = time.perf_counter()
            O.SVST(X, beta)
            parts[f"svst_scale{k + 1}"] = (time.perf_counter() - t) * p / rows
        else:
            ns = min(Np, 8 if p > 1 else 4000)
            # sample patches: first ns patches of the scale
            t = time.perf_counter()
            Psub = np.empty((p, Nt, ns), dtype=np.complex64)
            it = O._patch_starts(N, psc, ss, nst)
            for i in range(ns):
                sx, sy, sz = next(it)
                Psub[:, :, i] = img[sx:sx + psc[0], sy:sy + psc[1], sz:sz + psc[2], :].reshape(p, Nt, order="F")
            O.svst_batched(Psub, beta)
            parts[f"svst_scale{k + 1}"] = (time.perf_counter() - t) * Np / ns
    # POGM update on a tenth of the state
    S = len(cfg["patch_sizes"])
    n = V * Nt * S // 10
    xs = np.ones(n, dtype=np.complex128)
    us = np.ones(n, dtype=np.complex128)
    g = np.ones(n, dtype=np.complex64)
    t = time.perf_counter()
    u = xs - 0.3 * g
    z = u + 0.1 * (u - us) + 0.6 * (u - xs)
    float(np.sum(np.real(-g * (u - us)), dtype=np.float64))
    del z
    parts["pogm_update"] = (time.perf_counter() - t) * 10
    tot = sum(parts.values())
    sample = (f"oracle (NumPy/SciPy restatement of the reference CPU path) on samples of one iteration, "
              f"scaled: A+A' on {fs}/{Nt} frames; global-patch SVD on 1/10 of the rows; "
              f"8 of the 10^3 patches and 4000 of the voxel patches (LAPACK gesdd per patch, as the reference); "
              f"POGM update on 1/10 of the state; parts (s/iter): "
              + ", ".join(f"{k}={v:.2f}" for k, v in parts.items()))
    return 1.0 / tot, cores, sample


# --------------------------------------------------------------------------- counter calibration
def copy_calibration(mib: int) -> int:
    """A plain elementwise copy (dst = src * 1) of a known size: each launch reads and
    writes exactly mib MiB, so its FETCH_SIZE / WRITE_SIZE give the counters' scale."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib, "launches": 5, "bytes_read_per_launch": n * 4,
                      "bytes_written_per_launch": n * 4}), flush=True)
    return 0


# --------------------------------------------------------------------------- main
def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--form", default="plain", choices=["plain", "batched"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-prof", action="store_true", help="skip per-kernel event timing")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (read + write) five times, to scale "
                         "the FETCH_SIZE / WRITE_SIZE counters, and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(f"bench.py: this build runs the MSLR hot path on one GPU; --gpus {args.gpus} is not "
              f"supported (the path would shard frames for A/A' and space for the prox).", file=sys.stderr)
        return 2

    from mslr_hip import synth
    from mslr_hip.device import DeviceArray, default_context
    from mslr_hip.mirt_mod import POGM
    from mslr_hip.reconstruct import lambdas_ong_lustig, normalise_smaps, prepare_flat
    from mslr_hip.sense import SenseOp

    name = args.config
    cfg = synth.CONFIGS[name]
    seed = synth.SEEDS[name]
    N, Nt, Nc = tuple(cfg["N"]), cfg["Nt"], cfg["Nc"]
    S = len(cfg["patch_sizes"])
    t = time.time()
    pb = synth.generate(name)
    log(f"generated {name}: N={N} Nt={Nt} Nc={Nc} lines/frame={pb.lines} in {time.time() - t:.1f}s")
    ctx = default_context()
    info = ctx.info()
    log(f"device: {info['name']} total={info['total_bytes'] / 2**30:.0f} GiB")
    t = time.time()
    smaps = normalise_smaps(pb.smaps)
    A0 = SenseOp(pb.mask, smaps, ctx=ctx)
    yclean = A0.forward(DeviceArray.from_host(pb.x_true, ctx)).to_host()
    A0.close()
    y = synth.add_noise(yclean, seed)
    del yclean
    A, smaps, yd, sf, R = prepare_flat(pb.smaps, pb.mask, y, ctx)
    log(f"k-space {y.shape} scale factor {sf:.4g} R={R:.2f} in {time.time() - t:.1f}s")
    lams = lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    L = S * 1.0
    W, K = args.warmup, args.steps
    st = POGM(A, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=lams, L=L,
              niter=W + K, form=args.form)
    st.init(yd)
    ctx.sync()
    t = time.time()
    st.run(W)
    ctx.sync()
    log(f"warmup {W} iterations in {time.time() - t:.2f}s")
    prof = not args.no_prof
    if prof:
        ctx.prof_enable(True)
    ctx.sync()
    t0 = time.perf_counter()
    st.run(K)
    ctx.sync()
    t1 = time.perf_counter()
    kern = ctx.prof_report() if prof else {}
    if prof:
        ctx.prof_enable(False)
    elapsed = t1 - t0
    value = K / elapsed
    _, restarts = st.log()
    log(f"timed {K} iterations in {elapsed:.3f}s -> {value:.3f} iter/s; restarts {restarts}")
    for kname, (cnt, ms) in sorted(kern.items(), key=lambda kv: -kv[1][1]):
        log(f"  {kname:14s} launches={cnt:6d} total={ms:9.2f} ms  ({100 * ms / (elapsed * 1e3):5.1f}%)")

    roof = None
    if kern:
        lines_total = int(pb.mask.reshape(N[0], N[1] * N[2], Nt, order="F").any(axis=0).sum())
        budget = 4 << 30
        nchunk = max(math.ceil(Nc * N[0] * lines_total * 8 / budget),
                      math.ceil(int(np.prod(N)) * Nt * 8 / budget))
        by = kernel_bytes_per_iter(cfg, A.K, lines_total, nchunk)
        cand = {k: v for k, v in kern.items() if k in by}
        # per-scale tags ("tridiag@s2") summed per kernel for the dominant-kernel figure
        per_kernel = {}
        for kname, (cnt, ms) in kern.items():
            base = kname.split("@")[0]
            c0, m0 = per_kernel.get(base, (0, 0.0))
            per_kernel[base] = (c0 + cnt, m0 + ms)
        dom = max(per_kernel.items(), key=lambda kv: kv[1][1])[0]
        hb = max(cand.items(), key=lambda kv: kv[1][1])[0] if cand else None
        if hb is not None:
            cnt, ms = kern[hb]
            per_launch = by[hb] * K / cnt
            ach = per_launch / (ms / cnt / 1e3) / 1e9
            roof = {"bound": "hbm", "kernel": hb, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                    "bytes_per_launch": int(per_launch), "avg_launch_ms": round(ms / cnt, 4),
                    "dominant_kernel": dom,
                    "dominant_share": round(per_kernel[dom][1] / (elapsed * 1e3), 3)}

    cpu = None
    if not args.no_cpu_baseline:
        t = time.time()
        y_norm = (y / np.float32(sf)).astype(np.complex64)
        v, cores, sample = cpu_baseline(cfg, pb, y_norm, smaps)
        cpu = {"value": round(v, 6), "unit": "iter/s", "cores": cores, "kind": "port", "sample": sample}
        log(f"cpu baseline {v:.5f} iter/s ({cores} threads) in {time.time() - t:.1f}s")

    line = {
        "metric": "POGM iter/s @90x90x60,Nt=300,18ch: mixed+batched vs FP32 plain, ΔNRMSE<1e-4",
        "value": round(value, 4),
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * elapsed / K, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"{name}: {N[0]}x{N[1]}x{N[2]} Nt={Nt} {Nc} coils R={cfg['R']} scales={cfg['patch_sizes']}",
                   "form": args.form, "parallelism": "single-gpu", "restarts_timed": sum(restarts[W:])},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels_ms_per_iter": {k: round(v[1] / K, 3) for k, v in kern.items()},
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
