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
# HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated) of the kernel
# families, from the committed rocprofv3 counter summary of this workload (mixed
# form; the SENSE, update and patch kernels are the same in every form)
TRAFFIC_BYTES = {
    "fwd_yz": 4455800000,
    "adj_yz": 4359800000,
    "pogm_update": 15326200000,
    "x_resid": 8946600000,
    "sum_planes": 4666000000,
    "sy2sb": 5725100000,
    "img2patches": 8589500000,
    "prox_voxel": 2333500000,
    "split": 2473000000,
    "gram_tiles": 1822200000,
    "backtransform": 1406400000,
    "apply": 2125900000,
    "apply_lr": 1842100000,
    "sb2st": 240800000,
}
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
        # lines in, y in, sample table in, lines out (x FFT, residual, inverse x FFT fused)
        "x_resid": 2 * lines_b + K * Nc * Nt * b + K * Nt * 4,
        # lines in, coil maps in, image planes out
        "adj_yz": lines_b + nchunk * Nc * V * b + V * Nt * b,
        # X, U in and out (all scales), g in (plane-major, transposed in LDS)
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


def kernel_flops_per_iter(cfg, split=False):
    """Algorithmic floating-point work (real flops) of the dense SVST kernels per
    POGM iteration, per kernel family.  split: the Gram and apply products run as
    split BF16 and are counted as issued: four BF16 products per useful Gram product
    (hi.hi, hi.lo, lo.hi, lo.lo of the two-slice operands) and three per useful apply
    product (hi.hi, hi.lo, lo.hi)."""
    N, Nt = cfg["N"], cfg["Nt"]
    out = {}
    for ps, ss in zip(cfg["patch_sizes"], cfg["strides"]):
        psz = [min(ps[a], N[a]) for a in range(3)]
        ns = [math.ceil((N[a] - psz[a]) / ss[a]) + 1 for a in range(3)]
        Np = ns[0] * ns[1] * ns[2]
        p = psz[0] * psz[1] * psz[2]
        m = min(p, Nt)
        if m <= 1:
            continue
        K = max(p, Nt)
        def add(k, v):
            out[k] = out.get(k, 0) + v
        fg, fa = (4, 3) if split else (1, 1)
        add("gram_tiles", fg * 8 * K * m * (m + 1) // 2 * Np)     # Hermitian X^H X (FP64 or split BF16)
        add("sy2sb", 16 * m ** 3 // 3 * Np)                        # Hermitian reduction to band, FP64
        add("apply", fa * 8 * p * Nt * m * Np)                     # Y = X W (FP32 or split BF16)
    return out


def sense_fft_flops(cfg):
    """FP32 arithmetic of the SENSE FFT passes per POGM iteration, 5 n log2(n) per
    length-n transform plus 6 per coil-map product: the yz passes transform every
    (y, z) plane of every coil and frame along y and z; the x passes transform the
    sampled (ky, kz) lines (1/R of the plane) of every coil and frame along x, forward
    and inverse (the fused residual pass)."""
    N, Nt, Nc, R = cfg["N"], cfg["Nt"], cfg["Nc"], cfg["R"]
    nx, ny, nz = N
    yz_plane = nz * 5 * ny * math.log2(ny) + ny * 5 * nz * math.log2(nz) + 6 * ny * nz
    lines = round(ny * nz / R)
    x_lines = lines * 5 * nx * math.log2(nx)
    return {"fwd_yz": yz_plane * nx * Nc * Nt, "adj_yz": yz_plane * nx * Nc * Nt,
            "x_resid": 2 * x_lines * Nc * Nt}


# FP32 vector peak without packed FP32 instructions (the build disables them):
# 256 CUs x 64 lanes x 2 flops (FMA) x 2.4 GHz; 157.3 with packed FP32
FP32_VALU_PEAK_TFLOPS = 78.6
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
# kernel family -> (bound, unit, peak) of its roofline
BOUNDS = {
    "gram_tiles": ("mfma", "TFLOP/s", FP64_PEAK_TFLOPS),
    "sy2sb": ("mfma", "TFLOP/s", FP64_PEAK_TFLOPS),
    "apply": ("mfma", "TFLOP/s", FP32_PEAK_TFLOPS),
}
BOUNDS_SPLIT = dict(BOUNDS, gram_tiles=("mfma", "TFLOP/s", BF16_PEAK_TFLOPS),
                    apply=("mfma", "TFLOP/s", BF16_PEAK_TFLOPS))


# --------------------------------------------------------------------------- CPU baseline
# Worker state of the CPU baseline's process pool (inherited through fork; the pool
# is created before this process touches the GPU).
_CPU = {}


def _cpu_worker_init():
    from threadpoolctl import threadpool_limits
    threadpool_limits(1)  # one BLAS thread per worker, as Julia's @threads loop over patches


def _cpu_sense_frames(frames):
    """A and A' (residual form) of the given frames, one after another (FFT workers=1)."""
    import oracle as O
    import scipy.fft as sfft
    d = _CPU
    t = time.perf_counter()
    with sfft.set_workers(1):
        for f in frames:
            x = d["img"][..., f:f + 1]
            r = O.A_forward(x, d["mask"][..., f:f + 1], d["smaps"]) - d["y"][:, :, f:f + 1]
            O.A_adjoint(r, d["mask"][..., f:f + 1], d["smaps"])
    return time.perf_counter() - t


def _cpu_patches(job):
    """SVST (gesdd) and, separately timed, the singular values of patches [i0, i1) of a scale."""
    import oracle as O
    k, i0, i1 = job
    d = _CPU
    ps, ss, beta = d["scales"][k]
    N = d["img"].shape[:3]
    Nt = d["img"].shape[3]
    psc, nst, _ = O.patch_geometry(N, ps, ss)
    p = int(np.prod(psc))
    starts = list(O._patch_starts(N, psc, ss, nst))[i0:i1] if p > 1 else None
    img = d["img"]
    t_svst = t_vals = 0.0
    for i in range(i0, i1):
        if p > 1:
            sx, sy, sz = starts[i - i0]
            P = img[sx:sx + psc[0], sy:sy + psc[1], sz:sz + psc[2], :].reshape(p, Nt, order="F")
        else:
            P = img.reshape(-1, Nt, order="F")[i:i + 1]
        t = time.perf_counter()
        O.SVST(np.array(P), beta)
        t1 = time.perf_counter()
        np.linalg.svd(P, compute_uv=False)
        t_svst += t1 - t
        t_vals += time.perf_counter() - t1
    return t_svst, t_vals


def cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, pb, smaps, workers=16, part_s=5.0):
    """The oracle (a restatement of the reference CPU path) timed on samples of one POGM
    iteration on `workers` host processes and scaled to the whole iteration.  Like the
    reference's CPU run (julia -t auto: @threads over patches with one LAPACK gesdd per
    patch, src/recon.jl:217-221; FFTW per frame), patches and frames are spread over the
    processes with one BLAS thread each; the whole-image patch (one gesdd) uses all
    `workers` BLAS threads.  Must run before this process initialises the GPU (the pool
    forks).  Returns (iter/s, cores, sample text, reference-faithful iter/s).

    One iteration is the gradient (1 A + 1 A'), the update with its restart test and
    the prox.  The reference's own loop also evaluates the unused Fcost (1 A and the
    singular values of every patch, src/mirt_mod.jl:161) and the logger (the same
    again, scripts/reconstruct.jl:205): 3 A + 1 A' + the prox + 2 singular-value
    passes, reported as the reference-faithful rate from the same samples."""
    import multiprocessing as mp

    import oracle as O
    from threadpoolctl import threadpool_limits
    N, Nt, Nc = tuple(cfg["N"]), cfg["Nt"], cfg["Nc"]
    V = int(np.prod(N))
    lams = O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    img = pb.x_true
    # k-space of the sampled frames (the SENSE part's input), made here by the oracle
    nfs = 2 * workers
    y = np.zeros((int(pb.mask[..., 0].sum()), Nc, Nt), dtype=np.complex64, order="F")
    y[:, :, :nfs] = O.A_forward(img[..., :nfs], pb.mask[..., :nfs], smaps)
    _CPU.update(img=img, mask=pb.mask, smaps=smaps, y=y,
                scales=[(ps, ss, 0.5 * lams[k]) for k, (ps, ss) in
                        enumerate(zip(cfg["patch_sizes"], cfg["strides"]))])
    parts, notes, vals = {}, [], {}
    ctx = mp.get_context("fork")
    with ctx.Pool(workers, initializer=_cpu_worker_init) as pool:
        pool.map(_cpu_sense_frames, [[0]] * workers)  # warm-up (imports, FFT plans)
        t = time.perf_counter()
        pool.map(_cpu_sense_frames, [[f, f + workers] for f in range(workers)])
        wall = time.perf_counter() - t
        parts["sense"] = wall * Nt / nfs
        notes.append(f"A+A' of {nfs}/{Nt} frames on {workers} processes")
        for k, (ps, ss) in enumerate(zip(cfg["patch_sizes"], cfg["strides"])):
            psc, _, Np = O.patch_geometry(N, ps, ss)
            p = int(np.prod(psc))
            if Np == 1:
                # one global patch: gesdd with all BLAS threads on a row sample (linear in rows)
                rows = max(Nt, p // 8)
                X = img.reshape(V, Nt, order="F")[:rows]
                with threadpool_limits(workers):
                    t = time.perf_counter()
                    O.SVST(X, 0.5 * lams[k])
                    parts[f"svst_scale{k + 1}"] = (time.perf_counter() - t) * p / rows
                    t = time.perf_counter()
                    np.linalg.svd(X, compute_uv=False)
                    vals[k] = (time.perf_counter() - t) * p / rows
                notes.append(f"scale {k + 1}: gesdd of {rows}/{p} rows, {workers} BLAS threads")
                continue
            # per-patch gesdd in the pool: grow the sample until it takes part_s of wall time
            per = 1 if p > 1 else 2000
            done = 0
            while True:
                n = min(per * workers, Np - done)
                jobs = [(k, done + i * n // workers, done + (i + 1) * n // workers) for i in range(workers)]
                t = time.perf_counter()
                res = pool.map(_cpu_patches, jobs)
                wall = time.perf_counter() - t
                done += n
                if wall > part_s / 5 or done >= Np:
                    break
                per *= 4
            ts = sum(r[0] for r in res)
            tv = sum(r[1] for r in res)
            # wall time of the sample split between SVST and the singular values
            parts[f"svst_scale{k + 1}"] = wall * ts / (ts + tv) * Np / n
            vals[k] = wall * tv / (ts + tv) * Np / n
            notes.append(f"scale {k + 1}: {n}/{Np} patches on {workers} processes")
    # POGM update on a tenth of the state (a Julia broadcast: one thread)
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
    notes.append("POGM update (one thread) on 1/10 of the state")
    tot = sum(parts.values())
    t_fwd = parts["sense"] / 2
    sample = (f"oracle (NumPy/SciPy restatement of the reference CPU path) on {workers} processes of "
              f"'{cpu_model()}', one BLAS thread each, LAPACK gesdd per patch as the reference, on "
              "samples of one iteration, scaled: " + "; ".join(notes) +
              "; parts (s/iter): " + ", ".join(f"{k}={v:.2f}" for k, v in parts.items()))
    faithful = tot + 2 * t_fwd + 2 * sum(vals.values())
    sample += (f"; reference-faithful iteration adds 2 A (~{2 * t_fwd:.2f} s, half of A+A') and 2 "
               f"singular-value passes ({2 * sum(vals.values()):.2f} s)")
    _CPU.clear()
    return 1.0 / tot, workers, sample, 1.0 / faithful


# --------------------------------------------------------------------------- counter calibration
def differentiation_run(ctx, niter: int = 10, name: str = "c4") -> dict:
    """The differentiation workload (64x64x40, Nt=120, 12 coils, scales
    [[64,64,40],[8,8,8]]): MSE of the image sum after `niter` unrolled POGM
    iterations to the synthetic truth and its gradient with respect to the per-scale
    lambda_k (checkpointed forward, backward pass on the device); timed separately
    from the headline loop."""
    from mslr_hip import synth
    from mslr_hip.device import DeviceArray
    from mslr_hip.mirt_mod import POGM
    from mslr_hip.reconstruct import lambdas_ong_lustig, normalise_smaps, prepare_flat
    from mslr_hip.sense import SenseOp
    cfg = synth.CONFIGS[name]
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    S = len(cfg["patch_sizes"])
    pb = synth.generate(name)
    smaps = normalise_smaps(pb.smaps)
    A0 = SenseOp(pb.mask, smaps, ctx=ctx)
    y = synth.add_noise(A0.forward(DeviceArray.from_host(pb.x_true, ctx)).to_host(), synth.SEEDS[name])
    A0.close()
    A, _, yd, sf, _ = prepare_flat(pb.smaps, pb.mask, y, ctx)
    lams = lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    x_ref = (pb.x_true / np.float32(sf)).astype(np.complex64)
    xr_d = DeviceArray.from_host(x_ref, ctx)
    times = []
    for rep in range(2):  # the first run sizes the scratch
        ctx.sync()
        t0 = time.perf_counter()
        st = POGM(A, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=lams, L=S * 1.0,
                  niter=niter, checkpoint=True)
        st.init(yd)
        st.run(niter)
        ctx.sync()
        t1 = time.perf_counter()
        xs = st.X.to_host().sum(axis=4)
        r = xs.astype(np.complex128) - x_ref.astype(np.complex128)
        loss = float(np.vdot(r.ravel(), r.ravel()).real) / r.size
        xr_d.upload(np.asfortranarray((2.0 / r.size) * r).astype(np.complex64))
        t2 = time.perf_counter()
        grads = st.lambda_vjp(xr_d)
        ctx.sync()
        t3 = time.perf_counter()
        _, restarts = st.log()
        st.close()
        times.append((t1 - t0, t3 - t2))
    fwd, bwd = times[-1]
    return {"workload": f"{name}: {N[0]}x{N[1]}x{N[2]} Nt={Nt} {cfg['Nc']} coils scales={cfg['patch_sizes']}",
            "niter": niter, "loss": loss, "dloss_dlambda": grads, "restarts": restarts,
            "forward_checkpointed_s": round(fwd, 4), "backward_s": round(bwd, 4),
            "unrolled_iter_per_s_fwd_bwd": round(niter / (fwd + bwd), 3)}


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


# --------------------------------------------------------------------------- accuracy
def load_golden(name):
    """(record, voxel indices, image-sum subset) of the FP64 oracle fixture of this
    workload (tests/golden/recon_<name>.json + _subset.npz), or None."""
    js = ROOT / "tests" / "golden" / f"recon_{name}.json"
    npz = ROOT / "tests" / "golden" / f"recon_{name}_subset.npz"
    if not (js.exists() and npz.exists()):
        return None
    rec = json.loads(js.read_text())
    with np.load(npz, allow_pickle=False) as z:
        return rec, z["voxels"], z["re"] + 1j * z["im"]


def accuracy_vs_oracle(gold, y, A, cfg, lams, L, pb, nrmse_truth, form, ctx):
    """The measured form and the plain form run for the fixture's iteration count on the
    fixture's inputs (this run's k-space divided by the fixture's scale factor), compared
    with the FP64 oracle's run: final-image NRMSE to the truth (FP64 norms), the image sum
    at the fixture's voxels, the restart flags."""
    from mslr_hip.device import DeviceArray
    from mslr_hip.mirt_mod import POGM
    rec, vox, sub = gold
    niter, sf_o = rec["niter"], rec["scale_factor"]
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    V = int(np.prod(N))
    yd = DeviceArray.from_host(np.asfortranarray(y / np.float32(sf_o)), ctx)
    out = {"fixture": f"tests/golden/recon_{rec['config']}.json ({rec['mode']}, {niter} iterations)",
           "oracle_nrmse_vs_truth": rec["nrmse_vs_truth"][-1], "forms": {}}
    for f in dict.fromkeys([form, "plain"]):
        st = POGM(A, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=lams, L=L,
                  niter=niter, form=f)
        st.init(yd)
        st.run(niter)
        _, rs = st.log()
        xs = st.X.to_host().astype(np.complex128).sum(axis=-1)
        st.close()
        nr = nrmse_truth(xs, sf_o)
        got = xs.reshape(V, Nt, order="F")[vox, :]
        out["forms"][f] = {"nrmse_vs_truth": nr, "delta_nrmse_vs_oracle": abs(nr - rec["nrmse_vs_truth"][-1]),
                           "subset_rel_err": float(np.linalg.norm(got - sub) / np.linalg.norm(sub)),
                           "restarts_match": rs == rec["restarts"]}
    return out


# --------------------------------------------------------------------------- main
def metric_label(cfg, name):
    """The headline metric's label for the 90x90x60, Nt=300, 18-coil workload; other
    workloads get the same form of label with their own sizes."""
    N, Nt, Nc = cfg["N"], cfg["Nt"], cfg["Nc"]
    if name == "c1":
        return "POGM iter/s @90x90x60,Nt=300,18ch: mixed+batched vs FP32 plain, ΔNRMSE<1e-4"
    return (f"POGM iter/s @{N[0]}x{N[1]}x{N[2]},Nt={Nt},{Nc}ch,{len(cfg['patch_sizes'])} scales: "
            f"mixed+batched vs FP32 plain, ΔNRMSE<1e-4")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--form", default="mixed", choices=["plain", "batched", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-prof", action="store_true", help="skip the pass with per-kernel event timing")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the batched FP32 and plain runs (rates and result differences)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the built-in SVST launch shapes (no autotuning of the measured form)")
    ap.add_argument("--no-grad", action="store_true",
                    help="skip the differentiation workload (c4 lambda-gradient, timed apart)")
    ap.add_argument("--logger", default="off", choices=["on", "off"],
                    help="run_recon's logger (dc_cost, reg_cost per iterate, on the device) inside "
                         "the timed loop; off by default (one POGM iteration is gradient + update "
                         "+ restart test + prox), the rate with it is reported as with_logging")
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
    cpu = None
    if not args.no_cpu_baseline:
        # first: its process pool forks, which must happen before the GPU is initialised
        t = time.time()
        import oracle as O
        v, cores, sample, vf = cpu_baseline(cfg, pb, O.normalise_smaps(pb.smaps))
        cpu = {"value": round(v, 6), "unit": "iter/s", "cores": cores, "kind": "port",
               "cpu_model": cpu_model(), "sample": sample, "reference_faithful_value": round(vf, 6)}
        log(f"cpu baseline {v:.5f} iter/s ({cores} processes) in {time.time() - t:.1f}s")
    ctx = default_context()
    info = ctx.info()
    log(f"device: {info['name']} total={info['total_bytes'] / 2**30:.0f} GiB")
    t = time.time()
    smaps = normalise_smaps(pb.smaps)
    # k-space: A x_true in FP64 rounded to ComplexF32, plus the seeded noise (the
    # inputs of the FP64 oracle fixture tests/golden/recon_<config>.json)
    A0 = SenseOp(pb.mask, smaps, ctx=ctx)
    xd64 = DeviceArray.from_host(pb.x_true, ctx, dtype=np.complex128)
    yclean = A0.forward(xd64, f64=True).to_host().astype(np.complex64)
    del xd64
    A0.close()
    y = synth.add_noise(yclean, seed)
    del yclean
    A, smaps, yd, sf, R = prepare_flat(pb.smaps, pb.mask, y, ctx)
    log(f"k-space {y.shape} scale factor {sf:.4g} R={R:.2f} in {time.time() - t:.1f}s")
    lams = lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    L = S * 1.0
    W, K = args.warmup, args.steps
    V = int(np.prod(N))

    def make(form, niter, logger=None):
        logger = args.logger == "on" if logger is None else logger
        st = POGM(A, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=lams, L=L,
                  niter=niter, form=form, log_costs=logger)
        st.init(yd)
        ctx.sync()
        return st

    def timed(st, n):
        ctx.sync()
        t0 = time.perf_counter()
        st.run(n)
        ctx.sync()
        return time.perf_counter() - t0

    xt_norm = float(np.linalg.norm(pb.x_true))

    def nrmse_truth(x, scale=None):
        """NRMSE of a reconstruction (scale sum) against the synthetic truth, in the
        truth's units (the k-space was divided by sf), with FP64 norms."""
        d = np.asarray(x, dtype=np.complex128) * (sf if scale is None else scale) - pb.x_true
        return float(np.linalg.norm(d) / xt_norm)

    # launch shapes of the measured form from the autotuner (timed POGM iterations
    # of this workload per candidate; untimed here, like the warmup)
    tune = None
    tuned = {}
    if args.form != "plain" and not args.no_tune:
        from mslr_hip.autotune import autotune_pogm
        t = time.time()
        tune = autotune_pogm(A, yd, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"],
                             lambdas=lams, L=L, form=args.form, log=log)
        tuned = tune["tuning"]
        log(f"autotune: {tuned} ({tune['builtin_ms_per_iter']:.2f} -> {tune['ms_per_iter']:.2f} ms/iter "
            f"over {len(tune['trials'])} trials, {time.time() - t:.1f}s)")

    # the measured form: W untimed iterations, then K timed ones
    st = make(args.form, W + K)
    t = time.time()
    st.run(W)
    ctx.sync()
    log(f"{args.form}: warmup {W} iterations in {time.time() - t:.2f}s")
    elapsed = timed(st, K)
    value = K / elapsed
    _, restarts = st.log()
    x_form = st.X.to_host().sum(axis=-1)
    log(f"{args.form}: timed {K} iterations in {elapsed:.3f}s -> {value:.3f} iter/s; restarts {restarts}")
    st.close()
    # the same form with the logger kept on the device (dc_cost, reg_cost of every
    # iterate, as run_recon collects them)
    sl = make(args.form, W + K, logger=True)
    sl.run(W)
    el_l = timed(sl, K)
    dc, reg = sl.costs()
    sl.close()
    costs = {"dc_cost": [round(v, 6) for v in dc], "reg_cost": [round(v, 6) for v in reg]}
    with_logging = {"value": round(K / el_l, 4), "ms_per_step": round(1e3 * el_l / K, 3)}
    log(f"{args.form} with the logger: {K / el_l:.3f} iter/s; dc_cost {dc[0]:.6g} -> {dc[-1]:.6g}, "
        f"reg_cost {reg[0]:.6g} -> {reg[-1]:.6g}")

    # the other forms on the same problem with the same iteration count (POGM's
    # last iteration differs): their rates and the difference of the results;
    # then a run of the measured form with per-kernel event timing (scales run
    # one after another on one stream, so each kernel's events bracket it alone)
    plain = None
    others = {}
    kern = {}
    nr_form = nrmse_truth(x_form)
    if args.form != "plain" and not args.no_compare:
        # the other forms with the built-in launch shapes
        ctx.set_tuning()
        for other in (["batched"] if args.form == "mixed" else []) + ["plain"]:
            sp = make(other, W + K)
            sp.run(W)
            el_o = timed(sp, K)
            x_o = sp.X.to_host().sum(axis=-1)
            sp.close()
            rd = float(np.linalg.norm(x_form - x_o) / np.linalg.norm(x_o))
            nr_o = nrmse_truth(x_o)
            others[other] = {"value": round(K / el_o, 4), "ms_per_step": round(1e3 * el_o / K, 3),
                             "nrmse_vs_truth": nr_o, "rel_diff_of_measured_form": rd}
            log(f"{other}: {K / el_o:.3f} iter/s; NRMSE vs truth {nr_o:.5e}; "
                f"||x_{args.form} - x_{other}|| / ||x_{other}|| = {rd:.3e}")
        plain = others["plain"]
    log(f"{args.form}: NRMSE vs truth {nr_form:.5e}")
    ctx.set_tuning(**tuned)
    if not args.no_prof:
        sp = make(args.form, W + K)
        sp.run(W)
        ctx.prof_enable(True)
        el_k = timed(sp, K)
        kern = ctx.prof_report()
        ctx.prof_enable(False)
        sp.close()
        for kname, (cnt, ms) in sorted(kern.items(), key=lambda kv: -kv[1][1]):
            log(f"  {kname:18s} launches={cnt:6d} total={ms:9.2f} ms  ({100 * ms / (el_k * 1e3):5.1f}%)")

    roof = None
    roof_valu = {}
    if kern:
        lines_total = int(pb.mask.reshape(N[0], N[1] * N[2], Nt, order="F").any(axis=0).sum())
        budget = 4 << 30
        nchunk = max(math.ceil(Nc * N[0] * lines_total * 8 / budget), math.ceil(V * Nt * 8 / budget))
        by = kernel_bytes_per_iter(cfg, A.K, lines_total, nchunk)
        split = args.form == "mixed"
        fl = kernel_flops_per_iter(cfg, split)
        bounds = BOUNDS_SPLIT if split else BOUNDS
        # per-scale tags ("sy2sb@s2") summed per kernel family
        fam = {}
        for kname, (cnt, ms) in kern.items():
            base = kname.split("@")[0]
            c0, m0 = fam.get(base, (0, 0.0))
            fam[base] = (c0 + cnt, m0 + ms)
        total_ms = sum(ms for _, ms in fam.values())
        dom = max(fam.items(), key=lambda kv: kv[1][1])[0]
        cnt, ms = fam[dom]
        per_launch_ms = ms / cnt
        if dom in fl:
            bound, unit, peak = bounds[dom]
            work = fl[dom] * K / cnt
            ach = work / (per_launch_ms / 1e3) / 1e12
            roof = {"bound": bound, "achieved": round(ach, 2), "peak": peak, "unit": unit,
                    "frac": round(ach / peak, 4), "traffic": TRAFFIC_BYTES.get(dom),
                    "kernel": dom, "flops_per_launch": int(work)}
        elif dom in by:
            work = by[dom] * K / cnt
            ach = work / (per_launch_ms / 1e3) / 1e9
            roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": TRAFFIC_BYTES.get(dom),
                    "kernel": dom, "bytes_per_launch": int(work)}
        if roof is not None:
            roof.update({"avg_launch_ms": round(per_launch_ms, 4), "launches_per_iter": cnt / K,
                         "share_of_kernel_time": round(ms / total_ms, 3)})
        # the yz FFT passes against the FP32 vector roofline as well (they are bound by
        # their arithmetic, not by HBM)
        ff = sense_fft_flops(cfg)
        for kf in ("fwd_yz", "adj_yz", "x_resid"):
            if kf in fam and kf in ff:
                c_, m_ = fam[kf]
                t_ = m_ / c_ / 1e3
                roof_valu[kf] = {"bound": "valu", "flops_per_launch": int(ff[kf]),
                                 "achieved": round(ff[kf] / t_ / 1e12, 2), "unit": "TFLOP/s",
                                 "peak": FP32_VALU_PEAK_TFLOPS,
                                 "frac": round(ff[kf] / t_ / 1e12 / FP32_VALU_PEAK_TFLOPS, 4),
                                 "avg_launch_ms": round(1e3 * t_, 4)}

    accuracy = None
    gold = load_golden(name)
    if gold is not None and not args.no_compare:
        accuracy = accuracy_vs_oracle(gold, y, A, cfg, lams, L, pb, nrmse_truth, args.form, ctx)
        for f, r in accuracy["forms"].items():
            log(f"accuracy ({gold[0]['niter']} iterations) {f}: NRMSE {r['nrmse_vs_truth']:.7f} vs oracle "
                f"{accuracy['oracle_nrmse_vs_truth']:.7f} (|d| {r['delta_nrmse_vs_oracle']:.2e}), image "
                f"subset {r['subset_rel_err']:.2e}, restarts match {r['restarts_match']}")

    grad = None
    if not args.no_grad:
        t = time.time()
        ctx.set_tuning()
        grad = differentiation_run(ctx)
        log(f"differentiation (c4, {grad['niter']} iterations): forward {grad['forward_checkpointed_s']:.3f}s "
            f"backward {grad['backward_s']:.3f}s dloss/dlambda {grad['dloss_dlambda']} ({time.time() - t:.1f}s)")

    line = {
        "metric": metric_label(cfg, name),
        "value": round(value, 4),
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * elapsed / K, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64/f64" + ("/bf16x2" if args.form == "mixed" else ""),
        "data": "synthetic",
        "config": {"workload": f"{name}: {N[0]}x{N[1]}x{N[2]} Nt={Nt} {Nc} coils R={cfg['R']} scales={cfg['patch_sizes']}",
                   "form": args.form, "parallelism": "single-gpu", "restarts_timed": sum(restarts[W:]),
                   "logger": args.logger, "svst_tuning": tuned or "built-in"},
        "nrmse_vs_truth": nr_form,
        "delta_nrmse": (accuracy["forms"][args.form]["delta_nrmse_vs_oracle"] if accuracy else None),
        "delta_nrmse_vs_plain": abs(nr_form - plain["nrmse_vs_truth"]) if plain else None,
        "accuracy_vs_fp64_oracle": accuracy,
        "plain": plain,
        "batched_fp32": others.get("batched"),
        "speedup_vs_plain": round(value / plain["value"], 3) if plain else None,
        "roofline": roof,
        "roofline_valu": roof_valu or None,
        "cpu_baseline": cpu,
        "kernels_ms_per_iter": {k: round(v[1] / K, 3) for k, v in kern.items()},
        "with_logging": with_logging,
        "costs": costs,
        "differentiation": grad,
        "autotune": None if tune is None else {
            "tuning": tune["tuning"], "ms_per_iter": tune["ms_per_iter"],
            "builtin_ms_per_iter": tune["builtin_ms_per_iter"],
            "trials": [{"tuning": t, "ms_per_iter": ms, "iterate_rel_diff": err, "kept": ok}
                       for t, ms, err, ok in tune["trials"]]},
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
