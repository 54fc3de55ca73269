"""Autotuner for the launch shapes of the batched SVST in a POGM workload.

The shapes (mslr_ctx_set_tuning) are the Gram contraction chunk (rows per
workgroup: more chunks fill the device for tall patches but add partial Grams
to sum), the eigenvector lanes (16 per eigenvalue hide the bisection latency of
small batches, 4 suit batches that fill the device), the apply form (low-rank
X U_r products or the dense W) and the frame groups of the gradient (more groups
overlap more of the memory-bound kernels with the FFT passes on two streams, at
smaller launches), the thread counts of the compiled y-z FFT passes, the update's
grid and the wave schedule of the global-band tridiagonalisation (Gram matrices
above 300).  Which is fastest depends on the patch shapes,
the patch count and what the other scales run beside them, so the tuner times
whole POGM iterations of the workload itself: coordinate descent from the
built-in choice, one shape at a time, each candidate timed over `iters`
iterations after `warmup`, a candidate kept only when it is faster by more than
`min_gain` and its iterate (the middle frames of every scale) stays within
`rtol` of the built-in choice's.  The winner is set on the context and cached
per workload in a JSON file.
"""
from __future__ import annotations

import json
import os
import time
from pathlib import Path

import numpy as np

# candidate values per shape, the built-in choice (0) first
SPACE: dict[str, tuple[int, ...]] = {
    "gram_chunk": (0, 2048, 8192, 16384),
    "eig_lanes_limit": (0, -1, 1 << 20),
    "svst_apply": (0, 1),
    "sense_chunks": (0, 2, 4, 8),
    "yz_threads_fwd": (0, 512, 1024),
    "yz_threads_adj": (0, 512, 1024),
    "update_lines": (0, 32, 16),
    "sb2st_schedule": (0, 1, 2, 3),
    "apply_fp32": (0, 8),
}
BUILTIN = {k: 0 for k in SPACE}
# plane sizes (Ny, Nz) with compiled y-z FFT passes of several thread counts
YZ_COMPILED = {(90, 60)}
SBR_MAXM = 300  # Gram matrices above this keep their band in global memory


def relevant(name: str, N, Nt: int, patch_sizes, form: str | None = None) -> bool:
    """Whether a shape can change anything for this workload (the others are skipped):
    the yz thread counts only for the compiled planes above 2048 points, the global-band
    schedule only when some patch has a Gram matrix above 300, the reconstruction
    products' arithmetic only in the split-BF16 (mixed) form."""
    if name == "apply_fp32":
        return form in (None, "mixed")
    if name in ("yz_threads_fwd", "yz_threads_adj"):
        return (int(N[1]), int(N[2])) in YZ_COMPILED
    if name == "sb2st_schedule":
        return any(min(int(np.prod([min(int(a), int(n)) for a, n in zip(ps, N)])), int(Nt)) > SBR_MAXM
                   for ps in patch_sizes)
    return True


def workload_key(A, patch_sizes, strides, form: str) -> str:
    """Cache key: the device, the problem shape, the scales and the form."""
    ps = ";".join("x".join(str(int(v)) for v in p) for p in patch_sizes)
    ss = ";".join("x".join(str(int(v)) for v in s) for s in strides)
    dev = A.ctx.info()["name"] if hasattr(A.ctx, "info") else "device"
    n = "x".join(str(int(v)) for v in A.N)
    return f"{dev}|N={n}|Nt={A.Nt}|Nc={A.Nc}|K={A.K}|ps={ps}|ss={ss}|form={form}"


def descent_order() -> list[tuple[str, int]]:
    """(shape, value) pairs the descent tries after the built-in choice."""
    return [(k, v) for k, vals in SPACE.items() for v in vals if v != BUILTIN[k]]


def load_cache(path: str | os.PathLike | None) -> dict:
    if path is None or not Path(path).exists():
        return {}
    return json.loads(Path(path).read_text())


def store_cache(path: str | os.PathLike | None, key: str, entry: dict) -> None:
    if path is None:
        return
    cache = load_cache(path)
    cache[key] = entry
    Path(path).write_text(json.dumps(cache, indent=1, sort_keys=True) + "\n")


def iterate_sample(st, nframes: int = 2) -> np.ndarray:
    """The middle `nframes` frames of every scale of the POGM iterate (contiguous
    slabs of the (Nx, Ny, Nz, Nt, S) state): the tuner's check of a candidate's
    result without downloading the whole state."""
    N, Nt, S = st.A.N, st.A.Nt, st.S
    V = int(np.prod(N))
    nf = min(nframes, Nt)
    t0 = (Nt - nf) // 2
    X = st.X
    return np.concatenate([X.offset_view((s * Nt + t0) * V, (V * nf,)).to_host() for s in range(S)])


def _rel(a: np.ndarray, b: np.ndarray) -> float:
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return float(d / n) if n > 0 else float(d)


def autotune_pogm(A, y, *, patch_sizes, strides, lambdas, L: float, form: str = "mixed",
                  warmup: int = 2, iters: int = 4, rtol: float = 1e-4, min_gain: float = 0.01,
                  cache: str | os.PathLike | None = None, log=None) -> dict:
    """Tunes the context of `A` for POGM on (A, y) and leaves the winner set on it.

    Returns {"tuning", "ms_per_iter", "builtin_ms_per_iter", "trials", "cached"};
    each trial is (tuning, ms per iteration, relative difference of the iterate
    from the built-in choice's, accepted)."""
    from .mirt_mod import POGM
    ctx = A.ctx
    key = workload_key(A, patch_sizes, strides, form)
    hit = load_cache(cache).get(key)
    if hit is not None:
        ctx.set_tuning(**hit["tuning"])
        return dict(hit, cached=True)
    niter = warmup + iters

    def trial(t: dict):
        ctx.set_tuning(**t)
        st = POGM(A, patch_sizes=patch_sizes, strides=strides, lambdas=lambdas, L=L,
                  niter=niter, form=form)
        try:
            st.init(y)
            st.run(warmup)
            ctx.sync()
            t0 = time.perf_counter()
            st.run(iters)
            ctx.sync()
            ms = 1e3 * (time.perf_counter() - t0) / iters
            X = iterate_sample(st)
        finally:
            st.close()
        return ms, X

    best = dict(BUILTIN)
    best_ms, X_ref = trial(best)
    base_ms = best_ms
    trials = [(dict(best), round(best_ms, 4), 0.0, True)]
    if log:
        log(f"tune {best}: {best_ms:.3f} ms/iter")
    for name, value in descent_order():
        if not relevant(name, A.N, A.Nt, patch_sizes, form):
            continue
        cand = dict(best, **{name: value})
        ms, X = trial(cand)
        err = _rel(X, X_ref)
        ok = err <= rtol and ms < best_ms * (1.0 - min_gain)
        trials.append((cand, round(ms, 4), err, ok))
        if log:
            log(f"tune {cand}: {ms:.3f} ms/iter, iterate {err:.1e} from built-in"
                f"{' (kept)' if ok else ''}")
        if ok:
            best, best_ms = cand, ms
    ctx.set_tuning(**best)
    entry = {"tuning": best, "ms_per_iter": round(best_ms, 4),
             "builtin_ms_per_iter": round(base_ms, 4), "trials": trials}
    store_cache(cache, key, entry)
    return dict(entry, cached=False)
