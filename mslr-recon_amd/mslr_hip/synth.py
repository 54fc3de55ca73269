"""Deterministic synthetic MSLR problems (fMRI-like 3D+time images, coil maps, masks).

All randomness comes from a counter-based SplitMix64 stream (key = seed and a
per-quantity stream id; counter = element index), so the same inputs are
produced on every machine and NumPy version.

Truth image = global low-rank term (smooth spatial maps x temporal sinusoids at
0.01-0.1 cycles/frame plus linear drift) + low-rank terms in a subset of the
aligned B^3 blocks + sparse voxel activations (1% of voxels).  Coil maps are
Gaussian bumps centred on a sphere around the volume with random phase,
normalised to unit root-sum-of-squares.  Each frame has its own random
Cartesian ky-kz mask (kx fully sampled) with variable density, a fully sampled
6x6 centre and the same line count in every frame.
"""
from __future__ import annotations

import concurrent.futures as cf
import math
from dataclasses import dataclass, field

import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = z.copy()
    z ^= z >> np.uint64(30)
    z *= _M1
    z ^= z >> np.uint64(27)
    z *= _M2
    z ^= z >> np.uint64(31)
    return z


def _key(seed: int, stream: int) -> np.uint64:
    k = np.array([(int(seed) * 0x100000001B3 + int(stream) * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)
    return _mix(k + _GOLD)[0]


def uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n doubles in [0, 1): SplitMix64(key + i) >> 11 * 2^-53."""
    with np.errstate(over="ignore"):
        ctr = np.arange(offset, offset + n, dtype=np.uint64) * _GOLD + _key(seed, stream)
        v = _mix(ctr)
    return (v >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def cnormal(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n complex normals with E|z|^2 = 1 (Box-Muller on pairs of uniforms)."""
    u = uniform(seed, stream, 2 * n, 2 * offset)
    r = np.sqrt(-np.log1p(-u[0::2]))  # sqrt(-2 ln(1-u)) / sqrt(2)
    th = 2.0 * math.pi * u[1::2]
    return r * np.cos(th) + 1j * (r * np.sin(th))


def cnormal_parallel(seed: int, stream: int, n: int, workers: int = 8, chunk: int = 1 << 22) -> np.ndarray:
    out = np.empty(n, dtype=np.complex64)

    def job(o):
        m = min(chunk, n - o)
        out[o:o + m] = cnormal(seed, stream, m, o)

    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(job, range(0, n, chunk)))
    return out


# --------------------------------------------------------------------------- configs

CONFIGS = {
    # 32x32x24, Nt=40, 8 coils, 6x; rank-3 global, rank-1 in 8^3 blocks
    "c0": dict(N=(32, 32, 24), Nt=40, Nc=8, R=6, rank=3, block=8, block_rank=1, nblock=5,
               patch_sizes=[[32, 32, 24], [8, 8, 8]], strides=[[32, 32, 24], [8, 8, 8]], niter=20),
    # 90x90x60, Nt=300, 18 coils, 8x; rank-5 global, rank-2 in 10^3 blocks
    "c1": dict(N=(90, 90, 60), Nt=300, Nc=18, R=8, rank=5, block=10, block_rank=2, nblock=48,
               patch_sizes=[[90, 90, 60], [10, 10, 10], [1, 1, 1]],
               strides=[[90, 90, 60], [10, 10, 10], [1, 1, 1]], niter=50),
    # 90x90x60, Nt=387, four half-overlapping scales
    "c2": dict(N=(90, 90, 60), Nt=387, Nc=18, R=8, rank=5, block=10, block_rank=2, nblock=48,
               patch_sizes=[[90, 90, 60], [30, 30, 20], [10, 10, 10], [6, 6, 6]],
               strides=[[90, 90, 60], [15, 15, 10], [5, 5, 5], [3, 3, 3]], niter=50),
    # 64x64x40, Nt=120, 12 coils, 6x (427 lines per frame)
    "c4": dict(N=(64, 64, 40), Nt=120, Nc=12, R=6, rank=3, block=8, block_rank=1, nblock=32,
               patch_sizes=[[64, 64, 40], [8, 8, 8]], strides=[[64, 64, 40], [8, 8, 8]], niter=10),
}
CONFIGS["c3"] = dict(CONFIGS["c2"])
SEEDS = {"c0": 0, "c1": 1, "c2": 2, "c3": 3, "c4": 4}


def small_config(N=(12, 12, 10), Nt=10, Nc=3, R=2, rank=2, block=4, block_rank=1, nblock=3,
                 patch_sizes=([12, 12, 10], [4, 4, 4]), strides=([12, 12, 10], [4, 4, 4]), niter=5):
    return dict(N=tuple(N), Nt=Nt, Nc=Nc, R=R, rank=rank, block=block, block_rank=block_rank,
                nblock=nblock, patch_sizes=[list(p) for p in patch_sizes],
                strides=[list(s) for s in strides], niter=niter)


@dataclass
class Problem:
    cfg: dict
    seed: int
    smaps: np.ndarray          # (Nx,Ny,Nz,Nc) complex64, unit RSS
    mask: np.ndarray           # (Nx,Ny,Nz,Nt) bool
    x_true: np.ndarray | None = None  # (Nx,Ny,Nz,Nt) complex64
    lines: int = 0
    extra: dict = field(default_factory=dict)


def _gauss1(n, c, w):
    x = np.arange(n, dtype=np.float64)
    return np.exp(-0.5 * ((x - c) / w) ** 2)


def smooth_maps(N, nmaps, seed, stream, nblob=4):
    """(V, nmaps) complex maps, each a sum of separable Gaussian blobs (F order)."""
    V = int(np.prod(N))
    out = np.zeros((V, nmaps), dtype=np.complex128)
    u = uniform(seed, stream, nmaps * nblob * 6)
    amp = cnormal(seed, stream + 1, nmaps * nblob)
    j = 0
    for m in range(nmaps):
        acc = np.zeros(N, dtype=np.complex128, order="F")
        for b in range(nblob):
            c = [u[j + a] * N[a] for a in range(3)]
            w = [(0.15 + 0.2 * u[j + 3 + a]) * N[a] for a in range(3)]
            j += 6
            g = np.einsum("i,j,k->ijk", _gauss1(N[0], c[0], w[0]), _gauss1(N[1], c[1], w[1]),
                          _gauss1(N[2], c[2], w[2]))
            acc += amp[m * nblob + b] * g
        out[:, m] = acc.ravel(order="F")
    return out


def temporal(Nt, nterm, seed, stream):
    u = uniform(seed, stream, nterm * 3)
    t = np.arange(Nt, dtype=np.float64)
    T = np.zeros((nterm, Nt))
    for r in range(nterm):
        f = 0.01 + 0.09 * u[3 * r]
        ph = 2 * math.pi * u[3 * r + 1]
        drift = (u[3 * r + 2] - 0.5)
        T[r] = np.sin(2 * math.pi * f * t + ph) + drift * (t / max(Nt - 1, 1) - 0.5)
    return T


def coil_maps(N, Nc, seed):
    """Gaussian bumps on a sphere around the volume, random phase, unit RSS."""
    u = uniform(seed, 50, Nc)
    maps = np.zeros((*N, Nc), dtype=np.complex128, order="F")
    ga = math.pi * (3.0 - math.sqrt(5.0))
    for c in range(Nc):
        zc = 1.0 - 2.0 * (c + 0.5) / Nc
        rc = math.sqrt(max(0.0, 1.0 - zc * zc))
        d = (rc * math.cos(ga * c), rc * math.sin(ga * c), zc)
        cen = [N[a] / 2.0 + 0.75 * N[a] * d[a] for a in range(3)]
        wid = [0.6 * N[a] for a in range(3)]
        g = np.einsum("i,j,k->ijk", _gauss1(N[0], cen[0], wid[0]), _gauss1(N[1], cen[1], wid[1]),
                      _gauss1(N[2], cen[2], wid[2]))
        maps[..., c] = g * np.exp(2j * math.pi * u[c])
    rss = np.sqrt(np.sum(np.abs(maps) ** 2, axis=3, keepdims=True))
    return np.asfortranarray((maps / rss).astype(np.complex64))


def masks(N, Nt, R, seed):
    """(Nx,Ny,Nz,Nt) bool: per-frame variable-density ky-kz lines, 6x6 centre, equal counts."""
    Ny, Nz = N[1], N[2]
    L = int(round(Ny * Nz / R))
    ky = np.arange(Ny) - Ny // 2
    kz = np.arange(Nz) - Nz // 2
    KY, KZ = np.meshgrid(ky, kz, indexing="ij")
    r2 = (KY / (Ny / 2.0)) ** 2 + (KZ / (Nz / 2.0)) ** 2
    w = 1.0 / (1.0 + r2 / 0.05)
    centre = (KY >= -3) & (KY < 3) & (KZ >= -3) & (KZ < 3)
    nc = int(centre.sum())
    if L < nc:
        raise ValueError("acceleration too high for a fully sampled 6x6 centre")
    out = np.zeros((*N, Nt), dtype=bool, order="F")
    flat_w = w.ravel(order="F")
    flat_c = centre.ravel(order="F")
    cand = np.flatnonzero(~flat_c)
    for t in range(Nt):
        u = uniform(seed, 1000 + t, cand.size)
        keys = -np.log1p(-u) / flat_w[cand]  # Efraimidis-Spirakis weighted sampling
        pick = cand[np.argsort(keys, kind="stable")[: L - nc]]
        sel = flat_c.copy()
        sel[pick] = True
        out[:, :, :, t] = sel.reshape(Ny, Nz, order="F")[None, :, :]
    return out, L


def truth_image(cfg, seed):
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    V = int(np.prod(N))
    rank = cfg["rank"]
    M = smooth_maps(N, rank, seed, 10)
    M[:, 0] = np.abs(M[:, 0]) * np.exp(1j * np.angle(smooth_maps(N, 1, seed, 12)[:, 0]))
    M[:, 0] /= np.max(np.abs(M[:, 0]))
    M[:, 1:] *= 0.1 / max(np.max(np.abs(M[:, 1:])), 1e-30)
    T = np.ones((rank, Nt))
    T[1:] = temporal(Nt, rank - 1, seed, 20)
    t = np.arange(Nt) / max(Nt - 1, 1) - 0.5
    T[0] = 1.0 + 0.05 * (uniform(seed, 21, 1)[0] - 0.5) * t
    X = (M @ T).astype(np.complex64)  # (V, Nt)
    # low-rank terms in a subset of aligned blocks
    B = cfg["block"]
    nb = [N[a] // B for a in range(3)]
    nblocks_total = nb[0] * nb[1] * nb[2]
    ub = uniform(seed, 30, nblocks_total)
    chosen = np.argsort(ub, kind="stable")[: min(cfg["nblock"], nblocks_total)]
    br = cfg["block_rank"]
    for j, bidx in enumerate(sorted(chosen.tolist())):
        bx, by, bz = bidx % nb[0], (bidx // nb[0]) % nb[1], bidx // (nb[0] * nb[1])
        Ub = cnormal(seed, 100000 + j, B ** 3 * br).reshape(B ** 3, br)
        Tb = temporal(Nt, br, seed, 200000 + j)
        blk = (0.05 * (Ub @ Tb)).astype(np.complex64)
        xs = np.arange(bx * B, bx * B + B)
        ys = np.arange(by * B, by * B + B)
        zs = np.arange(bz * B, bz * B + B)
        vi = (xs[:, None, None] + N[0] * (ys[None, :, None] + N[1] * zs[None, None, :])).ravel(order="F")
        X[vi, :] += blk
    # sparse activations: 1% of voxels, boxcar time courses
    nsp = max(1, V // 100)
    us = uniform(seed, 40, V)
    vox = np.sort(np.argsort(us, kind="stable")[:nsp])
    up = uniform(seed, 41, 2 * nsp)
    tt = np.arange(Nt)
    period = 10 + (up[0::2] * 30).astype(int)
    phase = (up[1::2] * 40).astype(int)
    act = (((tt[None, :] + phase[:, None]) // period[:, None]) % 2).astype(np.float32)
    X[vox, :] += (0.05 * act).astype(np.complex64)
    return np.asfortranarray(X.reshape(*N, Nt, order="F"))


def generate(cfg: dict | str, seed: int | None = None, with_truth: bool = True) -> Problem:
    if isinstance(cfg, str):
        name = cfg
        cfg = CONFIGS[name]
        seed = SEEDS[name] if seed is None else seed
    seed = 0 if seed is None else seed
    N, Nt, Nc = tuple(cfg["N"]), cfg["Nt"], cfg["Nc"]
    smaps = coil_maps(N, Nc, seed)
    mask, L = masks(N, Nt, cfg["R"], seed)
    x = truth_image(cfg, seed) if with_truth else None
    return Problem(cfg=dict(cfg), seed=seed, smaps=smaps, mask=mask, x_true=x, lines=L)


def add_noise(y: np.ndarray, seed: int, rel: float = 1e-3, workers: int = 8) -> np.ndarray:
    """y + complex Gaussian noise with std rel * RMS(y) (in place on a complex64 copy)."""
    y = np.asfortranarray(y, dtype=np.complex64)
    flat = y.reshape(-1, order="F")
    rms = math.sqrt(float(np.mean(np.abs(flat.astype(np.complex128)) ** 2)))
    n = cnormal_parallel(seed, 60, flat.size, workers=workers)
    flat += np.complex64(rel * rms) * n
    return y
