"""The mixed-precision form: split-BF16 Gram and reconstruction products on the
matrix cores (FP64 eigensolve), against LAPACK SVD and the FP64 oracle.

A split operand v = hi + lo (two BF16) is off by at most 2^-16 |v|, so the
bounds here sit between the FP32 path's (2e-5) and that per-element error
scaled by the patch norm."""
import numpy as np
import pytest

import oracle as O
from conftest import crandn, rel_err
from problems import prepared
from test_gpu_svst import _beta_between, lowrank

pytestmark = pytest.mark.gpu


@pytest.fixture
def split(ctx):
    ctx.set_products("split_bf16")
    yield ctx
    ctx.set_products("fp32")


SHAPES = [(40, 12, 3), (12, 40, 3), (1000, 30, 4), (30, 200, 2), (216, 387, 5), (300, 300, 6),
          (1000, 300, 20), (5000, 64, 3), (64, 64, 1), (2, 3, 1), (130, 50, 20), (2000, 387, 40),
          (700, 387, 300)]


@pytest.mark.parametrize("p,n,r", SHAPES)
def test_split_svst_matches_lapack(split, rng, p, n, r):
    from mslr_hip import recon
    X = lowrank(rng, p, n, r)
    beta = _beta_between(X, r)
    Y = recon.SVST(X, beta).to_host()
    Yo = O.SVST(X.astype(np.complex128), np.float32(beta))
    assert rel_err(Y, Yo) < 1e-4


def test_split_long_contraction(split, rng):
    """p = 40000 rows: ten contraction chunks of 4096 rows, exact FP32 sums per
    32-row stage, FP64 over stages and chunks."""
    from mslr_hip import recon
    X = lowrank(rng, 40000, 48, 4)
    beta = _beta_between(X, 4)
    Y = recon.SVST(X, beta).to_host()
    Yo = O.SVST(X.astype(np.complex128), np.float32(beta))
    assert rel_err(Y, Yo) < 1e-4


@pytest.mark.parametrize("p,n,Np", [(48, 10000, 1), (216, 387, 5), (100, 130, 3)])
def test_split_left_patches(split, rng, p, n, Np):
    """Patches with fewer voxels than frames: the Gram X X^H contracts over time,
    split from time-contiguous words (k_split_cols) and conjugated; 10000 frames are
    three contraction chunks summed by the reduce."""
    from mslr_hip import recon
    P = np.asfortranarray(np.stack([lowrank(rng, p, n, 2 + i) for i in range(Np)], axis=2))
    betas = np.array([_beta_between(P[:, :, i], 2 + i) for i in range(Np)])
    Y = recon.svst_batched(P, betas).to_host()
    for i in range(Np):
        Yo = O.SVST(P[:, :, i].astype(np.complex128), np.float32(betas[i]))
        assert rel_err(Y[:, :, i], Yo) < 1e-4, i


def test_split_batch_per_patch_beta(split, rng):
    from mslr_hip import recon
    p, n, Np = 1000, 60, 9
    P = np.asfortranarray(np.stack([lowrank(rng, p, n, 1 + i % 4) for i in range(Np)], axis=2))
    betas = np.array([_beta_between(P[:, :, i], 1 + i % 4) for i in range(Np)])
    Y = recon.svst_batched(P, betas).to_host()
    for i in range(Np):
        Yo = O.SVST(P[:, :, i].astype(np.complex128), np.float32(betas[i]))
        assert rel_err(Y[:, :, i], Yo) < 1e-4


def test_split_and_fp32_products_agree(ctx, rng):
    """The same batch through both product forms."""
    from mslr_hip import recon
    p, n, Np = 1000, 300, 4
    P = np.asfortranarray(np.stack([lowrank(rng, p, n, 10 + 5 * i) for i in range(Np)], axis=2))
    betas = np.array([_beta_between(P[:, :, i], 10 + 5 * i) for i in range(Np)])
    Y32 = recon.svst_batched(P, betas).to_host()
    ctx.set_products("split_bf16")
    try:
        Ys = recon.svst_batched(P, betas).to_host()
    finally:
        ctx.set_products("fp32")
    assert rel_err(Ys, Y32) < 1e-4


@pytest.mark.parametrize("p,n", [(4000, 60), (40, 500)])
def test_split_high_dynamic_range(split, rng, p, n):
    """A dominant rank-1 term (s1/s_noise ~ 1e4, as the production global patch) with
    every noise singular value kept.  The split Gram sums without rounding, so the
    noise eigenvalues (1e-8 of the dominant one) only see the 2^-17 slice residual
    of the data, ~1e-2 of the noise per entry here; the apply adds the same order."""
    from mslr_hip import recon
    u = crandn(rng, (p, 1), np.complex128)
    v = crandn(rng, (n, 1), np.complex128)
    noise = crandn(rng, (p, n), np.complex128)
    X = (1e4 * (u @ v.conj().T) / np.sqrt(p * n) * np.sqrt(p + n) + noise).astype(np.complex64)
    s = np.linalg.svd(X.astype(np.complex128), compute_uv=False)
    beta = 0.5 * s[-1]
    Y = recon.SVST(X, beta).to_host()
    Yo = O.SVST(X.astype(np.complex128), np.float32(beta))
    e = rel_err(Y, Yo)
    U, S, Vh = np.linalg.svd(X.astype(np.complex128), full_matrices=False)
    Pn = np.eye(p) - np.outer(U[:, 0], U[:, 0].conj())
    en = rel_err(Pn @ Y, Pn @ Yo)
    print(f"split high dynamic range {p}x{n}: total {e:.3e}, noise part {en:.3e}")
    assert e < 1e-4
    assert en < 0.1


def test_split_zero_and_identity(split, rng):
    from mslr_hip import recon
    X = lowrank(rng, 300, 40, 3)
    assert np.all(recon.SVST(np.zeros_like(X), 1.0).to_host() == 0)
    assert rel_err(recon.SVST(X, 0.0).to_host(), X) < 1e-4
    assert np.all(recon.SVST(X, 1e6).to_host() == 0)


def test_products_setting_rejects_unknown(ctx):
    from mslr_hip import _lib
    with pytest.raises(ValueError):
        _lib.check(ctx._lib.mslr_ctx_set_products(ctx.handle, 7))


def _gpu_run(name, niter, form):
    from mslr_hip.mirt_mod import pogm_mslr
    from mslr_hip.sense import SenseOp
    cfg, smaps, ksp, omega, sf = prepared(name)
    S = len(cfg["patch_sizes"])
    A = SenseOp(omega, smaps)
    lams = O.lambdas_ong_lustig(tuple(cfg["N"]), cfg["Nt"], cfg["patch_sizes"])
    X, restarts, _ = pogm_mslr(A, ksp, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"],
                               lambdas=lams, L=S * 1.0, niter=niter, form=form)
    return X.to_host(), restarts


@pytest.mark.parametrize("name,niter", [("small", 6), ("c0", 4)])
def test_mixed_form_vs_fp64_oracle(name, niter):
    """Final image of the mixed form within 1e-4 of the FP64 oracle, same restarts."""
    cfg, smaps, ksp, omega, sf = prepared(name)
    X, restarts = _gpu_run(name, niter, "mixed")
    Xo, ro, _ = O.run_recon_arrays(ksp, omega, smaps, N=tuple(cfg["N"]), Nt=cfg["Nt"],
                                   PATCH_SIZES=cfg["patch_sizes"], STRIDES=cfg["strides"],
                                   NITERS=niter, sigma1A=1.0, mode="f64")
    assert restarts == ro
    assert rel_err(X.sum(axis=4), Xo.sum(axis=4)) < 1e-4


def test_mixed_form_vs_plain():
    X1, r1 = _gpu_run("small", 6, "plain")
    X2, r2 = _gpu_run("small", 6, "mixed")
    assert r1 == r2
    assert rel_err(X2.sum(axis=4), X1.sum(axis=4)) < 1e-4


@pytest.mark.parametrize("N,ps,ss,Nt", [((12, 12, 8), (4, 4, 4), (4, 4, 4), 20),      # exact tiling
                                        ((12, 12, 10), (12, 12, 10), (12, 12, 10), 30),  # one patch
                                        ((12, 12, 10), (6, 6, 6), (3, 3, 3), 20),      # overlapping
                                        ((12, 12, 8), (4, 4, 4), (4, 4, 4), 80)])      # p < Nt
def test_split_patch_svst(split, rng, N, ps, ss, Nt):
    """patchSVST in split BF16 for the layouts whose products write the image
    directly (an exact tiling's patches scattered to their voxels, the single patch)
    and for those that go through patches2img, against the FP64 oracle."""
    from mslr_hip import recon
    V = int(np.prod(N))
    L = crandn(rng, (V, 2)) @ crandn(rng, (2, Nt))
    img = np.asfortranarray((4.0 * L.reshape(*N, Nt, order="F") + crandn(rng, (*N, Nt)))
                            .astype(np.complex64))
    beta = 9.0
    out = recon.patchSVST(img, beta, ps, ss).to_host()
    want = O.patchSVST(img.astype(np.complex128), np.float32(beta), ps, ss, dtype=np.complex128)
    assert rel_err(out, want) < 1e-4
    # the input is left alone
    img2 = recon.patchSVST(img, 0.0, ps, ss).to_host()
    want0 = O.patchSVST(img.astype(np.complex128), np.float32(0.0), ps, ss, dtype=np.complex128)
    assert rel_err(img2, want0) < 1e-4


def test_split_apply_forms_at_the_boundary(split, rng):
    from mslr_hip import recon
    from test_gpu_svst import APPLY_RANKS, mixed_rank_batch
    P, betas = mixed_rank_batch(rng, APPLY_RANKS)
    Y = recon.svst_batched(P, betas).to_host()
    for i, k in enumerate(APPLY_RANKS):
        Yo = O.SVST(P[:, :, i].astype(np.complex128), np.float32(betas[i]))
        assert rel_err(Y[:, :, i], Yo) < 1e-4, (i, k)
