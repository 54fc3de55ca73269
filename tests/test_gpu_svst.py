"""Batched SVST on the GPU vs LAPACK SVD (oracle) and closed forms (src/recon.jl:183-269)."""
import numpy as np
import pytest

import oracle as O
from conftest import crandn, rel_err

pytestmark = pytest.mark.gpu


def lowrank(rng, p, n, r, noise=1e-3, scale=10.0):
    U = crandn(rng, (p, r), np.complex128)
    V = crandn(rng, (n, r), np.complex128)
    s = scale * (1.0 + np.arange(r)[::-1])
    X = (U * s) @ V.conj().T / np.sqrt(max(p, n))
    X = X + noise * crandn(rng, (p, n), np.complex128)
    return np.asfortranarray(X.astype(np.complex64))


def _beta_between(X, k):
    s = np.linalg.svd(X.astype(np.complex128), compute_uv=False)
    return 0.5 * (s[k - 1] + s[k])


SHAPES = [(40, 12, 3), (12, 40, 3), (1000, 30, 4), (30, 200, 2), (216, 387, 5), (300, 300, 6),
          (64, 64, 1), (2, 3, 1)]


@pytest.mark.parametrize("p,n,r", SHAPES)
def test_svst_matches_lapack(ctx, rng, p, n, r):
    from mslr_hip import recon
    X = lowrank(rng, p, n, r)
    beta = _beta_between(X, r)
    Y = recon.SVST(X, beta).to_host()
    Yo = O.SVST(X.astype(np.complex128), np.float32(beta))
    assert rel_err(Y, Yo) < 2e-5


def test_svst_batched_per_patch_beta(ctx, rng):
    from mslr_hip import recon
    p, n, Np = 50, 20, 7
    P = np.stack([lowrank(rng, p, n, 1 + i % 3) for i in range(Np)], axis=2)
    P = np.asfortranarray(P)
    betas = np.array([_beta_between(P[:, :, i], 1 + i % 3) for i in range(Np)])
    Y = recon.svst_batched(P, betas).to_host()
    for i in range(Np):
        Yo = O.SVST(P[:, :, i].astype(np.complex128), np.float32(betas[i]))
        assert rel_err(Y[:, :, i], Yo) < 2e-5


def test_rank1_closed_form(ctx, rng):
    """X = u v^H: sigma = |u||v|, SVST = X (1 - beta/sigma)."""
    from mslr_hip import recon
    u = crandn(rng, (80, 1), np.complex128)
    v = crandn(rng, (25, 1), np.complex128)
    X = (u @ v.conj().T).astype(np.complex64)
    s = np.linalg.norm(u) * np.linalg.norm(v)
    beta = 0.3 * s
    Y = recon.SVST(X, beta).to_host()
    assert rel_err(Y, X * (1 - np.float32(beta) / s)) < 1e-5


def test_vector_group_shrink(ctx, rng):
    """1 x Nt patch: v * max(|v| - beta, 0) / |v| (the voxel-scale prox)."""
    from mslr_hip import recon
    Np, Nt = 33, 17
    P = crandn(rng, (1, Nt, Np))
    nrm = np.linalg.norm(P.astype(np.complex128), axis=1)[0]
    beta = float(np.median(nrm))
    Y = recon.svst_batched(P, beta).to_host()
    f = np.maximum(nrm - np.float32(beta), 0) / nrm
    assert rel_err(Y, P * f[None, None, :]) < 1e-6


def test_orthogonal_columns(ctx, rng):
    """Orthogonal columns: each column is shrunk by max(|c| - beta, 0)/|c|."""
    from mslr_hip import recon
    Q, _ = np.linalg.qr(crandn(rng, (60, 6), np.complex128))
    d = np.array([9.0, 7.0, 5.0, 3.0, 2.0, 1.0])
    X = (Q * d).astype(np.complex64)
    beta = 4.0
    Y = recon.SVST(X, beta).to_host()
    ref = Q * np.maximum(d - beta, 0)
    assert rel_err(Y, ref) < 1e-5


def test_beta_zero_identity_and_large_beta_zero(ctx, rng):
    from mslr_hip import recon
    X = lowrank(rng, 30, 20, 4)
    assert np.array_equal(recon.SVST(X, 0.0).to_host(), X)
    s1 = np.linalg.svd(X.astype(np.complex128), compute_uv=False)[0]
    assert np.all(recon.SVST(X, 1.01 * s1).to_host() == 0)


def test_zero_patch(ctx):
    from mslr_hip import recon
    X = np.zeros((40, 10), dtype=np.complex64)
    assert np.all(recon.SVST(X, 1.0).to_host() == 0)


def test_negative_beta_rejected(ctx, rng):
    from mslr_hip import recon
    with pytest.raises(ValueError):
        recon.SVST(lowrank(rng, 10, 5, 1), -1.0)


def test_nonfinite_rejected(ctx, rng):
    from mslr_hip import recon
    from mslr_hip._lib import MSLRNumericError
    X = lowrank(rng, 10, 5, 1)
    X[3, 2] = np.nan
    with pytest.raises(MSLRNumericError):
        recon.SVST(X, 1.0)


def test_clustered_singular_values(ctx, rng):
    """Equal singular values above beta (a cluster) still reconstruct the right subspace."""
    from mslr_hip import recon
    Q1, _ = np.linalg.qr(crandn(rng, (70, 5), np.complex128))
    Q2, _ = np.linalg.qr(crandn(rng, (24, 5), np.complex128))
    d = np.array([6.0, 6.0, 6.0, 2.0, 1.0])
    X = ((Q1 * d) @ Q2.conj().T).astype(np.complex64)
    Y = recon.SVST(X, 3.0).to_host()
    ref = (Q1 * np.maximum(d - 3.0, 0)) @ Q2.conj().T
    assert rel_err(Y, ref) < 1e-5


@pytest.mark.parametrize("p,n", [(4000, 60), (40, 500)])
def test_high_dynamic_range_full_rank(ctx, rng, p, n):
    """A dominant rank-1 term (s1/s_noise ~ 1e4, as the production global patch) with every
    noise singular value above beta: all eigenpairs are kept, and the noise subspace must
    still be resolved."""
    from mslr_hip import recon
    u = crandn(rng, (p, 1), np.complex128)
    v = crandn(rng, (n, 1), np.complex128)
    noise = crandn(rng, (p, n), np.complex128)
    X = (1e4 * (u @ v.conj().T) / np.sqrt(p * n) * np.sqrt(p + n) + noise).astype(np.complex64)
    s = np.linalg.svd(X.astype(np.complex128), compute_uv=False)
    beta = 0.5 * s[-1]
    Y = recon.SVST(X, beta).to_host()
    Yo = O.SVST(X.astype(np.complex128), np.float32(beta))
    assert rel_err(Y, Yo) < 2e-5
    # The part orthogonal to the dominant direction: an FP32 Gram would leave it O(1)
    # wrong; FP32 data and the FP32 apply bound it at ~eps32 * s1/s_noise (as the
    # reference's FP32 SVD).
    U, S, Vh = np.linalg.svd(X.astype(np.complex128), full_matrices=False)
    Pn = np.eye(p) - np.outer(U[:, 0], U[:, 0].conj())
    assert rel_err(Pn @ Y, Pn @ Yo) < 5e-3


def test_patch_nucnorm(ctx, rng):
    from mslr_hip import recon
    P = np.stack([lowrank(rng, 40, 15, 2) for _ in range(4)], axis=2)
    P = np.asfortranarray(P)
    v = recon.patch_nucnorm(P)
    vo = O.patch_nucnorm(P.astype(np.complex128))
    assert abs(v - vo) / vo < 1e-5


def test_patch_nucnorm_weighted(ctx, rng):
    """patch_nucnorm(P, lambdas) (src/recon.jl:157-166) in one batched pass: the mean
    over patches of lambda_i ||P_i||_*, a zero patch adding nothing."""
    from mslr_hip import recon
    P = np.stack([lowrank(rng, 40, 15, 2) for _ in range(5)], axis=2)
    P[:, :, 2] = 0
    P = np.asfortranarray(P)
    lams = [0.5, 2.0, 7.0, 1.25, 3.0]
    v = recon.patch_nucnorm(P, lams)
    vo = O.patch_nucnorm(P.astype(np.complex128), lams)
    assert abs(v - vo) / vo < 1e-5
    # the same as the per-patch sums weighted on the host
    per = [O.patch_nucnorm(P[:, :, i:i + 1].astype(np.complex128)) for i in range(5)]
    assert abs(v - sum(l * n for l, n in zip(lams, per)) / 5) / vo < 1e-5


@pytest.mark.parametrize("N,ps,ss", [((12, 12, 10), (4, 4, 4), (4, 4, 4)),
                                     ((12, 12, 10), (6, 6, 6), (3, 3, 3)),
                                     ((12, 12, 10), (12, 12, 10), (12, 12, 10)),
                                     ((12, 12, 10), (1, 1, 1), (1, 1, 1))])
def test_patch_svst_vs_oracle(ctx, rng, N, ps, ss):
    from mslr_hip import recon
    Nt = 9
    img = crandn(rng, (*N, Nt)) + 3.0 * np.ones((*N, Nt), dtype=np.complex64)
    img = np.asfortranarray(img)
    psc, _, _ = O.patch_geometry(N, ps, ss)
    beta = 0.6 * (np.sqrt(np.prod(psc)) + np.sqrt(Nt))
    out = recon.patchSVST(img, beta, ps, ss).to_host()
    ref = O.patchSVST(img.astype(np.complex128), beta, ps, ss, dtype=np.complex128)
    assert rel_err(out, ref) < 2e-5


# Gram sizes handled by the two-stage (dense -> band -> tridiagonal) reduction:
# 64 <= min(p, Nt) <= 480, the band in LDS up to 300 and in global memory above;
# the production scales use m = Nt = 300 and 387.
@pytest.fixture
def two_stage(monkeypatch):
    monkeypatch.setenv("MSLR_SVST_REDUCTION", "two")


@pytest.mark.parametrize("p,n,r", [(64, 200, 3), (100, 77, 4), (1000, 299, 7), (300, 1000, 5),
                                   (4000, 300, 12), (300, 300, 150), (301, 301, 9), (1000, 387, 9),
                                   (387, 2000, 6), (5000, 450, 5), (480, 480, 200)])
def test_two_stage_sizes_match_lapack(ctx, rng, two_stage, p, n, r):
    from mslr_hip import recon
    X = lowrank(rng, p, n, r, noise=1e-2)
    beta = _beta_between(X, r)
    Y = recon.SVST(X, beta).to_host()
    Yo = O.SVST(X.astype(np.complex128), np.float32(beta))
    assert rel_err(Y, Yo) < 2e-5


@pytest.mark.parametrize("n", [300, 387])
def test_two_stage_batch_with_mixed_ranks(ctx, rng, two_stage, n):
    """A batch of production-shaped patches (1000 x Nt) with different ranks kept."""
    from mslr_hip import recon
    p, Np = 1000, 5
    ranks = [0, 1, 3, 40, n - 1]
    P = np.stack([lowrank(rng, p, n, max(1, k), noise=1e-2) for k in ranks], axis=2)
    P = np.asfortranarray(P)
    betas = []
    for i, k in enumerate(ranks):
        s = np.linalg.svd(P[:, :, i].astype(np.complex128), compute_uv=False)
        betas.append(2.0 * s[0] if k == 0 else 0.5 * (s[k - 1] + s[k]))
    Y = recon.svst_batched(P, np.array(betas)).to_host()
    for i in range(Np):
        Yo = O.SVST(P[:, :, i].astype(np.complex128), np.float32(betas[i]))
        assert rel_err(Y[:, :, i], Yo) < 3e-5, (i, ranks[i])


@pytest.mark.parametrize("p,n", [(4000, 300), (500, 120), (3000, 387)])
def test_two_stage_and_one_stage_agree(ctx, rng, p, n, monkeypatch):
    """Both tridiagonalisations give the same SVST (high dynamic range input)."""
    from mslr_hip import recon
    u = crandn(rng, (p, 1), np.complex128)
    v = crandn(rng, (n, 1), np.complex128)
    X = (1e4 * (u @ v.conj().T) / np.sqrt(p * n) * np.sqrt(p + n)
         + crandn(rng, (p, n), np.complex128)).astype(np.complex64)
    s = np.linalg.svd(X.astype(np.complex128), compute_uv=False)
    beta = s[n // 3]
    monkeypatch.setenv("MSLR_SVST_REDUCTION", "two")
    Y2 = recon.SVST(X, beta).to_host()
    monkeypatch.setenv("MSLR_SVST_REDUCTION", "one")
    Y1 = recon.SVST(X, beta).to_host()
    Yo = O.SVST(X.astype(np.complex128), np.float32(beta))
    assert rel_err(Y2, Yo) < 2e-5
    assert rel_err(Y1, Yo) < 2e-5
    assert rel_err(Y2, Y1) < 2e-5


@pytest.mark.parametrize("p,n", [(400, 150), (1000, 387)])
def test_two_stage_nucnorm(ctx, rng, two_stage, p, n):
    from mslr_hip import recon
    P = np.asfortranarray(np.stack([lowrank(rng, p, n, 5, noise=1e-2) for _ in range(3)], axis=2))
    got = recon.patch_nucnorm(P)
    ref = sum(np.linalg.svd(P[:, :, i].astype(np.complex128), compute_uv=False).sum() for i in range(3))
    assert abs(got - ref) / ref < 1e-6


@pytest.mark.parametrize("reduction", ["one", "two"])
@pytest.mark.parametrize("p,n,r", [(300, 300, 6), (1000, 300, 20), (216, 387, 5), (1000, 387, 20)])
def test_both_reductions_match_lapack(ctx, rng, monkeypatch, reduction, p, n, r):
    """The one-stage (column by column) and two-stage (band) tridiagonalisations."""
    from mslr_hip import recon
    monkeypatch.setenv("MSLR_SVST_REDUCTION", reduction)
    X = lowrank(rng, p, n, r, noise=1e-2)
    beta = _beta_between(X, r)
    Y = recon.SVST(X, beta).to_host()
    Yo = O.SVST(X.astype(np.complex128), np.float32(beta))
    assert rel_err(Y, Yo) < 2e-5


def mixed_rank_batch(rng, ranks, p=1000, n=300):
    """Patches with the given numbers of singular values above their thresholds."""
    P = np.asfortranarray(np.stack([lowrank(rng, p, n, max(1, k), noise=1e-2) for k in ranks], axis=2))
    betas = []
    for i, k in enumerate(ranks):
        s = np.linalg.svd(P[:, :, i].astype(np.complex128), compute_uv=False)
        betas.append(2.0 * s[0] if k == 0 else 0.5 * (s[k - 1] + s[k]))
    return P, np.array(betas)


# r <= m/2 takes Y = (X U_r f) U_r^H, r > m/2 forms W: both sides of the boundary
APPLY_RANKS = [0, 2, 149, 150, 151, 299]


def test_apply_forms_at_the_boundary(ctx, rng):
    from mslr_hip import recon
    P, betas = mixed_rank_batch(rng, APPLY_RANKS)
    Y = recon.svst_batched(P, betas).to_host()
    for i, k in enumerate(APPLY_RANKS):
        Yo = O.SVST(P[:, :, i].astype(np.complex128), np.float32(betas[i]))
        assert rel_err(Y[:, :, i], Yo) < 3e-5, (i, k)
