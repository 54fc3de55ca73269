"""SENSE A / A' on the GPU vs the oracle (src/sense_gpu.jl:56-108 convention)."""
import numpy as np
import pytest

import oracle as O
from conftest import crandn, rel_err

pytestmark = pytest.mark.gpu

# (N, Nc, Nt) cases: config-like, odd and mixed-radix sizes
CASES = [
    ((12, 12, 10), 3, 4),
    ((32, 32, 24), 8, 3),
    ((9, 15, 5), 2, 2),     # odd sizes: complex pre/post phases
    ((30, 20, 6), 4, 2),    # radix 2,3,5 mixtures
    ((90, 90, 60), 2, 1),   # production plane shape
]


def _problem(rng, N, Nc, Nt, frac=0.4):
    smaps = crandn(rng, (*N, Nc))
    smaps /= np.sqrt(np.sum(np.abs(smaps) ** 2, axis=3, keepdims=True))
    smaps = np.asfortranarray(smaps.astype(np.complex64))
    # same count per frame: random permutation of a fixed number of voxels
    V = int(np.prod(N))
    K = max(1, int(frac * V))
    mask = np.zeros((V, Nt), dtype=bool)
    for t in range(Nt):
        mask[rng.permutation(V)[:K], t] = True
    mask = mask.reshape(*N, Nt, order="F")
    return smaps, mask


@pytest.mark.parametrize("N,Nc,Nt", CASES)
def test_forward_adjoint_vs_oracle(ctx, rng, N, Nc, Nt):
    from mslr_hip.sense import SenseOp
    smaps, mask = _problem(rng, N, Nc, Nt)
    A = SenseOp(mask, smaps, ctx=ctx)
    x = crandn(rng, (*N, Nt))
    y = A.forward(x).to_host()
    yo = O.A_forward(x.astype(np.complex128), mask, smaps.astype(np.complex128))
    assert y.shape == yo.shape
    assert rel_err(y, yo) < 2e-6
    r = crandn(rng, y.shape)
    xa = A.adjoint(r).to_host()
    xo = O.A_adjoint(r.astype(np.complex128), mask, smaps.astype(np.complex128))
    assert rel_err(xa, xo) < 2e-6


@pytest.mark.parametrize("N,Nc,Nt", CASES[:4])
def test_f64_paths(ctx, rng, N, Nc, Nt):
    from mslr_hip.sense import SenseOp
    smaps, mask = _problem(rng, N, Nc, Nt)
    A = SenseOp(mask, smaps, ctx=ctx)
    x = crandn(rng, (*N, Nt), np.complex128)
    y = A.forward(x, f64=True).to_host()
    yo = O.A_forward(x, mask, smaps.astype(np.complex128))
    assert rel_err(y, yo) < 1e-12
    xa = A.adjoint(y, f64=True).to_host()
    xo = O.A_adjoint(yo, mask, smaps.astype(np.complex128))
    assert rel_err(xa, xo) < 1e-12


def test_dot_test_exact_adjoint(ctx, rng):
    from mslr_hip.sense import SenseOp
    N, Nc, Nt = (16, 12, 10), 4, 3
    smaps, mask = _problem(rng, N, Nc, Nt)
    A = SenseOp(mask, smaps, ctx=ctx)
    x = crandn(rng, (*N, Nt), np.complex128)
    y = crandn(rng, (A.K, Nc, Nt), np.complex128)
    lhs = np.vdot(A.forward(x, f64=True).to_host(), y)
    rhs = np.vdot(x, A.adjoint(y, f64=True).to_host())
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_full_mask_is_isometry(ctx, rng):
    """Full sampling + unit-RSS smaps: A'A = I."""
    from mslr_hip.sense import SenseOp
    N, Nc, Nt = (10, 12, 6), 3, 2
    smaps, _ = _problem(rng, N, Nc, Nt)
    mask = np.ones((*N, Nt), dtype=bool)
    A = SenseOp(mask, smaps, ctx=ctx)
    x = crandn(rng, (*N, Nt), np.complex128)
    back = A.adjoint(A.forward(x, f64=True), f64=True).to_host()
    rss2 = np.sum(np.abs(smaps.astype(np.complex128)) ** 2, axis=3)[..., None]
    assert rel_err(back, rss2 * x) < 1e-12


def test_single_coil_closed_form(ctx, rng):
    """S = 1, full mask: A is the centred unitary DFT."""
    from mslr_hip.sense import SenseOp
    N = (8, 6, 10)
    smaps = np.ones((*N, 1), dtype=np.complex64, order="F")
    mask = np.ones((*N, 1), dtype=bool)
    A = SenseOp(mask, smaps, ctx=ctx)
    x = crandn(rng, (*N, 1), np.complex128)
    y = A.forward(x, f64=True).to_host()[:, 0, 0]
    ref = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x[..., 0]))) / np.sqrt(np.prod(N))
    assert rel_err(y, ref.ravel(order="F")) < 1e-12


def test_grad_matches_oracle(ctx, rng):
    from mslr_hip.sense import SenseOp
    N, Nc, Nt, S = (12, 12, 10), 3, 4, 2
    smaps, mask = _problem(rng, N, Nc, Nt)
    A = SenseOp(mask, smaps, ctx=ctx)
    X = crandn(rng, (*N, Nt, S))
    y = crandn(rng, (A.K, Nc, Nt))
    g, dc = A.grad(X, y, want_cost=True)
    s = X.astype(np.complex128).sum(axis=4)
    sm = smaps.astype(np.complex128)
    r = O.A_forward(s, mask, sm) - y
    go = O.A_adjoint(r, mask, sm)
    assert rel_err(g.to_host(), go) < 3e-6
    assert abs(dc - 0.5 * np.linalg.norm(r) ** 2) / dc < 1e-5


def test_unequal_sample_counts_rejected(ctx, rng):
    from mslr_hip.sense import SenseOp
    N = (6, 6, 4)
    smaps, mask = _problem(rng, N, 2, 2)
    mask[0, 0, 0, 1] = not mask[0, 0, 0, 1]
    with pytest.raises(AssertionError):
        SenseOp(mask, smaps, ctx=ctx)


def test_unsupported_size_rejected(ctx, rng):
    from mslr_hip._lib import MSLRUnsupported
    from mslr_hip.sense import SenseOp
    N = (7, 6, 4)
    smaps, mask = _problem(rng, N, 2, 1)
    with pytest.raises(MSLRUnsupported):
        SenseOp(mask, smaps, ctx=ctx)


def test_poweriter_full_mask_is_one(ctx, rng):
    from mslr_hip.sense import SenseOp
    N, Nc, Nt = (8, 8, 6), 3, 3
    smaps, _ = _problem(rng, N, Nc, Nt)
    mask = np.ones((*N, Nt), dtype=bool)
    A = SenseOp(mask, smaps, ctx=ctx)
    s1, it = A.poweriter(50, 1e-10)
    assert abs(s1 - 1.0) < 1e-6


def test_poweriter_matches_oracle(ctx, rng):
    from mslr_hip.sense import SenseOp
    N, Nc, Nt = (8, 10, 6), 3, 2
    smaps, mask = _problem(rng, N, Nc, Nt, frac=0.3)
    A = SenseOp(mask, smaps, ctx=ctx)
    s1, it = A.poweriter(30, 1e-6)
    sm = smaps.astype(np.complex128)
    n = int(np.prod(N)) * Nt
    _, so, ito = O.poweriter_mod(lambda v: O.A_forward(v.reshape(*N, Nt, order="F"), mask, sm).ravel(order="F"),
                                 lambda w: O.A_adjoint(w.reshape(A.K, Nc, Nt, order="F"), mask, sm).ravel(order="F"),
                                 n, niter=30, tol=1e-6)
    assert it == ito
    assert abs(s1 - so) / so < 1e-10


@pytest.mark.parametrize("N", [(90, 90, 60), (32, 32, 24), (12, 12, 10)])
def test_fixed_and_runtime_fft_paths_agree(ctx, rng, N, monkeypatch):
    """The compile-time-length kernels and the runtime-length kernels compute the same operator."""
    from mslr_hip.sense import SenseOp
    Nc, Nt = 2, 2
    smaps, mask = _problem(rng, N, Nc, Nt)
    x = crandn(rng, (*N, Nt))
    A1 = SenseOp(mask, smaps, ctx=ctx)
    y1 = A1.forward(x).to_host()
    x1 = A1.adjoint(y1).to_host()
    monkeypatch.setenv("MSLR_FFT_GENERIC", "1")
    A2 = SenseOp(mask, smaps, ctx=ctx)
    y2 = A2.forward(x).to_host()
    x2 = A2.adjoint(y2).to_host()
    assert rel_err(y1, y2) < 2e-6
    assert rel_err(x1, x2) < 2e-6


def _line_mask(rng, N, Nt, nlines):
    """Fully sampled kx lines at nlines random (ky, kz) per frame."""
    Nx, Ny, Nz = N
    mask = np.zeros((Nx, Ny * Nz, Nt), dtype=bool)
    for t in range(Nt):
        mask[:, rng.permutation(Ny * Nz)[:nlines], t] = True
    return mask.reshape(Nx, Ny, Nz, Nt, order="F")


@pytest.mark.parametrize("N", [(90, 90, 60), (12, 12, 10), (9, 15, 5)])
def test_line_mask_vs_oracle(ctx, rng, N):
    """Readout-line masks (the acquisition's own pattern): every sampled line is whole."""
    from mslr_hip.sense import SenseOp
    Nc, Nt = 3, 2
    smaps, _ = _problem(rng, N, Nc, Nt)
    mask = _line_mask(rng, N, Nt, max(1, N[1] * N[2] // 8))
    A = SenseOp(mask, smaps, ctx=ctx)
    x = crandn(rng, (*N, Nt))
    sm = smaps.astype(np.complex128)
    assert rel_err(A.forward(x).to_host(), O.A_forward(x.astype(np.complex128), mask, sm)) < 2e-6
    r = crandn(rng, (A.K, Nc, Nt))
    assert rel_err(A.adjoint(r).to_host(), O.A_adjoint(r.astype(np.complex128), mask, sm)) < 2e-6


def test_frame_chunks_agree(ctx, rng, monkeypatch):
    """A 1 MB chunk budget (about one frame per chunk) gives the single-chunk operator."""
    from mslr_hip.sense import SenseOp
    N, Nc, Nt, S = (32, 32, 24), 3, 5, 2
    smaps, mask = _problem(rng, N, Nc, Nt, frac=0.3)
    X = crandn(rng, (*N, Nt, S))
    y = crandn(rng, (int(mask[..., 0].sum()), Nc, Nt))
    A1 = SenseOp(mask, smaps, ctx=ctx)
    g1, dc1 = A1.grad(X, y, want_cost=True)
    y1 = A1.forward(X[..., 0]).to_host()
    monkeypatch.setenv("MSLR_FFT_CHUNK_MB", "1")
    A2 = SenseOp(mask, smaps, ctx=ctx)
    g2, dc2 = A2.grad(X, y, want_cost=True)
    y2 = A2.forward(X[..., 0]).to_host()
    assert np.array_equal(y1, y2)
    assert np.array_equal(g1.to_host(), g2.to_host())
    assert dc1 == dc2


def test_frames_with_different_line_counts(ctx, rng):
    """Equal sample counts per frame spread over different numbers of lines."""
    from mslr_hip.sense import SenseOp
    N, Nc, Nt = (12, 12, 10), 2, 3
    Nx = N[0]
    smaps, _ = _problem(rng, N, Nc, Nt)
    K = 4 * Nx
    mask = np.zeros((Nx, N[1] * N[2], Nt), dtype=bool)
    mask[:, rng.permutation(N[1] * N[2])[:4], 0] = True          # 4 whole lines
    flat = np.zeros(Nx * N[1] * N[2], dtype=bool)                 # K scattered samples
    flat[rng.permutation(flat.size)[:K]] = True
    mask[..., 1] = flat.reshape(Nx, N[1] * N[2], order="F")
    mask[: Nx // 2, rng.permutation(N[1] * N[2])[:8], 2] = True   # 8 half lines
    mask = mask.reshape(*N, Nt, order="F")
    assert all(int(mask[..., t].sum()) == K for t in range(Nt))
    A = SenseOp(mask, smaps, ctx=ctx)
    x = crandn(rng, (*N, Nt))
    sm = smaps.astype(np.complex128)
    assert rel_err(A.forward(x).to_host(), O.A_forward(x.astype(np.complex128), mask, sm)) < 2e-6
    r = crandn(rng, (K, Nc, Nt))
    assert rel_err(A.adjoint(r).to_host(), O.A_adjoint(r.astype(np.complex128), mask, sm)) < 2e-6


def test_forward_residual_matches_oracle(ctx, rng):
    """||A sum_k X_k - y||^2 and the residual itself (the data term of dc_cost)."""
    from mslr_hip.sense import SenseOp
    from problems import prepared
    cfg, smaps, ksp, omega, sf = prepared("small")
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    X = crandn(rng, (*N, Nt, 2), np.complex64)
    A = SenseOp(omega, smaps)
    n2, r = A.residual(X, ksp, want_residual=True)
    ro = O.A_forward(X.astype(np.complex128).sum(axis=4), omega, smaps.astype(np.complex128)) - ksp
    assert n2 == pytest.approx(float(np.vdot(ro.ravel(), ro.ravel()).real), rel=1e-5)
    assert rel_err(r.to_host(), ro) < 2e-6
    assert A.residual(X, ksp) == pytest.approx(n2, rel=1e-12)
