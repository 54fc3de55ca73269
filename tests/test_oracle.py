"""Pins of the CPU oracle against closed forms, brute force and hand-traced values.

None of these retypes the oracle's own formula: each compares it with an
independent statement of the same quantity (explicit DFT sums, per-element
index loops, textbook SVD identities, hand-derived POGM numbers).
"""
import math

import numpy as np
import pytest

import oracle as O
from conftest import crandn, rel_err


# ----------------------------------------------------------------- SENSE
def brute_centered_dft3(x, sign=-1):
    """out[k] = sum_n x[n] exp(sign 2 pi i sum_a (k_a - c_a)(n_a - c_a)/N_a), c = N//2."""
    N = x.shape
    out = np.zeros(N, dtype=np.complex128)
    grids = [np.arange(n) - n // 2 for n in N]
    E = [np.exp(sign * 2j * np.pi * np.outer(g, g) / n) for g, n in zip(grids, N)]
    # separable explicit sums
    out = np.einsum("ai,bj,ck,ijk->abc", E[0], E[1], E[2], x)
    return out


@pytest.mark.parametrize("N", [(4, 6, 2), (5, 3, 4), (6, 5, 3)])
def test_forward_equals_bruteforce_dft(rng, N):
    Nc = 2
    smaps = crandn(rng, (*N, Nc), np.complex128)
    x = crandn(rng, N, np.complex128)
    mask = rng.random(N) < 0.5
    y = O.Asense_forward(x, mask, smaps)
    ref = np.stack([brute_centered_dft3(smaps[..., c] * x) for c in range(Nc)], axis=-1)
    ref = ref.reshape(-1, Nc, order="F")[np.flatnonzero(mask.ravel(order="F"))] / math.sqrt(np.prod(N))
    assert rel_err(y, ref) < 1e-13


@pytest.mark.parametrize("N", [(4, 6, 2), (5, 3, 4)])
def test_adjoint_is_conjugate_transpose_bruteforce(rng, N):
    """Build A as a dense matrix from the brute-force DFT; the oracle adjoint equals A^H."""
    Nc = 2
    smaps = crandn(rng, (*N, Nc), np.complex128)
    mask = rng.random(N) < 0.6
    V = int(np.prod(N))
    cols = []
    for j in range(V):
        e = np.zeros(V, dtype=np.complex128)
        e[j] = 1
        cols.append(O.Asense_forward(e.reshape(N, order="F"), mask, smaps).ravel(order="F"))
    Am = np.stack(cols, axis=1)
    y = crandn(rng, (int(mask.sum()), Nc), np.complex128)
    xa = O.Asense_adjoint(y, mask, smaps).ravel(order="F")
    assert rel_err(xa, Am.conj().T @ y.ravel(order="F")) < 1e-12


def test_single_coil_full_mask_is_unitary_centred_dft(rng):
    N = (8, 6, 4)
    x = crandn(rng, N, np.complex128)
    y = O.Asense_forward(x, np.ones(N, bool), np.ones((*N, 1), np.complex128))[:, 0]
    ref = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x))) / math.sqrt(np.prod(N))
    assert rel_err(y, ref.ravel(order="F")) < 1e-13


def test_even_size_sign_modulation_identity(rng):
    """For even N, fftshift(fft(ifftshift(x))) = (-1)^(k+N/2) fft((-1)^n x) per axis."""
    N = (8, 6, 4)
    x = crandn(rng, N, np.complex128)
    lhs = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x)))
    n = np.indices(N).sum(axis=0)
    half = sum(s // 2 for s in N)
    rhs = (-1.0) ** (n + half) * np.fft.fftn((-1.0) ** n * x)
    assert rel_err(lhs, rhs) < 1e-13


def test_full_mask_unit_rss_gives_identity(rng):
    N, Nc, Nt = (6, 4, 4), 3, 2
    smaps = crandn(rng, (*N, Nc), np.complex128)
    smaps /= np.sqrt(np.sum(np.abs(smaps) ** 2, axis=3, keepdims=True))
    mask = np.ones((*N, Nt), bool)
    x = crandn(rng, (*N, Nt), np.complex128)
    assert rel_err(O.A_adjoint(O.A_forward(x, mask, smaps), mask, smaps), x) < 1e-13


def test_reference_gpu_adjoint_scaling():
    """exact=False reproduces src/sense_gpu.jl:93: 1/prod(N) times the exact adjoint (even N)."""
    rng = np.random.default_rng(0)
    N = (4, 4, 2)
    smaps = crandn(rng, (*N, 2), np.complex128)
    mask = rng.random(N) < 0.5
    y = crandn(rng, (int(mask.sum()), 2), np.complex128)
    a = O.Asense_adjoint(y, mask, smaps, exact=True)
    b = O.Asense_adjoint(y, mask, smaps, exact=False)
    assert rel_err(b * np.prod(N), a) < 1e-13


# ----------------------------------------------------------------- patches
def brute_patches(img, ps, ss):
    N = img.shape[:3]
    psc = [min(ps[a], N[a]) for a in range(3)]
    ns = [-(-(N[a] - psc[a]) // ss[a]) + 1 for a in range(3)]
    Np = ns[0] * ns[1] * ns[2]
    p = psc[0] * psc[1] * psc[2]
    P = np.zeros((p, img.shape[3], Np), dtype=img.dtype)
    for ip in range(Np):
        ix, iy, iz = ip % ns[0], (ip // ns[0]) % ns[1], ip // (ns[0] * ns[1])
        for i in range(p):
            dx, dy, dz = i % psc[0], (i // psc[0]) % psc[1], i // (psc[0] * psc[1])
            x = min(ix * ss[0], N[0] - psc[0]) + dx
            y = min(iy * ss[1], N[1] - psc[1]) + dy
            z = min(iz * ss[2], N[2] - psc[2]) + dz
            P[i, :, ip] = img[x, y, z, :]
    return P


GEOMS = [((6, 5, 4), (2, 2, 2), (2, 2, 2)), ((6, 5, 4), (3, 3, 3), (2, 1, 2)),
         ((6, 5, 4), (9, 2, 4), (1, 2, 3)), ((7, 5, 3), (1, 1, 1), (1, 1, 1))]


@pytest.mark.parametrize("N,ps,ss", GEOMS)
def test_img2patches_bruteforce(rng, N, ps, ss):
    img = crandn(rng, (*N, 3))
    assert np.array_equal(O.img2patches(img, ps, ss), brute_patches(img, ps, ss))


@pytest.mark.parametrize("N,ps,ss", GEOMS)
def test_patches2img_adjoint_relation(rng, N, ps, ss):
    """patches2img = D^-1 E^T with E the extraction: <E x, P> = <x, D patches2img(P)>."""
    img = crandn(rng, (*N, 2), np.complex128)
    P = crandn(rng, brute_patches(img, ps, ss).shape, np.complex128)
    cnt = O.coverage_count(N, ps, ss)
    back = O.patches2img(P, ps, ss, N, dtype=np.complex128) * np.maximum(cnt, 1)[..., None]
    lhs = np.vdot(brute_patches(img, ps, ss), P)
    rhs = np.vdot(img, back)
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_coverage_closed_form():
    """Non-overlapping exact tiling covers once; half overlap covers interior voxels 2^3 times."""
    assert np.all(O.coverage_count((12, 12, 10), (4, 4, 5), (4, 4, 5)) == 1)
    c = O.coverage_count((12, 12, 12), (6, 6, 6), (3, 3, 3))
    assert c[3:9, 3:9, 3:9].min() == 8 and c[0, 0, 0] == 1


# ----------------------------------------------------------------- SVST
def test_svst_rank1(rng):
    u = crandn(rng, (30, 1), np.complex128)
    v = crandn(rng, (12, 1), np.complex128)
    X = u @ v.conj().T
    s = np.linalg.norm(u) * np.linalg.norm(v)
    assert rel_err(O.SVST(X, 0.25 * s), 0.75 * X) < 1e-13


def test_svst_vector_group_shrink(rng):
    v = crandn(rng, (1, 40), np.complex128)
    n = np.linalg.norm(v)
    assert rel_err(O.SVST(v, 0.4 * n), v * 0.6) < 1e-13


def test_svst_orthogonal_columns(rng):
    Q, _ = np.linalg.qr(crandn(rng, (20, 4), np.complex128))
    d = np.array([5.0, 3.0, 2.0, 0.5])
    X = Q * d
    assert rel_err(O.SVST(X, 2.5), Q * np.maximum(d - 2.5, 0)) < 1e-13


def test_svst_limits(rng):
    X = crandn(rng, (9, 7), np.complex128)
    assert rel_err(O.SVST(X, 0.0), X) < 1e-13
    s1 = np.linalg.norm(X, 2)
    assert np.all(O.SVST(X, 1.001 * s1) == 0)
    assert np.all(O.SVST(np.zeros((5, 3), np.complex64), 1.0) == 0)


def test_nucnorm_closed_forms(rng):
    Q, _ = np.linalg.qr(crandn(rng, (10, 3), np.complex128))
    P = np.stack([Q * np.array([3.0, 2.0, 1.0]), Q * np.array([1.0, 1.0, 1.0])], axis=2)
    assert abs(O.patch_nucnorm(P) - 9.0) < 1e-12
    assert abs(O.patch_nucnorm(P, [2.0, 1.0]) - (12.0 + 3.0) / 2) < 1e-12


# ----------------------------------------------------------------- POGM
def test_pogm_scalar_sequence_hand_values():
    """t_1 = (1+sqrt5)/2, beta_1 = 0, gamma_1 = 1/t_1; t_2 = (1+sqrt(1+4 t_1^2))/2 ...; the last
    iteration uses (1+sqrt(1+8 t^2))/2 (src/mirt_mod.jl:140-142)."""
    phi = (1 + math.sqrt(5)) / 2
    rows = O.pogm_scalars(3)
    assert rows[0][0] == pytest.approx(phi, rel=1e-15)
    assert rows[0][1] == 0.0
    assert rows[0][2] == pytest.approx(1 / phi, rel=1e-15)
    t2 = (1 + math.sqrt(1 + 4 * phi ** 2)) / 2
    assert rows[1][0] == pytest.approx(t2, rel=1e-15)
    assert rows[1][1] == pytest.approx((phi - 1) / t2, rel=1e-15)
    t3 = (1 + math.sqrt(1 + 8 * t2 ** 2)) / 2
    assert rows[2][0] == pytest.approx(t3, rel=1e-15)
    # a restart after iteration 1 resets t to 1: iteration 2 repeats iteration 1's coefficients
    rr = O.pogm_scalars(3, restarts=[1])
    assert rr[1][:3] == pytest.approx(rows[0][:3], rel=1e-15)


def test_pogm_aliased_toy_quadratic_hand_trace():
    """f(x) = 0.5 (x - 1)^2, L = 1, identity prox, x0 = 0: traced by hand.
    it1: g=-1, u1=1, t1=phi, beta=0, gamma=1/phi, z1 = 1 + 1/phi, no restart (dot = 1 > 0).
    it2: g = 1/phi, u2 = z1 - g = 1, t2, beta2=(phi-1)/t2, gamma2=phi/t2,
         z2 = 1 + beta2*(1-1) + gamma2*(1 - z1) = 1 - gamma2/phi; dot = -(1/phi)*(u2-u1)=0 -> restart."""
    phi = (1 + math.sqrt(5)) / 2
    x0 = np.zeros(1, dtype=np.complex128)

    def grad(x):
        return x - 1.0

    def prox(z, c):
        return z  # mutating identity

    x, out = O.pogm_mod(x0, lambda x: 0.0, grad, 1.0, niter=2, g_prox=prox,
                        fun=lambda it, xk, yk, r: (complex(xk[0]), r))
    assert out[1][0] == pytest.approx(1 + 1 / phi, rel=1e-15)
    assert out[1][1] is False
    t2 = (1 + math.sqrt(1 + 8 * phi ** 2)) / 2  # iteration 2 is the last one
    gamma2 = phi / t2
    assert out[2][0] == pytest.approx(1 - gamma2 / phi, rel=1e-14)
    assert out[2][1] is True  # equality (dot == 0) counts as a restart


def test_aliased_prox_differs_from_textbook():
    """With a non-mutating prox the zeta_old term survives and the iterates differ."""
    rng = np.random.default_rng(3)
    b = crandn(rng, (20,), np.complex128)

    def grad(x):
        return x - b

    def prox_inplace(z, c):
        z *= np.maximum(1 - 0.1 * c / np.maximum(np.abs(z), 1e-30), 0)
        return z

    def prox_copy(z, c):
        return z * np.maximum(1 - 0.1 * c / np.maximum(np.abs(z), 1e-30), 0)

    xa, _ = O.pogm_mod(np.zeros(20, np.complex128), lambda x: 0.0, grad, 2.0, niter=6, g_prox=prox_inplace)
    xb, _ = O.pogm_mod(np.zeros(20, np.complex128), lambda x: 0.0, grad, 2.0, niter=6, g_prox=prox_copy)
    assert rel_err(xa, xb) > 1e-6


def test_restart_is_unconjugated():
    Fg = np.array([1j])
    d = np.array([1j])
    # Re(-(1j)(1j)) = 1 > 0 -> no restart; the conjugated product would give -1 -> restart
    assert O._gr_restart(Fg, d, 0.0) is False


def test_pogm_argument_errors():
    x0 = np.zeros(2)
    with pytest.raises(ValueError):
        O.pogm_mod(x0, lambda x: 0, lambda x: x, 1.0, mom="foo")
    with pytest.raises(ValueError):
        O.pogm_mod(x0, lambda x: 0, lambda x: x, -1.0)
    with pytest.raises(ValueError):
        O.pogm_mod(x0, lambda x: 0, lambda x: x, 1.0, restart_cutoff=1.0)


# ----------------------------------------------------------------- power iteration, lambdas
def test_poweriter_diagonal():
    d = np.array([0.5, 3.0, 1.0, 2.0])
    _, s, _ = O.poweriter_mod(lambda x: d * x, lambda y: d * y, 4, niter=500, tol=1e-12)
    assert s == pytest.approx(3.0, rel=1e-9)


def test_lambdas_closed_form():
    # sqrt(p) + sqrt(Nt) + sqrt(log(V Nt / max(p, Nt))), computed by hand for 90x90x60, Nt=300
    V, Nt = 486000, 300
    for ps, lam in zip([[90, 90, 60], [10, 10, 10], [1, 1, 1]], O.lambdas_ong_lustig((90, 90, 60), 300, [[90, 90, 60], [10, 10, 10], [1, 1, 1]])):
        p = ps[0] * ps[1] * ps[2]
        assert lam == pytest.approx(math.sqrt(p) + math.sqrt(Nt) + math.sqrt(math.log(V * Nt / max(p, Nt))), rel=1e-15)
    assert O.lambdas_ong_lustig((90, 90, 60), 300, [[1, 1, 1]])[0] == pytest.approx(21.9391, abs=1e-4)
    assert O.lambdas_ong_lustig((32, 32, 24), 40, [[32, 32, 24]])[0] == pytest.approx(165.0125, abs=1e-4)


def test_quantile_is_type7():
    a = np.array([1.0, 2.0, 3.0, 4.0])
    # type 7: h = (n-1) q + 1 -> for q=0.99: 3.97 -> 3 + 0.97*(4-3)
    assert O.quantile_linear(a, 0.99) == pytest.approx(3.97, rel=1e-15)


def test_small_end_to_end_modes_agree():
    """The reference-faithful (FP32 SVD, FP64 state) and FP64 runs agree closely on a small problem."""
    from problems import prepared
    cfg, smaps, ksp, omega, sf = prepared("small")
    kw = dict(N=tuple(cfg["N"]), Nt=cfg["Nt"], PATCH_SIZES=cfg["patch_sizes"], STRIDES=cfg["strides"],
              NITERS=4, sigma1A=1.0)
    Xa, ra, _ = O.run_recon_arrays(ksp, omega, smaps, mode="f64", **kw)
    Xb, rb, _ = O.run_recon_arrays(ksp, omega, smaps, mode="ref", **kw)
    assert ra == rb
    assert rel_err(Xb, Xa) < 1e-5


# ----------------------------------------------------------------- logger costs
def _cost_problem(rng, N=(4, 6, 2), Nt=3, Nc=2):
    smaps = O.normalise_smaps(crandn(rng, (*N, Nc), np.complex128))
    omega = rng.random((*N, Nt)) < 0.6
    omega[..., :] = omega[..., :1]  # same sample count per frame
    return smaps, omega


def test_recon_costs_at_zero_and_at_truth(rng):
    N, Nt = (4, 6, 2), 3
    smaps, omega = _cost_problem(rng, N, Nt)
    ps, ss = [[4, 6, 2], [2, 2, 2]], [[4, 6, 2], [2, 2, 2]]
    lams = [1.5, 0.25]
    x = crandn(rng, (*N, Nt), np.complex128)
    y = O.A_forward(x, omega, smaps)
    X = np.zeros((*N, Nt, 2), np.complex128)
    dc, reg = O.recon_costs(X, y, omega, smaps, lams, ps, ss)
    half_sq = 0.5 * sum(abs(v) ** 2 for v in y.ravel())
    assert dc == pytest.approx(half_sq, rel=1e-13)
    assert reg == 0.0
    # the truth split over the two scales: the data term only sees the sum
    X[..., 0] = 0.3 * x
    X[..., 1] = 0.7 * x
    dc, _ = O.recon_costs(X, y, omega, smaps, lams, ps, ss)
    assert dc < 1e-20 * half_sq


def test_recon_costs_rank1_global_patch(rng):
    """One global patch holding u v^T: its nuclear norm is |u| |v|; the 2^3 scale holds
    one voxel per patch non-zero, each patch a 1 x Nt row c_i v^T (norm |c_i| |v|)."""
    N, Nt = (4, 6, 2), 3
    smaps, omega = _cost_problem(rng, N, Nt)
    ps, ss = [[4, 6, 2], [2, 2, 2]], [[4, 6, 2], [2, 2, 2]]
    lams = [1.5, 0.25]
    u = crandn(rng, N, np.complex128)
    v = crandn(rng, (Nt,), np.complex128)
    X = np.zeros((*N, Nt, 2), np.complex128)
    X[..., 0] = u[..., None] * v
    c = crandn(rng, (2, 3, 1), np.complex128)
    X[0::2, 0::2, 0::2, :, 1] = c[..., None] * v
    y = np.zeros((int(omega[..., 0].sum()), 2, Nt), np.complex128)
    _, reg = O.recon_costs(X, y, omega, smaps, lams, ps, ss)
    nv = math.sqrt(sum(abs(t) ** 2 for t in v))
    nu = math.sqrt(sum(abs(t) ** 2 for t in u.ravel()))
    want = 1.5 * nu * nv + 0.25 * nv * sum(abs(t) for t in c.ravel())
    assert reg == pytest.approx(want, rel=1e-12)


def test_recon_costs_scaling(rng):
    """dc is quadratic in X at y = 0, reg positively homogeneous (overlapping patches)."""
    N, Nt = (4, 6, 2), 3
    smaps, omega = _cost_problem(rng, N, Nt)
    ps, ss = [[4, 6, 2], [2, 3, 2]], [[4, 6, 2], [1, 2, 1]]
    lams = [1.0, 2.0]
    X = crandn(rng, (*N, Nt, 2), np.complex128)
    y = np.zeros((int(omega[..., 0].sum()), 2, Nt), np.complex128)
    d1, r1 = O.recon_costs(X, y, omega, smaps, lams, ps, ss)
    d3, r3 = O.recon_costs(3.0 * X, y, omega, smaps, lams, ps, ss)
    assert d3 == pytest.approx(9.0 * d1, rel=1e-12)
    assert r3 == pytest.approx(3.0 * r1, rel=1e-12)


# ----------------------------------------------------------------- derivatives
def _lr(rng, p, n, r, noise=1e-2):
    A = crandn(rng, (p, r), np.complex128) @ crandn(rng, (r, n), np.complex128)
    return A * np.linspace(3.0, 1.0, r).mean() + noise * crandn(rng, (p, n), np.complex128)


@pytest.mark.parametrize("p,n", [(40, 12), (12, 40)])
def test_svst_jvp_value_is_svst(rng, p, n):
    Z = _lr(rng, p, n, 3)
    s = np.linalg.svd(Z, compute_uv=False)
    for beta in [0.0, 0.5 * (s[2] + s[3]), 0.3 * s[0], 2 * s[0]]:
        Y, _ = O.svst_jvp(Z, beta, np.zeros_like(Z))
        assert rel_err(Y, O.SVST(Z, beta)) < 1e-12 or np.abs(Y).max() < 1e-12


@pytest.mark.parametrize("p,n", [(40, 12), (12, 40), (30, 30)])
def test_svst_jvp_central_differences(rng, p, n):
    Z = _lr(rng, p, n, 4, noise=0.3)
    s = np.linalg.svd(Z, compute_uv=False)
    beta = 0.5 * (s[3] + s[4])
    Zd = crandn(rng, (p, n), np.complex128)
    h = 1e-5
    for bd in [0.0, 1.0]:
        _, Yd = O.svst_jvp(Z, beta, Zd, bd)
        fd = (O.SVST(Z + h * Zd, beta + h * bd) - O.SVST(Z - h * Zd, beta - h * bd)) / (2 * h)
        assert rel_err(Yd, fd) < 1e-7, bd


@pytest.mark.parametrize("shape", [(1, 30), (30, 1)])
def test_svst_jvp_vector_central_differences(rng, shape):
    """A single row or column (the 1x1x1 patches): the tangent of the group shrink
    v max(1 - beta/|v|, 0) against FP64 central differences, kept and dropped."""
    v = crandn(rng, shape, np.complex128)
    nv = np.linalg.norm(v)
    vd = crandn(rng, shape, np.complex128)
    h = 1e-6
    for beta, bd in [(0.4 * nv, 0.0), (0.4 * nv, 1.0), (1.5 * nv, 1.0)]:
        _, Yd = O.svst_jvp(v, beta, vd, bd)
        fd = (O.SVST(v + h * vd, beta + h * bd) - O.SVST(v - h * vd, beta - h * bd)) / (2 * h)
        assert np.linalg.norm(Yd - fd) <= 1e-7 * max(np.linalg.norm(fd), 1.0), (beta, bd)


def test_svst_jvp_rank1_closed_forms(rng):
    u = crandn(rng, (30, 1), np.complex128)
    v = crandn(rng, (8, 1), np.complex128)
    u /= np.linalg.norm(u)
    v /= np.linalg.norm(v)
    Z = 5.0 * (u @ v.conj().T)
    _, Yd = O.svst_jvp(Z, 2.0, Z)  # radial: d/dt ((1+t) 5 - 2) u v^H = 5 u v^H
    assert rel_err(Yd, Z) < 1e-12
    _, Yb = O.svst_jvp(Z, 2.0, np.zeros_like(Z), 1.0)  # d/dbeta = -u v^H
    assert rel_err(Yb, -(u @ v.conj().T)) < 1e-12


def test_svst_divdiff_explicit():
    lam = np.array([9.0, 4.0, 1.0, 0.25])
    beta = 1.5  # above: 9, 4 ; below: 1, 0.25
    g = lambda l: max(1 - beta / math.sqrt(l), 0.0)
    D = O.svst_divdiff(lam, beta)
    for i in range(4):
        for j in range(4):
            if i == j:
                want = beta / (2 * lam[i] ** 1.5) if math.sqrt(lam[i]) > beta else 0.0
            else:
                want = (g(lam[i]) - g(lam[j])) / (lam[i] - lam[j])
            assert D[i, j] == pytest.approx(want, rel=1e-13, abs=1e-15)


def test_pogm_lambda_jvp_central_differences():
    """d loss / d lambda_k against central differences of the FP64 POGM run (small
    problem, 3 iterations, restart sequence unchanged by the step)."""
    from problems import prepared
    cfg, smaps, ksp, omega, sf = prepared("small")
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    kw = dict(N=N, Nt=Nt, PATCH_SIZES=cfg["patch_sizes"], STRIDES=cfg["strides"], NITERS=3)
    lams = O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    X0, r0, _ = O.run_recon_arrays(ksp, omega, smaps, mode="f64", **kw)
    rng = np.random.default_rng(3)
    x_ref = X0.sum(axis=4) + 0.1 * np.abs(X0).max() * crandn(rng, X0.shape[:4], np.complex128)
    loss, grads, rs, X = O.pogm_lambda_jvp(ksp, omega, smaps, x_ref, **kw)
    assert rs == r0
    assert rel_err(X, X0) < 1e-12

    def lossf(lm):
        Xl, rl, _ = O.run_recon_arrays(ksp, omega, smaps, mode="f64", lambdas=lm, **kw)
        assert rl == r0
        r = Xl.sum(axis=4) - x_ref
        return float(np.vdot(r.ravel(), r.ravel()).real) / r.size

    assert lossf(lams) == pytest.approx(loss, rel=1e-12)
    for k in range(len(lams)):
        h = 1e-4 * lams[k]
        lp = list(lams); lp[k] += h
        lm = list(lams); lm[k] -= h
        fd = (lossf(lp) - lossf(lm)) / (2 * h)
        assert grads[k] == pytest.approx(fd, rel=1e-5), k


# ----------------------------------------------------------------- analysis
def test_tsnr_closed_forms():
    Nt = 6
    img = np.zeros((2, 1, 1, Nt), np.complex64)
    img[0, 0, 0, :] = 3.0 * np.array([1, 1j, -1, -1j, 1, 1j])  # |x| = 3 exactly: std 0
    img[1, 0, 0, :] = np.array([1, 3, 1, 3, 1, 3]) * 1j  # mean 2, std sqrt(6/5)
    t = O.tSNR(img)
    assert t.shape == (2, 1, 1)
    assert t[0, 0, 0] == pytest.approx(3.0 / O.EPS32, rel=1e-6)
    assert t[1, 0, 0] == pytest.approx(2.0 / (math.sqrt(6 / 5) + O.EPS32), rel=1e-6)


def test_detrend_closed_forms(rng):
    Nt = 9
    t = np.arange(1, Nt + 1)
    a = crandn(rng, (3, 2, 2, 1), np.complex128)
    b = crandn(rng, (3, 2, 2, 1), np.complex128)
    out = O.detrend((a + b * t).astype(np.complex64))
    assert np.max(np.abs(out - np.broadcast_to(a, out.shape))) < 1e-5 * (np.abs(a).max() + np.abs(b).max() * Nt)
    c = np.broadcast_to(a, (3, 2, 2, Nt)).astype(np.complex64)
    assert np.array_equal(O.detrend(c), c) or np.max(np.abs(O.detrend(c) - c)) < 1e-6
    x = crandn(rng, (4, 3, 2, Nt), np.complex64)
    r = O.detrend(x).astype(np.complex128)
    # the residual has no linear component left: orthogonal to the centred ramp
    assert np.max(np.abs(r.reshape(-1, Nt) @ (t - t.mean()))) < 1e-4


# ----------------------------------------------------------------- memory-lean POGM driver
def _lean_kw(name, niter):
    from problems import prepared
    cfg, smaps, ksp, omega, sf = prepared(name)
    kw = dict(N=tuple(cfg["N"]), Nt=cfg["Nt"], PATCH_SIZES=cfg["patch_sizes"], STRIDES=cfg["strides"],
              NITERS=niter, sigma1A=1.0)
    return cfg, smaps, ksp, omega, kw


@pytest.mark.parametrize("name,restart,cutoff", [("small", "gr", 0.0), ("small_overlap", "gr", 0.0),
                                                 ("small", "none", 0.0), ("small", "gr", -0.5),
                                                 ("small_overlap", "gr", 0.3)])
def test_lean_driver_equals_pogm_mod(name, restart, cutoff):
    """run_recon_lean (two 5D arrays, scale by scale) reproduces the statement-by-statement
    pogm_mod run with the in-place prox: same restarts, iterates equal to FP64 rounding,
    and its logged costs equal recon_costs (an independent evaluation) of every iterate."""
    cfg, smaps, ksp, omega, kw = _lean_kw(name, 7)
    iters = []
    Xa, ra, ea = O.run_recon_arrays(ksp, omega, smaps, mode="f64", restart=restart, restart_cutoff=cutoff,
                                    fun=lambda it, xk, yk, r: (iters.append(np.array(xk)), r)[1], **kw)
    ra = [bool(v) for v in ra[1:]]
    seen = []
    Xb, rb, eb = O.run_recon_lean(ksp, omega, smaps, restart=restart, restart_cutoff=cutoff,
                                  callback=lambda it, X, r: seen.append(X.copy()), **kw)
    assert rb == ra
    assert len(seen) == len(iters) == kw["NITERS"] + 1
    for a, b in zip(iters, seen):
        assert rel_err(b, a) < 1e-12
    lams = ea["lambdas"]
    for k, X in enumerate(seen):
        dc, reg = O.recon_costs(X, ksp, omega, smaps, lams, cfg["patch_sizes"], cfg["strides"])
        assert eb["dc_costs"][k] == pytest.approx(dc, rel=1e-11)
        assert eb["reg_costs"][k] == pytest.approx(reg, rel=1e-11, abs=1e-12)


def test_lean_driver_restart_options_change_the_sequence():
    """The restart options reach the decision: with :none no iteration restarts; a cutoff
    of -0.5 restarts less often than 0 (src/mirt_mod.jl:26-29,163)."""
    cfg, smaps, ksp, omega, kw = _lean_kw("small", 8)
    _, r0, _ = O.run_recon_lean(ksp, omega, smaps, log_costs=False, **kw)
    _, rn, _ = O.run_recon_lean(ksp, omega, smaps, log_costs=False, restart="none", **kw)
    _, rc, _ = O.run_recon_lean(ksp, omega, smaps, log_costs=False, restart_cutoff=-0.5, **kw)
    assert not any(rn)
    assert sum(rc) < sum(r0)


# ----------------------------------------------------------------- setup (scripts/reconstruct.jl:86-118)
def test_normalise_smaps_sequential_float32():
    """Random maps: every value equals the per-voxel loop of scripts/reconstruct.jl:89 in
    Float32 (abs2 = re*re + im*im, coils summed in order, + eps(Float32), parts divided)."""
    rng = np.random.default_rng(8)
    raw = crandn(rng, (3, 2, 2, 5))
    out = O.normalise_smaps(raw)
    for idx in np.ndindex(3, 2, 2):
        acc = np.float32(0)
        for c in range(5):
            z = raw[idx + (c,)]
            acc = np.float32(acc + np.float32(np.float32(z.real * z.real) + np.float32(z.imag * z.imag)))
        den = np.float32(np.sqrt(acc) + np.float32(O.EPS32))
        for c in range(5):
            z = raw[idx + (c,)]
            assert out[idx + (c,)] == np.complex64(complex(np.float32(z.real / den), np.float32(z.imag / den)))


def test_sense_comb_is_scaled_adjoint_odd_sizes(rng):
    """sense_comb (fftshift o ifft o ifftshift, src/recon.jl:312-323) equals the exact
    adjoint A'y / (sqrt(V) (sum|S|^2 + eps)) also for odd sizes, the identity the device
    scale factor uses."""
    for N in [(5, 3, 4), (6, 5, 3), (4, 6, 2)]:
        V = int(np.prod(N))
        smaps = crandn(rng, (*N, 2), np.complex128)
        samp = rng.random(N) < 0.5
        y = crandn(rng, (int(samp.sum()), 2), np.complex128)
        k0 = np.zeros((V, 2), dtype=np.complex128)
        k0[O.sample_index(samp), :] = y
        comb = O.sense_comb(k0.reshape(*N, 2, 1, order="F"), smaps)[..., 0]
        den = np.sum(np.abs(smaps) ** 2, axis=3) + O.EPS64
        via = O.Asense_adjoint(y, samp, smaps) / (math.sqrt(V) * den)
        assert rel_err(via, comb) < 1e-13


def test_normalise_smaps_closed_form():
    """Constant maps S_c = c + 1: RSS = sqrt(sum (c+1)^2), S_c / (RSS + eps(Float32)); the
    result has unit root-sum-of-squares to FP32 rounding (scripts/reconstruct.jl:89)."""
    N, Nc = (3, 2, 2), 4
    raw = np.zeros((*N, Nc), dtype=np.complex64)
    for c in range(Nc):
        raw[..., c] = (c + 1) * np.exp(0.3j * c)
    rss = math.sqrt(sum((c + 1) ** 2 for c in range(Nc)))
    out = O.normalise_smaps(raw)
    for c in range(Nc):
        assert np.allclose(np.abs(out[..., c]), (c + 1) / (rss + O.EPS32), rtol=1e-6)
    assert np.allclose(np.sum(np.abs(out) ** 2, axis=3), 1.0, rtol=1e-6)


def test_sense_comb_closed_forms(rng):
    """(i) One coil, S = 1, a single sample at the centre of k-space (index N//2): the
    centred inverse DFT is the constant 1/V.  (ii) Full sampling of A x with unit-RSS maps:
    sense_comb returns x / sqrt(V) (the unitary forward's scale), src/recon.jl:312-323."""
    N, Nt = (4, 6, 2), 2
    V = int(np.prod(N))
    ksp = np.zeros((*N, 1, Nt), dtype=np.complex128)
    ksp[N[0] // 2, N[1] // 2, N[2] // 2, 0, :] = [1.0, 2.0 - 1j]
    img = O.sense_comb(ksp, np.ones((*N, 1), dtype=np.complex128))
    assert np.allclose(img[..., 0], 1.0 / V, rtol=1e-14)
    assert np.allclose(img[..., 1], (2.0 - 1j) / V, rtol=1e-14)
    Nc = 3
    smaps = O.normalise_smaps(crandn(rng, (*N, Nc), np.complex128)).astype(np.complex128)
    smaps /= np.sqrt(np.sum(np.abs(smaps) ** 2, axis=3, keepdims=True))
    x = crandn(rng, (*N, Nt), np.complex128)
    full = np.ones(N, dtype=bool)
    ksp = np.zeros((*N, Nc, Nt), dtype=np.complex128)
    for t in range(Nt):
        ksp[..., t] = O.Asense_forward(x[..., t], full, smaps).reshape(*N, Nc, order="F")
    assert rel_err(O.sense_comb(ksp, smaps), x / math.sqrt(V)) < 1e-13


def test_prepare_bruteforce(rng):
    """prepare (scripts/reconstruct.jl:86-118,145-147) against per-voxel loops: the mask is
    the nonzeros of coil 1, R = V / K, the flattened k-space lists the sampled voxels of each
    frame in column-major order divided by the type-7 0.99 quantile of |sense_comb|, and the
    two assertions fire on a coil with another pattern and a frame with another count."""
    N, Nc, Nt = (4, 6, 2), 2, 3
    V = int(np.prod(N))
    samp = rng.random((N[0], N[1], N[2])) < 0.5
    omega = np.stack([np.roll(samp, t, axis=1) for t in range(Nt)], axis=3)
    vals = crandn(rng, (*N, Nc, Nt))
    ksp0 = np.where(omega[:, :, :, None, :], vals, 0).astype(np.complex64)
    raw = crandn(rng, (*N, Nc))
    smaps, ksp, om, R, sf = O.prepare(ksp0, raw, N, Nc, Nt)
    assert np.array_equal(om, omega)
    K = int(samp.sum())
    assert R == V / K
    img = O.sense_comb(ksp0, smaps)
    mags = sorted(float(np.abs(np.complex64(v))) for v in img.ravel(order="F"))  # abs of ComplexF32
    h = (len(mags) - 1) * 0.99
    lo = int(math.floor(h))
    q = mags[lo] + (h - lo) * (mags[min(lo + 1, len(mags) - 1)] - mags[lo])
    assert sf == pytest.approx(q, rel=1e-12)
    for t in range(Nt):
        k = 0
        for iz in range(N[2]):
            for iy in range(N[1]):
                for ix in range(N[0]):
                    if omega[ix, iy, iz, t]:
                        for c in range(Nc):
                            v = complex(ksp0[ix, iy, iz, c, t])  # divided in Float64, stored Float32
                            assert ksp[k, c, t] == np.complex64(complex(v.real / sf, v.imag / sf))
                        k += 1
        assert k == K
    bad = ksp0.copy()
    bad[0, 0, 0, 1, 0] = 0 if omega[0, 0, 0, 0] else 1
    with pytest.raises(AssertionError, match="Coil 2"):
        O.prepare(bad, raw, N, Nc, Nt)
    bad = ksp0.copy()
    extra = np.argwhere(~omega[..., 1])[0]
    bad[extra[0], extra[1], extra[2], :, 1] = 1
    with pytest.raises(AssertionError, match="Frame 2"):
        O.prepare(bad, raw, N, Nc, Nt)
