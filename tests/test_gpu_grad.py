"""The differentiation form: the SVST pullback and the lambda-gradient of the unrolled
POGM on the device, against the FP64 forward-mode oracle (adjoint identity
<Ybar, Ydot> = <Zbar, Zdot> + betabar * betadot), central finite differences of the
device forward run, and the oracle's gradient at the differentiable workload (c4:
64x64x40, Nt=120, 12 coils, scales [[64,64,40],[8,8,8]], 10 iterations)."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from conftest import crandn
from problems import prepared, truth
from test_gpu_svst import lowrank

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _vdot(a, b):
    return float(np.vdot(np.asarray(a, np.complex128).ravel(), np.asarray(b, np.complex128).ravel()).real)


@pytest.mark.parametrize("p,n,r,Np", [(200, 40, 5, 1), (1000, 60, 8, 3), (64, 64, 20, 2), (4000, 30, 3, 1),
                                      (1000, 300, 12, 1), (1000, 387, 10, 2)])
def test_svst_vjp_adjoint_identity(ctx, rng, p, n, r, Np):
    from mslr_hip import recon
    Z = np.asfortranarray(np.stack([lowrank(rng, p, n, r, noise=0.1) for _ in range(Np)], axis=2))
    s = [np.linalg.svd(Z[:, :, i].astype(np.complex128), compute_uv=False) for i in range(Np)]
    betas = np.array([0.5 * (si[r - 1] + si[r]) for si in s])
    Yb = crandn(rng, Z.shape, np.complex64)
    Zb, bb = recon.svst_vjp(Z, Yb, betas)
    bb = np.atleast_1d(bb)
    Zb = Zb.to_host()
    assert np.all(np.isfinite(Zb))
    for trial in range(3):
        Zd = crandn(rng, Z.shape, np.complex128)
        bd = rng.standard_normal(Np)
        lhs = 0.0
        scale = 0.0
        for i in range(Np):
            _, Yd = O.svst_jvp(Z[:, :, i], float(np.float32(betas[i])), Zd[:, :, i], bd[i])
            lhs += _vdot(Yb[:, :, i], Yd)
            scale += np.linalg.norm(Yb[:, :, i]) * np.linalg.norm(Yd)
        rhs = _vdot(Zb, Zd) + float(np.dot(bb, bd))
        assert abs(lhs - rhs) <= 1e-4 * scale, (trial, lhs, rhs)
    # betabar alone: the threshold direction
    for i in range(Np):
        _, Yb1 = O.svst_jvp(Z[:, :, i], float(np.float32(betas[i])), np.zeros((p, n)), 1.0)
        want = _vdot(Yb[:, :, i], Yb1)
        assert bb[i] == pytest.approx(want, rel=1e-4, abs=1e-6 * np.linalg.norm(Yb[:, :, i]) * np.linalg.norm(Yb1))


def test_svst_vjp_zero_patch_and_all_kept(ctx, rng):
    from mslr_hip import recon
    Yb = crandn(rng, (300, 40), np.complex64)
    Zb, bb = recon.svst_vjp(np.zeros((300, 40), np.complex64), Yb, 1.0)
    assert np.all(Zb.to_host() == 0) and bb == 0.0
    # every singular value above beta: Zbar from the smooth part only
    Z = lowrank(rng, 300, 40, 40, noise=0.0)
    beta = 0.1 * np.linalg.svd(Z.astype(np.complex128), compute_uv=False)[-1]
    Zb, bb = recon.svst_vjp(Z, Yb, beta)
    Zd = crandn(rng, Z.shape, np.complex128)
    _, Yd = O.svst_jvp(Z, float(np.float32(beta)), Zd, 0.0)
    assert abs(_vdot(Zb.to_host(), Zd) - _vdot(Yb, Yd)) <= 1e-4 * np.linalg.norm(Yb) * np.linalg.norm(Yd)


@pytest.mark.parametrize("p,n,r,Np", [(20, 60, 3, 2), (216, 387, 8, 2), (2, 10, 1, 3),
                                      (1, 300, 1, 6), (40, 1, 1, 4)])
def test_svst_vjp_wide_and_vector_shapes(ctx, rng, p, n, r, Np):
    """Patches with fewer voxels than frames (through Y^H = SVST(Z^H)) and single rows or
    columns (the group shrink's closed form): the adjoint identity against the oracle's FP64
    forward mode, with thresholds that keep some singular values (and, for vectors, drop
    some vectors)."""
    from mslr_hip import recon
    Z = np.asfortranarray(np.stack([lowrank(rng, p, n, min(r, p, n), noise=0.1) for _ in range(Np)], axis=2))
    s = [np.linalg.svd(Z[:, :, i].astype(np.complex128), compute_uv=False) for i in range(Np)]
    if min(p, n) == 1:
        betas = np.array([si[0] * (0.5 if i % 2 == 0 else 1.5) for i, si in enumerate(s)])
    else:
        betas = np.array([0.5 * (si[r - 1] + si[r]) for si in s])
    Yb = crandn(rng, Z.shape, np.complex64)
    Zb, bb = recon.svst_vjp(Z, Yb, betas)
    bb = np.atleast_1d(bb)
    Zb = Zb.to_host()
    assert np.all(np.isfinite(Zb))
    for trial in range(3):
        Zd = crandn(rng, Z.shape, np.complex128)
        bd = rng.standard_normal(Np)
        lhs = scale = 0.0
        for i in range(Np):
            _, Yd = O.svst_jvp(Z[:, :, i], float(np.float32(betas[i])), Zd[:, :, i], bd[i])
            lhs += _vdot(Yb[:, :, i], Yd)
            scale += np.linalg.norm(Yb[:, :, i]) * np.linalg.norm(Yd)
        rhs = _vdot(Zb, Zd) + float(np.dot(bb, bd))
        assert abs(lhs - rhs) <= 1e-4 * scale, (trial, lhs, rhs)
    if min(p, n) == 1:  # the dropped vectors have no derivative
        for i in range(1, Np, 2):
            assert bb[i] == 0.0 and np.all(Zb[:, :, i] == 0)


def _device_loss(name, niter, lambdas):
    """The loss from a plain forward run (no checkpoints), for central differences."""
    from mslr_hip.mirt_mod import pogm_mslr
    from mslr_hip.sense import SenseOp
    cfg, smaps, ksp, omega, sf = prepared(name)
    S = len(cfg["patch_sizes"])
    A = SenseOp(omega, smaps)
    X, restarts, st = pogm_mslr(A, ksp, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"],
                                lambdas=lambdas, L=S * 1.0, niter=niter)
    r = X.to_host().sum(axis=4).astype(np.complex128) - truth(name).astype(np.complex128) / np.float32(sf)
    st.close()
    A.close()
    return float(np.vdot(r.ravel(), r.ravel()).real) / r.size, restarts


def _device_grad(name, niter, lambdas=None, form="plain"):
    from mslr_hip.mirt_mod import pogm_lambda_grad
    from mslr_hip.sense import SenseOp
    cfg, smaps, ksp, omega, sf = prepared(name)
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    S = len(cfg["patch_sizes"])
    lams = lambdas if lambdas is not None else O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    x_ref = (truth(name) / np.float32(sf)).astype(np.complex64)
    A = SenseOp(omega, smaps)
    loss, grads, X, st = pogm_lambda_grad(A, ksp, x_ref, patch_sizes=cfg["patch_sizes"],
                                          strides=cfg["strides"], lambdas=lams, L=S * 1.0,
                                          niter=niter, form=form)
    _, restarts = st.log()
    st.close()
    A.close()
    return loss, grads, restarts, lams


@pytest.mark.parametrize("niter", [3, 6])
def test_lambda_grad_matches_fp64_forward_mode(niter):
    cfg, smaps, ksp, omega, sf = prepared("small")
    print(f"lambda grad small niter={niter}", flush=True)
    loss, grads, rs, lams = _device_grad("small", niter)
    x_ref = truth("small").astype(np.complex128) / sf
    lo, go, ro, _ = O.pogm_lambda_jvp(ksp, omega, smaps, x_ref, N=tuple(cfg["N"]), Nt=cfg["Nt"],
                                      PATCH_SIZES=cfg["patch_sizes"], STRIDES=cfg["strides"],
                                      NITERS=niter)
    assert rs == ro
    assert loss == pytest.approx(lo, rel=1e-4)
    for k in range(len(lams)):
        assert grads[k] == pytest.approx(go[k], rel=1e-3), (k, grads, go)


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_lambda_grad_voxel_and_small_patch_scales(form):
    """Scales [[12,12,10],[2,2,2],[1,1,1]] over Nt = 10 frames: 2^3 patches have fewer
    voxels than frames and the last scale is 1x1x1 (the production scale sets carry both).
    The device gradient against the oracle's FP64 forward mode (3 iterations), 1e-3."""
    niter = 3
    cfg, smaps, ksp, omega, sf = prepared("small_vjp")
    loss, grads, rs, lams = _device_grad("small_vjp", niter, form=form)
    x_ref = truth("small_vjp").astype(np.complex128) / sf
    lo, go, ro, _ = O.pogm_lambda_jvp(ksp, omega, smaps, x_ref, N=tuple(cfg["N"]), Nt=cfg["Nt"],
                                      PATCH_SIZES=cfg["patch_sizes"], STRIDES=cfg["strides"],
                                      NITERS=niter)
    assert rs == ro
    assert loss == pytest.approx(lo, rel=1e-4)
    for k in range(len(lams)):
        assert grads[k] == pytest.approx(go[k], rel=1e-3, abs=1e-4 * max(abs(v) for v in go)), (k, grads, go)


def test_lambda_grad_graph_form_matches_plain():
    l1, g1, r1, _ = _device_grad("small", 5, form="plain")
    l2, g2, r2, _ = _device_grad("small", 5, form="batched")
    assert r1 == r2 and l1 == pytest.approx(l2, rel=1e-6)
    np.testing.assert_allclose(g2, g1, rtol=1e-5)


def test_lambda_grad_c4_central_differences_and_oracle():
    """c4, 10 iterations: the device gradient against the oracle's FP64 forward-mode
    gradient and the oracle's FP64 central differences (tests/golden/lambda_grad_c4.json),
    both to 1e-3; for each lambda_k the central difference with the smallest step the
    file holds (steps 1e-3 lambda_k for both, and 1e-4 lambda_2: at 1e-3 the FP64
    central difference for lambda_2 is itself 1.4e-3 from the forward-mode
    derivative, its truncation error); and against central differences of the
    device's own FP32 loss with steps 1e-3 lambda_k.  The FP32 loss
    carries a rounding jitter of ~1e-7 of its value from run to run of lambda (measured
    2^-23 L at c4), which a central difference divides by 2h: the device-side check
    allows max(1e-3 |fd|, 2^-21 L / h) (for lambda_2 = 36.8 the jitter term is ~1e-3
    absolute, 4% of the gradient)."""
    niter = 10
    print("lambda grad c4", flush=True)
    loss, grads, rs, lams = _device_grad("c4", niter)
    print(f"c4 loss {loss:.6e} grads {grads} restarts {rs}", flush=True)
    g = json.loads((GOLDEN / "lambda_grad_c4.json").read_text())
    assert g["niter"] == niter and rs == g["restarts"]
    assert loss == pytest.approx(g["loss"], rel=1e-4)
    by_step = g.get("fd_by_step", {})
    for k in range(len(lams)):
        assert grads[k] == pytest.approx(g["grads"][k], rel=1e-3), (k, grads, g["grads"])
        steps = sorted((float(r), v[k]) for r, v in by_step.items() if v[k] is not None)
        if steps:
            rel, fdk = steps[0]
            print(f"  lambda_{k}: FP64 central difference (step {rel:g} lambda) {fdk:.6e}", flush=True)
            assert grads[k] == pytest.approx(fdk, rel=1e-3), (k, rel, grads[k], fdk)
    for k in range(len(lams)):
        h = 1e-3 * lams[k]
        lp = list(lams); lp[k] += h
        lm = list(lams); lm[k] -= h
        fp, rp = _device_loss("c4", niter, lp)
        fm, rm = _device_loss("c4", niter, lm)
        assert rp == rs and rm == rs
        fd = (fp - fm) / (2 * h)
        tol = max(1e-3 * abs(fd), 2.0 ** -21 * loss / h)
        print(f"  lambda_{k}: adjoint {grads[k]:.6e} central difference {fd:.6e} (allowed {tol:.1e})",
              flush=True)
        assert abs(grads[k] - fd) <= tol, k
