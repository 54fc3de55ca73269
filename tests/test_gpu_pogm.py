"""The whole device POGM loop vs the oracle's run_recon restatement."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_err
from problems import prepared

pytestmark = pytest.mark.gpu


def _gpu_run(name, niter, form="plain"):
    from mslr_hip.mirt_mod import pogm_mslr
    from mslr_hip.sense import SenseOp
    cfg, smaps, ksp, omega, sf = prepared(name)
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    S = len(cfg["patch_sizes"])
    A = SenseOp(omega, smaps)
    lams = O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    X, restarts, st = pogm_mslr(A, ksp, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"],
                                lambdas=lams, L=S * 1.0, niter=niter, form=form)
    return X.to_host(), restarts, st


@pytest.mark.parametrize("name,niter", [("small", 6), ("c0", 4), ("small_odd", 5)])
def test_pogm_matches_oracle(name, niter):
    cfg, smaps, ksp, omega, sf = prepared(name)
    X, restarts, _ = _gpu_run(name, niter)
    Xo, ro, _ = O.run_recon_arrays(ksp, omega, smaps, N=tuple(cfg["N"]), Nt=cfg["Nt"],
                                   PATCH_SIZES=cfg["patch_sizes"], STRIDES=cfg["strides"],
                                   NITERS=niter, sigma1A=1.0, mode="f64")
    assert restarts == ro
    assert rel_err(X, Xo) < 1e-4
    # final image (scale sum)
    assert rel_err(X.sum(axis=4), Xo.sum(axis=4)) < 1e-4


def test_pogm_reference_faithful_mode():
    """Against the reference's own mixed precision (FP32 SVD, FP64 state)."""
    name, niter = "small", 5
    cfg, smaps, ksp, omega, sf = prepared(name)
    X, restarts, _ = _gpu_run(name, niter)
    Xo, ro, _ = O.run_recon_arrays(ksp, omega, smaps, N=tuple(cfg["N"]), Nt=cfg["Nt"],
                                   PATCH_SIZES=cfg["patch_sizes"], STRIDES=cfg["strides"],
                                   NITERS=niter, sigma1A=1.0, mode="ref")
    assert restarts == ro
    assert rel_err(X, Xo) < 1e-4


def test_graph_form_matches_plain():
    X1, r1, _ = _gpu_run("small", 6, form="plain")
    X2, r2, _ = _gpu_run("small", 6, form="batched")
    assert r1 == r2
    assert rel_err(X2, X1) < 1e-6


def test_thresholds_are_zeta_lambda():
    from mslr_hip.mirt_mod import POGM
    from mslr_hip.sense import SenseOp
    cfg, smaps, ksp, omega, sf = prepared("small")
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    A = SenseOp(omega, smaps)
    lams = O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    st = POGM(A, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=lams, L=2.0, niter=3)
    st.init(ksp)
    st.run(1)
    rows = O.pogm_scalars(3)
    zeta = 0.5 * rows[0][3]
    np.testing.assert_allclose(st.thresholds(), [zeta * l for l in lams], rtol=1e-15)


def test_invalid_options():
    from mslr_hip.mirt_mod import POGM
    from mslr_hip.sense import SenseOp
    cfg, smaps, ksp, omega, sf = prepared("small")
    A = SenseOp(omega, smaps)
    kw = dict(patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=[1.0, 1.0], niter=2)
    with pytest.raises(ValueError):
        POGM(A, L=-1.0, **kw)
    with pytest.raises(ValueError):
        POGM(A, L=1.0, mom="foo", **kw)
    with pytest.raises(ValueError):
        POGM(A, L=1.0, restart_cutoff=1.5, **kw)


@pytest.mark.parametrize("name,niter", [("small", 6), ("c0", 4)])
def test_per_iterate_parity(name, niter):
    """Every iterate x_k (k = 1..niter) within 1e-4 of the FP64 oracle's, and the
    final image's NRMSE to the synthetic truth within 1e-5 of the oracle's."""
    from mslr_hip.mirt_mod import POGM
    from mslr_hip.sense import SenseOp
    from problems import truth
    cfg, smaps, ksp, omega, sf = prepared(name)
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    S = len(cfg["patch_sizes"])
    lams = O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    A = SenseOp(omega, smaps)
    st = POGM(A, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=lams, L=S * 1.0,
              niter=niter)
    st.init(ksp)
    got = []
    for _ in range(niter):
        st.run(1)
        got.append(st.X.to_host())
    iterates = []

    def keep(it, xk, yk, r):
        if it > 0:
            iterates.append(np.array(xk))
        return r

    Xo, _, _ = O.run_recon_arrays(ksp, omega, smaps, N=N, Nt=Nt, PATCH_SIZES=cfg["patch_sizes"],
                                  STRIDES=cfg["strides"], NITERS=niter, sigma1A=1.0, mode="f64",
                                  fun=keep)
    assert len(iterates) == niter
    for k in range(niter):
        assert rel_err(got[k], iterates[k]) < 1e-4, k
    xt = truth(name).astype(np.complex128) / sf
    n_gpu = O.nrmse(got[-1].sum(axis=4), xt)
    n_ora = O.nrmse(Xo.sum(axis=4), xt)
    assert abs(n_gpu - n_ora) < 1e-5, (n_gpu, n_ora)


@pytest.mark.parametrize("name,niter,restart,cutoff,form", [
    ("small", 12, "none", 0.0, "plain"),
    ("small", 12, "none", 0.0, "mixed"),
    ("c0", 10, "none", 0.0, "plain"),
    # cosines of the restart test -0.49, -0.19, -0.23, -0.19, -0.18, -0.11, ...: restarts
    # at iterations 1 and 3 only, so t grows over the last seven (margin to the cutoff 0.007)
    ("c0", 10, "gr", -0.2, "batched"),
])
def test_momentum_path_matches_oracle(name, niter, restart, cutoff, form):
    """The growing-t momentum (beta != 0, the 8 t^2 rule on the last iteration) against the
    oracle's pogm_mod with the same options (src/mirt_mod.jl:140-174): identical restart
    flags, every iterate within 1e-4, and the final thresholds zeta_k lambda_s equal to
    those of the hand recurrence for that restart sequence."""
    from mslr_hip.mirt_mod import POGM
    from mslr_hip.sense import SenseOp
    cfg, smaps, ksp, omega, sf = prepared(name)
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    S = len(cfg["patch_sizes"])
    lams = O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    A = SenseOp(omega, smaps)
    st = POGM(A, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=lams, L=S * 1.0,
              niter=niter, restart=restart, restart_cutoff=cutoff, form=form)
    st.init(ksp)
    got = []
    for _ in range(niter):
        st.run(1)
        got.append(st.X.to_host())
    _, restarts = st.log()
    iterates = []

    def keep(it, xk, yk, r):
        if it > 0:
            iterates.append(np.array(xk))
        return r

    out = []
    O.run_recon_arrays(ksp, omega, smaps, N=N, Nt=Nt, PATCH_SIZES=cfg["patch_sizes"],
                       STRIDES=cfg["strides"], NITERS=niter, sigma1A=1.0, mode="f64",
                       restart=restart, restart_cutoff=cutoff,
                       fun=lambda it, xk, yk, r: (out.append(r), keep(it, xk, yk, r))[1])
    ro = [bool(v) for v in out[1:]]
    assert restarts == ro
    if restart == "none":
        assert not any(restarts)
    else:
        assert any(restarts) and not all(restarts)
    for k in range(niter):
        assert rel_err(got[k], iterates[k]) < 1e-4, k
    rows = O.pogm_scalars(niter, restarts=[i + 1 for i, r in enumerate(restarts) if r])
    zeta = rows[-1][3] / (S * 1.0)
    np.testing.assert_allclose(st.thresholds(), [zeta * l for l in lams], rtol=1e-14)
    st.close()
