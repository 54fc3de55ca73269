"""The device cost log (run_recon's logger: dc_cost, reg_cost, restart flag of every
iterate, scripts/reconstruct.jl:176-191,205,228-230) against the FP64 oracle, and the
run_recon_arrays output dict against the oracle's setup + POGM.

reg_cost takes three device routes, each covered here: singular values from the
prox's eigenvalues (disjoint patches, the global patch, the 1x1x1 voxel scale)
and a separate eigenvalue-only pass (overlapping patches, `small_overlap`)."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_err
from problems import prepared, raw

pytestmark = pytest.mark.gpu

# The iterates agree with the FP64 oracle to ~1e-6 (FP32 state); both costs are sums
# of magnitudes of those iterates.
TOL = 2e-5


def _oracle_log(name, niter):
    cfg, smaps, ksp, omega, sf = prepared(name)
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    lams = O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])

    def logger(it, xk, yk, r):
        dc, reg = O.recon_costs(xk, ksp, omega, smaps, lams, cfg["patch_sizes"], cfg["strides"])
        return (dc, reg, bool(r))

    _, out, _ = O.run_recon_arrays(ksp, omega, smaps, N=N, Nt=Nt, PATCH_SIZES=cfg["patch_sizes"],
                                   STRIDES=cfg["strides"], NITERS=niter, sigma1A=1.0, mode="f64",
                                   fun=logger)
    return out


def _state(name, niter, form="plain", log_costs=True):
    from mslr_hip.mirt_mod import POGM
    from mslr_hip.sense import SenseOp
    cfg, smaps, ksp, omega, sf = prepared(name)
    S = len(cfg["patch_sizes"])
    A = SenseOp(omega, smaps)
    lams = O.lambdas_ong_lustig(tuple(cfg["N"]), cfg["Nt"], cfg["patch_sizes"])
    st = POGM(A, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=lams, L=S * 1.0,
              niter=niter, form=form, log_costs=log_costs)
    st.init(ksp)
    return st


def _check(got, want, tol=TOL):
    assert len(got) == len(want)
    assert [g[2] for g in got] == [w[2] for w in want]
    dc, dco = np.array([g[0] for g in got]), np.array([w[0] for w in want])
    reg, rego = np.array([g[1] for g in got]), np.array([w[1] for w in want])
    assert np.all(np.abs(dc - dco) <= tol * dco), (dc, dco)
    assert np.all(np.abs(reg - rego) <= tol * np.maximum(rego, rego.max() * 1e-3)), (reg, rego)


@pytest.mark.parametrize("name,niter", [("small", 6), ("small_overlap", 5), ("c0", 4)])
def test_costs_match_oracle(name, niter):
    print(f"cost log {name} niter={niter}", flush=True)
    st = _state(name, niter)
    st.run(niter)
    _check(st.logger_out(), _oracle_log(name, niter))


@pytest.mark.parametrize("form", ["batched", "mixed"])
def test_costs_graph_forms(form):
    """The captured forms log the same costs (the cost kernels index by the device
    iteration counter, so every graph replay writes its own entry)."""
    p = _state("small_overlap", 6, "plain")
    p.run(6)
    g = _state("small_overlap", 6, form)
    g.run(6)
    _check(g.logger_out(), p.logger_out(), 1e-6 if form == "batched" else 1e-4)


def test_costs_across_partial_runs():
    """Reading the log mid-run (one extra forward for the newest iterate) and
    continuing gives the same log as one run."""
    a = _state("small", 6, "batched")
    a.run(6)
    b = _state("small", 6, "batched")
    b.run(2)
    mid = b.logger_out()
    assert len(mid) == 3
    b.run(3)
    b.run(1)
    full = a.logger_out()
    _check(mid, full[:3], 1e-12)
    _check(b.logger_out(), full, 1e-12)


def test_costs_after_reinit():
    st = _state("small", 4, "batched")
    st.run(4)
    first = st.logger_out()
    cfg, smaps, ksp, omega, sf = prepared("small")
    st.init(ksp)
    assert len(st.logger_out()) == 1
    st.run(4)
    _check(st.logger_out(), first, 1e-12)


def test_costs_need_the_flag():
    st = _state("small", 2, log_costs=False)
    st.run(1)
    with pytest.raises(ValueError):
        st.costs()


@pytest.mark.parametrize("name", ["small", "small_overlap"])
def test_run_recon_arrays_outputs(name):
    """The host mirror of run_recon (setup on the device + POGM + logger) against the
    oracle's setup and POGM: the reference's output keys, shapes and values."""
    from mslr_hip.reconstruct import run_recon_arrays
    niter = 4
    cfg, ksp0, smaps_raw = raw(name)
    N, Nt, Nc = tuple(cfg["N"]), cfg["Nt"], cfg["Nc"]
    print(f"run_recon_arrays {name}", flush=True)
    out = run_recon_arrays(ksp0, smaps_raw, N=N, Nvc=Nc, Nt=Nt, PATCH_SIZES=cfg["patch_sizes"],
                           STRIDES=cfg["strides"], NITERS=niter, σ1A_PRECOMPUTED=1.0, verbose=False)
    keys = {"X", "X_recon", "omega", "dc_costs", "reg_costs", "restarts", "R", "sigma1A", "L",
            "Nscales", "patch_sizes", "strides", "lambdas", "Niters", "scale_factor", "used_gpu"}
    assert set(out) == keys
    smaps, ksp, omega, R, sf = O.prepare(ksp0, smaps_raw, N, Nc, Nt)
    S = len(cfg["patch_sizes"])
    assert out["X"].shape == (*N, Nt, S) and out["X_recon"].shape == (*N, Nt)
    assert np.array_equal(out["omega"], omega)
    assert out["R"] == pytest.approx(R, rel=1e-15)
    assert out["scale_factor"] == pytest.approx(sf, rel=1e-6)
    assert out["L"] == S * 1.0 and out["Nscales"] == S and out["Niters"] == niter
    np.testing.assert_allclose(out["lambdas"], O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"]),
                               rtol=1e-15)
    want = _oracle_log(name, niter)
    got = list(zip(out["dc_costs"], out["reg_costs"], out["restarts"]))
    assert len(got) == niter + 1 and not got[0][2]
    _check(got, want)
    Xo, _, _ = O.run_recon_arrays(ksp, omega, smaps, N=N, Nt=Nt, PATCH_SIZES=cfg["patch_sizes"],
                                  STRIDES=cfg["strides"], NITERS=niter, sigma1A=1.0, mode="f64")
    assert rel_err(out["X_recon"], Xo.sum(axis=4)) < 1e-4


def test_run_recon_writes_the_reference_keys(tmp_path):
    """run_recon from files: .npy k-space and MAT smaps in, the reference's .mat out."""
    import scipy.io as sio
    from mslr_hip.reconstruct import run_recon
    cfg, ksp0, smaps_raw = raw("small")
    N, Nt, Nc = tuple(cfg["N"]), cfg["Nt"], cfg["Nc"]
    np.save(tmp_path / "ksp.npy", ksp0)
    sio.savemat(tmp_path / "smaps.mat", {"smaps": smaps_raw})
    run_recon(fn_ksp=str(tmp_path / "ksp.npy"), fn_smaps=str(tmp_path / "smaps.mat"),
              fn_recon_base=str(tmp_path / "recon"), N=N, Nvc=Nc, Nt=Nt, FOV=(1, 1, 1),
              PATCH_SIZES=cfg["patch_sizes"], STRIDES=cfg["strides"], NITERS=3, σ1A_PRECOMPUTED=1.0)
    f = sio.loadmat(tmp_path / "recon_2scales.mat")
    for k in ["X", "X_recon", "omega", "dc_costs", "reg_costs", "restarts", "R", "sigma1A", "L",
              "Nscales", "patch_sizes", "strides", "lambdas", "Niters", "scale_factor", "used_gpu"]:
        assert k in f, k
    assert f["X"].shape == (*N, Nt, 2) and f["X_recon"].shape == (*N, Nt)
    assert f["dc_costs"].shape == (4, 1) and f["restarts"].shape == (4, 1)
    assert f["patch_sizes"].shape[0] == 2 and list(f["patch_sizes"][0].ravel()[0].ravel()) == cfg["patch_sizes"][0]
