"""The production workloads at their full size.

c1 (90x90x60, Nt=300, 18 coils, scales [[90,90,60],[10,10,10],[1,1,1]]): the plain,
batched and mixed-precision forms against the FP64 oracle's run of the same problem
(tests/golden/recon_c1.json and _subset.npz, written by make_recon_golden.py).
c2/c3 (Nt=387, four half-overlapping scales down to 6^3 patches): the three forms
against each other.  An FP64 oracle run of this shape holds ~45 GB of host arrays
and takes hours per iteration on an 8-core host (19,284 LAPACK SVDs per prox and
as many again for the logger), so it has no fixture.
"""
import math

import numpy as np
import pytest

from conftest import rel_err
from problems import production_inputs, recon_golden

pytestmark = pytest.mark.gpu


def _state(A, yd, cfg, lams, niter, form, log_costs=True):
    from mslr_hip.mirt_mod import POGM
    S = len(cfg["patch_sizes"])
    st = POGM(A, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=lams, L=S * 1.0,
              niter=niter, form=form, log_costs=log_costs)
    st.init(yd)
    return st


@pytest.fixture(scope="module")
def c1_setup():
    gold = recon_golden("c1")
    if gold is None:
        pytest.skip("tests/golden/recon_c1.json not present")
    from mslr_hip.device import DeviceArray, default_context
    from mslr_hip.reconstruct import lambdas_ong_lustig, sense_comb_scale
    from mslr_hip.sense import SenseOp
    ctx = default_context()
    rec, vox, sub = gold
    cfg, pb, smaps, y = production_inputs("c1", ctx)
    assert float(np.vdot(y.ravel(), y.ravel()).real) == pytest.approx(rec["y_norm2"], rel=1e-6)
    A = SenseOp(pb.mask, smaps, ctx=ctx)
    yd = DeviceArray.from_host(y, ctx)
    sf_dev = sense_comb_scale(A, yd)
    # the oracle's inputs: y / Float32(sf) with its scale factor
    yd.upload(np.asfortranarray(y / np.float32(rec["scale_factor"])))
    lams = lambdas_ong_lustig(tuple(cfg["N"]), cfg["Nt"], cfg["patch_sizes"])
    np.testing.assert_allclose(lams, rec["lambdas"], rtol=1e-15)
    yield dict(rec=rec, vox=vox, sub=sub, cfg=cfg, pb=pb, A=A, yd=yd, lams=lams, sf_dev=sf_dev)
    A.close()


def test_c1_scale_factor_matches_oracle(c1_setup):
    """The device setup's k-space scale (0.99 quantile of |sense_comb| through the GPU
    adjoint, scripts/reconstruct.jl:102-104) equals the oracle's at the production size."""
    assert c1_setup["sf_dev"] == pytest.approx(c1_setup["rec"]["scale_factor"], rel=1e-6)


@pytest.mark.parametrize("form,tol_img,tol_nrmse", [("plain", 1e-4, 1e-5), ("batched", 1e-4, 1e-5),
                                                    ("mixed", 1e-4, 1e-4)])
def test_c1_forms_match_fp64_oracle(c1_setup, form, tol_img, tol_nrmse):
    """Same restart flags as the FP64 oracle, final-image NRMSE to the synthetic truth
    within tol_nrmse of the oracle's, the image sum at the fixture's voxels within tol_img
    (relative), and dc_cost / reg_cost of every iterate within 2e-5 (src/mirt_mod.jl:111-185,
    scripts/reconstruct.jl:159-231)."""
    s = c1_setup
    rec, cfg, pb = s["rec"], s["cfg"], s["pb"]
    niter = rec["niter"]
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    V = int(np.prod(N))
    st = _state(s["A"], s["yd"], cfg, s["lams"], niter, form)
    st.run(niter)
    _, restarts = st.log()
    dc, reg = st.costs()
    xs = st.X.to_host().astype(np.complex128).sum(axis=4)
    st.close()
    assert restarts == rec["restarts"]
    sf = rec["scale_factor"]
    xt = pb.x_true.astype(np.complex128)
    nr = float(np.linalg.norm(xs * sf - xt) / np.linalg.norm(xt))
    assert abs(nr - rec["nrmse_vs_truth"][-1]) < tol_nrmse, (nr, rec["nrmse_vs_truth"][-1])
    got = xs.reshape(V, Nt, order="F")[s["vox"], :]
    assert rel_err(got, s["sub"]) < tol_img
    np.testing.assert_allclose(dc, rec["dc_costs"], rtol=2e-5)
    np.testing.assert_allclose(reg, rec["reg_costs"], rtol=2e-5)


@pytest.fixture(scope="module")
def c2_setup():
    from mslr_hip.device import DeviceArray, default_context
    from mslr_hip.reconstruct import lambdas_ong_lustig, sense_comb_scale
    from mslr_hip.sense import SenseOp
    ctx = default_context()
    cfg, pb, smaps, y = production_inputs("c2", ctx)
    A = SenseOp(pb.mask, smaps, ctx=ctx)
    yd = DeviceArray.from_host(y, ctx)
    sf = sense_comb_scale(A, yd)
    yd.upload(np.asfortranarray(y / np.float32(sf)))
    lams = lambdas_ong_lustig(tuple(cfg["N"]), cfg["Nt"], cfg["patch_sizes"])
    runs = {}
    niter = 3
    st = _state(A, yd, cfg, lams, niter, "plain", log_costs=False)
    x0 = st.X.to_host().astype(np.complex128).sum(axis=4)
    st.close()
    for form in ("plain", "batched", "mixed"):
        st = _state(A, yd, cfg, lams, niter, form)
        st.run(niter)
        _, rs = st.log()
        dc, reg = st.costs()
        runs[form] = dict(restarts=rs, dc=dc, reg=reg, X=st.X.to_host())
        st.close()
    yield dict(cfg=cfg, pb=pb, sf=sf, runs=runs, niter=niter, x0=x0)
    A.close()


def test_c2_batched_is_plain(c2_setup):
    """The graph form with a stream per scale (all four half-overlapping scales, 19,284
    patches of 387 frames) gives the plain form's iterates and logger."""
    r = c2_setup["runs"]
    assert r["batched"]["restarts"] == r["plain"]["restarts"]
    assert rel_err(r["batched"]["X"], r["plain"]["X"]) < 1e-6
    np.testing.assert_allclose(r["batched"]["dc"], r["plain"]["dc"], rtol=1e-6)
    np.testing.assert_allclose(r["batched"]["reg"], r["plain"]["reg"], rtol=1e-6)


def test_c3_mixed_matches_plain(c2_setup):
    """The mixed-precision form at the c2 shape (split-BF16 Gram and apply, FP64
    eigensolve): the plain form's restarts, iterate within 1e-4, costs within 2e-5, and
    finite costs for every iterate."""
    r = c2_setup["runs"]
    assert r["mixed"]["restarts"] == r["plain"]["restarts"]
    assert rel_err(r["mixed"]["X"], r["plain"]["X"]) < 1e-4
    np.testing.assert_allclose(r["mixed"]["dc"], r["plain"]["dc"], rtol=2e-5)
    np.testing.assert_allclose(r["mixed"]["reg"], r["plain"]["reg"], rtol=2e-5)
    for form in r:
        assert all(math.isfinite(v) for v in r[form]["dc"] + r[form]["reg"])
        assert len(r[form]["dc"]) == c2_setup["niter"] + 1


def test_c2_reconstruction_improves(c2_setup):
    """Each form's image sum is closer to the synthetic truth than X0 = A'y
    (scripts/reconstruct.jl:209-214) after the iterations."""
    s = c2_setup
    xt = s["pb"].x_true.astype(np.complex128)
    nr0 = float(np.linalg.norm(s["x0"] * s["sf"] - xt) / np.linalg.norm(xt))
    for form, run in s["runs"].items():
        xs = run["X"].astype(np.complex128).sum(axis=4) * s["sf"]
        nr = float(np.linalg.norm(xs - xt) / np.linalg.norm(xt))
        assert nr < 0.9 * nr0, (form, nr, nr0)
