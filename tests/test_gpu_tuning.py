"""SVST launch shapes (mslr_ctx_set_tuning) and the autotuner.  Every shape gives
LAPACK's SVST to the bound of its products (2e-5 FP32, 1e-4 split BF16) and the
built-in shape's result to rounding; the tuner times every candidate of its
descent, keeps only candidates whose iterate stays within its tolerance, leaves
the winner set on the context and caches it per workload."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_err
from problems import prepared
from test_gpu_svst import APPLY_RANKS, _beta_between, lowrank, mixed_rank_batch

pytestmark = pytest.mark.gpu

BUILTIN = {"gram_chunk": 0, "eig_lanes_limit": 0, "svst_apply": 0, "sense_chunks": 0,
           "yz_threads_fwd": 0, "yz_threads_adj": 0, "update_lines": 0, "sb2st_schedule": 0,
           "apply_fp32": 0}


@pytest.fixture
def shapes(ctx):
    ctx.set_tuning()
    yield ctx
    ctx.set_tuning()
    ctx.set_products("fp32")


def test_tuning_roundtrip_and_validation(shapes):
    assert shapes.tuning() == BUILTIN
    want = {"gram_chunk": 2048, "eig_lanes_limit": -1, "svst_apply": 1, "sense_chunks": 4,
            "yz_threads_fwd": 512, "yz_threads_adj": 1024, "update_lines": 32, "sb2st_schedule": 2,
            "apply_fp32": 5}
    shapes.set_tuning(**want)
    assert shapes.tuning() == want
    for bad in [dict(gram_chunk=100), dict(gram_chunk=128), dict(gram_chunk=1 << 17),
                dict(gram_chunk=-256), dict(eig_lanes_limit=-2), dict(svst_apply=3),
                dict(sense_chunks=-1), dict(sense_chunks=65), dict(yz_threads_fwd=256),
                dict(yz_threads_adj=640), dict(update_lines=100), dict(update_lines=12), dict(sb2st_schedule=4),
                dict(apply_fp32=9), dict(apply_fp32=-1)]:
        with pytest.raises(ValueError):
            shapes.set_tuning(**bad)
        assert shapes.tuning() == want


@pytest.mark.parametrize("products,bound", [("fp32", 2e-5), ("split_bf16", 1e-4)])
@pytest.mark.parametrize("chunk", [256, 2048, 16384])
def test_gram_chunk(shapes, rng, products, bound, chunk):
    """A tall patch (5000 rows: 20 chunks of 256 or one of 16384) and a batch of
    1000-row patches, against the built-in 4096-row chunks and LAPACK."""
    from mslr_hip import recon
    shapes.set_products(products)
    X = lowrank(rng, 5000, 60, 4)
    beta = _beta_between(X, 4)
    P, betas = mixed_rank_batch(rng, [0, 3, 40], p=1000, n=120)
    Y0 = recon.SVST(X, beta).to_host()
    B0 = recon.svst_batched(P, betas).to_host()
    shapes.set_tuning(gram_chunk=chunk)
    Y1 = recon.SVST(X, beta).to_host()
    B1 = recon.svst_batched(P, betas).to_host()
    assert rel_err(Y1, Y0) < 1e-5
    assert rel_err(B1, B0) < 1e-5
    assert rel_err(Y1, O.SVST(X.astype(np.complex128), np.float32(beta))) < bound
    for i in range(P.shape[2]):
        Yo = O.SVST(P[:, :, i].astype(np.complex128), np.float32(betas[i]))
        assert rel_err(B1[:, :, i], Yo) < bound, i


@pytest.mark.parametrize("tuning", [dict(eig_lanes_limit=-1), dict(eig_lanes_limit=1 << 30),
                                    dict(svst_apply=1), dict(eig_lanes_limit=-1, svst_apply=1)])
def test_eig_lanes_and_apply(shapes, rng, tuning):
    """Both eigenvector kernels and both apply forms on ranks either side of the
    low-rank boundary (2 r <= m)."""
    from mslr_hip import recon
    P, betas = mixed_rank_batch(rng, APPLY_RANKS)
    Y0 = recon.svst_batched(P, betas).to_host()
    shapes.set_tuning(**tuning)
    Y1 = recon.svst_batched(P, betas).to_host()
    assert rel_err(Y1, Y0) < 1e-5
    for i, k in enumerate(APPLY_RANKS):
        Yo = O.SVST(P[:, :, i].astype(np.complex128), np.float32(betas[i]))
        assert rel_err(Y1[:, :, i], Yo) < 3e-5, (i, k)


@pytest.mark.parametrize("amask", [1, 2, 5, 8])
def test_apply_fp32(shapes, rng, amask):
    """Split-BF16 products with the reconstruction products in either arithmetic
    (1: all split BF16, 8: all FP32, 2 and 5 mixed), on ranks either side of the
    low-rank boundary and on a wide patch (p < Nt, the W X form)."""
    from mslr_hip import recon
    shapes.set_products("split_bf16")
    P, betas = mixed_rank_batch(rng, APPLY_RANKS)
    X = lowrank(rng, 90, 300, 5)
    beta = _beta_between(X, 5)
    Y0 = recon.svst_batched(P, betas).to_host()
    W0 = recon.SVST(X, beta).to_host()
    shapes.set_tuning(apply_fp32=amask)
    Y1 = recon.svst_batched(P, betas).to_host()
    W1 = recon.SVST(X, beta).to_host()
    assert rel_err(Y1, Y0) < 3e-5
    assert rel_err(W1, W0) < 3e-5
    for i, k in enumerate(APPLY_RANKS):
        Yo = O.SVST(P[:, :, i].astype(np.complex128), np.float32(betas[i]))
        assert rel_err(Y1[:, :, i], Yo) < 1e-4, (i, k)
    assert rel_err(W1, O.SVST(X.astype(np.complex128), np.float32(beta))) < 1e-4


def test_autotune_pogm(shapes, tmp_path):
    from mslr_hip.autotune import autotune_pogm, descent_order, relevant
    from mslr_hip.mirt_mod import POGM
    from mslr_hip.sense import SenseOp
    cfg, smaps, ksp, omega, sf = prepared("small")
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    lams = O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    S = len(cfg["patch_sizes"])
    A = SenseOp(omega, smaps)
    kw = dict(patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=lams, L=S * 1.0)
    cache = tmp_path / "tuning.json"
    print("autotune small", flush=True)
    r = autotune_pogm(A, ksp, form="mixed", warmup=1, iters=2, cache=cache, **kw)
    cands = [c for c in descent_order() if relevant(c[0], A.N, A.Nt, cfg["patch_sizes"])]
    assert not r["cached"] and len(r["trials"]) == 1 + len(cands)
    # every candidate shape gives the built-in iterate to rounding; the dense-W and
    # FP32 applies round their products differently from the built-in apply (which
    # multiplies the split pass's block-exponent words): both stay within the split
    # form's bound of the FP64 result, so they differ by at most about twice it
    for t, ms, err, ok in r["trials"]:
        tol = 5e-5 if t["svst_apply"] or t["apply_fp32"] else 1e-5
        assert ms > 0 and err < tol, (t, err)
    assert shapes.tuning() == r["tuning"]
    kept = [t for t, _, _, ok in r["trials"][1:] if ok]
    assert r["tuning"] == (dict(BUILTIN, **kept[-1]) if kept else BUILTIN)
    # a second call on the same workload reads the cache and sets the same shapes
    shapes.set_tuning()
    r2 = autotune_pogm(A, ksp, form="mixed", cache=cache, **kw)
    assert r2["cached"] and r2["tuning"] == r["tuning"] and shapes.tuning() == r["tuning"]
    # a run under the tuned shapes against one under the built-in shapes
    out = []
    for t in (r["tuning"], BUILTIN):
        shapes.set_tuning(**t)
        st = POGM(A, niter=4, form="mixed", **kw)
        st.init(ksp)
        st.run(4)
        out.append((st.X.to_host(), st.log()))
        st.close()
    A.close()
    assert out[0][1] == out[1][1]
    tol = 5e-5 if r["tuning"]["svst_apply"] or r["tuning"]["apply_fp32"] else 1e-5
    assert rel_err(out[0][0], out[1][0]) < tol


@pytest.mark.parametrize("form", ["plain", "batched"])
def test_gradient_frame_groups(shapes, form):
    """The gradient and update in frame groups (sense_chunks; the batched form alternates
    them between two streams) give the one-group iterates: the same restarts and X to
    rounding of the restart test's FP64 sums only."""
    from mslr_hip.mirt_mod import POGM
    from mslr_hip.sense import SenseOp
    cfg, smaps, ksp, omega, sf = prepared("c0")
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    lams = O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    A = SenseOp(omega, smaps)
    out = {}
    for nch in (1, 3, 7):
        shapes.set_tuning(sense_chunks=nch)
        st = POGM(A, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=lams, L=2.0,
                  niter=5, form=form, log_costs=True)
        st.init(ksp)
        st.run(5)
        out[nch] = (st.X.to_host(), st.log()[1], st.costs())
        st.close()
    for nch in (3, 7):
        assert out[nch][1] == out[1][1]
        assert np.array_equal(out[nch][0], out[1][0])
        np.testing.assert_allclose(out[nch][2][0], out[1][2][0], rtol=1e-12)


@pytest.mark.parametrize("name", ["c0", "small_odd"])
def test_update_lines(shapes, name):
    """The update fused with the gradient's transpose, 8 / 16 / 32 yz-lines per workgroup
    against the built-in 64 (c0: 16-byte accesses, even Nx; small_odd: Nx = 15): the
    elementwise results are bit-identical and the restart tests agree."""
    from mslr_hip.mirt_mod import POGM
    from mslr_hip.sense import SenseOp
    cfg, smaps, ksp, omega, sf = prepared(name)
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    lams = O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    A = SenseOp(omega, smaps)
    out = {}
    for lines in (0, 8, 16, 32):
        shapes.set_tuning(update_lines=lines)
        st = POGM(A, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=lams, L=2.0,
                  niter=5, form="plain", log_costs=True)
        st.init(ksp)
        st.run(5)
        out[lines] = (st.X.to_host(), st.log()[1], st.costs())
        st.close()
    A.close()
    for lines in (8, 16, 32):
        assert out[lines][1] == out[0][1]
        assert np.array_equal(out[lines][0], out[0][0])
        np.testing.assert_allclose(out[lines][2][0], out[0][2][0], rtol=1e-12)


def test_update_writes_split_words(shapes):
    """Mixed form: with 64 or 32 yz-lines per update workgroup at Nx = 12 (768 and 384
    voxels, whole 128-row blocks) the update writes the whole-image patch's split words
    itself and its SVST runs no split pass; with 16 or 8 (192, 96 voxels) the SVST splits
    z.  The words are the same bits either way, so the iterates, restarts and costs are
    identical."""
    from mslr_hip.mirt_mod import POGM
    from mslr_hip.sense import SenseOp
    cfg, smaps, ksp, omega, sf = prepared("small")
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    assert N[0] == 12 and cfg["patch_sizes"][0] == list(N)
    lams = O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    A = SenseOp(omega, smaps)
    out = {}
    for lines in (0, 32, 16, 8):
        shapes.set_tuning(update_lines=lines)
        st = POGM(A, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=lams, L=2.0,
                  niter=6, form="mixed", log_costs=True)
        st.init(ksp)
        st.run(6)
        out[lines] = (st.X.to_host(), st.log()[1], st.costs())
        st.close()
    A.close()
    for lines in (32, 16, 8):
        assert out[lines][1] == out[0][1]
        assert np.array_equal(out[lines][0], out[0][0]), lines
        np.testing.assert_allclose(out[lines][2], out[0][2], rtol=1e-12)


def test_compiled_variants_are_bit_identical(shapes, rng):
    """The alternative compiled kernels change the launch shape only: the y-z FFT passes
    with 512 / 768 / 1024 threads at 90 x 90 x 60 (forward and adjoint bit-identical), and the
    four global-band schedules of the tridiagonalisation on 387 x 387 Gram matrices
    (bit-identical SVST output)."""
    from mslr_hip.device import DeviceArray
    from mslr_hip import recon, synth
    from mslr_hip.sense import SenseOp
    N, Nt, Nc = (90, 90, 60), 3, 4
    A = SenseOp(synth.masks(N, Nt, 8, 3)[0], synth.coil_maps(N, Nc, 3))
    x = (rng.standard_normal((*N, Nt)) + 1j * rng.standard_normal((*N, Nt))).astype(np.complex64)
    xd = DeviceArray.from_host(x)
    ref_f = A.forward(xd).to_host()
    ref_a = A.adjoint(ref_f).to_host()
    for fw, ad in [(768, 512), (1024, 1024), (512, 768)]:
        shapes.set_tuning(yz_threads_fwd=fw, yz_threads_adj=ad)
        f = A.forward(xd).to_host()
        assert np.array_equal(f, ref_f)
        assert np.array_equal(A.adjoint(ref_f).to_host(), ref_a)
    A.close()
    Z = np.asfortranarray(np.stack([lowrank(rng, 1000, 387, 9, noise=0.1) for _ in range(3)], axis=2))
    s = np.linalg.svd(Z[:, :, 0].astype(np.complex128), compute_uv=False)
    beta = 0.5 * (s[8] + s[9])
    outs = []
    for sched in range(4):
        shapes.set_tuning(sb2st_schedule=sched)
        outs.append(recon.svst_batched(Z, beta).to_host())
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    ref = np.stack([O.SVST(Z[:, :, i].astype(np.complex128), float(np.float32(beta))) for i in range(3)], axis=2)
    assert rel_err(outs[0], ref) < 2e-5


def test_fused_lowrank_apply_is_bit_identical(shapes, rng):
    """Split BF16: patches keeping r <= 64 singular values (and r <= m/2) reconstruct
    Y = (X U_r) f U_r^H in one kernel that holds X U_r on chip (svst_apply 0); with
    svst_apply 2 the two products go through memory.  The same products in the same
    order give the same bits: ranks either side of 64 and of m/2 = 150 in one batch, and
    a mixed POGM run whose whole-image patch is written straight into the image."""
    from mslr_hip import recon
    from mslr_hip.mirt_mod import POGM
    from mslr_hip.sense import SenseOp
    ranks = [0, 1, 2, 33, 64, 65, 149, 150, 151, 299]
    P, betas = mixed_rank_batch(rng, ranks)
    shapes.set_products("split_bf16")
    out = {}
    for a in (0, 2):
        shapes.set_tuning(svst_apply=a)
        out[a] = recon.svst_batched(P, betas).to_host()
    assert np.array_equal(out[0], out[2])
    for i, k in enumerate(ranks):
        Yo = O.SVST(P[:, :, i].astype(np.complex128), np.float32(betas[i]))
        assert rel_err(out[0][:, :, i], Yo) < 1e-4, (i, k)
    cfg, smaps, ksp, omega, sf = prepared("small")
    N, Nt = tuple(cfg["N"]), cfg["Nt"]
    lams = O.lambdas_ong_lustig(N, Nt, cfg["patch_sizes"])
    A = SenseOp(omega, smaps)
    runs = {}
    for a in (0, 2):
        shapes.set_tuning(svst_apply=a)
        st = POGM(A, patch_sizes=cfg["patch_sizes"], strides=cfg["strides"], lambdas=lams, L=2.0,
                  niter=6, form="mixed", log_costs=True)
        st.init(ksp)
        st.run(6)
        runs[a] = (st.X.to_host(), st.log()[1], st.costs())
        st.close()
    A.close()
    assert runs[2][1] == runs[0][1]
    assert np.array_equal(runs[2][0], runs[0][0])
    np.testing.assert_allclose(runs[2][2], runs[0][2], rtol=1e-12)
