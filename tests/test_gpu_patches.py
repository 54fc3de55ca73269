"""img2patches / patches2img on the GPU: bit-exact vs the oracle (src/recon.jl:53-126)."""
import numpy as np
import pytest

import oracle as O
from conftest import crandn, rel_err

pytestmark = pytest.mark.gpu

GEOMS = [
    ((12, 12, 10), (4, 4, 4), (4, 4, 4)),        # exact tiling, clipped last patch in z
    ((12, 12, 10), (12, 12, 10), (12, 12, 10)),  # global patch
    ((12, 12, 10), (1, 1, 1), (1, 1, 1)),        # voxel patches
    ((12, 12, 10), (6, 6, 6), (3, 3, 3)),        # half overlap
    ((10, 9, 7), (4, 5, 3), (3, 2, 2)),          # ragged overlap with edge clipping
    ((10, 9, 7), (20, 4, 3), (5, 6, 4)),         # patch larger than image -> clamped; stride > ps gaps
    ((30, 30, 20), (6, 6, 6), (3, 3, 3)),
]


@pytest.mark.parametrize("N,ps,ss", GEOMS)
def test_img2patches_bit_exact(ctx, rng, N, ps, ss):
    from mslr_hip import recon
    Nt = 5
    img = crandn(rng, (*N, Nt))
    P = recon.img2patches(img, ps, ss).to_host()
    Po = O.img2patches(img, ps, ss)
    assert P.shape == Po.shape
    assert np.array_equal(P, Po)


@pytest.mark.parametrize("N,ps,ss", GEOMS)
def test_patches2img_bit_exact(ctx, rng, N, ps, ss):
    from mslr_hip import recon
    Nt = 4
    psc, _, Np = O.patch_geometry(N, ps, ss)
    P = crandn(rng, (int(np.prod(psc)), Nt, Np))
    img = recon.patches2img(P, ps, ss, N).to_host()
    imgo = O.patches2img(P, ps, ss, N)
    assert np.array_equal(img, imgo)


@pytest.mark.parametrize("N,ps,ss", GEOMS)
def test_round_trip(ctx, rng, N, ps, ss):
    """patches2img(img2patches(x)) = x on covered voxels (exact for coverage 1, 1 ulp else), 0 elsewhere."""
    from mslr_hip import recon
    Nt = 3
    img = crandn(rng, (*N, Nt))
    back = recon.patches2img(recon.img2patches(img, ps, ss), ps, ss, N).to_host()
    cnt = O.coverage_count(N, ps, ss)
    cov = cnt > 0
    assert np.all(back[~cov] == 0)
    one = cnt == 1
    assert np.array_equal(back[one], img[one])
    np.testing.assert_allclose(back[cov], img[cov], rtol=4e-7, atol=0)


def test_geometry_matches_oracle():
    from mslr_hip import recon
    for N, ps, ss in GEOMS + [((90, 90, 60), (10, 10, 10), (10, 10, 10)),
                              ((90, 90, 60), (6, 6, 6), (3, 3, 3)),
                              ((90, 90, 60), (30, 30, 20), (15, 15, 10))]:
        psc, ns, Np = recon.patch_geometry(N, ps, ss)
        pso, nso, Npo = O.patch_geometry(N, ps, ss)
        assert list(psc) == pso and list(ns) == nso and Np == Npo


def test_production_patch_counts():
    """Patch counts quoted for the 90x90x60 configurations."""
    from mslr_hip import recon
    N = (90, 90, 60)
    assert recon.patch_geometry(N, (10, 10, 10), (10, 10, 10))[2] == 486
    assert recon.patch_geometry(N, (1, 1, 1), (1, 1, 1))[2] == 486000
    assert recon.patch_geometry(N, (30, 30, 20), (15, 15, 10))[2] == 125
    assert recon.patch_geometry(N, (10, 10, 10), (5, 5, 5))[2] == 3179
    assert recon.patch_geometry(N, (6, 6, 6), (3, 3, 3))[2] == 15979


def test_g_prox_all_scales_in_place(ctx, rng):
    """run_recon's g_prox over a 5D array (global, overlapping 4^3 with stride 2, 1x1x1)."""
    import oracle as O
    from mslr_hip.recon import g_prox
    from mslr_hip.device import DeviceArray
    N, Nt = (12, 12, 10), 10
    ps = [[12, 12, 10], [4, 4, 4], [1, 1, 1]]
    ss = [[12, 12, 10], [2, 2, 2], [1, 1, 1]]
    lams = [30.0, 6.0, 2.0]
    X = crandn(rng, (*N, Nt, 3), np.complex64)
    d = DeviceArray.from_host(X, ctx)
    g_prox(d, 0.5, lams, ps, ss)
    out = d.to_host()
    for k in range(3):
        want = O.patchSVST(X[..., k].astype(np.complex128), np.float32(0.5 * lams[k]), ps[k], ss[k],
                           dtype=np.complex128)
        assert rel_err(out[..., k], want) < 2e-5, k
