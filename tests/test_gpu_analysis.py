"""tSNR and detrend! (src/analysis.jl) on the device against the oracle, and the
closed forms at the production shape (90x90x60, Nt=300)."""
import numpy as np
import pytest

import oracle as O
from conftest import crandn, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(12, 12, 10, 10), (7, 5, 3, 33), (4, 4, 4, 1), (90, 3, 2, 300)])
def test_tsnr_matches_oracle(ctx, rng, shape):
    from mslr_hip.analysis import tSNR
    img = (crandn(rng, shape, np.complex64) + 2.0).astype(np.complex64)
    got = tSNR(img).to_host()
    want = O.tSNR(img)
    assert got.shape == want.shape
    if shape[-1] == 1:
        assert np.all(np.isnan(got)) and np.all(np.isnan(want))
    else:
        assert rel_err(got, want) < 1e-6


@pytest.mark.parametrize("shape", [(12, 12, 10, 10), (7, 5, 3, 33), (90, 3, 2, 300)])
def test_detrend_matches_oracle(ctx, rng, shape):
    from mslr_hip.analysis import detrend_
    img = crandn(rng, shape, np.complex64)
    got = detrend_(img).to_host()
    want = O.detrend(img)
    assert rel_err(got, want) < 1e-6


def test_analysis_at_production_shape(ctx, rng):
    """90x90x60, Nt=300: a per-voxel ramp a + b t detrends to a; tSNR of |a|-constant
    series is |a| / eps."""
    from mslr_hip.analysis import detrend_, tSNR
    from mslr_hip.device import DeviceArray
    N, Nt = (90, 90, 60), 300
    a = crandn(rng, N, np.complex64)
    b = (0.01 * crandn(rng, N, np.complex64)).astype(np.complex64)
    t = np.arange(1, Nt + 1, dtype=np.float32)
    img = np.asfortranarray(a[..., None] + b[..., None] * t)
    d = DeviceArray.from_host(img, ctx)
    detrend_(d)
    out = d.to_host()
    err = np.abs(out - a[..., None]).max()
    assert err < 1e-4 * (np.abs(a).max() + np.abs(b).max() * Nt)
    phase = np.exp(1j * np.pi / 2 * (np.arange(Nt) % 4)).round().astype(np.complex64)
    m = tSNR(np.asfortranarray(np.abs(a)[..., None] * phase)).to_host()
    assert rel_err(m, np.abs(a) / O.EPS32) < 1e-6
