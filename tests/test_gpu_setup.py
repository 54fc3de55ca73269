"""The setup of run_recon on the device (scripts/reconstruct.jl:86-118,145-147) against
the oracle's prepare: coil-map normalisation and the k-space flatten bit-exact, the mask
exact, the scale factor (0.99 quantile of |sense_comb|) within 1e-6, the division of the
k-space by it bit-exact, the reference's assertion messages, and odd image sizes."""
import numpy as np
import pytest

import oracle as O
from problems import raw

pytestmark = pytest.mark.gpu

NAMES = ["small", "small_odd", "c0"]


@pytest.mark.parametrize("name", NAMES)
def test_normalise_smaps_bit_exact(name):
    from mslr_hip.reconstruct import normalise_smaps
    cfg, ksp0, smaps_raw = raw(name)
    assert np.array_equal(normalise_smaps(smaps_raw), O.normalise_smaps(smaps_raw))


def test_normalise_smaps_random_maps_bit_exact(rng):
    from conftest import crandn
    from mslr_hip.reconstruct import normalise_smaps
    s = crandn(rng, (7, 5, 3, 11)) * np.float32(3.7)
    assert np.array_equal(normalise_smaps(s), O.normalise_smaps(s))


@pytest.mark.parametrize("name", NAMES)
def test_flatten_and_mask_bit_exact(name):
    from mslr_hip.reconstruct import flatten_kspace
    cfg, ksp0, smaps_raw = raw(name)
    N, Nt, Nc = tuple(cfg["N"]), cfg["Nt"], cfg["Nc"]
    omega, yd = flatten_kspace(ksp0, Nc, Nt)
    om_ref = ksp0[:, :, :, 0, :] != 0
    assert np.array_equal(omega, om_ref)
    flat = ksp0.reshape(-1, Nc, Nt, order="F")
    y_ref = np.stack([flat[O.sample_index(om_ref[..., t]), :, t] for t in range(Nt)], axis=2)
    assert np.array_equal(yd.to_host(), y_ref)


@pytest.mark.parametrize("name", NAMES)
def test_scale_factor_and_division(name):
    from mslr_hip.reconstruct import flatten_kspace, normalise_smaps, scale_kspace, sense_comb_scale
    from mslr_hip.sense import SenseOp
    cfg, ksp0, smaps_raw = raw(name)
    N, Nt, Nc = tuple(cfg["N"]), cfg["Nt"], cfg["Nc"]
    smaps_o, ksp_o, omega_o, R_o, sf_o = O.prepare(ksp0, smaps_raw, N, Nc, Nt)
    omega, yd = flatten_kspace(ksp0, Nc, Nt)
    A = SenseOp(omega, normalise_smaps(smaps_raw))
    sf = sense_comb_scale(A, yd)
    assert sf == pytest.approx(sf_o, rel=1e-6)
    scale_kspace(yd, sf_o)  # the oracle's factor: the division itself is exact
    assert np.array_equal(yd.to_host(), ksp_o)
    A.close()


def test_prepare_device_matches_oracle():
    from mslr_hip.reconstruct import prepare_device
    cfg, ksp0, smaps_raw = raw("c0")
    N, Nt, Nc = tuple(cfg["N"]), cfg["Nt"], cfg["Nc"]
    smaps_o, ksp_o, omega_o, R_o, sf_o = O.prepare(ksp0, smaps_raw, N, Nc, Nt)
    A, smaps, yd, sf, R, omega = prepare_device(ksp0, smaps_raw, N, Nc, Nt)
    assert np.array_equal(omega, omega_o) and np.array_equal(smaps, smaps_o)
    assert R == R_o and A.K == ksp_o.shape[0]
    assert sf == pytest.approx(sf_o, rel=1e-6)
    np.testing.assert_allclose(yd.to_host(), ksp_o, rtol=2e-6, atol=0)
    A.close()


def test_assertions_of_the_reference():
    """scripts/reconstruct.jl:113-118: the first coil with another pattern, then the first
    frame with another sample count, with the reference's messages (1-based)."""
    from mslr_hip.reconstruct import flatten_kspace
    cfg, ksp0, smaps_raw = raw("small")
    Nt, Nc = cfg["Nt"], cfg["Nc"]
    bad = ksp0.copy()
    i5 = np.argwhere(bad[:, :, :, 0, 4] != 0)[0]
    i8 = np.argwhere(bad[:, :, :, 0, 7] != 0)[0]
    bad[i5[0], i5[1], i5[2], 2, 4] = 0             # coil 3 loses a sample of frame 5
    bad[i8[0], i8[1], i8[2], 1, 7] = 0             # ...and coil 2 one of frame 8
    with pytest.raises(AssertionError, match="^Coil 2 has a different sampling pattern$"):
        flatten_kspace(bad, Nc, Nt)
    bad = ksp0.copy()
    extra = np.argwhere(bad[:, :, :, 0, 3] == 0)[0]
    bad[extra[0], extra[1], extra[2], :, 3] = 1.0   # frame 4 gains a sample in every coil
    with pytest.raises(AssertionError, match="^Frame 4 has a different sample count$"):
        flatten_kspace(bad, Nc, Nt)
