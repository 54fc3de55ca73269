"""mslr_sense_forward_noalloc (the allocation-free SENSE forward) against
mslr_sense_forward, through the C ABI: bit-identical output on two calls in a row,
the second on a smaller frame range, and the size limit reported before any work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan(N, Nt, Nc, R, seed):
    from mslr_hip import synth
    from mslr_hip.sense import SenseOp
    smaps = synth.coil_maps(N, Nc, seed)
    mask, _ = synth.masks(N, Nt, R, seed)
    return SenseOp(mask, smaps)


def test_noalloc_forward_equals_forward():
    from mslr_hip.device import DeviceArray
    N, Nt, Nc = (90, 90, 60), 8, 18
    A = _plan(N, Nt, Nc, 8, 11)
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((*N, Nt)) + 1j * rng.standard_normal((*N, Nt))).astype(np.complex64)
    # call 1: all frames (one frame needs ~12.6 MB, so several chunks of the scratch)
    xd = DeviceArray.from_host(x)
    y_ref = A.forward(xd).to_host()
    out = DeviceArray((A.K, Nc, Nt), np.complex64)
    out.upload(np.full((A.K, Nc, Nt), np.nan, dtype=np.complex64))
    y_na = A.forward_noalloc(xd, out).to_host()
    assert np.array_equal(y_na, y_ref)
    # call 2: a smaller input, frames 2..4 (the scratch still holds call 1's data)
    xs = DeviceArray.from_host(np.asfortranarray(x[..., 2:5]))
    y_ref2 = A.forward(xs, t0=2).to_host()
    out2 = DeviceArray((A.K, Nc, 3), np.complex64)
    out2.upload(np.full((A.K, Nc, 3), np.nan, dtype=np.complex64))
    y_na2 = A.forward_noalloc(xs, out2, t0=2).to_host()
    assert np.array_equal(y_na2, y_ref2)
    assert np.array_equal(y_na2, y_ref[:, :, 2:5])
    A.close()


def test_noalloc_forward_size_limit():
    """A frame whose image plus sampled lines exceed the 64 MiB scratch is refused
    (MSLR_ERR_UNSUPPORTED) before any kernel is queued."""
    from mslr_hip import _lib
    from mslr_hip.device import DeviceArray
    from mslr_hip.sense import SenseOp
    N, Nc = (128, 64, 128), 8
    rng = np.random.default_rng(2)
    smaps = (rng.standard_normal((*N, Nc)) + 1j * rng.standard_normal((*N, Nc))).astype(np.complex64)
    mask = np.ones((*N, 1), dtype=bool)  # full sampling: Nc * V line values (64 MiB) + the image
    A = SenseOp(mask, smaps)
    xd = DeviceArray((*N, 1), np.complex64, zero=True)
    out = DeviceArray((A.K, Nc, 1), np.complex64)
    with pytest.raises(_lib.MSLRUnsupported, match="allocation-free"):
        A.forward_noalloc(xd, out)
    A.close()
