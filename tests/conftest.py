import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "mslr-recon_amd"))
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope="session")
def ctx():
    from mslr_hip.device import default_context
    return default_context()


@pytest.fixture
def rng():
    return np.random.default_rng(1234)


def crandn(rng, shape, dtype=np.complex64):
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return np.asfortranarray((a / np.sqrt(2)).astype(dtype))


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
