import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def rng():
    return np.random.default_rng(20261202)


def random_density(nbf, nocc, seed):
    """D = C C^T with C (nbf x nocc) seeded Gaussian, orthonormalised (QR)."""
    r = np.random.default_rng(seed)
    C, _ = np.linalg.qr(r.normal(size=(nbf, nocc)))
    return np.asfortranarray(C @ C.T)
