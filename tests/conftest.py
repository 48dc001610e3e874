import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "oracle"), os.path.join(REPO, "open-qmin_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

REFERENCE_CHECKOUT = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return np.load(os.path.join(REPO, "tests", "golden", "reference_vectors.npz"))


@pytest.fixture(scope="session")
def phase():
    import pyoracle as po
    A, B, C = po.rescaled_phase()
    return A, B, C, po.uniaxial_s0(A, B, C)
