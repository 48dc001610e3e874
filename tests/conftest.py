import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "oracle"), os.path.join(REPO, "open-qmin_amd"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)

REFERENCE_CHECKOUT = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")
    markexpr = config.getoption("markexpr", "") or ""
    if "gpu" in markexpr and "not gpu" not in markexpr:
        # torch before libqmin.so, so that the tests that hand torch tensors or
        # streams to the library find one HIP runtime (qmin.check_torch_runtime_order)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    return np.load(os.path.join(REPO, "tests", "golden", "reference_vectors.npz"))


@pytest.fixture(scope="session")
def phase():
    import pyoracle as po
    A, B, C = po.rescaled_phase()
    return A, B, C, po.uniaxial_s0(A, B, C)


def example_directors(shape):
    """(nx, ny, nz, S0) per site, in site order, of the example director function
    in the reference's setInitialConditions.h (openQmin -z 3): director
    (x, x + y + cos z, 0) normalized as normalizeDirector does (inc/functions.h:
    193-205), S0 = 0.53 + 1e-10 (x + y + z).  math.cos is the C library's cos,
    as in the reference's build and in openQmin-hip's."""
    import math
    import numpy as np
    lx, ly, lz = shape
    out = np.zeros((lx * ly * lz, 4))
    i = 0
    for z in range(lz):
        for y in range(ly):
            for x in range(lx):
                nx, ny, nz = float(x), x + y + math.cos(z), 0.0
                a = math.sqrt(nx * nx + ny * ny + nz * nz)
                if a == 0:
                    nx, ny, nz = 1.0, 0.0, 0.0
                else:
                    nx, ny, nz = nx / a, ny / a, nz / a
                out[i] = (nx, ny, nz, 0.53 + 0.0000000001 * (x + y + z))
                i += 1
    return out
