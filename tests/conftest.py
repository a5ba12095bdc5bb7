import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


# Tests of host code in libmsx.so that both suites run: the CPU suite
# (-m "not gpu") and, through the gpu-marked copy, the GPU suite, so that
# these golden-vector checks are recorded wherever the library is tested.
BOTH_SUITES = pytest.mark.parametrize("suite", ["cpu", pytest.param("gpu", marks=pytest.mark.gpu)])


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def ctx():
    """msx context on GPU 0.  GPU tests fail (never skip) when it cannot be made."""
    import multislice_amd as msx
    c = msx.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def max_rel(a, b):
    """Largest pointwise error on the reference's own scale: max |a - b| / max |b|
    (the image bar of 1e-4 "max relative error"; a pointwise ratio is
    undefined at the zeros of an intensity image)."""
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())
