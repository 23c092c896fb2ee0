import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

CURVE_L = 9.017203500515143  # length of r = 1 + 0.3 cos(5 theta)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def box_problem(n, seed=0):
    """Config-0 geometry: n sources in [0,1]^2, n targets in [2,3]x[0,1]."""
    import oracle as orc
    orc.seed(seed)
    src = orc.sample_box(n, 0, 1, 0, 1)
    tgt = orc.sample_box(n, 2, 3, 0, 1)
    return src, tgt, 32 * np.pi


def curve_problem(n, ppw=10):
    """Config-1 geometry: n sources on r = 1 + 0.3 cos(5 theta), n targets on
    the circle of radius 3, k for `ppw` points per wavelength on the curve."""
    th = 2 * np.pi * np.arange(n) / n
    r = 1 + 0.3 * np.cos(5 * th)
    src = np.stack([r * np.cos(th), r * np.sin(th)], 1)
    tgt = np.stack([3 * np.cos(th), 3 * np.sin(th)], 1)
    return src, tgt, 2 * np.pi * n / (ppw * CURVE_L)


@pytest.fixture(scope="session")
def gpu_lib():
    import butterfly_amd as bf
    L = bf.lib()
    if L.bfhipDeviceCount() < 1:
        pytest.fail("no HIP device visible to the library")
    return bf
