import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device discovery only)
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def small_cloud():
    """A 48 x 14 NACA0012 O-cloud (672 points, far field at 4 chords), seeded."""
    from mfcfd_amd.cloud import naca_ogrid
    return naca_ogrid(48, 14, seed=3, first_spacing=0.01, far_radius=4.0)


@pytest.fixture(scope="session")
def c1_cloud():
    """The 160 x 60 NACA0012 cloud of the first-order FP64 workload (9,600 points)."""
    from mfcfd_amd import workloads as W
    return W.cloud_of(W.C1)


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.maximum(np.abs(b).max(), floor)
    return float(np.abs(a - b).max() / scale)
