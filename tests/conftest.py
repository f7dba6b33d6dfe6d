import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the engine")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def ico5():
    """Level-5 icosphere (10,242 V / 20,480 F) as float32/int32."""
    from lagrange_amd import workloads as w
    return w.icosphere(32).f32()


@pytest.fixture(scope="session")
def open_mesh():
    """Displaced sphere (nu=48, 46,080 F) with 5% of facets deleted: fractional winding numbers."""
    from lagrange_amd import workloads as w
    return w.displaced_sphere(48, 2, amplitude=0.15, delete_fraction=0.05).f32()


def rel_err_w(a, b):
    return np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)) / (4 * np.pi)
