import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle_py

    oracle_py.build()
    return oracle_py


@pytest.fixture(scope="session")
def mesh_c1():
    """Config-1 mesh: 2,500 seeded points in a 10 x 10 km square."""
    from chm_amd import meshgen

    return meshgen.synthetic_mesh(2500, seed=0)


@pytest.fixture(scope="session")
def mesh_small():
    from chm_amd import meshgen

    return meshgen.synthetic_mesh(300, seed=3)


@pytest.fixture(scope="session")
def mesh_tiny():
    """A handful of triangles (dense solves are exact references here)."""
    from chm_amd import meshgen

    return meshgen.synthetic_mesh(12, seed=5, extent=2000.0)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)
