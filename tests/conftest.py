import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
SMALL_MESH = os.path.join(GOLDEN, "small_mesh")  # data copy of meshes/small_mesh


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsem_hip.so)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.set_threads(min(8, os.cpu_count() or 1))
    return oracle


@pytest.fixture(scope="session")
def sem():
    import schrodingersem_amd
    return schrodingersem_amd


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def structured(n, box=(0.0, 1.0, 0.0, 1.0), jitter=0.0, seed=0):
    """Seeded structured mesh through the product's host-only generator."""
    import schrodingersem_amd as s
    return s.structured_mesh(n, box, jitter, seed)
