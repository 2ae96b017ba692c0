"""Shared fixtures and markers.

Fixture systems follow the reference suite's conftest (tests/conftest.py:11-88
of maroba/bhut) so the API tests read like the reference's own.
"""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: longer-running tests")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture
def simple_system():
    positions = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]], dtype=np.float64)
    masses = np.array([1.0, 1.0], dtype=np.float64)
    return positions, masses


@pytest.fixture
def random_system():
    np.random.seed(42)
    positions = np.random.randn(10, 3).astype(np.float64)
    masses = np.ones(10, dtype=np.float64)
    return positions, masses


@pytest.fixture
def hierarchical_system():
    positions = np.array([[-0.1, 0.0, 0.0], [0.1, 0.0, 0.0], [10.0, 0.0, 0.0]], dtype=np.float64)
    masses = np.array([1.0, 1.0, 0.1], dtype=np.float64)
    return positions, masses


@pytest.fixture
def planar_system():
    np.random.seed(123)
    positions = np.random.randn(8, 3).astype(np.float64)
    positions[:, 2] = 0.0
    masses = np.ones(8, dtype=np.float64)
    return positions, masses


@pytest.fixture
def varied_masses_system():
    np.random.seed(456)
    positions = np.random.randn(8, 3).astype(np.float64)
    masses = np.array([0.1, 1.0, 10.0, 0.5, 2.0, 0.01, 5.0, 0.2], dtype=np.float64)
    return positions, masses


def golden_files():
    d = os.path.join(ROOT, "tests", "golden")
    return sorted(os.path.join(d, f) for f in os.listdir(d) if f.startswith("golden_") and f.endswith(".npz"))
