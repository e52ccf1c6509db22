import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

from mlfma_pmchwt_amd import geometry  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU test")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def relerr(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))


@pytest.fixture(scope="session")
def sphere_small():
    """480-RWG icosphere, radius 0.5 lambda (one MLFMA level)."""
    return geometry.icosphere(2, 0.5)


@pytest.fixture(scope="session")
def sphere_ml():
    """480-RWG icosphere, radius 1.5 lambda (tree depth 4: translations on levels 2..4)."""
    return geometry.icosphere(2, 1.5)


def seeded_x(n, seed=12345):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)
