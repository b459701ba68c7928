"""Shared test set-up: the `gpu` marker, the repo root on sys.path, and
helpers that rebuild the golden problems from their stored inputs."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def load_golden(name: str):
    return np.load(GOLDEN / name)


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (n if n > 0 else 1.0)


@pytest.fixture(scope="session")
def gpu_available():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return True
