"""Shared pytest configuration: the `gpu` marker and common helpers."""
import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(REPO / "sph-code-simulation_amd"))
sys.path.insert(0, str(REPO / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (only to count devices)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")
