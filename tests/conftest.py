"""Shared fixtures: the package (imported by path: its name starts with a
digit), the CPU oracle (test infrastructure) and GPU availability."""
from __future__ import annotations

import importlib
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests" / "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the GPU)")


@pytest.fixture(scope="session")
def gs():
    return importlib.import_module("3dgs_tile_webgpu_amd")


@pytest.fixture(scope="session")
def oracle():
    import gs_oracle

    gs_oracle.lib()
    return gs_oracle


@pytest.fixture(scope="session")
def gpu_available():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no GPU is visible")
    return True
