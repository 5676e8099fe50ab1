"""Shared pytest configuration: ``gpu`` tests are skipped where no HIP device is visible."""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def _has_gpu() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


HAS_GPU = _has_gpu()


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X / gfx950)")
    config.addinivalue_line("markers", "slow: long-running")


def pytest_collection_modifyitems(config: pytest.Config, items: list) -> None:
    if HAS_GPU:
        return
    skip = pytest.mark.skip(reason="needs a HIP GPU")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
