"""Shared fixtures. ``gpu``-marked tests need an AMD GPU and the built libspp_hip.so."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
DATA = ROOT / "tests" / "golden" / "data"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config) -> None:
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP kernels)")


def _gpu_available() -> bool:
    try:
        import torch
    except ImportError:
        return False
    return bool(torch.cuda.is_available())


def pytest_collection_modifyitems(config, items) -> None:
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def datadir() -> Path:
    return DATA


@pytest.fixture(scope="session")
def filfiles() -> dict[str, object]:
    return {
        "1bit": str(DATA / "parkes_1bit.fil"),
        "2bit": str(DATA / "parkes_2bit.fil"),
        "4bit": str(DATA / "parkes_4bit.fil"),
        "8bit": [str(DATA / "parkes_8bit_1.fil"), str(DATA / "parkes_8bit_2.fil")],
        "tutorial": str(DATA / "tutorial.fil"),
    }


@pytest.fixture(scope="session")
def rng() -> np.random.Generator:
    return np.random.default_rng(42)
