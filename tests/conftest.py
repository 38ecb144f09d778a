"""Shared test setup: the `gpu` marker, the oracle and the product module."""

from __future__ import annotations

import importlib
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o

    o.lib()
    return o


@pytest.fixture(scope="session")
def ndic():
    """The product module (loads libndic_hip.so; every call runs on the GPU)."""
    return importlib.import_module("nd-image-codecs_amd")


@pytest.fixture(scope="session")
def golden():
    return ROOT / "tests" / "golden"
