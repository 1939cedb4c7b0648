"""Shared pytest configuration.

Tests marked ``gpu`` need an AMD Instinct GPU (gfx950) and the built
extension; everything else runs on the CPU against the oracle in ``oracle/``.
"""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP extension")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as module

    module.build()
    return module


@pytest.fixture(scope="session")
def backend_module():
    from gpu_simulation_amd import native

    return native.load_extension()
