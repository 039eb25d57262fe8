"""pytest configuration: the `gpu` marker and the two backends.

Tests parametrised with the `backend` fixture run once against the CPU
oracle (always) and once against the gfx950 library (marked `gpu`, run on an
MI355X).  The oracle is the checker; the library is the product.
"""
from __future__ import annotations

import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
for p in (REPO, REPO / "fluid-simulator_amd", REPO / "oracle", REPO / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(params=["oracle", pytest.param("gpu", marks=pytest.mark.gpu)])
def backend(request):
    import backends

    if request.param == "oracle":
        return backends.OracleBackend()
    return backends.GpuBackend()


@pytest.fixture(scope="session")
def oracle64():
    from oracle import Oracle

    return Oracle("f64")


@pytest.fixture(scope="session")
def gpu():
    import backends

    return backends.GpuBackend()
