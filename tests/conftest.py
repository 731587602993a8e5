import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle_lib import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def lib_f():
    import scarf3d_amd

    return scarf3d_amd.load("f")


@pytest.fixture(scope="session")
def lib_d():
    import scarf3d_amd

    return scarf3d_amd.load("d")


@pytest.fixture(autouse=True)
def _finalize_after(request):
    """Every test leaves the library finalized (one process-global object)."""
    yield
    for name in ("lib_f", "lib_d"):
        if name in request.fixturenames:
            request.getfixturevalue(name).finalize()
