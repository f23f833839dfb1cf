import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "tests"), os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def eng():
    """The GPU engine; fails loudly when the device layer is unusable."""
    from teide_amd.td import engine
    e = engine()
    assert e.L.td_dev_available() == 1, "GPU / HIP device layer unavailable"
    return e


@pytest.fixture(scope="session")
def ref_engine():
    """The reference engine built from source (oracle/_ref), when present."""
    from teide_amd.td import Engine
    path = os.path.join(REPO, "oracle", "_ref", "libteide_ref.so")
    if not os.path.exists(path):
        pytest.skip("reference build not present (oracle/_ref)")
    return Engine(path, device=False)
