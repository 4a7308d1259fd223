import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: long CPU-side checks")


@pytest.fixture(scope="session")
def ctx():
    """One device context shared by the GPU tests (one process, one card)."""
    from parallax_amd import splat as S

    ok, name, reason = S.availability()
    if not ok:
        pytest.fail(f"no usable gfx950 device: {reason}")
    c = S.Context(0)
    yield c
    c.close()
