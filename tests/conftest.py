import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context (device 0, own stream) shared by the GPU tests."""
    from seismod_amd.solver import Context
    c = Context(int(os.environ.get("SEISMOD_DEVICE", "0")))
    yield c
    c.close()
