import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP kernels")


@pytest.fixture(scope="session")
def gpu():
    from lightbaker_amd.gpu import Gpu
    g = Gpu()
    print(f"GPU: {g.adapter}", flush=True)
    yield g
    g.close()
