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
    from lightbaker_amd.gpu import Gpu, Pass
    g = Gpu()
    g.defaults = {p: g.tuning(p) for p in Pass}
    print(f"GPU: {g.adapter}", flush=True)
    yield g
    g.close()


@pytest.fixture(autouse=True)
def _library_defaults(request):
    """Every test that uses the GPU starts from the library's default form and tuning, whatever
    the tests before it set."""
    g = request.getfixturevalue("gpu") if "gpu" in request.fixturenames else None
    if g is not None:
        for p, t in g.defaults.items():
            g.set_tuning(pass_=p, **t)
        g.set_form(0)
    yield
