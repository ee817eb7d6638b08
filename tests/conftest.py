import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: long CPU-oracle cases")


@pytest.fixture(scope="session")
def tracer():
    import hdcodex_amd as hx

    t = hx.PathTracer(0)
    yield t
    t.close()
