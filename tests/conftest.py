import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the HIP kernels")


@pytest.fixture(scope="session")
def gpu():
    import cudart_amd as ca

    n = ca.device_count()
    if n < 1:
        raise RuntimeError("no HIP device visible to a gpu-marked test")
    return 0
