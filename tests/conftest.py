import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")


def _gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Loads libsifv_hip.so; GPU tests fail loudly when it is missing."""
    if not _gpu_available():
        pytest.skip("no GPU in this process")
    from semiimplicitfv_amd import _abi

    return _abi.load_library()


@pytest.fixture(scope="session")
def oracle_c():
    from oracle import pyoracle

    if not pyoracle.available("c"):
        assert pyoracle.build(ref=False) == 0
    return pyoracle


@pytest.fixture(scope="session")
def oracle_ref(oracle_c):
    if not oracle_c.available("ref"):
        pytest.skip("compiled reference (oracle/_ref) not built")
    return oracle_c
