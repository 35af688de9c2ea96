import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "city-scale-nlos-acoustic-engine-code_amd")
GOLDEN = os.path.join(REPO, "tests", "golden")
for p in (PKG, os.path.join(REPO, "tests"), GOLDEN):
    if p not in sys.path:
        sys.path.insert(0, p)

import nlos_engine as nlos  # noqa: E402,F401


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built engine library")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    if not os.path.exists(nlos.LIB_PATH):
        pytest.fail("engine library not built: " + nlos.LIB_PATH)
    return nlos.load_library()
