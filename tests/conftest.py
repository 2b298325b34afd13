import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))

KPP_HOME = Path(os.environ.get("KPP_HOME", "/root/reference"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def kpp_home():
    if not (KPP_HOME / "models" / "small_strato.def").exists():
        pytest.skip("KPP distribution (KPP_HOME) not available")
    return KPP_HOME


def have_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not have_gpu():
        pytest.skip("no GPU")
    import torch
    return torch.device("cuda:0")
