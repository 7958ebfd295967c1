import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
if str(ROOT / "oracle") not in sys.path:
    sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpm_gather.so)")


@pytest.fixture(scope="session")
def oracle():
    import pm_oracle

    pm_oracle.load()
    return pm_oracle


@pytest.fixture
def rng():
    return np.random.default_rng(12345)


def has_gpu() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skips nothing: a `gpu`-marked test that cannot see a GPU must fail."""
    import torch

    assert torch.cuda.is_available(), "gpu tests need a visible GPU"
    import photon_mapping_amd as pm

    return pm
