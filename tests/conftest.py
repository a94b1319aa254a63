import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
if str(ROOT / "tests") not in sys.path:
    sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def torch_dev():
    """torch on the GPU: device buffers and streams only."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("a gpu-marked test ran where no GPU is visible")
    return torch
