"""Test configuration: the `gpu` marker, import paths, shared fixtures."""
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
for p in (REPO / "oracle", REPO / "speo-gpu-single-control_amd", REPO):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def speo():
    import speo_de
    speo_de.lib()
    return speo_de


@pytest.fixture(scope="session")
def cuda(speo):
    """The GPU, with the HIP library checked to load and to see a gfx950."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    speo.lib()
    rc = speo.lib().speo_device_check()
    assert rc == 0, speo.last_error()
    return torch.device("cuda:0")
