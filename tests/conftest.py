import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (deselect with '-m \"not gpu\"')")


def _gpu_count():
    try:
        import torch
        return torch.cuda.device_count()
    except Exception:
        return 0


@pytest.fixture(scope="session")
def gpu():
    """Skips unless a HIP device is visible; fails if the HIP libraries are
    missing on a GPU host (the product path must not silently degrade)."""
    if _gpu_count() < 1:
        pytest.skip("no HIP device")
    import harness
    if not harness.have("hip"):
        pytest.fail("HIP libraries not built (python -m simwave_amd.build)")
    return True
