import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X / gfx950)")


def _has_gpu() -> bool:
    try:
        import sih26_amd
        return sih26_amd.gpu_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Every `gpu` test requires a device; without one it fails rather than skipping, so a
    GPU run can never pass silently on a CPU-only path."""
    if not _has_gpu():
        pytest.fail("no HIP device visible: `-m gpu` tests need the MI355X")
    import sih26_amd
    return sih26_amd


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    return oracle
