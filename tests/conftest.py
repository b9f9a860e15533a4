import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs librbn's HIP kernels)")
    config.addinivalue_line("markers", "slow: longer CPU-side oracle runs")


@pytest.fixture(scope="session")
def gpu_available():
    from lagrangianrbnudging_amd import rbn

    return rbn.device_count() > 0


@pytest.fixture(autouse=True)
def _require_gpu(request):
    if request.node.get_closest_marker("gpu") is not None:
        from lagrangianrbnudging_amd import rbn

        if rbn.device_count() == 0:
            pytest.fail("GPU test run without a visible GPU (librbn has no CPU fallback)")
    yield
