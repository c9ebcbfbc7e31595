import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def _has_gpu():
    try:
        from stablebear_amd import _lib

        return _lib.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: GPU tests must find a HIP device (fail loudly otherwise)."""
    if not _has_gpu():
        pytest.fail("no HIP device visible to a gpu-marked test")
    return True


@pytest.fixture(params=["pcf32", "pcf64"])
def pcf_dtype(request):
    import stablebear_amd as sb

    return getattr(sb, request.param)


def np_dtype_of(pcf_dtype):
    import stablebear_amd as sb

    return np.float32 if pcf_dtype is sb.pcf32 else np.float64
