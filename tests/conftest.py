import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "asc-cpp_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def _gpu_problem():
    """None when a HIP device is usable, else a description of why not."""
    try:
        import asc_sparse_hip as H
        n = H.device_count()
        return None if n > 0 else "asc_sparse_hip_device_count() returned 0"
    except Exception as e:  # report the real cause
        return f"{type(e).__name__}: {e}"


@pytest.fixture(scope="session")
def ctx():
    import asc_sparse_hip as H
    problem = _gpu_problem()
    if problem:
        pytest.fail(f"no usable HIP device ({problem}): GPU tests need an MI355X")
    c = H.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def vctx():
    """Validation-only context: every check that needs no GPU."""
    import asc_sparse_hip as H
    c = H.Context(-1)
    yield c
    c.close()
