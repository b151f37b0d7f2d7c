import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "asc-cpp_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def _has_gpu():
    try:
        import asc_sparse_hip as H
        return H.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    import asc_sparse_hip as H
    if not _has_gpu():
        pytest.fail("no HIP device visible: GPU tests must run on a GPU box")
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
