import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "hbp-isp_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _gpu_available():
    try:
        import torch  # noqa: F401  (device count only; does not initialise HIP)
        return torch.cuda.device_count() > 0
    except Exception:
        return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def ctx():
    from hbp_amd import lib
    c = lib.Ctx(0)
    yield c
    c.close()


def pytest_collection_modifyitems(session, config, items):
    # the pair generator's worker processes are forked now, before any test initialises HIP
    if any(item.fspath.basename == "test_gpu_workloads.py" for item in items):
        import pairgen
        pairgen.start()


def pytest_sessionfinish(session, exitstatus):
    import sys as _sys
    if "pairgen" in _sys.modules:
        _sys.modules["pairgen"].stop()
