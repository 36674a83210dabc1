import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def pytest_terminal_summary(terminalreporter, exitstatus, config):
    """The parity tests' near-tie margins and X differences, after the run."""
    from tests.parity import REPORT
    if REPORT:
        terminalreporter.section("parity margins and X differences")
        for line in REPORT:
            terminalreporter.write_line(line)


def _gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no HIP device visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def librbrp():
    """The C-ABI library, built in-tree first if it is missing (CPU hosts only)."""
    from robust_blockwise_random_pivoting_amd import _capi
    if not os.path.exists(_capi.LIB_PATH):
        if _gpu_available():
            raise RuntimeError("librbrp.so missing on a GPU host: build it before running tests")
        from robust_blockwise_random_pivoting_amd import build_lib
        build_lib.build()
    return _capi.lib()


@pytest.fixture(scope="session")
def hooks(librbrp):
    """The kernel test hooks (include/rbrp_testing.h), a library of their own."""
    from robust_blockwise_random_pivoting_amd import _capi
    return _capi.testing_lib()
