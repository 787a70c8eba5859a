import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libsqbem kernels)")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


@pytest.fixture(scope="session")
def sq():
    """The libsqbem Python binding, bound to device 0."""
    import __graft_entry__ as ge

    # torch ships its own HIP runtime; it has to initialise the device before libsqbem's
    # runtime does, so that both see it (the tests pass torch device tensors to the C ABI)
    try:
        import torch

        if torch.cuda.is_available():
            torch.cuda.init()
    except ImportError:
        pass
    pkg = ge.load_package()
    pkg.init(0)
    return pkg


@pytest.fixture(scope="session")
def sqhost():
    """The binding without touching the GPU (host-only entry points)."""
    import __graft_entry__ as ge

    pkg = ge.load_package()
    pkg.load()
    return pkg


@pytest.fixture(scope="session")
def O():
    import oracle

    oracle.lib()
    return oracle
