import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "faster-local-solver-for-gdes_amd")
for p in (PKG, REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: full-size parity against the oracle (tens of seconds of host work)")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def small_graph():
    from flsg import graph
    return graph.powerlaw_graph_small()


@pytest.fixture(scope="session")
def large_graph():
    from flsg import graph
    return graph.powerlaw_graph_large()
