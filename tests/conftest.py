import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: full-size parity cases")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O  # oracle/oracle.py

    O.lib()
    return O


@pytest.fixture(scope="session")
def sample_graph():
    from starplat_amd import graph

    return graph(os.path.join(GOLDEN, "sample_graph.txt")).parseGraph()
