import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def small():
    from chsga_amd import small_workload
    return small_workload()


@pytest.fixture(scope="session")
def c1():
    from chsga_amd import WORKLOADS
    w = WORKLOADS["c1"]
    ds = w.dataset()
    return w, ds, w.readset(ds)


@pytest.fixture(scope="session")
def c5_ds():
    """The c5 graph (100 Mbp backbone, ~117M chars), generated once per session."""
    from chsga_amd import WORKLOADS
    w = WORKLOADS["c5"]
    return w, w.dataset()
