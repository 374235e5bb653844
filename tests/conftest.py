import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def device():
    """The GPU. A GPU test fails (does not skip) when no device is visible."""
    import zig_ornament_amd as zo
    n = zo.Device.count()
    assert n > 0, "no HIP device visible: GPU tests must run on a GPU host"
    d = zo.Device(0)
    name, cus = d.info()
    print(f"\n[device] {name}, {cus} CUs", flush=True)
    yield d
    d.close()


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    oracle_lib.build()
    return oracle_lib
