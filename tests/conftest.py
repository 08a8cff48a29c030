import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libttgpu.so")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    oracle_lib.build()
    return oracle_lib


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library; a GPU test fails (does not skip) when it is missing or no device is visible."""
    from naivetrace_amd import abi
    lib = abi.load_library()
    info = lib.ttgpu_device_info().decode()
    assert "gfx" in info, f"no HIP device: {info}"
    print("device:", info, flush=True)
    return lib
