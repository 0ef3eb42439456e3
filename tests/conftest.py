"""Shared test setup: the `gpu` marker, import paths and on-demand builds.

CPU tests (`-m "not gpu"`) check the oracle against the reference's golden
vectors and independent brute force, and check the C-ABI library's exports
without calling into a GPU.  GPU tests (`-m gpu`) compare the HIP path with
the oracle through the C ABI.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long-running parity case")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def wrach_lib_path():
    """Path of libwrach_hip.so, building it first if it is absent."""
    from wrach_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from wrach_amd import build
        build.build()
    return _lib.LIB_PATH
