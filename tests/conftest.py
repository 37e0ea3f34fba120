"""pytest configuration: import paths and the `gpu` marker."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fast-pauli_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")
    config.addinivalue_line("markers", "slow: long-running CPU case")


@pytest.fixture(scope="session")
def hip():
    """The ctypes C-ABI binding, initialised on device 0 (GPU tests only)."""
    from fast_pauli import _abi

    _abi.init(0)
    return _abi
