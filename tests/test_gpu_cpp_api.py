"""Runs the C++ host API checks (tests/cpp/test_host_api.cpp) on the GPU."""

import os
import subprocess

import pytest

from tests import hiputil as H

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_host_api")


def test_cpp_host_api():
    assert os.path.exists(BIN), "build the C++ host API test first (fast-pauli_amd/build.py)"
    H.say("C++ host API program")
    proc = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(proc.stdout)
    assert proc.returncode == 0, proc.stdout + proc.stderr
