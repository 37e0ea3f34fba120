"""The host-side coset-tile table builder (fast-pauli_amd/csrc/fp_tiles_host.cpp)
under the address and undefined-behaviour sanitizers: every image the kernels
would read, for random, weight-limited, diagonal, shared-x and duplicated string
sets at the kernels' geometries, checked for layout, basis form, coverage,
coefficients, local coordinates and the kernels' sign split
(tests/cpp/test_tile_tables.cpp).  Host code only: no GPU."""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "test_tile_tables")


def test_tile_table_builder_sanitized():
    if not os.path.exists(BIN):
        sys.path.insert(0, os.path.join(os.path.dirname(HERE), "fast-pauli_amd"))
        import build as B

        B.build_tile_test()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=300, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert " 0 failures" in p.stdout, p.stdout
