"""The C-ABI library loads and exports every function include/*.h declares (no GPU needed)."""

import ctypes
import glob
import os
import re

import pytest

from fast_pauli import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(fp_hip_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_hot_path():
    names = declared_functions()
    for required in ["fp_hip_string_apply_batch", "fp_hip_string_expectation", "fp_hip_op_apply",
                     "fp_hip_op_expectation", "fp_hip_summed_expectation", "fp_hip_ops_create"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_the_header():
    assert set(declared_functions()) <= set(_abi._SIGS)


def test_errors_without_touching_compute():
    # argument validation happens before any device work
    lib = _abi.lib()
    assert lib.fp_hip_op_apply(None, 0, None, None, 4, 1, 0, None) == _abi.FP_ERR_INVALID_ARGUMENT
    assert "NULL" in _abi.last_error()
    with pytest.raises(ValueError):
        _abi.check(_abi.FP_ERR_INVALID_ARGUMENT)
    assert b"gfx950" in lib.fp_hip_version()
