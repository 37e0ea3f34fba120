"""CPU oracle for the operator-application hot path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker or the timed CPU
baseline.  The product (``fast-pauli_amd``) never routes through it.

Two independent restatements live here:

* ``libfp_oracle.so`` (``fp_oracle.cpp``): the reference's C++/OpenMP loops
  (qognitive/fast-pauli ``fast_pauli/cpp/include/__pauli_string.hpp:49-538``,
  ``__pauli_op.hpp:413-549``, ``__summed_pauli_op.hpp:287-614``) with the same
  schedules and reduction order.  It is also the timed CPU baseline.
* ``dense_*`` below: the brute-force Kronecker construction the reference's
  own tests use as ground truth (``fast_pauli/pypauli/helpers.py:40-85``,
  ``fast_pauli/cpp/tests/test_pauli_string.cpp:75-113``), written with numpy.

Parity status of every function is recorded in tests/test_oracle.py; none is
unpinned.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
import threading

import math

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfp_oracle.so")
_lib = None
_lock = threading.Lock()

C128, C64 = 0, 1

PAULI_CODES = {"I": 0, "X": 1, "Y": 2, "Z": 3}


def build() -> str:
    """Compile the C++ oracle with its Makefile; return the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    """Load (building on first use when absent) the C++ oracle."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(_LIB_PATH):
                build()
            L = ctypes.CDLL(_LIB_PATH)
            vp, sz, u8p = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p
            i, d = ctypes.c_int, ctypes.c_double
            L.fpo_max_threads.restype = i
            L.fpo_set_threads.argtypes = [i]
            L.fpo_sparse_repr.argtypes = [sz, u8p, vp, vp]
            L.fpo_string_apply.argtypes = [i, i, sz, u8p, d, d, vp, vp]
            L.fpo_string_apply_batch.argtypes = [i, i, sz, u8p, d, d, vp, vp, sz]
            L.fpo_string_expectation.argtypes = [i, i, sz, u8p, d, d, vp, vp, sz]
            L.fpo_op_apply.argtypes = [i, i, sz, sz, u8p, vp, vp, vp, sz]
            L.fpo_op_expectation.argtypes = [i, i, sz, sz, u8p, vp, vp, vp, sz]
            L.fpo_summed_expectation.argtypes = [i, i, sz, sz, sz, u8p, vp, vp, vp, sz]
            L.fpo_summed_apply.argtypes = [i, i, sz, sz, sz, u8p, vp, vp, vp, sz]
            L.fpo_summed_apply_weighted.argtypes = [i, i, sz, sz, sz, u8p, vp, vp, vp, vp, sz]
            L.fpo_strings_max_weight.argtypes = [sz, sz, u8p]
            L.fpo_strings_max_weight.restype = sz
            L.fpo_string_mul.argtypes = [sz, u8p, u8p, u8p, vp]
            L.fpo_string_mul.restype = None
            L.fpo_summed_square.argtypes = [i, sz, sz, sz, u8p, vp, sz, u8p, vp]
            L.fpo_summed_square.restype = sz
            L.fpo_op_mul.argtypes = [sz, sz, u8p, vp, sz, u8p, vp, u8p, vp]
            L.fpo_op_mul.restype = sz
            for f in (L.fpo_sparse_repr, L.fpo_string_apply, L.fpo_string_apply_batch,
                      L.fpo_string_expectation, L.fpo_op_apply, L.fpo_op_expectation,
                      L.fpo_summed_expectation, L.fpo_summed_apply,
                      L.fpo_summed_apply_weighted, L.fpo_set_threads):
                f.restype = None
            _lib = L
    return _lib


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------

def codes_of(strings) -> np.ndarray:
    """(n_strings, n_qubits) uint8 letter codes, character 0 first."""
    if isinstance(strings, str):
        strings = [strings]
    strings = list(strings)
    if not strings:
        return np.zeros((0, 0), dtype=np.uint8)
    n = len(strings[0])
    out = np.empty((len(strings), n), dtype=np.uint8)
    for r, s in enumerate(strings):
        if len(s) != n:
            raise ValueError("all strings must have the same length")
        out[r] = [PAULI_CODES[ch] for ch in s]
    return out


def _cdtype(dtype: int):
    return np.complex128 if dtype == C128 else np.complex64


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _check(a: np.ndarray, dtype) -> np.ndarray:
    if a.dtype != dtype or not a.flags.c_contiguous:
        raise TypeError(f"expected C-contiguous {np.dtype(dtype)} array")
    return a


# ---------------------------------------------------------------------------
# C++ oracle wrappers (accumulate into `out`, like the reference C++ API)
# ---------------------------------------------------------------------------

def sparse_repr(string: str) -> tuple[np.ndarray, np.ndarray]:
    """get_sparse_repr (__pauli_string.hpp:49-118): (k, m) for one string."""
    codes = codes_of(string)[0] if string else np.zeros(0, np.uint8)
    n = codes.size
    dim = 1 << n if n else 0
    k = np.zeros(dim, np.uint64)
    m = np.zeros(dim, np.complex128)
    lib().fpo_sparse_repr(n, _ptr(codes), _ptr(k), _ptr(m))
    return k, m


def string_apply(string, c, psi, out, par=False, dtype=C128):
    codes = codes_of(string)[0]
    dt = _cdtype(dtype)
    _check(psi, dt), _check(out, dt)
    assert psi.shape == out.shape == (1 << codes.size,)
    lib().fpo_string_apply(dtype, int(par), codes.size, _ptr(codes), complex(c).real,
                           complex(c).imag, _ptr(psi), _ptr(out))
    return out


def string_apply_batch(string, c, states, out, par=False, dtype=C128):
    codes = codes_of(string)[0]
    dt = _cdtype(dtype)
    _check(states, dt), _check(out, dt)
    assert states.shape == out.shape and states.shape[0] == 1 << codes.size
    lib().fpo_string_apply_batch(dtype, int(par), codes.size, _ptr(codes), complex(c).real,
                                 complex(c).imag, _ptr(states), _ptr(out), states.shape[1])
    return out


def string_expectation(string, c, states, out, par=False, dtype=C128):
    codes = codes_of(string)[0]
    dt = _cdtype(dtype)
    _check(states, dt), _check(out, dt)
    assert states.shape[0] == 1 << codes.size and out.shape == (states.shape[1],)
    lib().fpo_string_expectation(dtype, int(par), codes.size, _ptr(codes), complex(c).real,
                                 complex(c).imag, _ptr(states), _ptr(out), states.shape[1])
    return out


def op_apply(strings, coeffs, states, out, par=False, dtype=C128):
    codes = codes_of(strings)
    dt = _cdtype(dtype)
    coeffs = np.ascontiguousarray(coeffs, dtype=dt)
    _check(states, dt), _check(out, dt)
    assert states.shape == out.shape and states.shape[0] == 1 << codes.shape[1]
    lib().fpo_op_apply(dtype, int(par), codes.shape[1], codes.shape[0], _ptr(codes), _ptr(coeffs),
                       _ptr(states), _ptr(out), states.shape[1])
    return out


def op_expectation(strings, coeffs, states, out, par=False, dtype=C128):
    codes = codes_of(strings)
    dt = _cdtype(dtype)
    coeffs = np.ascontiguousarray(coeffs, dtype=dt)
    _check(states, dt), _check(out, dt)
    lib().fpo_op_expectation(dtype, int(par), codes.shape[1], codes.shape[0], _ptr(codes),
                             _ptr(coeffs), _ptr(states), _ptr(out), states.shape[1])
    return out


def summed_expectation(strings, coeffs_nk, states, out, par=False, dtype=C128):
    codes = codes_of(strings)
    dt = _cdtype(dtype)
    coeffs_nk = np.ascontiguousarray(coeffs_nk, dtype=dt)
    _check(states, dt), _check(out, dt)
    n_ops = coeffs_nk.shape[1]
    assert out.shape == (n_ops, states.shape[1])
    lib().fpo_summed_expectation(dtype, int(par), codes.shape[1], codes.shape[0], n_ops,
                                 _ptr(codes), _ptr(coeffs_nk), _ptr(states), _ptr(out),
                                 states.shape[1])
    return out


def summed_apply(strings, coeffs_nk, states, out, par=False, dtype=C128):
    codes = codes_of(strings)
    dt = _cdtype(dtype)
    coeffs_nk = np.ascontiguousarray(coeffs_nk, dtype=dt)
    _check(states, dt), _check(out, dt)
    lib().fpo_summed_apply(dtype, int(par), codes.shape[1], codes.shape[0], coeffs_nk.shape[1],
                           _ptr(codes), _ptr(coeffs_nk), _ptr(states), _ptr(out),
                           states.shape[1])
    return out


def summed_apply_weighted(strings, coeffs_nk, states, data, out, par=False, dtype=C128):
    codes = codes_of(strings)
    dt = _cdtype(dtype)
    coeffs_nk = np.ascontiguousarray(coeffs_nk, dtype=dt)
    data = np.ascontiguousarray(data, dtype=np.float64 if dtype == C128 else np.float32)
    _check(states, dt), _check(out, dt)
    lib().fpo_summed_apply_weighted(dtype, int(par), codes.shape[1], codes.shape[0],
                                    coeffs_nk.shape[1], _ptr(codes), _ptr(coeffs_nk),
                                    _ptr(states), _ptr(data), _ptr(out), states.shape[1])
    return out


def strings_max_weight(n: int, max_weight: int) -> list[str]:
    """calculate_pauli_strings_max_weight (__pauli_helpers.hpp:147-153) as strings."""
    count = lib().fpo_strings_max_weight(n, max_weight, None)
    codes = np.zeros((count, n), np.uint8)
    lib().fpo_strings_max_weight(n, max_weight, _ptr(codes))
    letters = np.array(list("IXYZ"))
    return ["".join(letters[row]) for row in codes]


def string_mul(a: str, b: str) -> tuple[complex, str]:
    """PauliString * PauliString (__pauli_string.hpp:228-247): (phase, product)."""
    ca, cb = codes_of(a)[0], codes_of(b)[0]
    out = np.zeros_like(ca)
    ph = np.zeros(2, np.float64)
    lib().fpo_string_mul(ca.size, _ptr(ca), _ptr(cb), _ptr(out), _ptr(ph))
    return complex(ph[0], ph[1]), "".join("IXYZ"[c] for c in out)


def summed_square(strings, coeffs_nk, dtype=C128) -> tuple[list[str], np.ndarray]:
    """SummedPauliOp::square (__summed_pauli_op.hpp:197-268): the squared strings (in
    calculate_pauli_strings_max_weight order) and their (A, K) coefficients."""
    codes = codes_of(strings)
    N, n = codes.shape
    dt = _cdtype(dtype)
    coeffs_nk = np.ascontiguousarray(coeffs_nk, dtype=dt)
    K = coeffs_nk.shape[1]
    # the squared strings are every string of weight <= min(n, 2 * max weight): count them
    # here so that the (costly) loop runs once; the call returns the count it produced
    W = min(n, 2 * int((codes != 0).sum(axis=1).max(initial=0)))
    A = sum(3 ** w * math.comb(n, w) for w in range(W + 1))
    out_codes = np.zeros((A, n), np.uint8)
    out = np.zeros((A, K), dt)
    got = lib().fpo_summed_square(dtype, n, N, K, _ptr(codes), _ptr(coeffs_nk), A, _ptr(out_codes), _ptr(out))
    assert got == A, (got, A)
    letters = np.array(list("IXYZ"))
    return ["".join(letters[row]) for row in out_codes], out


def op_mul(strings1, c1, strings2, c2) -> tuple[list[str], np.ndarray]:
    """PauliOp * PauliOp (__pauli_op.hpp:237-281) on one thread: merged product strings in
    order of first occurrence and their complex128 coefficients."""
    a, b = codes_of(strings1), codes_of(strings2)
    c1 = np.ascontiguousarray(c1, np.complex128)
    c2 = np.ascontiguousarray(c2, np.complex128)
    n = a.shape[1]
    args = (n, a.shape[0], _ptr(a), _ptr(c1), b.shape[0], _ptr(b), _ptr(c2))
    count = lib().fpo_op_mul(*args, None, None)
    out_codes = np.zeros((count, n), np.uint8)
    out = np.zeros(count, np.complex128)
    lib().fpo_op_mul(*args, _ptr(out_codes), _ptr(out))
    letters = np.array(list("IXYZ"))
    return ["".join(letters[row]) for row in out_codes], out


def set_threads(n: int) -> None:
    lib().fpo_set_threads(int(n))


def max_threads() -> int:
    return int(lib().fpo_max_threads())


# ---------------------------------------------------------------------------
# Dense Kronecker brute force (ground truth for small n)
# ---------------------------------------------------------------------------

PAULI_MATRICES = {
    "I": np.array([[1, 0], [0, 1]], dtype=np.complex128),
    "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "Z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def dense_string(string: str) -> np.ndarray:
    """sigma_{s0} (x) sigma_{s1} (x) ... : character 0 is the most significant factor."""
    if not string:
        return np.zeros((0, 0), dtype=np.complex128)
    mat = np.ones((1, 1), dtype=np.complex128)
    for ch in string:
        mat = np.kron(mat, PAULI_MATRICES[ch])
    return mat


def dense_op(coeffs, strings) -> np.ndarray:
    strings = list(strings)
    acc = np.zeros_like(dense_string(strings[0]))
    for c, s in zip(coeffs, strings):
        acc = acc + complex(c) * dense_string(s)
    return acc
