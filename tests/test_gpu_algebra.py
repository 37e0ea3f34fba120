"""SummedPauliOp::square on the device (fp_hip_summed_square) against the oracle's
restatement of the reference's loop (__summed_pauli_op.hpp:197-268) and against dense
matrix products.

Bars: the squared strings are the oracle's (calculate_pauli_strings_max_weight order) and
the coefficients are IEEE-equal to the oracle's (np.array_equal: the device sums each
(string, operator) over its pairs in the reference's (i, j) order with the same complex
multiplies), in complex128 and complex64; A_k^2 equals the dense product to 1e-12.
"""

import numpy as np
import pytest

import fast_pauli as fp
from fast_pauli import _abi
from fast_pauli import synthetic as syn
from oracle import oracle as O
from tests import hiputil as H

pytestmark = pytest.mark.gpu

C128, C64 = _abi.FP_C128, _abi.FP_C64


def device_square(codes, coeffs_nk, dtype):
    import ctypes

    ops = _abi.Ops(codes, matrix=np.asarray(coeffs_nk, np.complex128))
    count = ctypes.c_size_t()
    _abi.call("fp_hip_summed_square", ops.handle, dtype, None, ctypes.byref(count), None)
    K = coeffs_nk.shape[1]
    esz = 16 if dtype == C128 else 8
    d = _abi.DeviceBuffer(max(count.value * K * esz, 16))
    _abi.call("fp_hip_summed_square", ops.handle, dtype, d.ptr, ctypes.byref(count), None)
    return d.to_array(H.CDT[dtype], (count.value, K))


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("n,w,K", [(1, 1, 1), (2, 2, 3), (3, 2, 2), (4, 2, 4), (6, 1, 5), (5, 3, 2)])
def test_square_all_strings_ieee_equal(hip, dtype, n, w, K):
    strings = O.strings_max_weight(n, w)
    H.say(f"square all strings n={n} w={w} K={K} dtype={dtype} N={len(strings)}")
    rng = np.random.default_rng(n * 100 + w * 10 + K)
    c = rng.standard_normal((len(strings), K)) + 1j * rng.standard_normal((len(strings), K))
    got = device_square(O.codes_of(strings), c, dtype)
    sq, exp = O.summed_square(strings, c.astype(H.CDT[dtype]), dtype=O.C128 if dtype == C128 else O.C64)
    assert got.shape == exp.shape
    assert np.array_equal(got, exp)
    if dtype == C128 and n <= 4:
        for k in range(K):
            a = O.dense_op(c[:, k], strings)
            np.testing.assert_allclose(O.dense_op(got[:, k], sq), a @ a, atol=1e-12)


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("n,N,K,seed", [(5, 12, 3, 1), (8, 40, 7, 2), (10, 500, 16, 3), (12, 120, 4, 4)])
def test_square_random_strings_ieee_equal(hip, dtype, n, N, K, seed):
    """Random strings (duplicates allowed) of every letter, and weight <= 3 strings at larger n."""
    codes = syn.random_codes(N, n, seed=seed) if n <= 8 else syn.random_codes_max_weight(N, n, 3, seed=seed)
    strings = syn.codes_to_strings(codes)
    H.say(f"square random n={n} N={N} K={K} dtype={dtype}")
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((N, K)) + 1j * rng.standard_normal((N, K))
    got = device_square(codes, c, dtype)
    sq, exp = O.summed_square(strings, c.astype(H.CDT[dtype]), dtype=O.C128 if dtype == C128 else O.C64)
    assert np.array_equal(got, exp)
    if dtype == C128 and n <= 5:
        for k in range(K):
            a = O.dense_op(c[:, k], strings)
            np.testing.assert_allclose(O.dense_op(got[:, k], sq), a @ a, atol=1e-11)


def test_square_size_query_and_errors(hip):
    import ctypes

    H.say("square size query and errors")
    ops = _abi.Ops(O.codes_of(["XI", "IZ"]))
    count = ctypes.c_size_t()
    with pytest.raises(ValueError):  # no coefficient matrix attached
        _abi.call("fp_hip_summed_square", ops.handle, C128, None, ctypes.byref(count), None)
    ops.set_matrix(np.ones((2, 3), np.complex128))
    _abi.call("fp_hip_summed_square", ops.handle, C128, None, ctypes.byref(count), None)
    assert count.value == len(O.strings_max_weight(2, 2))


# ---------------------------------------------------------------------------
# The Python surface's SummedPauliOp.square (test_summed_pauli_op.py:227-241)
# ---------------------------------------------------------------------------
def test_surface_square_matches_dense_and_oracle(hip):
    rng = np.random.default_rng(7)
    strings = fp.helpers.calculate_pauli_strings_max_weight(3, 1)
    N, K = len(strings), 2
    c = rng.standard_normal((N, K)) + 0j
    op = fp.SummedPauliOp(strings, c)
    sq = op.square()
    dense, dsq = op.to_tensor(), sq.to_tensor()
    for k in range(K):
        np.testing.assert_allclose(dsq[k], dense[k] @ dense[k], atol=1e-12)
    names, exp = O.summed_square([str(s) for s in strings], c)
    assert [str(s) for s in sq.pauli_strings] == names
    assert np.array_equal(sq.coeffs.T, exp)


def test_surface_square_random_strings(hip):
    """Random strings with every letter (test_summed_pauli_op.py:227-241 with random sets)."""
    rng = np.random.default_rng(11)
    n, N, K = 5, 12, 3
    letters = np.array(list("IXYZ"))
    strings = sorted({"".join(rng.choice(letters, n)) for _ in range(N)})
    c = rng.standard_normal((len(strings), K)) + 1j * rng.standard_normal((len(strings), K))
    op = fp.SummedPauliOp(strings, c)
    sq = op.square()
    dense, dsq = op.to_tensor(), sq.to_tensor()
    for k in range(K):
        np.testing.assert_allclose(dsq[k], dense[k] @ dense[k], atol=1e-12)
    _, exp = O.summed_square(strings, c)
    assert np.array_equal(sq.coeffs.T, exp)
