"""GPU dense and sparse materialisation through the C ABI against the oracle.

fp_hip_string_sparse_repr, fp_hip_string_to_tensor, fp_hip_op_to_tensor and
fp_hip_summed_to_tensor restate get_sparse_repr and the three to_tensor helpers
(__pauli_string.hpp:49-118, 549-560; __pauli_op.hpp:558-571;
__summed_pauli_op.hpp:644-665).  Bar: bitwise equality (signed zeros included)
with the oracle's get_sparse_repr and the reference's loops written over it in
numpy, in complex128 and complex64; and IEEE equality with dense Kronecker
products of the 2x2 Pauli matrices, which pins the matrices independently of
the oracle.
"""

import functools

import numpy as np
import pytest

from fast_pauli import _abi
from fast_pauli import synthetic as syn
from oracle import oracle as O
from tests import hiputil as H

pytestmark = pytest.mark.gpu

C128, C64 = _abi.FP_C128, _abi.FP_C64
PAULI = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.diag([1, -1])}


def bits(a):
    """Bit patterns of a complex array (so that -0.0 differs from 0.0)."""
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.complex128 else np.uint32)


def kron(s):
    return functools.reduce(np.kron, [PAULI[c] for c in s], np.ones((1, 1))).astype(np.complex128)


def random_strings(n, N, seed):
    return syn.codes_to_strings(syn.random_codes(N, n, seed=seed))


def sparse_repr_dev(ops, idx, dtype):
    dim = ops.dim
    cols = _abi.DeviceBuffer(dim * 8)
    vals = _abi.DeviceBuffer(dim * np.dtype(H.CDT[dtype]).itemsize)
    _abi.call("fp_hip_string_sparse_repr", ops.handle, idx, dtype, cols.ptr, vals.ptr, dim, None)
    return cols.to_array(np.uint64, (dim,)), vals.to_array(H.CDT[dtype], (dim,))


@pytest.mark.parametrize("dtype", [C128, C64])
def test_sparse_repr_matches_oracle(hip, dtype):
    cdt = H.CDT[dtype]
    strings = ["X", "Y", "Z", "I", "XYZXYZ", "YYYY", "ZIXIZYXX"] + random_strings(12, 4, seed=5)
    for s in strings:
        H.say(f"sparse_repr {s} dtype={dtype}")
        ops = _abi.Ops(O.codes_of([s]))
        cols, vals = sparse_repr_dev(ops, 0, dtype)
        k, m = O.sparse_repr(s)
        assert np.array_equal(cols, k), s
        assert np.array_equal(bits(vals), bits(m.astype(cdt))), s


@pytest.mark.parametrize("dtype", [C128, C64])
def test_string_to_tensor_assigns_nonzeros_only(hip, dtype):
    cdt = H.CDT[dtype]
    for s in ["XYZ", "IIII", "YZYX", "ZXIY"] + random_strings(7, 3, seed=9):
        H.say(f"string_to_tensor {s} dtype={dtype}")
        ops = _abi.Ops(O.codes_of(["I" * len(s), s]))
        dim = 1 << len(s)
        sentinel = np.full((dim, dim), 7 - 3j, cdt)
        d = _abi.DeviceBuffer.from_array(sentinel)
        _abi.call("fp_hip_string_to_tensor", ops.handle, 1, dtype, d.ptr, dim, None)
        got = d.to_array(cdt, (dim, dim))
        k, m = O.sparse_repr(s)
        exp = sentinel.copy()
        exp[np.arange(dim), k.astype(np.int64)] = m.astype(cdt)
        assert np.array_equal(bits(got), bits(exp)), s
        # the dense matrix is the Kronecker product of the letters
        z = np.zeros((dim, dim), cdt)
        dz = _abi.DeviceBuffer.from_array(z)
        _abi.call("fp_hip_string_to_tensor", ops.handle, 1, dtype, dz.ptr, dim, None)
        assert np.array_equal(dz.to_array(cdt, (dim, dim)), kron(s).astype(cdt)), s


def op_expected(strings, coeffs, out0, cdt):
    """PauliOp::to_tensor (__pauli_op.hpp:558-571) over the oracle's sparse repr."""
    out = out0.astype(cdt).copy()
    dim = out.shape[0]
    for s, c in zip(strings, coeffs):
        k, m = O.sparse_repr(s)
        out[np.arange(dim), k.astype(np.int64)] += cdt(c) * m.astype(cdt)
    return out


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("accumulate", [0, 1])
def test_op_to_tensor(hip, dtype, accumulate):
    cdt = H.CDT[dtype]
    rng = np.random.default_rng(321)
    n = 6
    base = random_strings(n, 12, seed=3)
    # strings that share an x mask (same off-diagonal pattern) hit the same elements
    strings = base + ["XIIIIY", "YIIIIX", "XIIIIX", "ZZZZZZ", "IIIIII", "ZIZIZI"]
    coeffs = rng.normal(size=len(strings)) + 1j * rng.normal(size=len(strings))
    coeffs[3] = 0.5  # a real coefficient (signed zeros in the products)
    coeffs[4] = -2j
    dim = 1 << n
    out0 = (rng.normal(size=(dim, dim)) + 1j * rng.normal(size=(dim, dim))) if accumulate else np.zeros((dim, dim))
    H.say(f"op_to_tensor n={n} N={len(strings)} dtype={dtype} accumulate={accumulate}")
    ops = _abi.Ops(O.codes_of(strings), coeffs)
    # a non-zero start shows that accumulate=0 clears the output
    start = out0 if accumulate else np.full((dim, dim), 5 + 5j)
    d = _abi.DeviceBuffer.from_array(np.ascontiguousarray(start, cdt))
    _abi.call("fp_hip_op_to_tensor", ops.handle, dtype, d.ptr, dim, accumulate, None)
    got = d.to_array(cdt, (dim, dim))
    exp = op_expected(strings, coeffs, out0, cdt)
    assert np.array_equal(bits(got), bits(exp))
    if not accumulate and dtype == C128:
        dense = sum(c * kron(s) for s, c in zip(strings, coeffs))
        np.testing.assert_allclose(got, dense, rtol=0, atol=1e-13)


@pytest.mark.parametrize("dtype", [C128, C64])
def test_summed_to_tensor(hip, dtype):
    cdt = H.CDT[dtype]
    rng = np.random.default_rng(18)
    n, K = 5, 3
    strings = random_strings(n, 9, seed=11) + ["XXXXX", "YXXXX"]
    N = len(strings)
    h = rng.normal(size=(N, K)) + 1j * rng.normal(size=(N, K))
    h[:, 1] = h[:, 1].real  # real column: zero imaginary parts
    dim = 1 << n
    H.say(f"summed_to_tensor n={n} N={N} K={K} dtype={dtype}")
    ops = _abi.Ops(O.codes_of(strings), matrix=h)
    for accumulate in (0, 1):
        out0 = rng.normal(size=(K, dim, dim)) + 1j * rng.normal(size=(K, dim, dim))
        d = _abi.DeviceBuffer.from_array(np.ascontiguousarray(out0, cdt))
        _abi.call("fp_hip_summed_to_tensor", ops.handle, dtype, d.ptr, dim, accumulate, None)
        got = d.to_array(cdt, (K, dim, dim))
        start = out0 if accumulate else np.zeros_like(out0)
        exp = np.stack([op_expected(strings, h[:, k], start[k], cdt) for k in range(K)])
        assert np.array_equal(bits(got), bits(exp)), accumulate


def test_dense_argument_errors(hip):
    ops = _abi.Ops(O.codes_of(["XYZ", "ZZI"]))
    buf = _abi.DeviceBuffer(64 * 16)
    lib = _abi.lib()
    E = _abi.FP_ERR_INVALID_ARGUMENT
    assert lib.fp_hip_op_to_tensor(ops.handle, C128, buf.ptr, 4, 0, None) == E  # wrong dimension
    assert "dimension" in _abi.last_error()
    assert lib.fp_hip_string_to_tensor(ops.handle, 2, C128, buf.ptr, 8, None) == E  # string index
    assert lib.fp_hip_summed_to_tensor(ops.handle, C128, buf.ptr, 8, 0, None) == E  # no matrix
    assert lib.fp_hip_op_to_tensor(ops.handle, 7, buf.ptr, 8, 0, None) == E  # dtype
    assert lib.fp_hip_string_sparse_repr(ops.handle, 0, C128, None, buf.ptr, 8, None) == E
