"""The Python surface's compute methods on the GPU (through the C++ host API
and the C ABI), against the oracle — the reference's own Python test cases
(tests/fast_pauli/test_pauli_string.py, test_pauli_op.py,
test_summed_pauli_op.py) with the dense construction as ground truth, plus
bit-exact comparisons with the oracle's sequential loops and the complex64
fast path."""

import itertools as it

import numpy as np
import pytest

import fast_pauli as fp
from fast_pauli import synthetic as syn
from oracle import oracle as O
from tests import hiputil as H

pytestmark = pytest.mark.gpu

SAMPLE = (["I", "X", "Y", "Z"] + ["".join(p) for p in it.product("IXYZ", repeat=2)]
          + ["".join(p) for p in it.product("IXYZ", repeat=3)]
          + ["XYZXYZ", "ZZZIII", "XYIZXYZ", "XXIYYIZZ", "ZIXIZYXX"])


def rand_c(rng, *shape):
    return rng.random(shape) + 1j * rng.random(shape)


def test_string_apply_1d_and_batch():
    H.say("surface PauliString.apply")
    rng = np.random.default_rng(321)
    np.testing.assert_allclose(fp.PauliString("IXYZ").apply(np.zeros(16)), np.zeros(16), atol=1e-15)
    np.testing.assert_allclose(fp.PauliString("III").apply(np.arange(8)), np.arange(8), atol=1e-15)
    np.testing.assert_allclose(fp.PauliString("ZYX").apply(np.ones(8)), O.dense_string("ZYX").sum(1), atol=1e-15)
    np.testing.assert_allclose(fp.PauliString("ZYX").apply(np.eye(8)), O.dense_string("ZYX"), atol=1e-15)
    for s in SAMPLE:
        psi = rand_c(rng, 2 ** len(s))
        np.testing.assert_allclose(fp.PauliString(s).apply(psi), O.dense_string(s) @ psi, atol=1e-15)
        psis = rand_c(rng, 2 ** len(s), 42)
        got = fp.PauliString(s).apply(psis)
        np.testing.assert_allclose(got, O.dense_string(s) @ psis, atol=1e-15)
        assert np.array_equal(got, O.string_apply_batch(s, 1, psis, np.zeros_like(psis)))
    for s in ["".join(p) for p in it.product("IXYZ", repeat=4)]:
        coeff = rand_c(rng, 1)[0]
        psis = rand_c(rng, 16, 7)
        got = fp.PauliString(s).apply(psis, coeff)
        np.testing.assert_allclose(got, coeff * O.dense_string(s) @ psis, atol=1e-15)
        assert np.array_equal(got, O.string_apply_batch(s, coeff, psis, np.zeros_like(psis)))
    # 1-D input honours coeff
    psi = rand_c(rng, 8)
    np.testing.assert_allclose(fp.PauliString("XYZ").apply(psi, 2j), 2j * O.dense_string("XYZ") @ psi, atol=1e-15)


def test_string_expectation_value():
    H.say("surface PauliString.expectation_value")
    rng = np.random.default_rng(321)
    np.testing.assert_allclose(fp.PauliString("IXYZ").expectation_value(np.zeros(16)), 0, atol=1e-15)
    np.testing.assert_allclose(fp.PauliString("IXYZ").expectation_value(np.zeros((16, 16))), np.zeros(16),
                               atol=1e-15)
    np.testing.assert_allclose(fp.PauliString("III").expectation_value(np.arange(8)), 140, atol=1e-15)
    np.testing.assert_allclose(fp.PauliString("III").expectation_value(np.arange(64).reshape(8, 8)),
                               np.square(np.arange(64)).reshape(8, 8).sum(0), atol=1e-12)
    for s in SAMPLE:
        psi = rand_c(rng, 2 ** len(s))
        np.testing.assert_allclose(fp.PauliString(s).expectation_value(psi),
                                   O.dense_string(s).dot(psi).dot(psi.conj()), atol=1e-14)
        psis = rand_c(rng, 21, 2 ** len(s))
        exp = np.einsum("ti,ij,tj->t", psis.conj(), O.dense_string(s), psis)
        np.testing.assert_allclose(fp.PauliString(s).expectation_value(psis.T.copy()), exp, atol=1e-13)


def test_string_shape_errors():
    H.say("surface PauliString errors")
    with pytest.raises(ValueError):
        fp.PauliString("II").apply(np.array([0.1, 0.2, 0.3]))
    with pytest.raises(ValueError):
        fp.PauliString("XYZ").apply(np.eye(4))
    with pytest.raises(ValueError):
        fp.PauliString("XYZ").apply(np.zeros((8, 2, 2)))
    with pytest.raises(ValueError):
        fp.PauliString("XY").apply(np.zeros((8, 4), np.complex128)[::2])  # not C-contiguous


def test_op_apply_and_expectation():
    H.say("surface PauliOp.apply / expectation_value")
    rng = np.random.default_rng(321)
    np.testing.assert_allclose(fp.PauliOp([1, 2], ["IXYZ", "ZYXI"]).apply(np.zeros(16)), 0, atol=1e-15)
    np.testing.assert_allclose(fp.PauliOp([0.5, 0.5], ["III", "III"]).apply(np.arange(8)), np.arange(8), atol=1e-15)
    np.testing.assert_allclose(fp.PauliOp([1j], ["XYZ"]).apply(np.ones(8)), 1j * O.dense_string("XYZ").sum(1),
                               atol=1e-15)
    np.testing.assert_allclose(fp.PauliOp([1, 1], ["III", "III"]).expectation_value(np.arange(8)), 280, atol=1e-13)
    for n in (1, 2, 3, 4):
        strings = ["".join(p) for p in it.product("IXYZ", repeat=n)]
        c = rand_c(rng, len(strings))
        po = fp.PauliOp(c, strings)
        dense = O.dense_op(c, strings)
        psi = rand_c(rng, 2 ** n)
        np.testing.assert_allclose(po.apply(psi), dense @ psi, atol=1e-13)
        psis = rand_c(rng, 2 ** n, 42)
        got = po.apply(psis)
        np.testing.assert_allclose(got, dense @ psis, atol=1e-13)
        assert np.array_equal(got, O.op_apply(strings, c, psis, np.zeros_like(psis)))
        np.testing.assert_allclose(po.expectation_value(psi), psi.conj() @ dense @ psi, atol=1e-13)
        exp = np.einsum("it,ij,jt->t", psis.conj(), dense, psis)
        np.testing.assert_allclose(po.expectation_value(psis), exp, atol=1e-12)
    with pytest.raises(ValueError):
        fp.PauliOp([1, 1, 1], ["X", "Y", "Z"]).apply(np.ones(3))
    with pytest.raises(ValueError):
        fp.PauliOp([1, 1], ["XYZ", "ZYX"]).expectation_value(np.eye(16))


def test_op_cached_device_copy_tracks_mutation():
    H.say("surface PauliOp mutation")
    psi = syn.random_states(16, 3, seed=1)
    op = fp.PauliOp([1.0, 2.0], ["XYIZ", "ZZII"])
    a = op.apply(psi)
    op.scale(2.0)
    np.testing.assert_allclose(op.apply(psi), 2 * a, atol=1e-15)
    op += fp.PauliString("IIII")
    np.testing.assert_allclose(op.apply(psi), 2 * a + psi, atol=1e-15)


def test_complex64_fast_path():
    H.say("surface complex64")
    n, N, S = 10, 40, 8
    codes = syn.random_codes(N, n, seed=5)
    strings = syn.codes_to_strings(codes)
    c = syn.random_coeffs(N, seed=5)
    psi = syn.random_states(1 << n, S, seed=5).astype(np.complex64)
    op = fp.PauliOp(c, strings)
    got = op.apply(psi)
    assert got.dtype == np.complex64
    assert np.array_equal(got, O.op_apply(strings, c.astype(np.complex64), psi, np.zeros_like(psi), dtype=O.C64))
    ev = op.expectation_value(psi)
    assert ev.dtype == np.complex64
    ref = O.op_expectation(strings, c.astype(np.complex64).astype(np.complex128), psi.astype(np.complex128),
                           np.zeros(S, np.complex128))
    assert H.rel_err(ev, ref) <= 1e-5
    ps = fp.PauliString(strings[3])
    assert np.array_equal(ps.apply(psi), O.string_apply_batch(strings[3], 1, psi, np.zeros_like(psi), dtype=O.C64))


@pytest.mark.parametrize("n_states,n_operators,n_qubits",
                         [(s, o, q) for s in [1, 10, 1000] for o in [1, 10, 100] for q in [1, 2, 6]])
def test_summed_expectation_apply_weighted(n_states, n_operators, n_qubits):
    H.say(f"surface SummedPauliOp S={n_states} K={n_operators} n={n_qubits}")
    rng = np.random.default_rng(n_states * 7 + n_operators * 3 + n_qubits)
    strings = fp.helpers.calculate_pauli_strings_max_weight(n_qubits, 2)
    strs = [str(s) for s in strings]
    N = len(strings)
    c = rng.random((N, n_operators)).astype(np.complex128)
    op = fp.SummedPauliOp(strings, c)
    psi = rng.random((2 ** n_qubits, n_states)).astype(np.complex128)
    data = rng.random((n_operators, n_states))
    if n_states == 1:
        psi, data = psi[:, 0].copy(), data[:, 0].copy()
    psi2 = psi.reshape(2 ** n_qubits, -1)
    data2 = data.reshape(n_operators, -1)
    # expectation values (test_summed_pauli_op.py:135-180)
    ev = op.expectation_value(psi)
    exp = O.summed_expectation(strs, c, psi2, np.zeros((n_operators, psi2.shape[1]), np.complex128))
    np.testing.assert_allclose(ev.reshape(n_operators, -1), exp, atol=1e-13 * max(1, np.abs(exp).max()))
    # apply (test_summed_pauli_op.py:54-94): bit-exact with the sequential loops
    got = op.apply(psi)
    ref = O.summed_apply(strs, c, psi2, np.zeros_like(psi2)).reshape(psi.shape)
    assert np.array_equal(got, ref)
    # apply_weighted (test_summed_pauli_op.py:97-139): bit-exact with the sequential loops
    got = op.apply_weighted(psi, data)
    ref = O.summed_apply_weighted(strs, c, psi2, data2, np.zeros_like(psi2)).reshape(psi.shape)
    assert np.array_equal(got, ref)


def test_summed_fast_contractions():
    H.say("surface SummedPauliOp contraction modes")
    n, N, K, S = 8, 300, 50, 64
    codes = syn.random_codes_max_weight(N, n, 3, seed=9)
    strs = syn.codes_to_strings(codes)
    W = syn.random_real_matrix(N, K, seed=9)
    op = fp.SummedPauliOp(strs, W.astype(np.complex128))
    psi = syn.random_states(1 << n, S, seed=9)
    ref = O.summed_expectation(strs, W.astype(np.complex128), psi, np.zeros((K, S), np.complex128))
    assert H.rel_err(op.expectation_value(psi), ref) <= 1e-12
    psi64 = psi.astype(np.complex64)
    for mode, tol in ((1, 1e-5), (2, 1e-4)):  # FP32 MFMA, BF16x3 MFMA
        op.contraction = mode
        got = op.expectation_value(psi64)
        assert got.dtype == np.complex64
        assert H.rel_err(got, ref) <= tol, mode
