"""Pins the CPU oracle (oracle/fp_oracle.cpp) before anything is checked against it.

Every oracle entry point is checked against at least one of:
  * the dense Kronecker brute force, the ground truth of the reference's own
    tests (fast_pauli/pypauli/helpers.py:40-85, test_pauli_string.cpp:75-113);
  * known answers from the reference's tests (cited per test);
  * invariants fixed by the mathematics (P^2 = I, Hermiticity, the
    sum rule sum_P <P>^2 = 2^n for a pure normalised state).
A final group shows these checks are sensitive: plausible mistakes planted in
a copy of the oracle's inputs/outputs fail them.
"""

import itertools as it

import numpy as np
import pytest

from oracle import oracle as O
from fast_pauli import synthetic as syn

SAMPLE = (["I", "X", "Y", "Z"]
          + ["".join(p) for p in it.product("IXYZ", repeat=2)]
          + ["".join(p) for p in it.product("IXYZ", repeat=3)]
          + ["XYZXYZ", "ZZZIII", "XYIZXYZ", "XXIYYIZZ", "ZIXIZYXX"])  # tests/conftest.py:43-52


def rand_c(*shape, seed=321):
    rng = np.random.default_rng(seed)
    return rng.random(shape) + 1j * rng.random(shape)


def closed_form(s):
    n = len(s)
    x = z = 0
    for i, ch in enumerate(s):
        b = 1 << (n - 1 - i)
        if ch in "XY":
            x |= b
        if ch in "YZ":
            z |= b
    ny = bin(x & z).count("1")
    j = np.arange(1 << n, dtype=np.uint64)
    par = np.array([bin(int(v) & z).count("1") & 1 for v in j])
    m = ((-1j) ** ny) * np.where(par, -1.0, 1.0)
    return j ^ np.uint64(x), m


# ---------------------------------------------------------------------------
# Pauli matrices: the reference's Cayley table (test_pauli.cpp:67-87)
# ---------------------------------------------------------------------------
CAYLEY = {
    ("I", "I"): (1, "I"), ("I", "X"): (1, "X"), ("I", "Y"): (1, "Y"), ("I", "Z"): (1, "Z"),
    ("X", "I"): (1, "X"), ("X", "X"): (1, "I"), ("X", "Y"): (1j, "Z"), ("X", "Z"): (-1j, "Y"),
    ("Y", "I"): (1, "Y"), ("Y", "X"): (-1j, "Z"), ("Y", "Y"): (1, "I"), ("Y", "Z"): (1j, "X"),
    ("Z", "I"): (1, "Z"), ("Z", "X"): (1j, "Y"), ("Z", "Y"): (-1j, "X"), ("Z", "Z"): (1, "I"),
}


def test_dense_pauli_matrices_match_cayley_table():
    P = O.PAULI_MATRICES
    for (a, b), (ph, c) in CAYLEY.items():
        np.testing.assert_array_equal(P[a] @ P[b], ph * P[c])


def test_dense_known_answers():
    # test_pauli_string.py:31-59
    np.testing.assert_array_equal(O.dense_string("III"), np.eye(8))
    np.testing.assert_array_equal(O.dense_string("XXX"), np.fliplr(np.eye(8)))
    np.testing.assert_array_equal(O.dense_string("IZ"), np.kron(O.PAULI_MATRICES["I"], O.PAULI_MATRICES["Z"]))
    np.testing.assert_array_equal(
        O.dense_string("XYZ"),
        np.kron(O.PAULI_MATRICES["X"], np.kron(O.PAULI_MATRICES["Y"], O.PAULI_MATRICES["Z"])))


# ---------------------------------------------------------------------------
# get_sparse_repr
# ---------------------------------------------------------------------------
def test_sparse_repr_matches_dense_and_closed_form():
    strings = SAMPLE + ["".join(p) for p in it.product("IXYZ", repeat=4)]
    for s in strings:
        k, m = O.sparse_repr(s)
        dense = O.dense_string(s)
        rows = np.arange(dense.shape[0])
        np.testing.assert_array_equal(dense[rows, k.astype(np.int64)], m)
        # exactly one non-zero per row
        assert np.count_nonzero(dense) == dense.shape[0]
        kc, mc = closed_form(s)
        np.testing.assert_array_equal(k, kc)
        np.testing.assert_array_equal(m, mc)


def test_sparse_repr_empty():
    k, m = O.sparse_repr("")
    assert k.size == 0 and m.size == 0


# ---------------------------------------------------------------------------
# PauliString apply / apply_batch / expectation_value
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("par", [False, True])
def test_string_apply_1d(par):
    # known answers, test_pauli_string.py:184-198
    out = O.string_apply("IXYZ", 1, np.zeros(16, np.complex128), np.zeros(16, np.complex128), par)
    np.testing.assert_array_equal(out, 0)
    a = np.arange(8).astype(np.complex128)
    np.testing.assert_array_equal(O.string_apply("III", 1, a, np.zeros(8, np.complex128), par), a)
    ones = np.ones(8, np.complex128)
    np.testing.assert_allclose(O.string_apply("ZYX", 1, ones, np.zeros(8, np.complex128), par),
                               O.dense_string("ZYX").sum(1), atol=1e-15)
    for s in SAMPLE:
        psi = rand_c(1 << len(s))
        c = 0.3 - 1.7j
        got = O.string_apply(s, c, psi, np.zeros_like(psi), par)
        np.testing.assert_allclose(got, c * O.dense_string(s) @ psi, atol=1e-14)


@pytest.mark.parametrize("par", [False, True])
def test_string_apply_batch(par):
    np.testing.assert_array_equal(
        O.string_apply_batch("ZYX", 1, np.eye(8, dtype=np.complex128),
                             np.zeros((8, 8), np.complex128), par), O.dense_string("ZYX"))
    for s in SAMPLE:
        psis = rand_c(1 << len(s), 42)
        got = O.string_apply_batch(s, 1, psis, np.zeros_like(psis), par)
        np.testing.assert_allclose(got, O.dense_string(s) @ psis, atol=1e-15)
    for s in ["".join(p) for p in it.product("IXYZ", repeat=4)]:
        coeff = rand_c(1, seed=7)[0]
        psis = rand_c(16, 7)
        got = O.string_apply_batch(s, coeff, psis, np.zeros_like(psis), par)
        np.testing.assert_allclose(got, coeff * O.dense_string(s) @ psis, atol=1e-14)


def test_string_apply_batch_accumulates():
    # the C++ API adds into new_states (__pauli_string.hpp:419, 432)
    psis = rand_c(32, 5)
    base = rand_c(32, 5, seed=9)
    got = O.string_apply_batch("XYZZY", 2.0, psis, base.copy())
    np.testing.assert_allclose(got, base + 2.0 * O.dense_string("XYZZY") @ psis, atol=1e-14)


def test_string_apply_is_involution_bit_exact():
    psis = syn.random_states(1 << 8, 5, seed=3)
    for s in ["XYZIIXYZ", "YYYYYYYY", "IZIZXXYY"]:
        once = O.string_apply_batch(s, 1, psis, np.zeros_like(psis))
        twice = O.string_apply_batch(s, 1, once, np.zeros_like(psis))
        np.testing.assert_array_equal(twice, psis)


@pytest.mark.parametrize("par", [False, True])
def test_string_expectation(par):
    a = np.arange(8).astype(np.complex128).reshape(8, 1).copy()
    out = O.string_expectation("III", 1, a, np.zeros(1, np.complex128), par)
    np.testing.assert_allclose(out, [140], atol=1e-15)  # test_pauli_string.py:285-289
    a64 = np.arange(64).astype(np.complex128).reshape(8, 8)
    out = O.string_expectation("III", 1, a64, np.zeros(8, np.complex128), par)
    np.testing.assert_allclose(out, np.square(np.arange(64)).reshape(8, 8).sum(0), atol=1e-12)
    for s in SAMPLE:
        psis = rand_c(21, 1 << len(s))
        expected = np.einsum("ti,ij,tj->t", psis.conj(), O.dense_string(s), psis)
        got = O.string_expectation(s, 1, psis.T.copy(), np.zeros(21, np.complex128), par)
        np.testing.assert_allclose(got, expected, atol=1e-13)


def test_expectation_invariants():
    n = 3
    psi = syn.random_states(1 << n, 4, seed=11)
    total = np.zeros(4)
    for p in it.product("IXYZ", repeat=n):
        s = "".join(p)
        ev = O.string_expectation(s, 1, psi, np.zeros(4, np.complex128))
        assert np.all(np.abs(ev.imag) < 1e-15)  # Pauli strings are Hermitian
        total += ev.real ** 2
    np.testing.assert_allclose(total, 2 ** n, rtol=1e-13)  # sum rule for pure states
    # Z strings on basis states are +-1
    e5 = np.zeros((8, 1), np.complex128)
    e5[5, 0] = 1
    ev = O.string_expectation("ZIZ", 1, e5, np.zeros(1, np.complex128))
    assert ev[0] == 1  # bits 101: Z on bit 2 and bit 0 -> (-1)^2


# ---------------------------------------------------------------------------
# PauliOp
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("par", [False, True])
def test_op_apply(par):
    a = np.arange(8).astype(np.complex128).reshape(8, 1).copy()
    got = O.op_apply(["III", "III"], [0.5, 0.5], a, np.zeros_like(a), par)
    np.testing.assert_allclose(got, a, atol=1e-15)  # test_pauli_op.py:187-189
    strings = syn.codes_to_strings(syn.random_codes(24, 5, seed=4))
    coeffs = syn.random_coeffs(24, seed=4)
    psis = rand_c(32, 6)
    got = O.op_apply(strings, coeffs, psis, np.zeros_like(psis), par)
    np.testing.assert_allclose(got, O.dense_op(coeffs, strings) @ psis, atol=1e-13)


def test_op_apply_seq_equals_sum_of_string_applies_bit_exact():
    strings = syn.codes_to_strings(syn.random_codes(16, 6, seed=5))
    coeffs = syn.random_coeffs(16, seed=5)
    psis = syn.random_states(64, 3, seed=5)
    got = O.op_apply(strings, coeffs, psis, np.zeros_like(psis))
    acc = np.zeros_like(psis)
    for s, c in zip(strings, coeffs):
        O.string_apply_batch(s, c, psis, acc)
    np.testing.assert_array_equal(got, acc)


def test_op_apply_hardcoded_state():
    # test_pauli_op.cpp:159-162: PauliOp({"IXYZ"}, {1}) == PauliString("IXYZ").apply
    state = np.array([0.6514 + 0.8887j, 0.0903 + 0.8848j, 0.8130 + 0.6729j, 0.9854 + 0.9497j,
                      0.0410 + 0.5132j, 0.7920 + 0.6519j, 0.4683 + 0.5708j, 0.2148 + 0.4561j,
                      0.5400 + 0.9932j, 0.9736 + 0.0449j, 0.2979 + 0.9834j, 0.4346 + 0.5635j,
                      0.0767 + 0.3516j, 0.4873 + 0.4097j, 0.1543 + 0.7329j, 0.4121 + 0.5059j])
    res = O.op_apply(["IXYZ"], [1], state.reshape(16, 1).copy(), np.zeros((16, 1), np.complex128))
    exp = O.string_apply("IXYZ", 1, state, np.zeros(16, np.complex128))
    np.testing.assert_array_equal(res[:, 0], exp)
    np.testing.assert_allclose(exp, O.dense_string("IXYZ") @ state, atol=1e-15)


@pytest.mark.parametrize("par", [False, True])
def test_op_expectation(par):
    a = np.arange(8).astype(np.complex128).reshape(8, 1).copy()
    got = O.op_expectation(["III", "III"], [1, 1], a, np.zeros(1, np.complex128), par)
    np.testing.assert_allclose(got, [280], atol=1e-13)  # test_pauli_op.py:286-289
    strings = syn.codes_to_strings(syn.random_codes(30, 6, seed=8))
    coeffs = syn.random_coeffs(30, seed=8)
    psis = rand_c(64, 9)
    exp = np.einsum("it,ij,jt->t", psis.conj(), O.dense_op(coeffs, strings), psis)
    got = O.op_expectation(strings, coeffs, psis, np.zeros(9, np.complex128), par)
    np.testing.assert_allclose(got, exp, atol=1e-12)


# ---------------------------------------------------------------------------
# SummedPauliOp
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("par", [False, True])
def test_summed_expectation(par):
    n, K, S = 4, 5, 7
    codes = syn.random_codes_max_weight(40, n, 2, seed=12)
    strings = syn.codes_to_strings(codes)
    coeffs = rand_c(40, K, seed=12)
    psis = rand_c(1 << n, S, seed=13)
    got = O.summed_expectation(strings, coeffs, psis, np.zeros((K, S), np.complex128), par)
    for k in range(K):
        A = O.dense_op(coeffs[:, k], strings)
        np.testing.assert_allclose(got[k], np.einsum("it,ij,jt->t", psis.conj(), A, psis), atol=1e-12)


@pytest.mark.parametrize("par", [False, True])
def test_summed_apply_and_weighted(par):
    n, K, S = 3, 4, 5
    strings = syn.codes_to_strings(syn.random_codes(20, n, seed=14))
    coeffs = rand_c(20, K, seed=14)
    psis = rand_c(1 << n, S, seed=15)
    got = O.summed_apply(strings, coeffs, psis, np.zeros_like(psis), par)
    exp = sum(O.dense_op(coeffs[:, k], strings) for k in range(K)) @ psis
    np.testing.assert_allclose(got, exp, atol=1e-12)
    data = np.random.default_rng(16).random((K, S))
    got = O.summed_apply_weighted(strings, coeffs, psis, data, np.zeros_like(psis), par)
    exp = np.zeros_like(psis)
    for k in range(K):
        exp += (O.dense_op(coeffs[:, k], strings) @ psis) * data[k]
    np.testing.assert_allclose(got, exp, atol=1e-12)


# ---------------------------------------------------------------------------
# complex64 instantiation (the reference's float template)
# ---------------------------------------------------------------------------
def test_c64_paths_track_c128():
    strings = syn.codes_to_strings(syn.random_codes(8, 6, seed=21))
    coeffs = syn.random_coeffs(8, seed=21)
    psis = syn.random_states(64, 4, seed=21)
    ref = O.op_apply(strings, coeffs, psis, np.zeros_like(psis))
    p64 = psis.astype(np.complex64)
    got = O.op_apply(strings, coeffs.astype(np.complex64), p64, np.zeros_like(p64), dtype=O.C64)
    np.testing.assert_allclose(got, ref, atol=1e-5)
    ev = O.op_expectation(strings, coeffs, psis, np.zeros(4, np.complex128))
    ev32 = O.op_expectation(strings, coeffs.astype(np.complex64), p64, np.zeros(4, np.complex64),
                            dtype=O.C64)
    np.testing.assert_allclose(ev32, ev, atol=1e-5)


# ---------------------------------------------------------------------------
# the checks above are sensitive: planted mistakes are caught
# ---------------------------------------------------------------------------
def test_planted_mistakes_are_caught():
    s = "XYZY"
    psi = rand_c(16, 3)
    dense = O.dense_string(s) @ psi
    good = O.string_apply_batch(s, 1, psi, np.zeros_like(psi))
    np.testing.assert_allclose(good, dense, atol=1e-15)
    # reversed character order (wrong bit <-> character mapping)
    wrong = O.string_apply_batch(s[::-1], 1, psi, np.zeros_like(psi))
    assert not np.allclose(wrong, dense)
    # conjugated phase ((+i)^nY instead of (-i)^nY), nY odd here
    k, m = O.sparse_repr("XYZI")
    wrong_m = np.conj(m)
    assert not np.allclose(O.dense_string("XYZI")[np.arange(16), k.astype(int)], wrong_m)
    # the literal form i^{nY} (-1)^{popcount(j&z)} (sign of the phase flipped for odd nY)
    kc, mc = closed_form("YII")
    assert not np.allclose(np.conj(mc), mc)


# ---------------------------------------------------------------------------
# String algebra: products, the enumeration order, PauliOp * PauliOp, square
# ---------------------------------------------------------------------------
def test_string_mul_letters_match_cayley_table():
    for (a, b), (phase, c) in CAYLEY.items():
        assert O.string_mul(a, b) == (phase, c)


def test_string_mul_matches_dense_products():
    strings = ["".join(p) for p in it.product("IXYZ", repeat=3)]
    for a in strings:
        for b in strings[::3]:
            phase, c = O.string_mul(a, b)
            np.testing.assert_array_equal(phase * O.dense_string(c), O.dense_string(a) @ O.dense_string(b))


def test_strings_max_weight_known_answers():
    # fast_pauli/cpp/tests/test_pauli_helpers.cpp:42-51, 110-166; counts sum_w 3^w C(n, w)
    assert O.strings_max_weight(2, 1) == ["II", "XI", "IX", "YI", "IY", "ZI", "IZ"]
    r = O.strings_max_weight(4, 2)
    w2 = r[1 + 12:]
    assert len(w2) == 54 and w2[0] == "XXII" and w2[1] == "XIXI" and w2[53] == "IIZZ"
    assert len(O.strings_max_weight(12, 2)) == 631
    from math import comb
    for n in range(1, 7):
        for w in range(0, n + 1):
            v = O.strings_max_weight(n, w)
            assert len(v) == sum(3 ** k * comb(n, k) for k in range(w + 1))
            assert len(set(v)) == len(v)
            assert [s.count("I") for s in v] == sorted(s.count("I") for s in v)[::-1]


@pytest.mark.parametrize("n,K", [(1, 1), (2, 3), (3, 2), (4, 4)])
def test_summed_square_matches_dense(n, K):
    """A_k^2 from the oracle's square equals the dense product (test_summed_pauli_op.py:227-241,
    test_summed_pauli_op.cpp:411-470)."""
    strings = O.strings_max_weight(n, 2)
    c = rand_c(len(strings), K, seed=n * 10 + K)
    sq, csq = O.summed_square(strings, c)
    assert sq == O.strings_max_weight(n, min(n, 4))
    for k in range(K):
        a = O.dense_op(c[:, k], strings)
        np.testing.assert_allclose(O.dense_op(csq[:, k], sq), a @ a, atol=1e-12)


def test_summed_square_random_strings_and_c64():
    rng = np.random.default_rng(5)
    n, N, K = 5, 14, 3
    strings = ["".join(rng.choice(list("IXYZ"), n)) for _ in range(N)]
    c = rng.standard_normal((N, K)) + 1j * rng.standard_normal((N, K))
    sq, csq = O.summed_square(strings, c)
    for k in range(K):
        a = O.dense_op(c[:, k], strings)
        np.testing.assert_allclose(O.dense_op(csq[:, k], sq), a @ a, atol=1e-11)
    sq32, c32 = O.summed_square(strings, c.astype(np.complex64), dtype=O.C64)
    assert sq32 == sq and c32.dtype == np.complex64
    np.testing.assert_allclose(c32, csq, atol=1e-4)


def test_op_mul_matches_dense():
    rng = np.random.default_rng(9)
    s1 = ["".join(rng.choice(list("IXYZ"), 3)) for _ in range(7)]
    s2 = ["".join(rng.choice(list("IXYZ"), 3)) for _ in range(5)]
    c1, c2 = rand_c(7, seed=1), rand_c(5, seed=2)
    strs, c = O.op_mul(s1, c1, s2, c2)
    assert len(set(strs)) == len(strs)
    np.testing.assert_allclose(O.dense_op(c, strs), O.dense_op(c1, s1) @ O.dense_op(c2, s2), atol=1e-13)


def test_square_with_planted_phase_mistake_fails_dense():
    """The dense check is sensitive to the phases: the same T_aij contraction with the
    phases dropped (a plausible mistake) is caught.  (Conjugated phases would not be: in
    a square, P_j P_i carries the conjugate phase of P_i P_j with the same h_i h_j.)"""
    strings = O.strings_max_weight(2, 2)
    c = rand_c(len(strings), 1, seed=4)[:, 0]
    a = O.dense_op(c, strings)
    index = {s: i for i, s in enumerate(O.strings_max_weight(2, 2))}
    wrong = np.zeros(len(index), np.complex128)
    for i, si in enumerate(strings):
        for j, sj in enumerate(strings):
            phase, p = O.string_mul(si, sj)
            wrong[index[p]] += c[i] * c[j]
    assert not np.allclose(O.dense_op(wrong, list(index)), a @ a, atol=1e-6)
