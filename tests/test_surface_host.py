"""The Python surface, host-side behaviour (no GPU): construction, properties,
errors, string algebra, dense forms, helpers and pickling.

Modelled on the reference's tests (tests/fast_pauli/test_pauli.py,
test_pauli_string.py, test_pauli_op.py, test_summed_pauli_op.py,
test_helpers.py); the dense Kronecker construction of oracle/oracle.py is the
ground truth.
"""

import itertools as it
import pickle

import numpy as np
import pytest

import fast_pauli as fp
from oracle import oracle as O

P = O.PAULI_MATRICES
SAMPLE = (["I", "X", "Y", "Z"] + ["".join(p) for p in it.product("IXYZ", repeat=2)]
          + ["".join(p) for p in it.product("IXYZ", repeat=3)]
          + ["XYZXYZ", "ZZZIII", "XYIZXYZ", "XXIYYIZZ", "ZIXIZYXX"])


def lex_strings(size, limit=1000):
    out = []
    for s in it.product("IXYZ", repeat=size):
        if limit and len(out) >= limit:
            break
        out.append("".join(s))
    return out


# ---------------------------------------------------------------- Pauli
def test_pauli_basics():
    np.testing.assert_array_equal(fp.Pauli().to_tensor(), P["I"])
    for i, c in enumerate("IXYZ"):
        np.testing.assert_array_equal(fp.Pauli(code=i).to_tensor(), P[c])
        np.testing.assert_array_equal(fp.Pauli(symbol=c).to_tensor(), P[c])
        assert str(fp.Pauli(c)) == c


def test_pauli_multiply_matches_dense():
    for a, b in it.product("IXYZ", repeat=2):
        ph, p = fp.Pauli(a) @ fp.Pauli(b)
        np.testing.assert_allclose(ph * p.to_tensor(), P[a] @ P[b], atol=1e-15)


def test_pauli_errors():
    with pytest.raises(TypeError):
        fp.Pauli("II")
    with pytest.raises(ValueError):
        fp.Pauli("A")
    with pytest.raises(ValueError):
        fp.Pauli(-1)
    with pytest.raises(ValueError):
        fp.Pauli(5)


def test_pauli_clone_and_pickle():
    for c in "IXYZ":
        p = fp.Pauli(c)
        q = p.clone()
        assert q is not p
        np.testing.assert_array_equal(q.to_tensor(), p.to_tensor())
        np.testing.assert_array_equal(pickle.loads(pickle.dumps(p)).to_tensor(), p.to_tensor())


# ---------------------------------------------------------- PauliString
def test_string_trivial():
    for e in (fp.PauliString(), fp.PauliString([]), fp.PauliString("")):
        assert e.weight == 0 and e.dim == 0 and e.n_qubits == 0
        assert str(e) == ""
        assert len(e.to_tensor()) == 0
    ps = fp.PauliString("III")
    assert ps.dim == 8 and ps.weight == 0
    np.testing.assert_array_equal(ps.to_tensor(), np.eye(8))


def test_string_basics_and_dense():
    for s in SAMPLE:
        ps = fp.PauliString(s)
        assert ps.weight == len(s) - s.count("I")
        assert ps.dim == 2 ** len(s) and ps.n_qubits == len(s)
        assert str(ps) == s
        np.testing.assert_allclose(ps.to_tensor(), O.dense_string(s), atol=1e-15)
    for c in "IXYZ":
        ps = fp.PauliString([fp.Pauli(c)])
        assert ps.dim == 2 and str(ps) == c
    assert fp.PauliString("ZIXIZYXXY").weight == 7


def test_string_multiplication():
    for full in lex_strings(2) + lex_strings(4) + lex_strings(8, 100) + lex_strings(12, 10):
        mid = len(full) // 2
        a, b = full[:mid], full[mid:]
        ph, p = fp.PauliString(a) @ fp.PauliString(b)
        np.testing.assert_allclose(ph * p.to_tensor(), O.dense_string(a) @ O.dense_string(b), atol=1e-15)


def test_string_errors():
    with pytest.raises(ValueError):
        fp.PauliString("ABC")
    with pytest.raises(ValueError):
        fp.PauliString("xyz")
    with pytest.raises(TypeError):
        fp.PauliString([1, 2])
    with pytest.raises(AttributeError):
        fp.PauliString("XYZ").dim = 99
    with pytest.raises(AttributeError):
        fp.PauliString("XYZ").weight = 99


def test_string_clone_pickle_and_sparse_repr():
    for s in SAMPLE:
        ps = fp.PauliString(s)
        np.testing.assert_array_equal(ps.clone().to_tensor(), ps.to_tensor())
        np.testing.assert_array_equal(pickle.loads(pickle.dumps(ps)).to_tensor(), ps.to_tensor())
    for s in ["I", "X", "Y", "Z"] + lex_strings(3):
        k, m = fp.helpers.pauli_string_sparse_repr([fp.Pauli(c) for c in s])
        ko, mo = O.sparse_repr(s)
        np.testing.assert_array_equal(np.asarray(k), ko)
        np.testing.assert_array_equal(np.asarray(m), mo)


# -------------------------------------------------------------- PauliOp
def test_op_trivial_and_dense():
    po = fp.PauliOp()
    assert po.dim == 0 and po.n_qubits == 0 and po.n_pauli_strings == 0
    for c in "IXYZ":
        np.testing.assert_allclose(fp.PauliOp([c]).to_tensor(), P[c], atol=1e-15)
    np.testing.assert_allclose(fp.PauliOp([1j, 1j, 1j, 1j], ["III"] * 4).to_tensor() * 2,
                               4j * 2 * np.eye(8), atol=1e-15)
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 4):
        strings = lex_strings(n)
        c = rng.random(len(strings)) + 1j * rng.random(len(strings))
        po = fp.PauliOp(c, strings)
        assert po.n_pauli_strings == len(strings) and po.dim == 2 ** n
        np.testing.assert_allclose(po.to_tensor(), O.dense_op(c, strings), atol=1e-13)
        assert po.pauli_strings_as_str == strings
        np.testing.assert_allclose(po.coeffs, c)


def test_op_scaling():
    strings = lex_strings(3)
    rng = np.random.default_rng(4)
    c = rng.random(len(strings)) + 1j * rng.random(len(strings))
    po = fp.PauliOp(c, strings)
    ref = O.dense_op(c, strings)
    np.testing.assert_allclose((po * 2.5).to_tensor(), 2.5 * ref, atol=1e-13)
    np.testing.assert_allclose((1j * po).to_tensor(), 1j * ref, atol=1e-13)
    f = rng.random(len(strings))
    q = po.clone()
    q.scale(f)
    np.testing.assert_allclose(q.to_tensor(), O.dense_op(c * f, strings), atol=1e-13)
    q = po.clone()
    q *= 3
    np.testing.assert_allclose(q.to_tensor(), 3 * ref, atol=1e-13)


def test_op_products_and_sums():
    rng = np.random.default_rng(5)
    for n in (1, 2, 3):
        sa, sb = lex_strings(n), lex_strings(n)[::-1][: max(1, len(lex_strings(n)) // 2)]
        ca = rng.random(len(sa)) + 1j * rng.random(len(sa))
        cb = rng.random(len(sb)) + 1j * rng.random(len(sb))
        A, B = fp.PauliOp(ca, sa), fp.PauliOp(cb, sb)
        dA, dB = O.dense_op(ca, sa), O.dense_op(cb, sb)
        np.testing.assert_allclose((A @ B).to_tensor(), dA @ dB, atol=1e-12)
        np.testing.assert_allclose((A + B).to_tensor(), dA + dB, atol=1e-13)
        np.testing.assert_allclose((A - B).to_tensor(), dA - dB, atol=1e-13)
        s = fp.PauliString(sb[0])
        ds = O.dense_string(sb[0])
        np.testing.assert_allclose((A @ s).to_tensor(), dA @ ds, atol=1e-12)
        np.testing.assert_allclose((s @ A).to_tensor(), ds @ dA, atol=1e-12)
        np.testing.assert_allclose((A + s).to_tensor(), dA + ds, atol=1e-13)
        np.testing.assert_allclose((s + A).to_tensor(), dA + ds, atol=1e-13)
        np.testing.assert_allclose((A - s).to_tensor(), dA - ds, atol=1e-13)
        np.testing.assert_allclose((s - A).to_tensor(), ds - dA, atol=1e-13)
        C = A.clone()
        C += B
        np.testing.assert_allclose(C.to_tensor(), dA + dB, atol=1e-13)
        C = A.clone()
        C -= s
        np.testing.assert_allclose(C.to_tensor(), dA - ds, atol=1e-13)
        C = A.clone()
        C.extend(s, 2j, dedupe=True)
        np.testing.assert_allclose(C.to_tensor(), dA + 2j * ds, atol=1e-13)
    sum_ps = fp.PauliString("XY") + fp.PauliString("ZZ")
    np.testing.assert_allclose(sum_ps.to_tensor(), O.dense_string("XY") + O.dense_string("ZZ"))
    diff = fp.PauliString("XY") - fp.PauliString("ZZ")
    np.testing.assert_allclose(diff.to_tensor(), O.dense_string("XY") - O.dense_string("ZZ"))


def test_op_errors():
    with pytest.raises(ValueError):
        fp.PauliOp(["ABC"])
    with pytest.raises(ValueError):
        fp.PauliOp(["XI", "XY", "XZ", "III"])
    with pytest.raises(ValueError):
        fp.PauliOp([1, 1], ["X", "YZ"])
    with pytest.raises(ValueError):
        fp.PauliOp([1, 2, 3], ["I", "X", "Y", "Z"])
    op = fp.PauliOp([-1, 1], ["IX", "YZ"])
    with pytest.raises(ValueError):
        op @ fp.PauliOp([1], ["III"])
    with pytest.raises(ValueError):
        op @ fp.PauliString("X")
    with pytest.raises(ValueError):
        op + fp.PauliOp([1], ["X"])
    with pytest.raises(ValueError):
        op - fp.PauliString("XXXX")
    with pytest.raises(ValueError):
        op.clone().extend(fp.PauliString("XYZ"), 1j, dedupe=False)
    with pytest.raises(ValueError):
        op.clone().extend(fp.PauliOp([1j], ["XXX"]))


def test_op_pickle():
    op = fp.PauliOp([1, 2j, -3], ["XYZ", "ZZI", "IXI"])
    q = pickle.loads(pickle.dumps(op))
    np.testing.assert_allclose(q.to_tensor(), op.to_tensor())


# --------------------------------------------------------- SummedPauliOp
def test_summed_ctor_coeffs_split_and_dense():
    rng = np.random.default_rng(6)
    for n in (1, 2, 3):
        for K in (1, 3):
            strings = fp.helpers.calculate_pauli_strings_max_weight(n, 2)
            N = len(strings)
            c = rng.random((N, K)) + 1j * rng.random((N, K))
            op = fp.SummedPauliOp(strings, c)
            assert op.dim == 2 ** n and op.n_operators == K and op.n_pauli_strings == N
            np.testing.assert_array_equal(op.coeffs, c.T)  # property is (n_operators, n_pauli_strings)
            dense = op.to_tensor()
            strs = [str(s) for s in strings]
            for k in range(K):
                np.testing.assert_allclose(dense[k], O.dense_op(c[:, k], strs), atol=1e-13)
            parts = op.split()
            assert len(parts) == K
            for k, p in enumerate(parts):
                np.testing.assert_allclose(p.to_tensor(), dense[k], atol=1e-13)
            op2 = fp.SummedPauliOp(strs, c)
            np.testing.assert_allclose(op2.to_tensor(), dense, atol=1e-15)
            new = rng.random((K, N)) + 0j
            op.coeffs = new
            np.testing.assert_array_equal(op.coeffs, new)


def test_summed_pickle_clone():
    strings = ["XI", "IZ", "YY"]
    c = np.arange(6, dtype=np.complex128).reshape(3, 2)
    op = fp.SummedPauliOp(strings, c)
    for q in (op.clone(), pickle.loads(pickle.dumps(op))):
        np.testing.assert_array_equal(q.coeffs, op.coeffs)
        np.testing.assert_array_equal(q.to_tensor(), op.to_tensor())
    with pytest.raises(ValueError):
        fp.SummedPauliOp(["XI", "IZ"], c)
    with pytest.raises(ValueError):
        op.coeffs = np.zeros((3, 3))


# --------------------------------------------------------------- helpers
def trusted_strings(n, w):
    out = []
    for idx in it.combinations(range(n), w):
        for letters in it.product("XYZ", repeat=w):
            s = ["I"] * n
            for i, l in zip(idx, letters):
                s[i] = l
            out.append("".join(s))
    return out


@pytest.mark.parametrize("weight", [0, 1, 2, 3])
def test_nontrivial_paulis(weight):
    res = fp.helpers.get_nontrivial_paulis(weight)
    assert res == ["".join(x) for x in it.product("XYZ", repeat=weight) if x]


@pytest.mark.parametrize("n", [1, 2, 3, 4])
@pytest.mark.parametrize("w", [1, 2, 3])
def test_calculate_pauli_strings(n, w):
    if n < w:
        pytest.skip("weight above qubit count")
    res = [str(x) for x in fp.helpers.calculate_pauli_strings(n, w)]
    assert sorted(res) == sorted(trusted_strings(n, w))
    res_max = [str(x) for x in fp.helpers.calculate_pauli_strings_max_weight(n, w)]
    trusted = ["I" * n] + sum((trusted_strings(n, k) for k in range(1, w + 1)), [])
    assert sorted(res_max) == sorted(trusted)


def test_enumeration_order_known_answers():
    # fast_pauli/cpp/tests/test_pauli_helpers.cpp:42-51, 110-166
    assert fp.helpers.get_nontrivial_paulis(2) == ["XX", "XY", "XZ", "YX", "YY", "YZ", "ZX", "ZY", "ZZ"]
    assert [str(s) for s in fp.helpers.calculate_pauli_strings(4, 0)] == ["IIII"]
    assert [str(s) for s in fp.helpers.calculate_pauli_strings(2, 1)] == ["XI", "IX", "YI", "IY", "ZI", "IZ"]
    r = [str(s) for s in fp.helpers.calculate_pauli_strings(4, 2)]
    assert len(r) == 54 and r[0] == "XXII" and r[1] == "XIXI" and r[53] == "IIZZ"
    assert len(fp.helpers.calculate_pauli_strings_max_weight(2, 1)) == 7
    assert len(fp.helpers.calculate_pauli_strings_max_weight(12, 2)) == 631
