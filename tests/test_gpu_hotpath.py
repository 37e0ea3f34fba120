"""GPU parity of the hot path, through the C ABI, against the CPU oracle.

Bars:
  * apply / apply_batch / PauliOp::apply: bit-exact with the oracle's
    sequential loops (same FP operations, no fused multiply-add), in both
    complex128 and complex64.  Bit-exact here means IEEE equality
    (np.array_equal); only the sign of an exact zero could differ.
  * expectation values, complex128: <= 1e-12 normwise relative to the oracle.
  * expectation values, complex64 (FP32 accumulation): <= 1e-5 normwise
    relative to the FP64 oracle evaluated on the same complex64 inputs.
"""

import itertools as it

import numpy as np
import pytest

from fast_pauli import _abi
from fast_pauli import synthetic as syn
from oracle import oracle as O
from tests import hiputil as H

pytestmark = pytest.mark.gpu

C128, C64 = _abi.FP_C128, _abi.FP_C64
SAMPLE = (["I", "X", "Y", "Z"] + ["".join(p) for p in it.product("IXYZ", repeat=2)]
          + ["".join(p) for p in it.product("IXYZ", repeat=3)]
          + ["XYZXYZ", "ZZZIII", "XYIZXYZ", "XXIYYIZZ", "ZIXIZYXX"])


def ops_of(strings, coeffs=None):
    return _abi.Ops(O.codes_of(strings), coeffs)


def odt(dtype):
    return O.C128 if dtype == C128 else O.C64


# ---------------------------------------------------------------------------
# PauliString::apply_batch / apply
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [C128, C64])
def test_string_apply_batch_sample_bit_exact(hip, dtype):
    cdt = H.CDT[dtype]
    for n_states in (1, 3, 16, 42):
        for s in SAMPLE:
            H.say(f"string_apply {s} S={n_states} dtype={dtype}")
            ops = ops_of([s])
            psi = syn.random_states(1 << len(s), n_states, seed=len(s) + n_states).astype(cdt)
            for c in (1.0, 0.3 - 1.7j):
                got = H.string_apply(ops, 0, psi, c, dtype)
                exp = O.string_apply_batch(s, c, psi, np.zeros_like(psi), dtype=odt(dtype))
                assert np.array_equal(got, exp), (s, n_states, c)


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("n,S", [(10, 1), (10, 16), (12, 7), (14, 128), (9, 1000)])
def test_string_apply_batch_random_bit_exact(hip, dtype, n, S):
    cdt = H.CDT[dtype]
    codes = syn.random_codes(6, n, seed=n * 31 + S)
    strings = syn.codes_to_strings(codes)
    ops = _abi.Ops(codes)
    psi = syn.random_states(1 << n, S, seed=S).astype(cdt)
    base = syn.random_states(1 << n, S, seed=S + 1).astype(cdt)
    for i, s in enumerate(strings):
        H.say(f"string_apply random n={n} S={S} idx={i} dtype={dtype}")
        c = complex(syn.random_coeffs(1, seed=i)[0])
        got = H.string_apply(ops, i, psi, c, dtype)
        exp = O.string_apply_batch(s, c, psi, np.zeros_like(psi), dtype=odt(dtype))
        assert np.array_equal(got, exp)
        # accumulate (+=) semantics of the C++ API
        got = H.string_apply(ops, i, psi, c, dtype, out0=base)
        exp = O.string_apply_batch(s, c, psi, base.copy(), dtype=odt(dtype))
        assert np.array_equal(got, exp)


def test_string_apply_1d_known_answers(hip):
    H.say("string_apply known answers")
    # test_pauli_string.py:184-198 (1-D inputs are the S = 1 case)
    got = H.string_apply(ops_of(["IXYZ"]), 0, np.zeros((16, 1)), 1.0)
    assert np.array_equal(got, np.zeros((16, 1)))
    a = np.arange(8).astype(np.complex128).reshape(8, 1)
    assert np.array_equal(H.string_apply(ops_of(["III"]), 0, a), a)
    got = H.string_apply(ops_of(["ZYX"]), 0, np.ones((8, 1)))
    np.testing.assert_array_equal(got[:, 0], O.dense_string("ZYX").sum(1))
    got = H.string_apply(ops_of(["ZYX"]), 0, np.eye(8))
    np.testing.assert_array_equal(got, O.dense_string("ZYX"))


# ---------------------------------------------------------------------------
# PauliOp::apply
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("n,N,S", [(1, 3, 1), (5, 7, 5), (10, 64, 128), (12, 33, 3), (8, 512, 2)])
def test_op_apply_bit_exact(hip, dtype, n, N, S):
    H.say(f"op_apply n={n} N={N} S={S} dtype={dtype}")
    cdt = H.CDT[dtype]
    codes = syn.random_codes(N, n, seed=N + n)
    strings = syn.codes_to_strings(codes)
    coeffs = syn.random_coeffs(N, seed=N)
    ops = _abi.Ops(codes, coeffs)
    psi = syn.random_states(1 << n, S, seed=n).astype(cdt)
    got = H.op_apply(ops, psi, dtype)
    exp = O.op_apply(strings, coeffs.astype(cdt), psi, np.zeros_like(psi), dtype=odt(dtype))
    assert np.array_equal(got, exp)
    base = syn.random_states(1 << n, S, seed=n + 7).astype(cdt)
    got = H.op_apply(ops, psi, dtype, out0=base)
    exp = O.op_apply(strings, coeffs.astype(cdt), psi, base.copy(), dtype=odt(dtype))
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("kernel", ["stream", "tiled", "tiled_wide", "tiled_kb128"])
@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("kind,n,N,S", [
    ("random", 2, 5, 1), ("random", 8, 40, 3), ("random", 11, 64, 16), ("random", 13, 200, 7),
    ("random", 15, 50, 33), ("random", 16, 300, 2), ("weight2", 14, 300, 5), ("diagonal", 12, 40, 16),
    ("shared_x", 12, 64, 4), ("duplicates", 10, 33, 1), ("random", 13, 80, 6), ("random", 12, 64, 34),
    ("random", 14, 100, 58), ("random", 13, 120, 24), ("random", 12, 64, 40),
])
def test_op_apply_kernels(hip, monkeypatch, kernel, dtype, kind, n, N, S):
    """PauliOp::apply through the streaming kernel and the coset-tiled kernel
    (groups of consecutive strings, one launch each): IEEE-equal to the oracle's
    sequential loop, overwrite and accumulate."""
    H.say(f"op_apply kernel={kernel} {kind} n={n} N={N} S={S} dtype={dtype}")
    monkeypatch.setenv("FAST_PAULI_APPLY_KERNEL", kernel.split("_")[0])
    if kernel == "tiled_wide":
        monkeypatch.setenv("FAST_PAULI_TILE_WIDE", "1")
    if kernel == "tiled_kb128":
        # the expectation kernel's tile-size knob leaves the apply tiles at 64 KiB
        monkeypatch.setenv("FAST_PAULI_TILE_KB", "128")
    cdt = H.CDT[dtype]
    codes = _structured_codes(kind, N, n, seed=7 * n + N)
    strings = syn.codes_to_strings(codes)
    coeffs = syn.random_coeffs(N, seed=N + 3)
    ops = _abi.Ops(codes, coeffs)
    psi = syn.random_states(1 << n, S, seed=n + 5 * S).astype(cdt)
    got = H.op_apply(ops, psi, dtype)
    exp = O.op_apply(strings, coeffs.astype(cdt), psi, np.zeros_like(psi), dtype=odt(dtype))
    assert np.array_equal(got, exp)
    base = syn.random_states(1 << n, S, seed=n + 11).astype(cdt)
    got = H.op_apply(ops, psi, dtype, out0=base)
    exp = O.op_apply(strings, coeffs.astype(cdt), psi, base.copy(), dtype=odt(dtype))
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("dtype", [C128, C64])
def test_op_apply_tiled_is_default_for_large_states(hip, monkeypatch, dtype):
    """At 2^21 x 8 states (128 / 64 MiB) with 96 random strings the automatic
    choice takes the tiled kernel; it equals the streaming kernel bit for bit."""
    n, N, S = 21, 96, 8
    H.say(f"op_apply large state auto vs stream dtype={dtype}")
    cdt = H.CDT[dtype]
    ops = _abi.Ops(syn.random_codes(N, n, seed=41), syn.random_coeffs(N, seed=41))
    psi = syn.random_states(1 << n, S, seed=41).astype(cdt)
    auto = H.op_apply(ops, psi, dtype)
    B = psi.nbytes
    p = _abi.op_apply_plan(ops, dtype, 1 << n, S)
    assert p["kernel"] == _abi.FP_EV_TILED and 0 < p["n_groups"] < N
    G = p["n_groups"]
    assert p["hbm_bytes"] == (3 * G - 1) * B  # state + out per group; the first group overwrites
    assert _abi.op_apply_plan(ops, dtype, 1 << n, S, accumulate=True)["hbm_bytes"] == 3 * G * B
    monkeypatch.setenv("FAST_PAULI_APPLY_KERNEL", "stream")
    ref = H.op_apply(ops, psi, dtype)
    assert np.array_equal(auto, ref)
    p = _abi.op_apply_plan(ops, dtype, 1 << n, S)
    assert p["kernel"] == _abi.FP_EV_STREAM and p["hbm_bytes"] == (N + 1) * B


def test_op_apply_known_answers(hip):
    H.say("op_apply known answers")
    a = np.arange(64).astype(np.complex128).reshape(8, 8)
    got = H.op_apply(ops_of(["III", "III"], [0.5, 0.5]), a)
    np.testing.assert_array_equal(got, a)  # test_pauli_op.py:234-236
    got = H.op_apply(ops_of(["XYZ"], [1j]), np.eye(8))
    np.testing.assert_allclose(got, 1j * O.dense_string("XYZ"), atol=1e-15)
    got = H.op_apply(ops_of(["IXYZ", "ZYXI"], [1, 2]), np.zeros((16, 16)))
    assert np.array_equal(got, np.zeros((16, 16)))


# ---------------------------------------------------------------------------
# expectation values
# ---------------------------------------------------------------------------
def test_string_expectation_sample(hip):
    for s in SAMPLE:
        H.say(f"string_expectation {s}")
        psi = syn.random_states(1 << len(s), 21, seed=len(s))
        ops = ops_of([s])
        for c in (1.0, 2.5 - 0.5j):
            got = H.string_expectation(ops, 0, psi, c)
            exp = O.string_expectation(s, c, psi, np.zeros(21, np.complex128))
            np.testing.assert_allclose(got, exp, rtol=0, atol=1e-14)
    # known answer: <arange|III|arange> = 140 (test_pauli_string.py:285-289)
    got = H.string_expectation(ops_of(["III"]), 0, np.arange(8).reshape(8, 1))
    np.testing.assert_allclose(got, [140], rtol=1e-15)


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("n,N,S", [(1, 4, 1), (6, 30, 9), (10, 64, 16), (14, 256, 16), (12, 100, 1000)])
def test_op_expectation(hip, dtype, n, N, S):
    H.say(f"op_expectation n={n} N={N} S={S} dtype={dtype}")
    cdt = H.CDT[dtype]
    codes = syn.random_codes(N, n, seed=3 * N + n)
    strings = syn.codes_to_strings(codes)
    coeffs = syn.random_coeffs(N, seed=N + 1)
    ops = _abi.Ops(codes, coeffs)
    psi = syn.random_states(1 << n, S, seed=n + S).astype(cdt)
    got = H.op_expectation(ops, psi, dtype)
    # FP64 oracle on the same inputs
    exp = O.op_expectation(strings, coeffs.astype(cdt).astype(np.complex128),
                           psi.astype(np.complex128), np.zeros(S, np.complex128))
    tol = 1e-12 if dtype == C128 else 1e-5
    err = H.rel_err(got, exp)
    assert err <= tol, err
    assert H.ev_scaled_err(got, exp, coeffs, psi) <= tol
    base = syn.random_coeffs(S, seed=99).astype(cdt)
    got2 = H.op_expectation(ops, psi, dtype, out0=base)
    assert H.rel_err(got2 - base, exp) <= 10 * tol


def _structured_codes(kind, N, n, seed):
    if kind == "random":
        return syn.random_codes(N, n, seed=seed)
    if kind == "weight2":
        return syn.random_codes_max_weight(N, n, 2, seed=seed)
    if kind == "diagonal":  # I/Z only: every x mask is 0
        c = syn.random_codes(N, n, seed=seed)
        return np.where(c == 0, 0, 3).astype(np.uint8)
    if kind == "shared_x":  # 4 distinct x masks, many z masks
        c = syn.random_codes(N, n, seed=seed)
        xs = (c == 1) | (c == 2)
        xs = xs[np.arange(N) % 4]
        zs = syn.random_codes(N, n, seed=seed + 1) >= 2
        return np.select([xs & zs, xs & ~zs, ~xs & zs], [2, 1, 3], 0).astype(np.uint8)
    if kind == "duplicates":
        c = syn.random_codes(max(1, N // 4), n, seed=seed)
        return np.concatenate([c] * (N // len(c) + 1))[:N]
    raise ValueError(kind)


@pytest.mark.parametrize("kernel", ["stream", "tiled", "tiled_wide", "tiled_kb128", "auto"])
@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("kind,n,N,S", [
    ("random", 2, 5, 1), ("random", 8, 40, 3), ("random", 11, 64, 16), ("random", 13, 200, 7),
    ("random", 15, 50, 33), ("random", 16, 300, 2), ("random", 18, 64, 16), ("weight2", 14, 300, 5),
    ("diagonal", 12, 40, 16), ("shared_x", 12, 64, 4), ("duplicates", 10, 33, 1), ("random", 20, 24, 1),
    ("random", 8, 40, 16), ("random", 9, 40, 32), ("random", 14, 120, 64), ("random", 12, 90, 4),
])
def test_op_expectation_kernels(hip, monkeypatch, kernel, dtype, kind, n, N, S):
    """PauliOp::expectation_value through the streaming kernel, the coset-tiled
    kernel and the automatic choice, on random and structured string sets, at
    ragged state counts; same bars as above.  Repeated calls are identical."""
    H.say(f"op_expectation kernel={kernel} {kind} n={n} N={N} S={S} dtype={dtype}")
    if kernel != "auto":
        monkeypatch.setenv("FAST_PAULI_EV_KERNEL", kernel.split("_")[0])
    if kernel == "tiled_wide":  # 64-bit element offsets, normally used from 2^32 elements up
        monkeypatch.setenv("FAST_PAULI_TILE_WIDE", "1")
    if kernel == "tiled_kb128":  # 128 KiB tiles: 1024-thread blocks, shared groups of up to 10
        monkeypatch.setenv("FAST_PAULI_TILE_KB", "128")
    cdt = H.CDT[dtype]
    codes = _structured_codes(kind, N, n, seed=5 * n + N)
    strings = syn.codes_to_strings(codes)
    coeffs = syn.random_coeffs(N, seed=N + 7)
    ops = _abi.Ops(codes, coeffs)
    psi = syn.random_states(1 << n, S, seed=n + 3 * S).astype(cdt)
    got = H.op_expectation(ops, psi, dtype)
    exp = O.op_expectation(strings, coeffs.astype(cdt).astype(np.complex128),
                           psi.astype(np.complex128), np.zeros(S, np.complex128))
    tol = 1e-12 if dtype == C128 else 1e-5
    assert H.rel_err(got, exp) <= tol
    assert H.ev_scaled_err(got, exp, coeffs, psi) <= tol
    assert np.array_equal(H.op_expectation(ops, psi, dtype), got)
    base = syn.random_coeffs(S, seed=98).astype(cdt)
    got2 = H.op_expectation(ops, psi, dtype, out0=base)
    assert H.rel_err(got2 - base, exp) <= 10 * tol


def test_op_expectation_plan(hip, monkeypatch):
    """fp_hip_op_expectation_plan: at 16 complex128 states, tiles take 128-byte
    row slices (8 states, 2^9-row tiles) and random strings group about 9 per
    tile; complex64 takes 16 states per block (2^9 rows, 128-byte row segments, about 9
    strings per group, shared leaders with a power-of-two state count);
    the bytes follow the groups; a forced streaming kernel reports one state
    read per string."""
    H.say("op_expectation plan")
    n, N, S = 16, 256, 16
    ops = _abi.Ops(syn.random_codes(N, n, seed=4), syn.random_coeffs(N, seed=4))
    p = _abi.op_expectation_plan(ops, C128, 1 << n, S)
    assert p["kernel"] == _abi.FP_EV_TILED
    assert p["tile_rows_log2"] == 9 and p["states_per_block"] == 8
    assert N // 9 <= p["n_groups"] <= N // 4
    assert p["hbm_bytes"] == p["n_groups"] * (1 << n) * S * 16
    assert N // 18 <= p["shared_groups"] and p["shared_groups"] + p["hoisted_groups"] <= p["n_groups"]
    p64 = _abi.op_expectation_plan(ops, C64, 1 << n, S)
    assert p64["tile_rows_log2"] == 9 and p64["states_per_block"] == 16 and N // 9 <= p64["n_groups"] <= N // 4
    assert p64["hbm_bytes"] == p64["n_groups"] * (1 << n) * S * 8
    assert N // 18 <= p64["shared_groups"] and p64["shared_groups"] + p64["hoisted_groups"] <= p64["n_groups"]
    # complex64 shared leaders need the paired loads: a non-power-of-two state count
    # runs the other instances
    p64b = _abi.op_expectation_plan(ops, C64, 1 << n, 12)
    assert p64b["shared_groups"] == 0
    single = _abi.Ops(syn.random_codes(1, n, seed=5))
    assert _abi.op_expectation_plan(single, C128, 1 << n, S)["kernel"] == _abi.FP_EV_STREAM
    monkeypatch.setenv("FAST_PAULI_EV_KERNEL", "stream")
    q = _abi.op_expectation_plan(ops, C128, 1 << n, S)
    assert q["kernel"] == _abi.FP_EV_STREAM and q["hbm_bytes"] == N * (1 << n) * S * 16
    with pytest.raises(ValueError):
        _abi.op_expectation_plan(ops, C128, 1 << (n - 1), S)


@pytest.mark.parametrize("pack", ["1", "2"])
@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("S", [1, 2, 16, 24])
def test_tiled_expectation_both_instances(hip, monkeypatch, dtype, S, pack):
    """The tiled kernel's instances: groups whose strings all share leader rows (at
    most 9 strings, rebased so that every local x mask has the lowest span bit) run in
    the shared-leader instance (complex64: with a power-of-two state count and full
    tiles, two states per lane), other groups whose waves
    each own one single-string unit in the set-up-once instance (two states per
    lane for complex64, which needs 2 states or more per block), the rest in the
    per-tile instance; launches mixing them equal the oracle, and so do launches
    of one kind alone.  pack 1: the shared-leader instance with each string's partner
    offset and coset sign mask in one word (the default below 2^16 cosets); 2: apart."""
    H.say(f"tiled expectation instances dtype={dtype} S={S} pack={pack}")
    monkeypatch.setenv("FAST_PAULI_EV_KERNEL", "tiled")
    monkeypatch.setenv("FAST_PAULI_EV_PACK", pack)
    n = 16
    cdt = H.CDT[dtype]
    psi = syn.random_states(1 << n, S, seed=41).astype(cdt)
    mixed = np.concatenate([syn.random_codes(64, n, seed=41), _structured_codes("shared_x", 12, n, seed=42)])
    cases = [("random", syn.random_codes(64, n, seed=43)), ("shared_x", _structured_codes("shared_x", 40, n, seed=44)),
             ("mixed", mixed)]
    for name, codes in cases:
        coeffs = syn.random_coeffs(codes.shape[0], seed=45)
        ops = _abi.Ops(codes, coeffs)
        p = _abi.op_expectation_plan(ops, dtype, 1 << n, S)
        assert p["kernel"] == _abi.FP_EV_TILED
        h = p["hoisted_groups"] + p["shared_groups"]
        assert h <= p["n_groups"], (name, p)
        # a random group of up to 8 strings shares leaders; one that took more than
        # 8 strings (x masks falling in its span) may still set up once when it has
        # at most 8 chunks of one string each
        if dtype == C64 and S == 1:
            assert h == 0, name  # one state per block: no two-state lanes
        elif name == "random" and (dtype == C128 or S & (S - 1) == 0):
            assert p["shared_groups"] > 0, (name, p)
        elif name == "random":  # complex64 shared leaders need a power-of-two state count
            assert p["shared_groups"] == 0 and p["hoisted_groups"] > 0, (name, p)
        elif name == "shared_x":
            assert h == 0, (name, p)  # 40 strings of 4 x masks: one group, chunks of 4
        else:
            assert 0 < h < p["n_groups"], (name, p)
        got = H.op_expectation(ops, psi, dtype)
        exp = O.op_expectation(syn.codes_to_strings(codes), coeffs.astype(cdt).astype(np.complex128),
                               psi.astype(np.complex128), np.zeros(S, np.complex128))
        tol = 1e-12 if dtype == C128 else 1e-5
        assert H.ev_scaled_err(got, exp, coeffs, psi) <= tol, name
        assert np.array_equal(H.op_expectation(ops, psi, dtype), got), name


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("n,S", [(14, 16), (18, 8), (20, 8)])
def test_tiled_expectation_packed_equals_unpacked(hip, monkeypatch, dtype, n, S):
    """The shared-leader instance with packed per-string words gives the same bits as
    the unpacked one (same arithmetic in the same order), and both match the oracle."""
    H.say(f"tiled expectation packed vs unpacked dtype={dtype} n={n} S={S}")
    monkeypatch.setenv("FAST_PAULI_EV_KERNEL", "tiled")
    cdt = H.CDT[dtype]
    codes = syn.random_codes(32, n, seed=n + S)
    coeffs = syn.random_coeffs(32, seed=n + S)
    ops = _abi.Ops(codes, coeffs)
    psi = syn.random_states(1 << n, S, seed=n * S).astype(cdt)
    assert _abi.op_expectation_plan(ops, dtype, 1 << n, S)["shared_groups"] > 0
    monkeypatch.setenv("FAST_PAULI_EV_PACK", "1")
    packed = H.op_expectation(ops, psi, dtype)
    monkeypatch.setenv("FAST_PAULI_EV_PACK", "2")
    unpacked = H.op_expectation(ops, psi, dtype)
    assert np.array_equal(packed, unpacked)
    exp = O.op_expectation(syn.codes_to_strings(codes), coeffs.astype(cdt).astype(np.complex128),
                           psi.astype(np.complex128), np.zeros(S, np.complex128))
    tol = 1e-12 if dtype == C128 else 1e-5
    assert H.ev_scaled_err(packed, exp, coeffs, psi) <= tol


@pytest.mark.parametrize("dtype", [C128, C64])
def test_tiled_tables_follow_assign_and_coeffs(hip, monkeypatch, dtype):
    """The tiled kernel's cached string groups are rebuilt after
    fp_hip_ops_assign (new strings, device memory reused) and after
    fp_hip_ops_set_coeffs (new scales): results equal a fresh operator set."""
    H.say(f"tiled tables after assign / set_coeffs dtype={dtype}")
    monkeypatch.setenv("FAST_PAULI_EV_KERNEL", "tiled")
    n, S = 12, 8
    cdt = H.CDT[dtype]
    psi = syn.random_states(1 << n, S, seed=31).astype(cdt)
    ops = _abi.Ops(syn.random_codes(64, n, seed=31), syn.random_coeffs(64, seed=31))
    H.op_expectation(ops, psi, dtype)
    for N, seed in ((100, 32), (40, 33), (100, 34)):
        codes = syn.random_codes(N, n, seed=seed)
        coeffs = syn.random_coeffs(N, seed=seed)
        ops.assign(codes, coeffs)
        _abi.synchronize()
        got = H.op_expectation(ops, psi, dtype)
        ref = H.op_expectation(_abi.Ops(codes, coeffs), psi, dtype)
        assert np.array_equal(got, ref), (N, seed)
    new = syn.random_coeffs(100, seed=35)
    ops.set_coeffs(new)
    got = H.op_expectation(ops, psi, dtype)
    ref = H.op_expectation(_abi.Ops(syn.random_codes(100, n, seed=34), new), psi, dtype)
    assert np.array_equal(got, ref)


def test_tiled_kernels_random_shapes(hip, monkeypatch):
    """Seeded random shapes (n 1..15, strings 1..200, states 1..40, both dtypes,
    random / weight-2 / diagonal / shared-x / duplicate string sets), forced
    through the tiled expectation and apply kernels, against the oracle."""
    rng = np.random.default_rng(2026)
    monkeypatch.setenv("FAST_PAULI_EV_KERNEL", "tiled")
    monkeypatch.setenv("FAST_PAULI_APPLY_KERNEL", "tiled")
    kinds = ["random", "weight2", "diagonal", "shared_x", "duplicates"]
    for case in range(40):
        n = int(rng.integers(1, 16))
        N = int(rng.integers(1, 201))
        S = int(rng.integers(1, 41))
        kind = kinds[int(rng.integers(0, len(kinds)))] if n >= 2 else "random"
        dtype = C128 if rng.random() < 0.5 else C64
        H.say(f"random tiled case {case}: {kind} n={n} N={N} S={S} dtype={dtype}")
        cdt = H.CDT[dtype]
        codes = _structured_codes(kind, N, n, seed=1000 + case)
        N = codes.shape[0]
        strings = syn.codes_to_strings(codes)
        coeffs = syn.random_coeffs(N, seed=2000 + case)
        ops = _abi.Ops(codes, coeffs)
        psi = syn.random_states(1 << n, S, seed=3000 + case).astype(cdt)
        got = H.op_expectation(ops, psi, dtype)
        exp = O.op_expectation(strings, coeffs.astype(cdt).astype(np.complex128),
                               psi.astype(np.complex128), np.zeros(S, np.complex128))
        tol = 1e-12 if dtype == C128 else 1e-5
        assert H.ev_scaled_err(got, exp, coeffs, psi) <= tol, case
        got = H.op_apply(ops, psi, dtype)
        exp = O.op_apply(strings, coeffs.astype(cdt), psi, np.zeros_like(psi), dtype=odt(dtype))
        assert np.array_equal(got, exp), case


def test_expectation_known_answers_and_invariants(hip):
    H.say("expectation known answers")
    a = np.arange(8).astype(np.complex128).reshape(8, 1)
    np.testing.assert_allclose(H.op_expectation(ops_of(["III", "III"], [1, 1]), a), [280], rtol=1e-15)
    assert np.all(H.op_expectation(ops_of(["IXYZ", "ZYXI"], [1, 2]), np.zeros((16, 16))) == 0)
    # sum rule over all 4^n strings for normalised states
    n = 4
    strings = ["".join(p) for p in it.product("IXYZ", repeat=n)]
    psi = syn.random_states(1 << n, 5, seed=5)
    ev = H.ev_matrix(ops_of(strings), psi)
    np.testing.assert_allclose((ev ** 2).sum(axis=0), 2 ** n, rtol=1e-13)


@pytest.mark.parametrize("dtype", [C128, C64])
def test_strings_expectation_matrix(hip, dtype):
    n, N, S = 8, 200, 33
    H.say(f"ev matrix n={n} N={N} S={S} dtype={dtype}")
    codes = syn.random_codes_max_weight(N, n, 3, seed=2)
    strings = syn.codes_to_strings(codes)
    psi = syn.random_states(1 << n, S, seed=2).astype(H.CDT[dtype])
    got = H.ev_matrix(_abi.Ops(codes), psi, dtype)
    p128 = psi.astype(np.complex128)
    exp = np.stack([O.string_expectation(s, 1, p128, np.zeros(S, np.complex128)) for s in strings])
    tol = 1e-12 if dtype == C128 else 1e-5
    assert np.abs(exp.imag).max() < 1e-15
    assert H.rel_err(got, exp.real) <= tol


# ---------------------------------------------------------------------------
# SummedPauliOp::expectation_value and the MFMA contraction
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode", [_abi.FP_CONTRACT_FP64, _abi.FP_CONTRACT_FP32, _abi.FP_CONTRACT_BF16X3])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (16, 16, 4), (200, 1000, 2000), (33, 65, 129)])
def test_gemm(hip, mode, m, n, k):
    H.say(f"gemm mode={mode} m={m} n={n} k={k}")
    a = syn.random_real_matrix(m, k, seed=m)
    b = syn.random_real_matrix(k, n, seed=n)
    got = H.gemm(mode, a, b)
    dt = np.float64 if mode == _abi.FP_CONTRACT_FP64 else np.float32
    exp = a.astype(dt).astype(np.float64) @ b.astype(dt).astype(np.float64)
    tol = {_abi.FP_CONTRACT_FP64: 1e-13, _abi.FP_CONTRACT_FP32: 1e-6, _abi.FP_CONTRACT_BF16X3: 3e-5}[mode]
    assert H.rel_err(got, exp) <= tol


@pytest.mark.parametrize("dtype,mode,tol", [
    (C128, _abi.FP_CONTRACT_FP64, 1e-12),
    (C128, _abi.FP_CONTRACT_FP32, 1e-5),
    (C64, _abi.FP_CONTRACT_FP32, 1e-5),
    (C64, _abi.FP_CONTRACT_BF16X3, 1e-4),
])
def test_summed_expectation(hip, dtype, mode, tol):
    n, N, K, S = 8, 150, 20, 40
    H.say(f"summed_expectation n={n} N={N} K={K} S={S} dtype={dtype} mode={mode}")
    codes = syn.random_codes_max_weight(N, n, 3, seed=7)
    strings = syn.codes_to_strings(codes)
    W = syn.random_real_matrix(K, N, seed=7)  # (operators x strings), real weights
    coeffs_nk = W.T.astype(np.complex128)
    ops = _abi.Ops(codes, matrix=coeffs_nk)
    psi = syn.random_states(1 << n, S, seed=8).astype(H.CDT[dtype])
    got = H.summed_expectation(ops, psi, dtype, mode)
    exp = O.summed_expectation(strings, coeffs_nk, psi.astype(np.complex128),
                               np.zeros((K, S), np.complex128))
    assert H.rel_err(got, exp) <= tol


def test_summed_expectation_complex_coeffs(hip):
    n, N, K, S = 5, 40, 6, 7
    H.say("summed_expectation complex coefficients")
    codes = syn.random_codes(N, n, seed=17)
    strings = syn.codes_to_strings(codes)
    rng = np.random.default_rng(17)
    coeffs_nk = rng.standard_normal((N, K)) + 1j * rng.standard_normal((N, K))
    ops = _abi.Ops(codes, matrix=coeffs_nk)
    psi = syn.random_states(1 << n, S, seed=18)
    got = H.summed_expectation(ops, psi, C128, _abi.FP_CONTRACT_FP64)
    exp = O.summed_expectation(strings, coeffs_nk, psi, np.zeros((K, S), np.complex128))
    assert H.rel_err(got, exp) <= 1e-12
    got1 = H.summed_expectation(ops, psi[:, :1].copy(), C128, _abi.FP_CONTRACT_FP64)
    assert H.rel_err(got1[:, 0], exp[:, 0]) <= 1e-12


# ---------------------------------------------------------------------------
# synthetic inputs on the device
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [C128, C64])
def test_fill_states_normalised(hip, dtype):
    dim, S = 1 << 12, 5
    H.say(f"fill_states dtype={dtype}")
    d = _abi.DeviceBuffer(dim * S * np.dtype(H.CDT[dtype]).itemsize)
    _abi.call("fp_hip_fill_states", dtype, d.ptr, dim, S, syn.gpu_seed(18), 1, None)
    psi = d.to_array(H.CDT[dtype], (dim, S)).astype(np.complex128)
    np.testing.assert_allclose(np.linalg.norm(psi, axis=0), 1.0, rtol=1e-12 if dtype == C128 else 1e-6)
    # the device stream follows the host generator up to libm rounding
    host = syn.random_states(dim, S, seed=18)
    assert H.rel_err(psi, host) <= (1e-12 if dtype == C128 else 1e-6)


# ---------------------------------------------------------------------------
# errors
# ---------------------------------------------------------------------------
def test_shape_errors(hip):
    H.say("shape errors")
    ops = ops_of(["XYZ"])
    d = _abi.DeviceBuffer(16 * 4)
    with pytest.raises(ValueError):
        _abi.call("fp_hip_string_apply_batch", ops.handle, 0, C128, d.ptr, d.ptr + 64, 4, 1, 1.0, 0.0, 0, None)
    with pytest.raises(ValueError):
        _abi.call("fp_hip_op_expectation", ops.handle, C128, d.ptr, d.ptr, 8, 0, 0, None)
    with pytest.raises(ValueError):
        _abi.call("fp_hip_string_apply_batch", ops.handle, 5, C128, d.ptr, d.ptr, 8, 1, 1.0, 0.0, 0, None)
    with pytest.raises(ValueError):
        _abi.Ops(np.array([[0, 4]], np.uint8))
    # partially overlapping input and output (out starts one element into states)
    d2 = _abi.DeviceBuffer(16 * 9)
    with pytest.raises(ValueError):
        _abi.call("fp_hip_string_apply_batch", ops.handle, 0, C128, d2.ptr, d2.ptr + 16, 8, 1, 1.0, 0.0, 0, None)
    with pytest.raises(ValueError):
        _abi.call("fp_hip_op_apply", ops.handle, C128, d2.ptr + 16, d2.ptr, 8, 1, 0, None)
    with pytest.raises(ValueError):
        _abi.call("fp_hip_op_apply", ops.handle, C64, d2.ptr, d2.ptr + 56, 8, 1, 0, None)


@pytest.mark.parametrize("dtype,n,S,kernel", [
    (C64, 6, 3, "auto"), (C64, 13, 40, "auto"), (C64, 14, 17, "auto"), (C128, 13, 5, "auto"),
    (C128, 9, 1100, "auto"), (C64, 15, 4, "auto"), (C128, 14, 9, "auto"), (C128, 14, 9, "stream"),
    (C128, 15, 1, "auto"), (C128, 16, 3, "auto"), (C64, 17, 2, "auto"), (C64, 15, 4, "stream")])
def test_strings_expectation_matrix_all_row_modes(hip, monkeypatch, dtype, n, S, kernel):
    """Strings with x = 0, with all X/Y letters on the low 6 row bits, and with
    higher X/Y letters (the three row enumerations of the LDS kernels).  Whole
    columns fit in LDS up to n = 13 (complex128) / 14 (complex64); above that
    the coset kernel runs, or the streaming kernel when forced."""
    H.say(f"ev matrix modes n={n} S={S} dtype={dtype} kernel={kernel}")
    if kernel != "auto":
        monkeypatch.setenv("FAST_PAULI_EV_MATRIX_KERNEL", kernel)
    rng = np.random.default_rng(n * 100 + S)
    strings = []
    for _ in range(40):  # diagonal strings
        strings.append("".join(rng.choice(list("IZ"), n)))
    for _ in range(40):  # X/Y only on the last 6 characters (row bits 0..5)
        strings.append("".join(rng.choice(list("IZ"), n - 6)) + "".join(rng.choice(list("IXYZ"), 6)))
    for _ in range(80):  # anything
        strings.append("".join(rng.choice(list("IXYZ"), n)))
    strings += ["X" + "I" * (n - 1), "Y" * n, "I" * n]
    codes = O.codes_of(strings)
    psi = syn.random_states(1 << n, S, seed=n + S).astype(H.CDT[dtype])
    got = H.ev_matrix(_abi.Ops(codes), psi, dtype)
    p128 = psi.astype(np.complex128)
    exp = np.stack([O.string_expectation(s, 1, p128, np.zeros(S, np.complex128)).real for s in strings])
    tol = 1e-12 if dtype == C128 else 1e-5
    assert H.rel_err(got, exp) <= tol
    # each row against its own scale: |<P>| <= 1 for normalised states
    assert np.abs(got - exp).max() <= (1e-13 if dtype == C128 else 2e-6)


@pytest.mark.parametrize("dtype,n,S", [(C128, 10, 7), (C64, 12, 5), (C128, 13, 33), (C64, 14, 9),
                                       (C128, 14, 3), (C128, 15, 2), (C64, 16, 4), (C128, 11, 1)])
def test_strings_expectation_matrix_transform_classes(hip, dtype, n, S):
    """Classes of >= 12 strings with one x (evaluated through a Walsh-Hadamard transform of the
    pair values): x = 0, x on the lowest bits, around bit 6 and on the top bit, both halves of w
    (X or Y letters on x), and a few small classes beside them, against the oracle."""
    H.say(f"ev matrix transform classes n={n} S={S} dtype={dtype}")
    rng = np.random.default_rng(1000 + n * 10 + S)
    strings = []
    x_sets = [[], [n - 1], [n - 2, n - 1], [n - 7], [n - 8, n - 3], [0], [0, n // 2, n - 1]]
    for pos in x_sets:
        for _ in range(30):
            s = list(rng.choice(list("IZ"), n))
            for p in pos:
                s[p] = rng.choice(list("XY"))
            strings.append("".join(s))
    strings += ["".join(rng.choice(list("IXYZ"), n)) for _ in range(25)]
    codes = O.codes_of(strings)
    psi = syn.random_states(1 << n, S, seed=n + S).astype(H.CDT[dtype])
    got = H.ev_matrix(_abi.Ops(codes), psi, dtype)
    p128 = psi.astype(np.complex128)
    exp = np.stack([O.string_expectation(s, 1, p128, np.zeros(S, np.complex128)).real for s in strings])
    tol = 1e-12 if dtype == C128 else 1e-5
    assert H.rel_err(got, exp) <= tol
    assert np.abs(got - exp).max() <= (1e-13 if dtype == C128 else 2e-6)
