"""Full-size checks of the headline workload (n = 24, 16 states).

Against the oracle, on host copies of the device-generated states:
  * apply_batch of the benchmark's string, IEEE-equal (complex128, complex64);
  * PauliOp::expectation_value of the first 16 of the benchmark's 256 strings
    on the default (coset-tiled) plan, <= 1e-12 / <= 1e-5;
  * PauliOp::apply of 16 strings on the default (coset-tiled) plan, IEEE-equal.
The full 256-string expectation (about 16x the oracle's time of the 16-string
one) is checked against the per-string streaming kernel instead, and 64-string
PauliOp::apply tiled against streaming, IEEE-equal.

Size-independent properties besides:
  * P^2 = I: applying a string twice returns the input bit-exactly (each
    element is only sign-flipped / re-im swapped);
  * <psi|I|psi> = 1 for states normalised on the device;
  * linearity: the PauliOp expectation equals the coefficient-weighted sum of
    the single-string expectations;
  * at n = 22, <psi|P|psi> from the expectation kernel equals the inner
    product of psi with the apply kernel's P psi (computed on the host).
"""

import numpy as np
import pytest

from fast_pauli import _abi
from fast_pauli import synthetic as syn
from oracle import oracle as O
from tests import hiputil as H

pytestmark = pytest.mark.gpu

C128, C64 = _abi.FP_C128, _abi.FP_C64


@pytest.mark.parametrize("dtype", [C128, C64])
def test_apply_twice_is_identity_n24(hip, dtype):
    n, S = 24, 16
    dim = 1 << n
    esz = np.dtype(H.CDT[dtype]).itemsize
    H.say(f"fullsize involution n={n} S={S} dtype={dtype}")
    a = _abi.DeviceBuffer(dim * S * esz)
    b = _abi.DeviceBuffer(dim * S * esz)
    c = _abi.DeviceBuffer(dim * S * esz)
    _abi.call("fp_hip_fill_states", dtype, a.ptr, dim, S, syn.gpu_seed(18), 1, None)
    ops = _abi.Ops(syn.random_codes(1, n, seed=24))
    _abi.call("fp_hip_string_apply_batch", ops.handle, 0, dtype, a.ptr, b.ptr, dim, S, 1.0, 0.0, 0, None)
    _abi.call("fp_hip_string_apply_batch", ops.handle, 0, dtype, b.ptr, c.ptr, dim, S, 1.0, 0.0, 0, None)
    ha = a.to_array(H.CDT[dtype], (dim, S))
    hc = c.to_array(H.CDT[dtype], (dim, S))
    assert np.array_equal(ha, hc)
    hb = b.to_array(H.CDT[dtype], (dim, S))
    assert not np.array_equal(ha, hb)


@pytest.mark.parametrize("dtype", [C128, C64])
def test_expectation_linearity_n24(hip, dtype):
    n, S, N = 24, 16, 4
    dim = 1 << n
    esz = np.dtype(H.CDT[dtype]).itemsize
    H.say(f"fullsize linearity n={n} S={S} N={N} dtype={dtype}")
    a = _abi.DeviceBuffer(dim * S * esz)
    _abi.call("fp_hip_fill_states", dtype, a.ptr, dim, S, syn.gpu_seed(18), 1, None)
    codes = np.vstack([np.zeros((1, n), np.uint8), syn.random_codes(N - 1, n, seed=24)])
    coeffs = syn.random_coeffs(N, seed=24)
    ops = _abi.Ops(codes, coeffs)
    d_ev = _abi.DeviceBuffer(S * esz)
    singles = []
    for i in range(N):
        _abi.call("fp_hip_string_expectation", ops.handle, i, dtype, a.ptr, d_ev.ptr, dim, S, 1.0, 0.0, 0, None)
        singles.append(d_ev.to_array(H.CDT[dtype], (S,)).astype(np.complex128))
    tol = 1e-12 if dtype == C128 else 1e-5
    np.testing.assert_allclose(singles[0], 1.0, rtol=tol)  # identity string on normalised states
    _abi.call("fp_hip_op_expectation", ops.handle, dtype, a.ptr, d_ev.ptr, dim, S, 0, None)
    got = d_ev.to_array(H.CDT[dtype], (S,))
    exp = sum(c * e for c, e in zip(coeffs, singles))
    assert H.ev_scaled_err(got, exp, coeffs, np.ones((1, S))) <= tol


def test_expectation_equals_inner_product_n22(hip):
    n, S = 22, 16
    dim = 1 << n
    H.say(f"fullsize <psi|P psi> n={n} S={S}")
    a = _abi.DeviceBuffer(dim * S * 16)
    b = _abi.DeviceBuffer(dim * S * 16)
    _abi.call("fp_hip_fill_states", C128, a.ptr, dim, S, syn.gpu_seed(7), 1, None)
    ops = _abi.Ops(syn.random_codes(1, n, seed=22))
    _abi.call("fp_hip_string_apply_batch", ops.handle, 0, C128, a.ptr, b.ptr, dim, S, 1.0, 0.0, 0, None)
    d_ev = _abi.DeviceBuffer(S * 16)
    _abi.call("fp_hip_string_expectation", ops.handle, 0, C128, a.ptr, d_ev.ptr, dim, S, 1.0, 0.0, 0, None)
    ev = d_ev.to_array(np.complex128, (S,))
    psi = a.to_array(np.complex128, (dim, S))
    ppsi = b.to_array(np.complex128, (dim, S))
    ref = np.einsum("jt,jt->t", psi.conj(), ppsi)
    assert np.abs(ev - ref).max() <= 1e-12


# ---------------------------------------------------------------------------
# The headline workload at full size against the oracle: n = 24, 16 states,
# the benchmark's strings (one apply string from seed 18, the 256 expectation
# strings and coefficients from seed 19).  The oracle runs its OpenMP loops on
# host copies of the device-generated states (4 GiB in complex128).
# Bars: apply IEEE-equal; expectation values <= 1e-12 (complex128) and <= 1e-5
# (complex64, against the FP64 oracle on the same complex64 inputs).
# ---------------------------------------------------------------------------
N_Q, N_S = 24, 16


@pytest.fixture(scope="module")
def headline(hip):
    dim = 1 << N_Q
    H.say(f"fullsize headline states n={N_Q} S={N_S}: fill, convert, copy to host")
    d128 = _abi.DeviceBuffer(dim * N_S * 16)
    d64 = _abi.DeviceBuffer(dim * N_S * 8)
    _abi.call("fp_hip_fill_states", C128, d128.ptr, dim, N_S, syn.gpu_seed(18), 1, None)
    _abi.call("fp_hip_convert", C128, d128.ptr, C64, d64.ptr, dim * N_S, None)
    data = {
        C128: (d128, d128.to_array(np.complex128, (dim, N_S))),
        C64: (d64, d64.to_array(np.complex64, (dim, N_S))),
        "codes_apply": syn.random_codes(1, N_Q, seed=18),
        "codes_ev": syn.random_codes(256, N_Q, seed=19),
        "coeffs_ev": syn.random_coeffs(256, seed=19),
    }
    yield data
    del data


@pytest.mark.parametrize("dtype", [C128, C64])
def test_headline_apply_batch_bit_exact_n24(headline, dtype):
    dim = 1 << N_Q
    d_in, psi = headline[dtype]
    codes = headline["codes_apply"]
    H.say(f"fullsize apply_batch vs oracle n={N_Q} S={N_S} dtype={dtype} string={syn.codes_to_strings(codes)[0]}")
    ops = _abi.Ops(codes)
    d_out = _abi.DeviceBuffer(psi.nbytes)
    _abi.call("fp_hip_string_apply_batch", ops.handle, 0, dtype, d_in.ptr, d_out.ptr, dim, N_S, 1.0, 0.0, 0, None)
    got = d_out.to_array(H.CDT[dtype], psi.shape)
    del d_out
    exp = O.string_apply_batch(syn.codes_to_strings(codes)[0], 1, psi, np.zeros_like(psi), par=True,
                               dtype=O.C128 if dtype == C128 else O.C64)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("dtype", [C128, C64])
def test_headline_expectation_16_strings_vs_oracle_n24(headline, dtype):
    dim = 1 << N_Q
    d_in, psi = headline[dtype]
    codes, coeffs = headline["codes_ev"][:16], headline["coeffs_ev"][:16]
    ops = _abi.Ops(codes, coeffs)
    plan = _abi.op_expectation_plan(ops, dtype, dim, N_S)
    H.say(f"fullsize expectation x16 vs oracle n={N_Q} S={N_S} dtype={dtype} plan={plan}")
    assert plan["kernel"] == _abi.FP_EV_TILED
    d_ev = _abi.DeviceBuffer(N_S * 16)
    _abi.call("fp_hip_op_expectation", ops.handle, dtype, d_in.ptr, d_ev.ptr, dim, N_S, 0, None)
    got = d_ev.to_array(H.CDT[dtype], (N_S,)).astype(np.complex128)
    strings = syn.codes_to_strings(codes)
    if dtype == C128:
        exp = O.op_expectation(strings, coeffs, psi, np.zeros(N_S, np.complex128), par=True)
        tol = 1e-12
    else:
        exp = O.op_expectation(strings, coeffs.astype(np.complex64).astype(np.complex128),
                               psi.astype(np.complex128), np.zeros(N_S, np.complex128), par=True)
        tol = 1e-5
    assert H.rel_err(got, exp) <= tol
    assert H.ev_scaled_err(got, exp, coeffs, np.ones((1, N_S))) <= tol


@pytest.mark.parametrize("dtype", [C128, C64])
def test_headline_expectation_256_default_vs_streaming_n24(headline, monkeypatch, dtype):
    """The 256-string default plan (string groups on coset tiles; shared leaders
    in complex128) against the per-string streaming kernel, whose arithmetic the
    oracle pins at smaller sizes.  complex64 is compared with the complex128
    streaming kernel on the same complex64 inputs (FP64 arithmetic)."""
    dim = 1 << N_Q
    d_in, _ = headline[dtype]
    ops = _abi.Ops(headline["codes_ev"], headline["coeffs_ev"])
    plan = _abi.op_expectation_plan(ops, dtype, dim, N_S)
    H.say(f"fullsize expectation x256 default vs streaming n={N_Q} S={N_S} dtype={dtype} plan={plan}")
    assert plan["kernel"] == _abi.FP_EV_TILED
    d_ev = _abi.DeviceBuffer(N_S * 16)
    _abi.call("fp_hip_op_expectation", ops.handle, dtype, d_in.ptr, d_ev.ptr, dim, N_S, 0, None)
    got = d_ev.to_array(H.CDT[dtype], (N_S,)).astype(np.complex128)
    monkeypatch.setenv("FAST_PAULI_EV_KERNEL", "stream")
    if dtype == C128:
        src = d_in
    else:
        src = _abi.DeviceBuffer(dim * N_S * 16)
        _abi.call("fp_hip_convert", C64, d_in.ptr, C128, src.ptr, dim * N_S, None)
    _abi.call("fp_hip_op_expectation", ops.handle, C128, src.ptr, d_ev.ptr, dim, N_S, 0, None)
    ref = d_ev.to_array(np.complex128, (N_S,))
    tol = 1e-12 if dtype == C128 else 1e-5
    assert H.rel_err(got, ref) <= tol
    assert H.ev_scaled_err(got, ref, headline["coeffs_ev"], np.ones((1, N_S))) <= tol


@pytest.mark.parametrize("dtype,n_strings", [(C128, 16), (C64, 16)])
def test_headline_op_apply_vs_oracle_n24(headline, dtype, n_strings):
    """PauliOp::apply at n = 24 on its default (coset-tiled) plan against the
    reference's sequential loop (string by string, accumulating), IEEE-equal."""
    dim = 1 << N_Q
    d_in, psi = headline[dtype]
    codes, coeffs = headline["codes_ev"][:n_strings], headline["coeffs_ev"][:n_strings]
    ops = _abi.Ops(codes, coeffs)
    plan = _abi.op_apply_plan(ops, dtype, dim, N_S)
    H.say(f"fullsize PauliOp.apply x{n_strings} vs oracle n={N_Q} S={N_S} dtype={dtype} plan={plan}")
    assert plan["kernel"] == _abi.FP_EV_TILED
    d_out = _abi.DeviceBuffer(psi.nbytes)
    _abi.call("fp_hip_op_apply", ops.handle, dtype, d_in.ptr, d_out.ptr, dim, N_S, 0, None)
    got = d_out.to_array(H.CDT[dtype], psi.shape)
    del d_out
    cdt = H.CDT[dtype]
    exp = np.zeros_like(psi)
    # PauliOp::apply's sequential branch (__pauli_op.hpp:459-467) is apply_batch per string,
    # accumulating; apply_batch's OpenMP branch performs the same element operations
    for s, c in zip(syn.codes_to_strings(codes), coeffs.astype(cdt)):
        O.string_apply_batch(s, c, psi, exp, par=True, dtype=O.C128 if dtype == C128 else O.C64)
    assert np.array_equal(got, exp)


def test_headline_op_apply_64_default_vs_streaming_n24(headline, monkeypatch):
    """The benchmark's PauliOp.apply (64 strings, 8 tiled group launches) against
    the streaming kernel, IEEE-equal."""
    dim = 1 << N_Q
    d_in, psi = headline[C128]
    ops = _abi.Ops(headline["codes_ev"][:64], headline["coeffs_ev"][:64])
    H.say(f"fullsize PauliOp.apply x64 default vs streaming n={N_Q} S={N_S}")
    d_a = _abi.DeviceBuffer(psi.nbytes)
    _abi.call("fp_hip_op_apply", ops.handle, C128, d_in.ptr, d_a.ptr, dim, N_S, 0, None)
    a = d_a.to_array(np.complex128, psi.shape)
    monkeypatch.setenv("FAST_PAULI_APPLY_KERNEL", "stream")
    _abi.call("fp_hip_op_apply", ops.handle, C128, d_in.ptr, d_a.ptr, dim, N_S, 0, None)
    b = d_a.to_array(np.complex128, psi.shape)
    assert np.array_equal(a, b)
