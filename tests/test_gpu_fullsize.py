"""Full-size (n = 24, 16 states) checks through size-independent properties.

The oracle cannot run these sizes in seconds, so the checks use identities
that hold exactly or to rounding at any size:
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
