"""fp_hip_op_expectation_noalloc (the allocation-free variant of
fp_hip_op_expectation, its buffers in a static device array) against
fp_hip_op_expectation, through the C ABI.

Bars: <= 1e-12 (complex128) / <= 1e-5 (complex64) normwise relative, as for
fp_hip_op_expectation against the oracle; at these shapes the partials fit in
the static workspace without fewer blocks, so the results are also identical.
Each case runs twice in a row on one GPU, the second call with a smaller
input, so stale contents of the shared array would show.
"""

import numpy as np
import pytest

from fast_pauli import _abi
from fast_pauli import synthetic as syn
from tests import hiputil as H

pytestmark = pytest.mark.gpu

C128, C64 = _abi.FP_C128, _abi.FP_C64


def _both(ops, d_in, dim, S, dtype, acc0=None):
    esz = 16 if dtype == C128 else 8
    res = []
    for entry in ("fp_hip_op_expectation", "fp_hip_op_expectation_noalloc"):
        if acc0 is None:
            d_out = _abi.DeviceBuffer(S * esz)
            _abi.call(entry, ops.handle, dtype, d_in.ptr, d_out.ptr, dim, S, 0, None)
        else:
            d_out = _abi.DeviceBuffer.from_array(acc0.astype(H.CDT[dtype]))
            _abi.call(entry, ops.handle, dtype, d_in.ptr, d_out.ptr, dim, S, 1, None)
        res.append(d_out.to_array(H.CDT[dtype], (S,)))
    return res


@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("kernel,n,N,S_big,S_small", [
    ("tiled", 20, 200, 16, 3),     # coset-tiled groups (the default for many strings)
    ("tiled", 14, 64, 40, 7),      # ragged state counts
    ("stream", 12, 1, 9, 2),       # one string: the streaming kernel
])
def test_noalloc_matches_op_expectation(hip, dtype, kernel, n, N, S_big, S_small):
    dim = 1 << n
    tol = 1e-12 if dtype == C128 else 1e-5
    codes = syn.random_codes(N, n, seed=n + N)
    coeffs = syn.random_coeffs(N, seed=n + N)
    ops = _abi.Ops(codes, coeffs)
    for S in (S_big, S_small):  # the second call has the smaller input
        plan = _abi.op_expectation_plan(ops, dtype, dim, S)
        H.say(f"noalloc vs op_expectation kernel={kernel} n={n} N={N} S={S} dtype={dtype} plan={plan}")
        assert plan["kernel"] == (_abi.FP_EV_TILED if kernel == "tiled" else _abi.FP_EV_STREAM)
        d_in = _abi.DeviceBuffer(dim * S * (16 if dtype == C128 else 8))
        _abi.call("fp_hip_fill_states", dtype, d_in.ptr, dim, S, syn.gpu_seed(S), 1, None)
        ref, got = _both(ops, d_in, dim, S, dtype)
        assert H.rel_err(got, ref) <= tol
        assert np.array_equal(got, ref)
        base = syn.random_coeffs(S, seed=S)
        ref, got = _both(ops, d_in, dim, S, dtype, acc0=base)
        assert H.rel_err(got, ref) <= tol
        assert np.array_equal(got, ref)


def test_noalloc_rejects_oversized_partials(hip):
    """n_states x 16 B of partials per block above the static workspace: a clear
    error before any kernel runs."""
    n, S = 1, (32 << 20) // 16 + 1
    H.say(f"noalloc oversized partials n={n} S={S}")
    ops = _abi.Ops(np.array([[3]], np.uint8))
    d_in = _abi.DeviceBuffer(2 * S * 16)
    d_out = _abi.DeviceBuffer(S * 16)
    with pytest.raises(ValueError, match="workspace"):
        _abi.call("fp_hip_op_expectation_noalloc", ops.handle, C128, d_in.ptr, d_out.ptr, 2, S, 0, None)
