"""GPU parity at the shapes of the other benchmark workloads, through the C ABI.

* SummedPauliOp::expectation_value at n = 14 with 2000 strings of weight <= 3,
  a real (200 operators x 2000 strings) weight matrix and 1000 normalised
  states, in every contraction precision, against the oracle's OpenMP loops
  (__summed_pauli_op.hpp:530-614) on host copies of the device states.
  Bars: FP64 <= 1e-12, FP32 <= 1e-5, BF16x3 <= 1e-4 (normwise relative, against
  the FP64 oracle on the same inputs).
* SummedPauliOp::apply and apply_weighted at n = 12..14 in complex128 and
  complex64, IEEE-equal to the oracle's sequential loops
  (__summed_pauli_op.hpp:287-349, 378-503).
* The serving loop over its full request ranges (n 10..22, 8..1024 strings,
  1..32 states, complex64, 200 GB budget): every result equals the same request
  run alone (apply IEEE-equal, expectation values <= 1e-5), requests share
  batches, and small expectation requests match the FP64 oracle.
"""

import numpy as np
import pytest

from fast_pauli import _abi
from fast_pauli import synthetic as syn
from fast_pauli.serving import Server, request_mix
from oracle import oracle as O
from tests import hiputil as H

pytestmark = pytest.mark.gpu

C128, C64 = _abi.FP_C128, _abi.FP_C64


# ---------------------------------------------------------------------------
# SummedPauliOp::expectation_value at n = 14, 2000 strings, 200 operators, 1000 states
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def summed14(hip):
    n, N, K, S = 14, 2000, 200, 1000
    dim = 1 << n
    H.say(f"summed n={n} N={N} K={K} S={S}: states on the device, oracle on the host")
    codes = syn.random_codes_max_weight(N, n, 3, seed=18)
    W = syn.random_real_matrix(K, N, seed=18)
    coeffs_nk = np.ascontiguousarray(W.T.astype(np.complex128))
    d128 = _abi.DeviceBuffer(dim * S * 16)
    d64 = _abi.DeviceBuffer(dim * S * 8)
    _abi.call("fp_hip_fill_states", C128, d128.ptr, dim, S, syn.gpu_seed(18), 1, None)
    _abi.call("fp_hip_convert", C128, d128.ptr, C64, d64.ptr, dim * S, None)
    psi128 = d128.to_array(np.complex128, (dim, S))
    psi64w = d64.to_array(np.complex64, (dim, S)).astype(np.complex128)
    strings = syn.codes_to_strings(codes)
    exp128 = O.summed_expectation(strings, coeffs_nk, psi128, np.zeros((K, S), np.complex128), par=True)
    exp64 = O.summed_expectation(strings, coeffs_nk, psi64w, np.zeros((K, S), np.complex128), par=True)
    ops = _abi.Ops(codes, matrix=coeffs_nk)
    yield {"n": n, "N": N, "K": K, "S": S, "ops": ops, C128: (d128, exp128), C64: (d64, exp64)}


@pytest.mark.parametrize("dtype,mode,tol", [
    (C128, _abi.FP_CONTRACT_FP64, 1e-12),
    (C128, _abi.FP_CONTRACT_FP32, 1e-5),
    (C64, _abi.FP_CONTRACT_FP32, 1e-5),
    (C64, _abi.FP_CONTRACT_BF16X3, 1e-4),
])
def test_summed_expectation_n14_full_shape(summed14, dtype, mode, tol):
    c = summed14
    dim, K, S = 1 << c["n"], c["K"], c["S"]
    d_in, exp = c[dtype]
    H.say(f"summed n=14 N=2000 K=200 S=1000 dtype={dtype} mode={mode}")
    d_out = _abi.DeviceBuffer(K * S * 16)
    _abi.call("fp_hip_summed_expectation", c["ops"].handle, dtype, mode, d_in.ptr, d_out.ptr, dim, S, 0, None)
    got = d_out.to_array(H.CDT[dtype], (K, S))
    err = H.rel_err(got, exp)
    print(f"rel err vs FP64 oracle: {err:.3e}")
    assert err <= tol
    # the per-operator error too: every row within 10x the bar of its own norm
    row = np.linalg.norm(got.astype(np.complex128) - exp, axis=1) / np.linalg.norm(exp, axis=1)
    assert row.max() <= 10 * tol


def test_summed_ev_matrix_n14_vs_oracle_sample(summed14):
    """Stage 1 alone (the per-string expectation matrix, the dominant kernel of
    this workload) on 64 strings spread over the 2000, FP64, against the oracle's
    PauliString::expectation_value per string."""
    c = summed14
    dim, S = 1 << c["n"], c["S"]
    d_in, _ = c[C128]
    H.say("summed n=14 stage-1 matrix (complex128) vs oracle sample")
    d_ev = _abi.DeviceBuffer(c["N"] * S * 8)
    _abi.call("fp_hip_strings_expectation_matrix", c["ops"].handle, C128, d_in.ptr, d_ev.ptr, dim, S, None)
    ev = d_ev.to_array(np.float64, (c["N"], S))
    psi = d_in.to_array(np.complex128, (dim, S))
    codes = syn.random_codes_max_weight(c["N"], c["n"], 3, seed=18)
    strings = syn.codes_to_strings(codes)
    for j in range(0, c["N"], c["N"] // 64):
        ref = O.string_expectation(strings[j], 1, psi, np.zeros(S, np.complex128), par=True)
        assert np.abs(ref.imag).max() <= 1e-15
        assert np.abs(ev[j] - ref.real).max() <= 1e-13, j


# ---------------------------------------------------------------------------
# SummedPauliOp::apply / apply_weighted at n >= 12 in both precisions
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kernel", ["stream", "tiled"])
@pytest.mark.parametrize("dtype", [C128, C64])
@pytest.mark.parametrize("n,N,K,S", [(12, 200, 20, 16), (13, 97, 7, 5), (14, 60, 33, 32)])
def test_summed_apply_and_weighted_bit_exact(hip, monkeypatch, kernel, dtype, n, N, K, S):
    """SummedPauliOp::apply (the row-summed coefficients) and apply_weighted (per-state
    weights) through the streaming kernels and (forced) the coset-tiled kernel: IEEE-equal
    to the oracle's sequential loops."""
    monkeypatch.setenv("FAST_PAULI_APPLY_KERNEL", kernel)
    H.say(f"summed apply / apply_weighted kernel={kernel} n={n} N={N} K={K} S={S} dtype={dtype}")
    cdt = H.CDT[dtype]
    rdt = np.float64 if dtype == C128 else np.float32
    odt = O.C128 if dtype == C128 else O.C64
    codes = syn.random_codes(N, n, seed=n * 7 + N)
    strings = syn.codes_to_strings(codes)
    rng = np.random.default_rng(n + N + K + S)
    coeffs_nk = rng.standard_normal((N, K)) + 1j * rng.standard_normal((N, K))
    ops = _abi.Ops(codes, matrix=coeffs_nk)
    psi = syn.random_states(1 << n, S, seed=n + S).astype(cdt)
    data = rng.standard_normal((K, S)).astype(rdt)
    dim = 1 << n
    d_in = _abi.DeviceBuffer.from_array(psi)
    d_data = _abi.DeviceBuffer.from_array(data)
    d_out = _abi.DeviceBuffer(psi.nbytes)
    c_dt = coeffs_nk.astype(cdt)
    want = _abi.FP_EV_TILED if kernel == "tiled" else _abi.FP_EV_STREAM
    assert _abi.summed_apply_plan(ops, dtype, dim, S)["kernel"] == want
    assert _abi.summed_apply_weighted_plan(ops, dtype, dim, S)["kernel"] == want
    _abi.call("fp_hip_summed_apply", ops.handle, dtype, d_in.ptr, d_out.ptr, dim, S, 0, None)
    got = d_out.to_array(cdt, psi.shape)
    exp = O.summed_apply(strings, c_dt, psi, np.zeros_like(psi), dtype=odt)
    assert np.array_equal(got, exp)
    _abi.call("fp_hip_summed_apply_weighted", ops.handle, dtype, d_in.ptr, d_data.ptr, d_out.ptr, dim, S, 0, None)
    got = d_out.to_array(cdt, psi.shape)
    exp = O.summed_apply_weighted(strings, c_dt, psi, data, np.zeros_like(psi), dtype=odt)
    assert np.array_equal(got, exp)
    # accumulate into a non-zero output
    base = syn.random_states(dim, S, seed=99, normalise=False).astype(cdt)
    d_out.upload(base)
    _abi.call("fp_hip_summed_apply", ops.handle, dtype, d_in.ptr, d_out.ptr, dim, S, 1, None)
    got = d_out.to_array(cdt, psi.shape)
    exp = O.summed_apply(strings, c_dt, psi, base.copy(), dtype=odt)
    assert np.array_equal(got, exp)
    d_out.upload(base)
    _abi.call("fp_hip_summed_apply_weighted", ops.handle, dtype, d_in.ptr, d_data.ptr, d_out.ptr, dim, S, 1, None)
    got = d_out.to_array(cdt, psi.shape)
    exp = O.summed_apply_weighted(strings, c_dt, psi, data, base.copy(), dtype=odt)
    assert np.array_equal(got, exp)


# ---------------------------------------------------------------------------
# Serving over the full request ranges
# ---------------------------------------------------------------------------
def _alone(kind, codes, coeffs, n, S, seed, dtype):
    esz = 16 if dtype == C128 else 8
    d_in = _abi.DeviceBuffer((1 << n) * S * esz)
    _abi.call("fp_hip_fill_states", dtype, d_in.ptr, 1 << n, S, seed, 1, None)
    ops = _abi.Ops(codes, coeffs)
    if kind == _abi.FP_BATCH_APPLY:
        d_out = _abi.DeviceBuffer((1 << n) * S * esz)
        _abi.call("fp_hip_op_apply", ops.handle, dtype, d_in.ptr, d_out.ptr, 1 << n, S, 0, None)
        return d_out.to_array(H.CDT[dtype], (1 << n, S)), d_in
    d_out = _abi.DeviceBuffer(S * esz)
    _abi.call("fp_hip_op_expectation", ops.handle, dtype, d_in.ptr, d_out.ptr, 1 << n, S, 0, None)
    return d_out.to_array(H.CDT[dtype], (S,)), d_in


def test_serving_full_ranges_alone_equals_mix(hip):
    """A burst of 40 requests from the serving workload's default ranges, all
    submitted before the first completes, so that requests share batches."""
    reqs = request_mix(40, seed=18)
    H.say(f"serving full ranges: {len(reqs)} requests, n {min(r['n'] for r in reqs)}..{max(r['n'] for r in reqs)}, "
          f"strings {min(r['strings'] for r in reqs)}..{max(r['strings'] for r in reqs)}")
    with Server(hbm_budget_bytes=200 << 30, n_slots=4, max_batch=32) as srv:
        ids = []
        for r in reqs:
            codes = syn.random_codes(r["strings"], r["n"], seed=r["seed"])
            coeffs = syn.random_coeffs(r["strings"], seed=r["seed"])
            rid = srv.submit(r["kind"], codes, coeffs, r["states"], r["seed"], C64, keep_output=True)
            ids.append((rid, r, codes, coeffs))
        batch_of = {}
        for rid, r, codes, coeffs in ids:
            res = srv.wait(rid, timeout_s=120)
            assert res is not None and res["status"] == 0, r
            batch_of[rid] = res["batch_index"]
            got = srv.fetch(rid)
            srv.release(rid)
            ref, d_in = _alone(r["kind"], codes, coeffs, r["n"], r["states"], r["seed"], C64)
            if r["kind"] == _abi.FP_BATCH_APPLY:
                assert np.array_equal(got, ref), r
            else:
                assert H.rel_err(got, ref) <= 1e-5, r
                if r["n"] <= 16:
                    psi = d_in.to_array(np.complex64, (1 << r["n"], r["states"])).astype(np.complex128)
                    exp = O.op_expectation(syn.codes_to_strings(codes),
                                           coeffs.astype(np.complex64).astype(np.complex128), psi,
                                           np.zeros(r["states"], np.complex128), par=True)
                    assert H.rel_err(got, exp) <= 1e-5, r
            del d_in
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        sizes = np.bincount(np.unique(list(batch_of.values()), return_inverse=True)[1])
        print(f"batches {st['batches']} for {len(reqs)} requests, largest {sizes.max()}")
        assert st["batches"] < len(reqs)
        assert sizes.max() > 1
