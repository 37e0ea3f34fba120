"""Serving loop: each served result equals the same request run alone (apply
bit-exact, expectation values within the stated tolerance), admission honours
the HBM budget, batches overlap on the device, and a Poisson run reports
latency percentiles."""

import numpy as np
import pytest

from fast_pauli import _abi
from fast_pauli import synthetic as syn
from fast_pauli.serving import Server, request_mix, run_poisson
from oracle import oracle as O
from tests import hiputil as H

pytestmark = pytest.mark.gpu

C128, C64 = _abi.FP_C128, _abi.FP_C64


def alone(kind, codes, coeffs, n, S, seed, dtype):
    """The request through the single-operator entry points, on the same device states."""
    esz = 16 if dtype == C128 else 8
    d_in = _abi.DeviceBuffer((1 << n) * S * esz)
    _abi.call("fp_hip_fill_states", dtype, d_in.ptr, 1 << n, S, seed, 1, None)
    ops = _abi.Ops(codes, coeffs)
    if kind == _abi.FP_BATCH_APPLY:
        d_out = _abi.DeviceBuffer((1 << n) * S * esz)
        _abi.call("fp_hip_op_apply", ops.handle, dtype, d_in.ptr, d_out.ptr, 1 << n, S, 0, None)
        return d_out.to_array(H.CDT[dtype], (1 << n, S)), d_in.to_array(H.CDT[dtype], (1 << n, S))
    d_out = _abi.DeviceBuffer(S * esz)
    _abi.call("fp_hip_op_expectation", ops.handle, dtype, d_in.ptr, d_out.ptr, 1 << n, S, 0, None)
    return d_out.to_array(H.CDT[dtype], (S,)), d_in.to_array(H.CDT[dtype], (1 << n, S))


@pytest.mark.parametrize("dtype", [C64, C128])
def test_served_equals_alone(hip, dtype):
    H.say(f"serving alone-vs-mix dtype={dtype}")
    reqs = request_mix(48, seed=5, n_range=(4, 12), strings_range=(2, 64), states_range=(1, 9))
    with Server(hbm_budget_bytes=1 << 30, n_slots=3, max_batch=16) as srv:
        ids = []
        for r in reqs:
            codes = syn.random_codes(r["strings"], r["n"], seed=r["seed"])
            coeffs = syn.random_coeffs(r["strings"], seed=r["seed"])
            rid = srv.submit(r["kind"], codes, coeffs, r["states"], r["seed"], dtype,
                             keep_output=True)
            ids.append((rid, r, codes, coeffs))
        for rid, r, codes, coeffs in ids:
            res = srv.wait(rid, timeout_s=60)
            assert res is not None and res["status"] == 0
            got = srv.fetch(rid)
            ref, psi = alone(r["kind"], codes, coeffs, r["n"], r["states"], r["seed"], dtype)
            if r["kind"] == _abi.FP_BATCH_APPLY:
                assert np.array_equal(got, ref)
            else:
                tol = 1e-12 if dtype == C128 else 1e-5
                assert H.rel_err(got, ref) <= tol
                # and against the FP64 oracle on the same states
                strings = syn.codes_to_strings(codes)
                exp = O.op_expectation(strings, coeffs.astype(H.CDT[dtype]).astype(np.complex128),
                                       psi.astype(np.complex128), np.zeros(r["states"], np.complex128))
                assert H.rel_err(got, exp) <= tol
            srv.release(rid)
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert st["batches"] < len(reqs)  # requests were batched together


def test_admission_budget_and_oversize_requests(hip):
    H.say("serving admission budget")
    with Server(hbm_budget_bytes=4 << 20, n_slots=2, max_batch=8) as srv:
        big = srv.submit_apply(syn.random_codes(4, 20, seed=1), n_states=4, seed=1)  # 32 MiB states
        small = [srv.submit_expectation(syn.random_codes(8, 10, seed=i), n_states=4, seed=i) for i in range(20)]
        assert srv.wait(big, 60)["status"] == _abi.FP_ERR_OUT_OF_MEMORY
        for rid in small:
            assert srv.wait(rid, 60)["status"] == 0
        st = srv.stats()
        assert st["peak_bytes_in_use"] <= 4 << 20
        assert st["failed"] == 1


def test_batches_overlap_on_device(hip):
    H.say("serving overlap")
    with Server(hbm_budget_bytes=8 << 30, n_slots=4, max_batch=4) as srv:
        ids = [srv.submit_expectation(syn.random_codes(256, 18, seed=i), n_states=16, seed=i) for i in range(32)]
        for rid in ids:
            assert srv.wait(rid, 120)["status"] == 0
        assert srv.stats()["max_concurrent_batches"] >= 2


def test_poisson_run_reports_percentiles(hip):
    H.say("serving poisson")
    reqs = request_mix(200, seed=9, n_range=(10, 16), strings_range=(8, 256), states_range=(1, 16))
    with Server(hbm_budget_bytes=16 << 30, n_slots=4, max_batch=32) as srv:
        out = run_poisson(srv, reqs, qps=400.0, seed=9)
    assert out["failed"] == 0
    assert 0 < out["p50_ms"] <= out["p99_ms"]
    print(out)


def test_oversized_request_rejected_at_submit(hip):
    """State or output bytes that do not fit in size_t (2^40 rows x 2^20 states x
    16 B = 2^64) are refused at submission with a clear error; nothing is queued
    and the loop keeps serving."""
    H.say("serving oversized request at submit")
    with Server(hbm_budget_bytes=1 << 30, n_slots=1, max_batch=4) as srv:
        codes = np.zeros((1, 40), np.uint8)
        for n_states in (1 << 20, (1 << 24) + 3):
            with pytest.raises(ValueError):
                srv.submit(_abi.FP_BATCH_EXPECT, codes, None, n_states, 1, C128)
        with pytest.raises(ValueError):
            srv.submit(_abi.FP_BATCH_APPLY, codes, None, 1 << 20, 1, C128)
        rid = srv.submit_expectation(syn.random_codes(4, 8, seed=1), n_states=2, seed=1)
        assert srv.wait(rid, 60)["status"] == 0
        assert srv.stats()["submitted"] == 1
