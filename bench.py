#!/usr/bin/env python3
"""Benchmark of the operator-application hot path on one MI355X.

Workload (one "step"), n = 24 qubits, 16 normalised random states
(2^24 x 16 row-major, device-resident before timing starts):
  1. PauliString::apply_batch of one random string   (out = P psi)
  2. PauliOp::expectation_value of 256 random strings with N(0,1) complex
     coefficients                                     (out[t] = sum_s c_s <psi_t|P_s|psi_t>)
Strings have every site iid uniform over {I,X,Y,Z}; everything is seeded
(splitmix64, seed 18).  The headline runs in complex128 (the reference's
precision); the complex64 / FP32-accumulation form is measured on the same
states rounded to complex64 and its error is stated against FP64 arithmetic
on those same inputs.

value = strings x states per second over the whole step = (1 + 256) * 16 / t_step.

Roofline: algorithmic HBM bytes of each kernel launch divided by its average
duration, measured with HIP events on the stream the kernels run on:
  expectation_value (coset-tiled, the default): groups x 2^n x S x 16 B
                     (each group of strings reads the state once; the plan
                     reports the group count)
  expectation_value (streaming form, reported beside it): N x 2^n x S x 16 B
  apply_batch:       2 x 2^n x S x 16 B   (read psi, write P psi)
against the MI355X HBM3E peak of 8.0 TB/s.

Parity is checked in the same run: the default expectation kernel against the
streaming kernel (<= 1e-12), the complex64 form against FP64 arithmetic on the
same inputs (<= 1e-5), and, on the CPU baseline's sample of strings, both
precisions against the oracle (<= 1e-12 / <= 1e-5) and apply_batch against the
oracle (IEEE-equal).  The script exits with status 1 when a check fails.

The CPU baseline is the repository's OpenMP restatement of the reference's
loops (oracle/), timed on this host on a bounded sample of the same workload
with OMP_PROC_BIND=close and every available core unless OMP_NUM_THREADS says
otherwise.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "fast-pauli_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
AFFINITY_CPUS = None  # CPUs this process may run on, read at start-up


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=24)
    ap.add_argument("--states", type=int, default=16)
    ap.add_argument("--strings", type=int, default=256)
    ap.add_argument("--seed", type=int, default=18)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-c64", action="store_true", help="skip the complex64 leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other workloads (n=10 PauliOp.apply, n=14 SummedPauliOp, uneven batch, serving)")
    ap.add_argument("--qps", type=float, default=50.0, help="serving: Poisson arrival rate")
    ap.add_argument("--serve-requests", type=int, default=300, help="serving: number of requests")
    ap.add_argument("--qps-sweep", default="50,150,300,450,600",
                    help="serving: comma-separated arrival rates for the saturation sweep ('' to skip)")
    ap.add_argument("--pmc-pass", action="store_true",
                    help="only run the profiled kernels once each (a plain copy of the state, apply_batch, "
                         "expectation_value) for hardware-counter runs")
    ap.add_argument("--pmc-summed", action="store_true",
                    help="only run SummedPauliOp.expectation_value at n=14 (2000 strings, 200 operators, 1000 "
                         "states) once in complex128/FP64 and once in complex64/FP32, for profiler runs")
    ap.add_argument("--pmc-summed-apply", action="store_true",
                    help="only run SummedPauliOp.apply and apply_weighted at n=14 (2000 strings, 200 operators, "
                         "1000 states, complex128) once each, for profiler runs")
    ap.add_argument("--pmc-batch", action="store_true",
                    help="only build the 256-operator uneven batch and launch it twice, for profiler runs")
    ap.add_argument("--pmc-gemm", action="store_true",
                    help="only run the SummedPauliOp contraction (200 x 2000) x (2000 x 1000) once per MFMA mode "
                         "(FP64, FP32, BF16x3) for hardware-counter runs")
    return ap.parse_args()


def pmc_gemm() -> int:
    """The n=14 SummedPauliOp contraction shape, once per contraction mode."""
    from fast_pauli import _abi
    from fast_pauli import synthetic as syn

    _abi.init(0)
    K, N, S = 200, 2000, 1000
    for mode, dt in ((_abi.FP_CONTRACT_FP64, np.float64), (_abi.FP_CONTRACT_FP32, np.float32),
                     (_abi.FP_CONTRACT_BF16X3, np.float32)):
        A = _abi.DeviceBuffer.from_array(syn.random_real_matrix(K, N, seed=18).astype(dt))
        B = _abi.DeviceBuffer.from_array(syn.random_real_matrix(N, S, seed=1).astype(dt))
        C = _abi.DeviceBuffer(K * S * np.dtype(dt).itemsize)
        _abi.call("fp_hip_gemm", mode, K, S, N, A.ptr, B.ptr, C.ptr, None)
        _abi.synchronize()
    log("pmc gemm pass done: FP64, FP32, BF16x3 contraction 200 x 2000 x 1000")
    return 0


def pmc_summed(args) -> int:
    """The n=14 SummedPauliOp workload, once per state precision (stage 1 + contraction)."""
    from fast_pauli import _abi
    from fast_pauli import synthetic as syn

    _abi.init(0)
    n, N, K, S = 14, 2000, 200, 1000
    dim = 1 << n
    codes = syn.random_codes_max_weight(N, n, 3, seed=args.seed)
    W = syn.random_real_matrix(K, N, seed=args.seed)
    ops = _abi.Ops(codes, matrix=W.T.astype(np.complex128))
    d128 = _abi.DeviceBuffer(dim * S * 16)
    d64 = _abi.DeviceBuffer(dim * S * 8)
    _abi.call("fp_hip_fill_states", _abi.FP_C128, d128.ptr, dim, S, syn.gpu_seed(args.seed), 1, None)
    _abi.call("fp_hip_convert", _abi.FP_C128, d128.ptr, _abi.FP_C64, d64.ptr, dim * S, None)
    out = _abi.DeviceBuffer(K * S * 16)
    for dt, mode, src in ((_abi.FP_C128, _abi.FP_CONTRACT_FP64, d128), (_abi.FP_C64, _abi.FP_CONTRACT_FP32, d64)):
        _abi.call("fp_hip_summed_expectation", ops.handle, dt, mode, src.ptr, out.ptr, dim, S, 0, None)
        _abi.synchronize()
    log("pmc summed pass done: n=14, 2000 strings, 200 operators, 1000 states, complex128/FP64 then complex64/FP32")
    return 0


def _summed_apply_operands(args, n, N, K, S):
    """The SummedPauliOp.apply workload's operator set (N strings of weight <= 3, real h of
    N x K) and apply_weighted data x (K x S): (ops, x, codes, h)."""
    from fast_pauli import _abi
    from fast_pauli import synthetic as syn

    codes = syn.random_codes_max_weight(N, n, 3, seed=args.seed)
    W = syn.random_real_matrix(K, N, seed=args.seed)
    h = W.T.astype(np.complex128)
    ops = _abi.Ops(codes, matrix=h)
    x = syn.random_real_matrix(K, S, seed=args.seed + 1)
    return ops, x, codes, h


def pmc_summed_apply(args) -> int:
    """SummedPauliOp.apply and apply_weighted on the n=14 operator set, complex128, once each."""
    from fast_pauli import _abi
    from fast_pauli import synthetic as syn

    _abi.init(0)
    n, N, K, S = 14, 2000, 200, 1000
    dim = 1 << n
    ops, x, _, _ = _summed_apply_operands(args, n, N, K, S)
    d_in = _abi.DeviceBuffer(dim * S * 16)
    _abi.call("fp_hip_fill_states", _abi.FP_C128, d_in.ptr, dim, S, syn.gpu_seed(args.seed), 1, None)
    d_x = _abi.DeviceBuffer.from_array(np.ascontiguousarray(x, np.float64))
    d_out = _abi.DeviceBuffer(dim * S * 16)
    _abi.call("fp_hip_summed_apply", ops.handle, _abi.FP_C128, d_in.ptr, d_out.ptr, dim, S, 0, None)
    _abi.call("fp_hip_summed_apply_weighted", ops.handle, _abi.FP_C128, d_in.ptr, d_x.ptr, d_out.ptr, dim, S, 0,
              None)
    _abi.synchronize()
    log("pmc summed apply pass done: n=14, 2000 strings, 200 operators, 1000 states, complex128, "
        f"{_abi.summed_apply_plan(ops, _abi.FP_C128, dim, S)['n_groups']} string groups")
    return 0


def pmc_batch(args) -> int:
    """The 256-operator uneven batch, default schedule: one build launch, one replay."""
    from fast_pauli import _abi, synthetic as syn
    from fast_pauli.batch import Batch, uneven_workload

    _abi.init(0)
    b = Batch()
    keep = []
    for sp in uneven_workload(256, seed=args.seed):
        n, N, S = sp["n"], sp["strings"], sp["states"]
        ops = _abi.Ops(syn.random_codes(N, n, seed=sp["seed"]), syn.random_coeffs(N, seed=sp["seed"]))
        d_in = _abi.DeviceBuffer((1 << n) * S * 8)
        d_out = _abi.DeviceBuffer((1 << n) * S * 8)
        _abi.call("fp_hip_fill_states", _abi.FP_C64, d_in.ptr, 1 << n, S, sp["seed"], 1, None)
        b.add_apply(ops, d_in.ptr, d_out.ptr, 1 << n, S, _abi.FP_C64)
        keep.append((ops, d_in, d_out))
    b.launch()
    b.launch()
    _abi.synchronize()
    log(f"pmc batch pass done: {len(keep)} operators, {b.info()}")
    return 0


def main() -> int:
    args = parse()
    # the CPU baseline's OpenMP team: every available core unless the caller says otherwise,
    # threads bound close (read when the oracle library loads)
    global AFFINITY_CPUS
    AFFINITY_CPUS = len(os.sched_getaffinity(0))  # before OpenMP binds this thread to one place
    os.environ.setdefault("OMP_NUM_THREADS", str(AFFINITY_CPUS))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    if args.pmc_gemm:
        return pmc_gemm()
    if args.pmc_summed:
        return pmc_summed(args)
    if args.pmc_summed_apply:
        return pmc_summed_apply(args)
    if args.pmc_batch:
        return pmc_batch(args)
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; sharding over several GPUs "
                          "is not implemented", "n_gpus": args.gpus}), flush=True)
        return 2
    from fast_pauli import _abi
    from fast_pauli import synthetic as syn

    _abi.init(0)
    n, S, N = args.n, args.states, args.strings
    dim = 1 << n
    C128, C64 = _abi.FP_C128, _abi.FP_C64
    call = _abi.call
    stream = None  # the null stream: every kernel and event below is on it

    log(f"allocating n={n} S={S} complex128: {dim * S * 16 / 2**30:.1f} GiB per buffer")
    psi = _abi.DeviceBuffer(dim * S * 16)
    out = _abi.DeviceBuffer(dim * S * 16)
    ev = _abi.DeviceBuffer(S * 16)
    call("fp_hip_fill_states", C128, psi.ptr, dim, S, syn.gpu_seed(args.seed), 1, stream)

    codes_apply = syn.random_codes(1, n, seed=args.seed)
    codes_ev = syn.random_codes(N, n, seed=args.seed + 1)
    coeffs_ev = syn.random_coeffs(N, seed=args.seed + 1)
    ops_apply = _abi.Ops(codes_apply)
    ops_ev = _abi.Ops(codes_ev, coeffs_ev)

    def step(dtype, src, dst, evb, marks=None):
        if marks:
            marks[0].record(stream)
        call("fp_hip_string_apply_batch", ops_apply.handle, 0, dtype, src.ptr, dst.ptr, dim, S,
             1.0, 0.0, 0, stream)
        if marks:
            marks[1].record(stream)
        call("fp_hip_op_expectation", ops_ev.handle, dtype, src.ptr, evb.ptr, dim, S, 0, stream)
        if marks:
            marks[2].record(stream)

    def timed(dtype, src, dst, evb):
        for _ in range(args.warmup):
            step(dtype, src, dst, evb)
        _abi.synchronize()
        marks = [[_abi.Event() for _ in range(3)] for _ in range(args.steps)]
        t0, t1 = _abi.Event(), _abi.Event()
        _abi.synchronize()
        t0.record(stream)
        for k in range(args.steps):
            step(dtype, src, dst, evb, marks[k])
        t1.record(stream)
        t1.synchronize()
        _abi.synchronize()
        total_ms = t0.elapsed_ms(t1)
        apply_ms = [m[0].elapsed_ms(m[1]) for m in marks]
        ev_ms = [m[1].elapsed_ms(m[2]) for m in marks]
        return total_ms / args.steps, float(np.mean(apply_ms)), float(np.mean(ev_ms))

    plan128 = _abi.op_expectation_plan(ops_ev, C128, dim, S)
    plan64 = _abi.op_expectation_plan(ops_ev, C64, dim, S)
    ev_kernel = {_abi.FP_EV_TILED: "fpt::k_op_expect_tiled", _abi.FP_EV_STREAM: "fpk::k_op_expect"}

    def stream_ev_once(dtype, src, evb):
        """One launch of the streaming expectation kernel (every string reads the state)."""
        os.environ["FAST_PAULI_EV_KERNEL"] = "stream"
        try:
            call("fp_hip_op_expectation", ops_ev.handle, dtype, src.ptr, evb.ptr, dim, S, 0, stream)
        finally:
            os.environ.pop("FAST_PAULI_EV_KERNEL", None)

    if args.pmc_pass:
        # one launch of each kernel whose HBM bytes the counters measure
        call("fp_hip_copy", out.ptr, psi.ptr, dim * S * 16, stream)
        call("fp_hip_string_apply_batch", ops_apply.handle, 0, C128, psi.ptr, out.ptr, dim, S, 1.0, 0.0, 0, stream)
        call("fp_hip_op_expectation", ops_ev.handle, C128, psi.ptr, ev.ptr, dim, S, 0, stream)
        stream_ev_once(C128, psi, ev)
        # PauliOp.apply of the n=24 workload (64 strings): one coset-tiled launch per group
        ops_op = _op_apply_n24_ops(args)
        call("fp_hip_op_apply", ops_op.handle, C128, psi.ptr, out.ptr, dim, S, 0, stream)
        _abi.synchronize()
        log(f"pmc pass done: copy 4 GiB, apply_batch, expectation_value x{N} ({ev_kernel[plan128['kernel']]}, "
            f"then fpk::k_op_expect), PauliOp.apply x{ops_op.n_strings}, n={n}, {S} states, complex128")
        return 0

    # achievable copy bandwidth on this device (read + write of one 4 GiB buffer)
    ms_copy = _time_ms(lambda: call("fp_hip_copy", out.ptr, psi.ptr, dim * S * 16, stream), 5)
    copy_gbs = 2 * dim * S * 16 / (ms_copy / 1e3) / 1e9
    log(f"copy kernel: {ms_copy:.3f} ms for 2 x 4 GiB -> {copy_gbs:.0f} GB/s")

    log(f"complex128: warmup {args.warmup}, timing {args.steps} steps")
    ms_step, ms_apply, ms_ev = timed(C128, psi, out, ev)
    ev128 = ev.to_array(np.complex128, (S,))
    bc = 16
    bytes_apply = 2 * dim * S * bc
    bytes_ev = plan128["hbm_bytes"]
    value = (1 + N) * S / (ms_step / 1e3)
    log(f"complex128: step {ms_step:.3f} ms (apply {ms_apply:.3f} ms, EV {ms_ev:.3f} ms, "
        f"{ev_kernel[plan128['kernel']]}, {plan128['n_groups']} groups)")
    # the streaming kernel on the same inputs, for reference (roofline of the per-string form)
    ms_ev_stream = _time_ms(lambda: stream_ev_once(C128, psi, ev), 2)
    ev_stream = ev.to_array(np.complex128, (S,))
    bytes_ev_stream = N * dim * S * bc
    log(f"complex128 streaming EV kernel: {ms_ev_stream:.3f} ms")

    failures: list[str] = []
    rel_stream = float(np.linalg.norm(ev_stream - ev128) / np.linalg.norm(ev128))
    if not rel_stream <= 1e-12:
        failures.append(f"default expectation kernel differs from the streaming kernel by {rel_stream:.3e} > 1e-12")
    result = {
        "metric": "strings\u00d7states/s & % HBM peak at n=24; p50/p99 latency at fixed QPS",
        "value": value,
        "unit": "strings*states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded splitmix64 + Box-Muller, normalised states, uniform IXYZ letters)",
        "config": {"workload": f"PauliString.apply_batch x1 + PauliOp.expectation_value x{N}, "
                               f"n={n}, {S} states, complex128",
                   "n_qubits": n, "n_states": S, "n_strings_ev": N, "n_strings_apply": 1,
                   "parallelism": "single GPU"},
        "roofline": {
            "kernel": f"{ev_kernel[plan128['kernel']]} (PauliOp::expectation_value)",
            "bound": "hbm",
            "achieved": bytes_ev / (ms_ev / 1e3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": bytes_ev / (ms_ev / 1e3) / 1e9 / HBM_PEAK_GBS,
            "traffic": measured_traffic(ev_kernel[plan128["kernel"]], n, S, N),
            "algorithmic_bytes_per_launch": bytes_ev,
            "algorithmic_bytes_rule": "groups x 2^n x S x 16 B: each string group reads the state once"
                                      if plan128["kernel"] == _abi.FP_EV_TILED else
                                      "strings x 2^n x S x 16 B: each string reads the state once",
            "plan": plan128,
            "avg_launch_ms": ms_ev,
            "frac_of_measured_copy_bw": bytes_ev / (ms_ev / 1e3) / 1e9 / copy_gbs,
            "measured_copy_gbs": copy_gbs,
        },
        "roofline_ev_streaming": {
            "kernel": "fpk::k_op_expect (PauliOp::expectation_value, one state read per string)",
            "bound": "hbm",
            "achieved": bytes_ev_stream / (ms_ev_stream / 1e3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": bytes_ev_stream / (ms_ev_stream / 1e3) / 1e9 / HBM_PEAK_GBS,
            "traffic": measured_traffic("fpk::k_op_expect", n, S, N),
            "algorithmic_bytes_per_launch": bytes_ev_stream,
            "avg_launch_ms": ms_ev_stream,
            "speedup_of_tiled": ms_ev_stream / ms_ev,
            "rel_diff_vs_default_kernel": rel_stream,
        },
        "roofline_apply_batch": {
            "kernel": "k_string_apply (PauliString::apply_batch)",
            "bound": "hbm",
            "achieved": bytes_apply / (ms_apply / 1e3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": bytes_apply / (ms_apply / 1e3) / 1e9 / HBM_PEAK_GBS,
            "algorithmic_bytes_per_launch": bytes_apply,
            "avg_launch_ms": ms_apply,
            "strings_x_states_per_s": S / (ms_apply / 1e3),
            "frac_of_measured_copy_bw": bytes_apply / (ms_apply / 1e3) / 1e9 / copy_gbs,
            "traffic": measured_traffic("fpk::k_string_apply", n, S, 1),
        },
    }

    if not args.no_c64:
        # complex64 form on the same states rounded to complex64; FP64 reference on the same inputs
        psi64 = _abi.DeviceBuffer(dim * S * 8)
        out64 = _abi.DeviceBuffer(dim * S * 8)
        ev64 = _abi.DeviceBuffer(S * 8)
        call("fp_hip_convert", C128, psi.ptr, C64, psi64.ptr, dim * S, stream)
        log(f"complex64: warmup {args.warmup}, timing {args.steps} steps")
        ms_step64, ms_apply64, ms_ev64 = timed(C64, psi64, out64, ev64)
        got64 = ev64.to_array(np.complex64, (S,)).astype(np.complex128)
        # FP64 on the complex64 inputs: widen them into `out` and run the complex128 kernel
        call("fp_hip_convert", C64, psi64.ptr, C128, out.ptr, dim * S, stream)
        call("fp_hip_op_expectation", ops_ev.handle, C128, out.ptr, ev.ptr, dim, S, 0, stream)
        ref = ev.to_array(np.complex128, (S,))
        scale = np.abs(coeffs_ev).sum()
        err_norm = float(np.linalg.norm(got64 - ref) / np.linalg.norm(ref))
        err_state = float(np.max(np.abs(got64 - ref)) / scale)
        err_input = float(np.linalg.norm(ref - ev128) / np.linalg.norm(ev128))
        result["c64"] = {
            "value": (1 + N) * S / (ms_step64 / 1e3),
            "ms_per_step": ms_step64,
            "speedup_vs_c128": ms_step / ms_step64,
            "apply_batch_ms": ms_apply64,
            "apply_batch_frac_hbm": bytes_apply / 2 / (ms_apply64 / 1e3) / 1e9 / HBM_PEAK_GBS,
            "ev_ms": ms_ev64,
            "ev_kernel": ev_kernel[plan64["kernel"]],
            "ev_plan": plan64,
            "ev_frac_hbm": plan64["hbm_bytes"] / (ms_ev64 / 1e3) / 1e9 / HBM_PEAK_GBS,
            "ev_rel_err_vs_fp64_same_inputs": err_norm,
            "ev_max_abs_err_over_sum_abs_coeffs": err_state,
            "input_rounding_rel_effect": err_input,
        }
        log(f"complex64: step {ms_step64:.3f} ms (apply {ms_apply64:.3f}, EV {ms_ev64:.3f}); "
            f"EV rel err {err_norm:.2e}")
        if not err_norm <= 1e-5:
            failures.append(f"complex64 expectation differs from FP64 on the same inputs by {err_norm:.3e} > 1e-5")
        del out64, ev64
    else:
        psi64 = None

    if not args.no_cpu:
        cpu, apply_ref, ev_ref, n_sample = cpu_baseline(args, psi, codes_apply, codes_ev, coeffs_ev)
        result["cpu_baseline"] = cpu
        result["parity_vs_oracle"] = oracle_parity(args, psi, out, psi64, ops_apply, codes_ev, coeffs_ev, apply_ref,
                                                   ev_ref, n_sample, failures)
        del apply_ref
    del psi, out, psi64

    if not args.no_extra:
        result["workloads"] = {
            "pauli_op_apply_n10": bench_op_apply_n10(args),
            "pauli_op_apply_n24": bench_op_apply_n24(args),
            "summed_expectation_n14": bench_summed_n14(args),
            "summed_square_n14": bench_square_n14(args, failures),
            "summed_apply_n14": bench_summed_apply_n14(args, failures),
            "uneven_batch_256": bench_uneven_batch(args),
            "serving": bench_serving(args),
        }
        serving = result["workloads"]["serving"]
        result["p50_ms"] = serving["p50_ms"]
        result["p99_ms"] = serving["p99_ms"]
        result["serving_qps"] = serving["offered_qps"]

    result["parity_checks_failed"] = failures
    print(json.dumps(result), flush=True)
    if failures:
        for f in failures:
            log(f"PARITY FAILURE: {f}")
        return 1
    return 0


def measured_traffic(kernel: str, n: int, S: int, n_strings: int):
    """HBM bytes per call of `kernel` (summed over its launches when one call makes
    several, as the tiled PauliOp.apply does per string group) from the committed counter summary
    (profiles/pmc_traffic_n24_s16.json: read bytes from the TCC-to-memory read requests by size,
    32/64/128 B, write bytes from the 32/64-byte write requests, no scaling factor), for the
    shape it was measured at; None otherwise."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_n24_s16.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    entry = data.get("kernels", {}).get(kernel)
    if not entry or entry.get("n") != n or entry.get("states") != S or entry.get("strings") != n_strings:
        return None
    return entry["bytes_per_launch"]


def _time_ms(fn, reps):
    from fast_pauli import _abi

    fn()
    _abi.synchronize()
    a, b = _abi.Event(), _abi.Event()
    a.record(None)
    for _ in range(reps):
        fn()
    b.record(None)
    b.synchronize()
    return a.elapsed_ms(b) / reps


def bench_op_apply_n10(args) -> dict:
    """PauliOp::apply, n=10, 64 strings, 128 states, complex128 (cache-resident, latency bound)."""
    from fast_pauli import _abi, synthetic as syn

    n, N, S = 10, 64, 128
    log("workload: PauliOp.apply n=10, 64 strings, 128 states")
    codes = syn.random_codes(N, n, seed=args.seed)
    coeffs = syn.random_coeffs(N, seed=args.seed)
    ops = _abi.Ops(codes, coeffs)
    psi_h = syn.random_states(1 << n, S, seed=args.seed)
    d_in = _abi.DeviceBuffer.from_array(psi_h)
    d_out = _abi.DeviceBuffer(psi_h.nbytes)
    ms = _time_ms(lambda: _abi.call("fp_hip_op_apply", ops.handle, _abi.FP_C128, d_in.ptr, d_out.ptr, 1 << n, S, 0,
                                    None), 200)
    out = {"ms": ms, "strings_x_states_per_s": N * S / (ms / 1e3), "dtype": "c128",
           "config": "n=10, 64 strings (uniform IXYZ), 128 normalised states"}
    if not args.no_cpu:
        from oracle import oracle as O

        strings = syn.codes_to_strings(codes)
        o = np.zeros_like(psi_h)
        O.op_apply(strings, coeffs, psi_h, o, par=True)
        reps = 5
        t = time.perf_counter()
        for _ in range(reps):
            o[:] = 0
            O.op_apply(strings, coeffs, psi_h, o, par=True)
        t_cpu = (time.perf_counter() - t) / reps
        got = d_out.to_array(np.complex128, psi_h.shape)
        ref = O.op_apply(strings, coeffs, psi_h, np.zeros_like(psi_h))
        out["cpu_par_ms"] = t_cpu * 1e3
        out["cpu_threads"] = O.max_threads()
        out["bit_exact_vs_oracle_seq"] = bool(np.array_equal(got, ref))
    return out


def _op_apply_n24_ops(args):
    """The operator of the n=24 PauliOp.apply workload: 64 random strings."""
    from fast_pauli import _abi, synthetic as syn

    return _abi.Ops(syn.random_codes(64, 24, seed=args.seed + 1), syn.random_coeffs(64, seed=args.seed + 1))


def bench_op_apply_n24(args) -> dict:
    """PauliOp::apply, n=24, 64 strings, 16 states (HBM-resident): the plan's bytes against time."""
    from fast_pauli import _abi

    n, N, S = 24, 64, 16
    log("workload: PauliOp.apply n=24, 64 strings, 16 states")
    ops = _op_apply_n24_ops(args)
    res = {"config": "n=24, 64 strings (uniform IXYZ), 16 normalised states, overwrite"}
    for dt, name, esz in ((_abi.FP_C128, "c128", 16), (_abi.FP_C64, "c64", 8)):
        d_in = _abi.DeviceBuffer((1 << n) * S * esz)
        d_out = _abi.DeviceBuffer((1 << n) * S * esz)
        _abi.call("fp_hip_fill_states", dt, d_in.ptr, 1 << n, S, args.seed, 1, None)
        plan = _abi.op_apply_plan(ops, dt, 1 << n, S)
        ms = _time_ms(lambda: _abi.call("fp_hip_op_apply", ops.handle, dt, d_in.ptr, d_out.ptr, 1 << n, S, 0, None),
                      3)
        gbs = plan["hbm_bytes"] / (ms / 1e3) / 1e9
        res[name] = {"ms": ms, "strings_x_states_per_s": N * S / (ms / 1e3), "plan": plan,
                     "achieved_gbs": gbs, "frac_hbm": gbs / HBM_PEAK_GBS,
                     "streaming_model_bytes": (N + 1) * (1 << n) * S * esz}
        if dt == _abi.FP_C128:
            # HBM bytes of all group launches from the committed counter summary
            res[name]["traffic"] = measured_traffic("fpa::k_op_apply_tiled", n, S, N)
        del d_in, d_out
    return res


def bench_summed_n14(args) -> dict:
    """SummedPauliOp::expectation_value, n=14, 2000 strings (weight <= 3), 200 operators, 1000 states."""
    from fast_pauli import _abi, synthetic as syn

    n, N, K, S = 14, 2000, 200, 1000
    dim = 1 << n
    log("workload: SummedPauliOp.expectation_value n=14, 2000 strings, 200 operators, 1000 states")
    codes = syn.random_codes_max_weight(N, n, 3, seed=args.seed)
    W = syn.random_real_matrix(K, N, seed=args.seed)
    ops = _abi.Ops(codes, matrix=W.T.astype(np.complex128))
    d128 = _abi.DeviceBuffer(dim * S * 16)
    d64 = _abi.DeviceBuffer(dim * S * 8)
    _abi.call("fp_hip_fill_states", _abi.FP_C128, d128.ptr, dim, S, syn.gpu_seed(args.seed), 1, None)
    _abi.call("fp_hip_convert", _abi.FP_C128, d128.ptr, _abi.FP_C64, d64.ptr, dim * S, None)
    out128 = _abi.DeviceBuffer(K * S * 16)
    out64 = _abi.DeviceBuffer(K * S * 8)
    ev = _abi.DeviceBuffer(N * S * 8)
    res = {"config": "n=14, 2000 strings of weight <= 3, W (200 x 2000) ~ N(0,1), 1000 normalised states"}
    modes = [("c128_fp64", _abi.FP_C128, _abi.FP_CONTRACT_FP64, d128, out128),
             ("c64_fp32_mfma", _abi.FP_C64, _abi.FP_CONTRACT_FP32, d64, out64),
             ("c64_bf16x3_mfma", _abi.FP_C64, _abi.FP_CONTRACT_BF16X3, d64, out64)]
    ref = None
    for name, dt, mode, src, dst in modes:
        # 20 timed calls: over 3 the first calls' clock ramp adds about 0.15 ms to a 2 ms call
        ms = _time_ms(lambda: _abi.call("fp_hip_summed_expectation", ops.handle, dt, mode, src.ptr, dst.ptr, dim, S,
                                        0, None), 20)
        ms1 = _time_ms(lambda: _abi.call("fp_hip_strings_expectation_matrix", ops.handle, dt, src.ptr, ev.ptr, dim,
                                         S, None), 20)
        got = dst.to_array(np.complex128 if dt == _abi.FP_C128 else np.complex64, (K, S)).astype(np.complex128)
        entry = {"ms": ms, "stage1_ms": ms1, "strings_x_states_per_s": N * S / (ms / 1e3)}
        if ref is None:
            ref = got
        else:
            entry["rel_err_vs_c128_fp64"] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        res[name] = entry
    # the contraction alone on MFMA (200 x 2000) x (2000 x 1000)
    for name, mode, dt, peak in (("gemm_fp64", _abi.FP_CONTRACT_FP64, np.float64, 78.6),
                                 ("gemm_fp32", _abi.FP_CONTRACT_FP32, np.float32, 157.3),
                                 ("gemm_bf16x3", _abi.FP_CONTRACT_BF16X3, np.float32, 2500.0)):
        A = _abi.DeviceBuffer.from_array(W.astype(dt))
        B = _abi.DeviceBuffer.from_array(syn.random_real_matrix(N, S, seed=1).astype(dt))
        C = _abi.DeviceBuffer(K * S * np.dtype(dt).itemsize)
        ms = _time_ms(lambda: _abi.call("fp_hip_gemm", mode, K, S, N, A.ptr, B.ptr, C.ptr, None), 20)
        flops = 2.0 * K * N * S * (3 if mode == _abi.FP_CONTRACT_BF16X3 else 1)
        res[name] = {"ms": ms, "tflops": flops / (ms / 1e3) / 1e12,
                     "mfma_peak_tflops": peak, "frac_of_peak": flops / (ms / 1e3) / 1e12 / peak}
    if not args.no_cpu:
        from oracle import oracle as O

        # bounded CPU sample: 200 of the 2000 strings, scaled by 10 (strings run independently per thread)
        psi = d128.to_array(np.complex128, (dim, S))
        strings = syn.codes_to_strings(codes[:200])
        o = np.zeros((K, S), np.complex128)
        t = time.perf_counter()
        O.summed_expectation(strings, W.T[:200].astype(np.complex128), psi, o, par=True)
        t_cpu = (time.perf_counter() - t) * (N / 200)
        res["cpu_par_ms_estimate"] = t_cpu * 1e3
        res["cpu_sample"] = "200 of 2000 strings, time x10"
        res["cpu_threads"] = O.max_threads()
    return res


def bench_summed_apply_n14(args, failures) -> dict:
    """SummedPauliOp::apply and apply_weighted on the n=14 operator set (2000 strings of
    weight <= 3, 200 operators) and 1000 states, complex128: every output element sums its
    2000 strings' terms in order, through the coset-tiled kernel (string groups sharing LDS
    tiles) where its plan pays, else the streaming kernels (one partner read per string)."""
    from fast_pauli import _abi, synthetic as syn

    n, N, K, S = 14, 2000, 200, 1000
    dim = 1 << n
    log("workload: SummedPauliOp.apply / apply_weighted n=14, 2000 strings, 200 operators, 1000 states")
    ops, x, codes, h = _summed_apply_operands(args, n, N, K, S)
    d_in = _abi.DeviceBuffer(dim * S * 16)
    _abi.call("fp_hip_fill_states", _abi.FP_C128, d_in.ptr, dim, S, syn.gpu_seed(args.seed), 1, None)
    d_x = _abi.DeviceBuffer.from_array(np.ascontiguousarray(x, np.float64))
    d_apply = _abi.DeviceBuffer(dim * S * 16)
    d_weighted = _abi.DeviceBuffer(dim * S * 16)
    ms_apply = _time_ms(lambda: _abi.call("fp_hip_summed_apply", ops.handle, _abi.FP_C128, d_in.ptr, d_apply.ptr, dim,
                                          S, 0, None), 3)
    ms_w = _time_ms(lambda: _abi.call("fp_hip_summed_apply_weighted", ops.handle, _abi.FP_C128, d_in.ptr, d_x.ptr,
                                      d_weighted.ptr, dim, S, 0, None), 3)
    state_bytes = dim * S * 16
    model = (N + 1) * state_bytes  # streaming: one partner read per string, one output write
    plan = _abi.summed_apply_plan(ops, _abi.FP_C128, dim, S)
    wplan = _abi.summed_apply_weighted_plan(ops, _abi.FP_C128, dim, S)
    res = {"config": "n=14, 2000 strings of weight <= 3, h (2000 x 200) ~ N(0,1) real, x (200 x 1000) ~ N(0,1), "
                     "1000 normalised states, complex128",
           "apply_ms": ms_apply, "apply_weighted_ms": ms_w,
           "apply_kernel": "coset-tiled" if plan["kernel"] == _abi.FP_EV_TILED else "streaming",
           "apply_groups": plan["n_groups"], "apply_plan_bytes": plan["hbm_bytes"],
           "apply_plan_gbs": plan["hbm_bytes"] / (ms_apply / 1e3) / 1e9,
           "apply_plan_frac_hbm": plan["hbm_bytes"] / (ms_apply / 1e3) / 1e9 / HBM_PEAK_GBS,
           "streaming_model_bytes": model, "streaming_bytes_rule": "(strings + 1) x 2^n x S x 16 B",
           "apply_weighted_kernel": "coset-tiled" if wplan["kernel"] == _abi.FP_EV_TILED else "streaming",
           "apply_weighted_plan_bytes": wplan["hbm_bytes"],
           "apply_weighted_plan_gbs": wplan["hbm_bytes"] / (ms_w / 1e3) / 1e9,
           "strings_x_states_per_s": N * S / (ms_apply / 1e3)}
    if not args.no_cpu:
        from oracle import oracle as O

        # bounded samples (state columns are independent): the reference's sequential loop
        # on 4 states is the bit-exact check; its parallel branch (per-thread partial sums,
        # another summation order) on 16 states is the timed baseline
        Sp, Ss = 4, 16
        psi = d_in.to_array(np.complex128, (dim, S))[:, :Ss].copy()
        xs = np.ascontiguousarray(x[:, :Ss])
        strings = syn.codes_to_strings(codes)
        got_a = d_apply.to_array(np.complex128, (dim, S))[:, :Ss]
        got_w = d_weighted.to_array(np.complex128, (dim, S))[:, :Ss]
        seq_a = O.summed_apply(strings, h, psi[:, :Sp].copy(), np.zeros((dim, Sp), np.complex128))
        seq_w = O.summed_apply_weighted(strings, h, psi[:, :Sp].copy(), np.ascontiguousarray(xs[:, :Sp]),
                                        np.zeros((dim, Sp), np.complex128))
        res["apply_ieee_equal_vs_oracle_seq_4_states"] = bool(np.array_equal(got_a[:, :Sp], seq_a))
        res["apply_weighted_ieee_equal_vs_oracle_seq_4_states"] = bool(np.array_equal(got_w[:, :Sp], seq_w))
        if not res["apply_ieee_equal_vs_oracle_seq_4_states"]:
            failures.append("SummedPauliOp.apply differs from the oracle's sequential loop")
        if not res["apply_weighted_ieee_equal_vs_oracle_seq_4_states"]:
            failures.append("SummedPauliOp.apply_weighted differs from the oracle's sequential loop")
        t = time.perf_counter()
        par_a = O.summed_apply(strings, h, psi, np.zeros_like(psi), par=True)
        t_a = time.perf_counter() - t
        t = time.perf_counter()
        par_w = O.summed_apply_weighted(strings, h, psi, xs, np.zeros_like(psi), par=True)
        t_w = time.perf_counter() - t
        res["apply_rel_diff_vs_oracle_par"] = float(np.linalg.norm(got_a - par_a) / np.linalg.norm(par_a))
        res["apply_weighted_rel_diff_vs_oracle_par"] = float(np.linalg.norm(got_w - par_w) / np.linalg.norm(par_w))
        res["cpu_apply_ms_estimate"] = t_a * 1e3 * S / Ss
        res["cpu_apply_weighted_ms_estimate"] = t_w * 1e3 * S / Ss
        res["cpu_sample"] = f"{Ss} of {S} states, time x{S // Ss} (states run independently)"
        res["cpu_threads"] = O.max_threads()
    return res


def bench_square_n14(args, failures) -> dict:
    """SummedPauliOp::square on the n=14 operator set (2000 strings of weight <= 3, 200
    operators with N(0,1) weights): the 2000^2 string pairs contracted into the 2.77 M strings
    of weight <= 6, written as a (2.77 M x 200) complex128 coefficient matrix."""
    import ctypes

    from fast_pauli import _abi, synthetic as syn

    n, N, K = 14, 2000, 200
    log("workload: SummedPauliOp.square n=14, 2000 strings, 200 operators")
    codes = syn.random_codes_max_weight(N, n, 3, seed=args.seed)
    W = syn.random_real_matrix(K, N, seed=args.seed)
    h = W.T.astype(np.complex128)  # h(j, k)
    ops = _abi.Ops(codes, matrix=h)
    count = ctypes.c_size_t()
    _abi.call("fp_hip_summed_square", ops.handle, _abi.FP_C128, None, ctypes.byref(count), None)
    A = count.value
    out = _abi.DeviceBuffer(A * K * 16)
    ms = _time_ms(lambda: _abi.call("fp_hip_summed_square", ops.handle, _abi.FP_C128, out.ptr, ctypes.byref(count),
                                    None), 5)
    out_bytes = A * K * 16
    res = {"config": "n=14, 2000 strings of weight <= 3, h (2000 x 200) ~ N(0,1) real, complex128",
           "ms": ms, "out_strings": A, "pairs": N * N, "out_bytes": out_bytes,
           "out_gbs": out_bytes / (ms / 1e3) / 1e9,
           "out_frac_hbm": out_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
           "bound": "hbm (the output matrix is written once; the pair sort moves 2000^2 x 12 B per pass)"}
    del out
    if not args.no_cpu:
        from oracle import oracle as O

        # the whole workload through the oracle's loop (one thread), and the device result of
        # the first 4 operators against its columns (each column's arithmetic is independent)
        strings = syn.codes_to_strings(codes)
        t = time.perf_counter()
        _, exp = O.summed_square(strings, h)
        t_cpu = time.perf_counter() - t
        Ks = 4
        small = _abi.Ops(codes, matrix=np.ascontiguousarray(h[:, :Ks]))
        d = _abi.DeviceBuffer(A * Ks * 16)
        _abi.call("fp_hip_summed_square", small.handle, _abi.FP_C128, d.ptr, ctypes.byref(count), None)
        got = d.to_array(np.complex128, (A, Ks))
        equal = bool(np.array_equal(got, exp[:, :Ks]))
        res["ieee_equal_vs_oracle_4_operators"] = equal
        if not equal:
            failures.append("SummedPauliOp.square differs from the oracle")
        res["cpu_ms"] = t_cpu * 1e3
        res["cpu_sample"] = "the whole workload, one call of the oracle's loop"
        res["cpu_threads"] = 1
        del got, exp, d
    return res


def bench_uneven_batch(args) -> dict:
    """256 PauliOps (n 8..18, 16..512 strings, 1..64 states, complex64) in one persistent launch."""
    from fast_pauli import _abi, synthetic as syn
    from fast_pauli.batch import Batch, uneven_workload

    log("workload: uneven batch of 256 PauliOp.apply")
    specs = uneven_workload(256, seed=args.seed)
    b = Batch()
    keep = []
    strst = 0
    model_bytes = 0
    for sp in specs:
        n, N, S = sp["n"], sp["strings"], sp["states"]
        ops = _abi.Ops(syn.random_codes(N, n, seed=sp["seed"]), syn.random_coeffs(N, seed=sp["seed"]))
        d_in = _abi.DeviceBuffer((1 << n) * S * 8)
        d_out = _abi.DeviceBuffer((1 << n) * S * 8)
        _abi.call("fp_hip_fill_states", _abi.FP_C64, d_in.ptr, 1 << n, S, sp["seed"], 1, None)
        b.add_apply(ops, d_in.ptr, d_out.ptr, 1 << n, S, _abi.FP_C64)
        keep.append((ops, d_in, d_out, n, S))
        strst += N * S
        model_bytes += (N + 1) * (1 << n) * S * 8
    tune = b.autotune(reps=2)
    ms = _time_ms(lambda: b.launch(), 5)
    d_ref = _abi.DeviceBuffer(max((1 << n) * S * 8 for *_, n, S in keep))

    def one_by_one():
        for ops, d_in, _, n, S in keep:
            _abi.call("fp_hip_op_apply", ops.handle, _abi.FP_C64, d_in.ptr, d_ref.ptr, 1 << n, S, 0, None)

    ms_seq = _time_ms(one_by_one, 2)
    # per-op ideal under the streaming model: every op's (strings + 1) state passes at the
    # HBM peak (small states are L2 / Infinity-Cache resident, so the batch can beat it)
    ideal_ms = model_bytes / (HBM_PEAK_GBS * 1e9) * 1e3
    info = b.info()
    b.check()
    # the chosen plan's bytes: coset entries move about 3 state passes per string group
    plan_bytes = info["plan_bytes"]
    return {"ms": ms, "strings_x_states_per_s": strst / (ms / 1e3), "one_launch_per_op_ms": ms_seq,
            "sum_of_per_op_ideal_ms": ideal_ms, "ideal_over_batch": ideal_ms / ms,
            "model_bytes_per_launch": model_bytes, "model_gbs": model_bytes / (ms / 1e3) / 1e9,
            "plan_bytes_per_launch": plan_bytes, "plan_gbs": plan_bytes / (ms / 1e3) / 1e9,
            "plan_frac_hbm": plan_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
            "schedule": tune["schedule"], "schedule_cached": tune["cached"], "info": info,
            "config": "256 PauliOp.apply, n ~ U{8..18}, strings log-uniform 16..512, states ~ U{1..64}, complex64, "
                      "one persistent hipGraph launch"}


def bench_serving(args) -> dict:
    """Poisson request stream: 70% expectation_value / 30% apply, n 10..22, 8..1024 strings, 1..32 states.

    The headline point runs at --qps; the sweep (--qps-sweep) runs the same request
    list at each rate on a fresh server and reports latency, mean batch size and
    whether the loop kept up (achieved >= 0.9 x offered)."""
    from fast_pauli.serving import Server, request_mix, run_poisson

    cfg = dict(hbm_budget_bytes=200 << 30, n_slots=4, max_batch=32)
    reqs = request_mix(args.serve_requests, seed=args.seed)

    def point(qps):
        with Server(**cfg) as srv:
            r = run_poisson(srv, reqs, qps=qps, seed=args.seed)
        st = r["stats"]
        r["mean_batch_size"] = st["batched_requests"] / max(st["batches"], 1)
        r["kept_up"] = r["achieved_qps"] >= 0.9 * qps
        return r

    log(f"workload: serving, {args.serve_requests} requests at {args.qps} QPS")
    out = point(args.qps)
    sweep = []
    for q in [float(v) for v in args.qps_sweep.split(",") if v.strip()]:
        log(f"workload: serving sweep, {args.serve_requests} requests at {q} QPS")
        r = point(q)
        sweep.append({k: r[k] for k in ("offered_qps", "achieved_qps", "p50_ms", "p99_ms", "queue_p50_ms",
                                        "mean_batch_size", "kept_up", "failed", "strings_x_states_per_s")})
        sweep[-1]["batches"] = r["stats"]["batches"]
        sweep[-1]["max_concurrent_batches"] = r["stats"]["max_concurrent_batches"]
    out["sweep"] = sweep
    kept = [p["offered_qps"] for p in sweep if p["kept_up"]]
    out["saturation_qps"] = {"highest_rate_kept_up": max(kept) if kept else None,
                             "lowest_rate_fell_behind": min((p["offered_qps"] for p in sweep if not p["kept_up"]),
                                                            default=None)}
    out["config"] = ("Poisson arrivals, 70% PauliOp.expectation_value / 30% PauliOp.apply, n ~ U{10..22}, "
                     "strings ~ U{8..1024}, states ~ U{1..32}, complex64 device-resident, 200 GB admission budget, "
                     "4 concurrent batch slots, up to 32 requests per batch")
    return out


def cpu_baseline(args, psi_dev, codes_apply, codes_ev, coeffs_ev):
    """The reference's OpenMP loops (oracle/, 'par' branches) on this host, bounded sample.

    Returns (record, apply output, expectation values of the sample, sample size): the
    outputs are the oracle's, kept for the parity checks."""
    from oracle import oracle as O
    from fast_pauli import synthetic as syn

    n, S, N = args.n, args.states, args.strings
    dim = 1 << n
    threads = O.max_threads()
    n_sample = int(min(N, max(32, threads)))
    log(f"cpu baseline: {threads} OpenMP threads (OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}), "
        f"apply_batch x1 + expectation_value x{n_sample} at n={n}")
    psi = psi_dev.to_array(np.complex128, (dim, S))
    s_apply = syn.codes_to_strings(codes_apply)[0]
    strings = syn.codes_to_strings(codes_ev[:n_sample])
    out = np.zeros_like(psi)
    t = time.perf_counter()
    O.string_apply_batch(s_apply, 1, psi, out, par=True)
    t_apply = time.perf_counter() - t
    evo = np.zeros(S, np.complex128)
    t = time.perf_counter()
    O.op_expectation(strings, coeffs_ev[:n_sample], psi, evo, par=True)
    t_ev = time.perf_counter() - t
    t_step = t_apply + t_ev * (N / n_sample)
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    record = {
        "value": (1 + N) * S / t_step,
        "unit": "strings*states/s",
        "cores": threads,
        "kind": "port",
        "sample": f"n={n}, {S} states (same device-generated states), apply_batch of 1 string "
                  f"({t_apply:.2f} s) + PauliOp.expectation_value of {n_sample} of the {N} strings "
                  f"({t_ev:.2f} s), par (OpenMP) branches of the reference loops; strings run "
                  f"independently per thread, so the EV time is scaled by {N}/{n_sample} to the full step",
        "cpu_model": cpu_model,
        "omp_num_threads": threads,
        "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
        "nproc": os.cpu_count(),
        "affinity_cpus": AFFINITY_CPUS,
        "apply_batch_s": t_apply,
        "expectation_value_s_per_string": t_ev / n_sample,
    }
    return record, out, evo, n_sample


def oracle_parity(args, psi_dev, out_dev, psi64_dev, ops_apply, codes_ev, coeffs_ev, apply_ref, ev_ref, n_sample,
                  failures) -> dict:
    """The GPU results on the CPU baseline's sample against the oracle's outputs:
    apply_batch IEEE-equal; PauliOp.expectation_value of the sample's strings
    <= 1e-12 (complex128) and <= 1e-5 (complex64, against the oracle in FP64 on the
    same complex64 inputs, and against the complex128 oracle result)."""
    from fast_pauli import _abi
    from fast_pauli import synthetic as syn
    from oracle import oracle as O

    n, S = args.n, args.states
    dim = 1 << n
    call = _abi.call
    res = {"sample_strings": n_sample}
    call("fp_hip_string_apply_batch", ops_apply.handle, 0, _abi.FP_C128, psi_dev.ptr, out_dev.ptr, dim, S,
         1.0, 0.0, 0, None)
    got = out_dev.to_array(np.complex128, (dim, S))
    res["apply_batch_ieee_equal"] = bool(np.array_equal(got, apply_ref))
    del got
    if not res["apply_batch_ieee_equal"]:
        failures.append("apply_batch differs from the oracle")
    ops = _abi.Ops(codes_ev[:n_sample], coeffs_ev[:n_sample])
    evb = _abi.DeviceBuffer(S * 16)
    plan = _abi.op_expectation_plan(ops, _abi.FP_C128, dim, S)
    call("fp_hip_op_expectation", ops.handle, _abi.FP_C128, psi_dev.ptr, evb.ptr, dim, S, 0, None)
    g128 = evb.to_array(np.complex128, (S,))
    scale = np.abs(coeffs_ev[:n_sample]).sum()
    e128 = float(np.linalg.norm(g128 - ev_ref) / np.linalg.norm(ev_ref))
    res["ev_c128"] = {"kernel": plan["kernel"], "groups": plan["n_groups"], "rel_err_vs_oracle": e128,
                      "max_abs_err_over_sum_abs_coeffs": float(np.abs(g128 - ev_ref).max() / scale)}
    if not e128 <= 1e-12:
        failures.append(f"complex128 expectation differs from the oracle by {e128:.3e} > 1e-12")
    log(f"parity: apply_batch IEEE-equal {res['apply_batch_ieee_equal']}, complex128 EV x{n_sample} "
        f"rel err vs oracle {e128:.2e}")
    if psi64_dev is not None:
        call("fp_hip_op_expectation", ops.handle, _abi.FP_C64, psi64_dev.ptr, evb.ptr, dim, S, 0, None)
        g64 = evb.to_array(np.complex64, (S,)).astype(np.complex128)
        psi64w = psi64_dev.to_array(np.complex64, (dim, S)).astype(np.complex128)
        ref64 = np.zeros(S, np.complex128)
        O.op_expectation(syn.codes_to_strings(codes_ev[:n_sample]),
                         coeffs_ev[:n_sample].astype(np.complex64).astype(np.complex128), psi64w, ref64, par=True)
        del psi64w
        e_same = float(np.linalg.norm(g64 - ref64) / np.linalg.norm(ref64))
        e_fp64 = float(np.linalg.norm(g64 - ev_ref) / np.linalg.norm(ev_ref))
        res["ev_c64"] = {"rel_err_vs_oracle_same_inputs": e_same, "rel_err_vs_oracle_complex128_inputs": e_fp64,
                         "max_abs_err_over_sum_abs_coeffs": float(np.abs(g64 - ev_ref).max() / scale)}
        if not e_same <= 1e-5:
            failures.append(f"complex64 expectation differs from the oracle by {e_same:.3e} > 1e-5")
        log(f"parity: complex64 EV rel err vs oracle {e_same:.2e} (same inputs), {e_fp64:.2e} (complex128 inputs)")
    return res


if __name__ == "__main__":
    sys.exit(main())
