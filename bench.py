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
  expectation_value: N * 2^n * S * 16 B   (each string reads the state once)
  apply_batch:       2 * 2^n * S * 16 B   (read psi, write P psi)
against the MI355X HBM3E peak of 8.0 TB/s.

The CPU baseline is the repository's OpenMP restatement of the reference's
loops (oracle/), timed on this host on a bounded sample of the same workload.
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
    return ap.parse_args()


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is "
                          "design-only (see DESIGN.md)", "n_gpus": args.gpus}), flush=True)
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

    log(f"complex128: warmup {args.warmup}, timing {args.steps} steps")
    ms_step, ms_apply, ms_ev = timed(C128, psi, out, ev)
    ev128 = ev.to_array(np.complex128, (S,))
    bc = 16
    bytes_apply = 2 * dim * S * bc
    bytes_ev = N * dim * S * bc
    value = (1 + N) * S / (ms_step / 1e3)
    log(f"complex128: step {ms_step:.3f} ms (apply {ms_apply:.3f} ms, EV {ms_ev:.3f} ms)")

    result = {
        "metric": "strings x states/s (apply_batch 1 string + expectation_value 256 strings) at n=24, "
                  "16 states; % of HBM peak",
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
            "kernel": "k_op_expect (PauliOp::expectation_value)",
            "bound": "hbm",
            "achieved": bytes_ev / (ms_ev / 1e3) / 1e9,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": bytes_ev / (ms_ev / 1e3) / 1e9 / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": bytes_ev,
            "avg_launch_ms": ms_ev,
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
            "ev_frac_hbm": bytes_ev / 2 / (ms_ev64 / 1e3) / 1e9 / HBM_PEAK_GBS,
            "ev_rel_err_vs_fp64_same_inputs": err_norm,
            "ev_max_abs_err_over_sum_abs_coeffs": err_state,
            "input_rounding_rel_effect": err_input,
        }
        log(f"complex64: step {ms_step64:.3f} ms (apply {ms_apply64:.3f}, EV {ms_ev64:.3f}); "
            f"EV rel err {err_norm:.2e}")
        del psi64, out64, ev64

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args, psi, codes_apply, codes_ev, coeffs_ev)

    print(json.dumps(result), flush=True)
    return 0


def cpu_baseline(args, psi_dev, codes_apply, codes_ev, coeffs_ev) -> dict:
    """The reference's OpenMP loops (oracle/, 'par' branches) on this host, bounded sample."""
    from oracle import oracle as O
    from fast_pauli import synthetic as syn

    n, S, N = args.n, args.states, args.strings
    dim = 1 << n
    threads = O.max_threads()
    n_sample = int(min(N, max(16, threads)))
    log(f"cpu baseline: {threads} OpenMP threads, apply_batch x1 + expectation_value x{n_sample} at n={n}")
    psi = psi_dev.to_array(np.complex128, (dim, S))
    s_apply = syn.codes_to_strings(codes_apply)[0]
    strings = syn.codes_to_strings(codes_ev[:n_sample])
    out = np.zeros_like(psi)
    t = time.perf_counter()
    O.string_apply_batch(s_apply, 1, psi, out, par=True)
    t_apply = time.perf_counter() - t
    del out
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
    return {
        "value": (1 + N) * S / t_step,
        "unit": "strings*states/s",
        "cores": threads,
        "kind": "port",
        "sample": f"n={n}, {S} states (same device-generated states), apply_batch of 1 string "
                  f"({t_apply:.2f} s) + PauliOp.expectation_value of {n_sample} of the {N} strings "
                  f"({t_ev:.2f} s), par (OpenMP) branches of the reference loops; EV time scaled "
                  f"by {N}/{n_sample} to the full step",
        "cpu_model": cpu_model,
        "apply_batch_s": t_apply,
        "expectation_value_s_per_string": t_ev / n_sample,
    }


if __name__ == "__main__":
    sys.exit(main())
