"""Continuous-batching request loop (C ABI fp_hip_server_*) and a Poisson load
generator that reports latency percentiles beside throughput.

    with Server(hbm_budget_bytes=200 << 30) as srv:
        rid = srv.submit_expectation(codes, coeffs, n_states=8, seed=1)
        res = srv.wait(rid)            # {'latency_s', 'queue_s', 'status', ...}
        ev = srv.fetch(rid)            # (n_states,) complex
"""

from __future__ import annotations

import ctypes
import threading
import time

import numpy as np

from . import _abi
from . import synthetic as syn

CDT = {_abi.FP_C128: np.complex128, _abi.FP_C64: np.complex64}


class Server:
    def __init__(self, hbm_budget_bytes: int = 200 << 30, n_slots: int = 4, max_batch: int = 64,
                 poll_us: int = 20):
        _abi.init()
        cfg = _abi.ServeConfig(int(hbm_budget_bytes), int(n_slots), int(max_batch), int(poll_us))
        h = ctypes.c_void_p()
        _abi.call("fp_hip_server_create", ctypes.byref(h), ctypes.byref(cfg))
        self.handle = h.value
        self._meta: dict[int, tuple] = {}
        self._lock = threading.Lock()

    # -- requests -----------------------------------------------------------
    def submit(self, kind: int, codes: np.ndarray, coeffs=None, n_states: int = 1, seed: int = 0,
               dtype: int = _abi.FP_C64, states_ptr=None, keep_output: bool = False) -> int:
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        n_strings, n_qubits = codes.shape
        c = None
        if coeffs is not None:
            c = np.ascontiguousarray(coeffs, dtype=np.complex128)
        req = _abi.ServeRequest(int(kind), int(dtype), n_qubits, n_strings, codes.ctypes.data,
                                c.ctypes.data if c is not None else None, int(n_states), int(seed),
                                states_ptr, int(bool(keep_output)))
        rid = ctypes.c_uint64()
        _abi.call("fp_hip_server_submit", self.handle, ctypes.byref(req), ctypes.byref(rid))
        with self._lock:
            self._meta[rid.value] = (int(kind), int(dtype), n_qubits, int(n_states))
        return rid.value

    def submit_expectation(self, codes, coeffs=None, n_states=1, seed=0, dtype=_abi.FP_C64, states_ptr=None):
        return self.submit(_abi.FP_BATCH_EXPECT, codes, coeffs, n_states, seed, dtype, states_ptr)

    def submit_apply(self, codes, coeffs=None, n_states=1, seed=0, dtype=_abi.FP_C64, states_ptr=None,
                     keep_output=False):
        return self.submit(_abi.FP_BATCH_APPLY, codes, coeffs, n_states, seed, dtype, states_ptr, keep_output)

    def wait(self, rid: int, timeout_s: float = -1.0) -> dict | None:
        r = _abi.ServeResult()
        _abi.call("fp_hip_server_wait", self.handle, int(rid), float(timeout_s), ctypes.byref(r))
        if not r.done:
            return None
        return {"status": int(r.status), "latency_s": float(r.latency_s), "queue_s": float(r.queue_s),
                "batch_index": int(r.batch_index)}

    def fetch(self, rid: int) -> np.ndarray:
        kind, dtype, n, S = self._meta[rid]
        shape = (S,) if kind == _abi.FP_BATCH_EXPECT else (1 << n, S)
        out = np.empty(shape, CDT[dtype])
        _abi.call("fp_hip_server_fetch", self.handle, int(rid), out.ctypes.data, out.nbytes)
        return out

    def release(self, rid: int) -> None:
        _abi.call("fp_hip_server_release", self.handle, int(rid))
        with self._lock:
            self._meta.pop(rid, None)

    def stats(self) -> dict:
        s = _abi.ServeStats()
        _abi.call("fp_hip_server_stats", self.handle, ctypes.byref(s))
        return s.as_dict()

    def close(self) -> None:
        if getattr(self, "handle", None):
            _abi.call("fp_hip_server_destroy", self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request_mix(n_requests: int, seed: int = 18, ev_fraction: float = 0.7, n_range=(10, 22),
                strings_range=(8, 1024), states_range=(1, 32)) -> list[dict]:
    """The serving workload: 70 % expectation / 30 % apply, n, string count and
    state count uniform in their ranges, uniform IXYZ letters."""
    w = syn.words(syn.stream_seed(seed, syn.TAG_MISC ^ 0x5E57), 0, 4 * n_requests).reshape(n_requests, 4)
    out = []
    for i in range(n_requests):
        u = (float(w[i, 0] >> np.uint64(11)) + 0.5) * 2.0 ** -53
        kind = _abi.FP_BATCH_EXPECT if u < ev_fraction else _abi.FP_BATCH_APPLY
        n = n_range[0] + int(w[i, 1] % np.uint64(n_range[1] - n_range[0] + 1))
        N = strings_range[0] + int(w[i, 2] % np.uint64(strings_range[1] - strings_range[0] + 1))
        S = states_range[0] + int(w[i, 3] % np.uint64(states_range[1] - states_range[0] + 1))
        out.append({"kind": kind, "n": n, "strings": N, "states": S, "seed": seed * 100003 + i})
    return out


def run_poisson(server: Server, requests: list[dict], qps: float, seed: int = 18,
                dtype: int = _abi.FP_C64) -> dict:
    """Submits `requests` with exponential inter-arrival times at rate `qps`
    (open loop: arrivals do not wait for completions), waits for all, and
    returns latency percentiles and throughput."""
    gaps = -np.log(syn.uniform01(syn.stream_seed(seed, 0xA77), 0, len(requests))) / qps
    payload = []
    for r in requests:
        codes = syn.random_codes(r["strings"], r["n"], seed=r["seed"])
        coeffs = syn.random_coeffs(r["strings"], seed=r["seed"])
        payload.append((r, codes, coeffs))
    ids = []
    t0 = time.perf_counter()
    t_next = t0
    for (r, codes, coeffs), gap in zip(payload, gaps):
        t_next += float(gap)
        delay = t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(server.submit(r["kind"], codes, coeffs, r["states"], r["seed"], dtype))
    t_submitted = time.perf_counter()
    lat, queue, failed = [], [], 0
    for rid in ids:
        res = server.wait(rid)
        if res["status"] != 0:
            failed += 1
        else:
            lat.append(res["latency_s"])
            queue.append(res["queue_s"])
        server.release(rid)
    t_end = time.perf_counter()
    lat_ms = np.array(lat) * 1e3
    str_st = sum(r["strings"] * r["states"] for r, _, _ in payload)
    return {
        "requests": len(requests),
        "failed": failed,
        "offered_qps": qps,
        "achieved_qps": len(requests) / (t_end - t0),
        "arrival_window_s": t_submitted - t0,
        "wall_s": t_end - t0,
        "p50_ms": float(np.percentile(lat_ms, 50)) if lat else None,
        "p99_ms": float(np.percentile(lat_ms, 99)) if lat else None,
        "mean_ms": float(lat_ms.mean()) if lat else None,
        "max_ms": float(lat_ms.max()) if lat else None,
        "queue_p50_ms": float(np.percentile(np.array(queue) * 1e3, 50)) if lat else None,
        "strings_x_states_per_s": str_st / (t_end - t0),
        "stats": server.stats(),
    }
