"""Uneven batches: many Pauli operators of different sizes in one persistent,
graph-captured launch (C ABI fp_hip_batch_*), with an autotuned schedule.

    b = Batch()
    b.add_apply(ops, states_buf, out_buf, dim, n_states, dtype=FP_C64)
    b.add_expectation(ops2, states2, ev_out, dim2, n_states2)
    b.autotune()          # times schedule candidates; cached by batch signature
    b.launch(); synchronize()

Each entry's result equals the same operator run alone: bit-exact for apply,
within rounding for expectation values (a different partial-sum grouping).
"""

from __future__ import annotations

import ctypes
import hashlib
import json
import os
import threading

import numpy as np

from . import _abi
from . import synthetic as syn

_cache_lock = threading.Lock()


def default_cache_path() -> str:
    return os.environ.get(
        "FAST_PAULI_AUTOTUNE_CACHE",
        os.path.join(os.path.expanduser("~"), ".cache", "fast_pauli_amd", "batch_schedules.json"))


class Batch:
    def __init__(self):
        _abi.init()
        h = ctypes.c_void_p()
        _abi.call("fp_hip_batch_create", ctypes.byref(h))
        self.handle = h.value
        self.entries: list[tuple] = []
        self._keep = []

    # -- building --------------------------------------------------------
    def add(self, kind, ops, states_ptr, out_ptr, dim, n_states, dtype=_abi.FP_C64, accumulate=False):
        _abi.call("fp_hip_batch_add", self.handle, int(kind), ops.handle, int(dtype), states_ptr, out_ptr,
                  int(dim), int(n_states), int(bool(accumulate)))
        self.entries.append((int(kind), int(dtype), int(ops.n_qubits), int(ops.n_strings), int(n_states)))
        self._keep.append(ops)

    def add_apply(self, ops, states_ptr, out_ptr, dim, n_states, dtype=_abi.FP_C64, accumulate=False):
        self.add(_abi.FP_BATCH_APPLY, ops, states_ptr, out_ptr, dim, n_states, dtype, accumulate)

    def add_expectation(self, ops, states_ptr, out_ptr, dim, n_states, dtype=_abi.FP_C64, accumulate=False):
        self.add(_abi.FP_BATCH_EXPECT, ops, states_ptr, out_ptr, dim, n_states, dtype, accumulate)

    def clear(self):
        _abi.call("fp_hip_batch_clear", self.handle)
        self.entries.clear()
        self._keep.clear()

    # -- schedule ----------------------------------------------------------
    @property
    def schedule(self) -> dict:
        s = _abi.BatchSchedule()
        _abi.call("fp_hip_batch_get_schedule", self.handle, ctypes.byref(s))
        return s.as_dict()

    @schedule.setter
    def schedule(self, d: dict) -> None:
        """Sets the given fields; the others keep their current values."""
        s = _abi.BatchSchedule(**{k: int(v) for k, v in {**self.schedule, **d}.items()})
        _abi.call("fp_hip_batch_set_schedule", self.handle, ctypes.byref(s))

    def info(self) -> dict:
        a, b, c = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        _abi.call("fp_hip_batch_info", self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        e, g, i = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        pb = ctypes.c_double()
        _abi.call("fp_hip_batch_coset_info", self.handle, ctypes.byref(e), ctypes.byref(g), ctypes.byref(i),
                  ctypes.byref(pb))
        te, tg = ctypes.c_size_t(), ctypes.c_size_t()
        _abi.call("fp_hip_batch_tiled_info", self.handle, ctypes.byref(te), ctypes.byref(tg))
        return {"entries": a.value, "tiles": b.value, "blocks": c.value,
                "coset_entries": e.value, "coset_groups": g.value, "coset_items": i.value,
                "tiled_entries": te.value, "tiled_groups": tg.value, "plan_bytes": pb.value}

    def check(self) -> None:
        """Waits for the last launch; raises if a coset item gave up waiting for its
        operator's previous string group (an internal error)."""
        _abi.call("fp_hip_batch_check", self.handle)

    def signature(self) -> str:
        # the schedule layout is part of the key: cached choices of another layout are not reused
        key = {"entries": sorted(self.entries), "schedule": [f for f, _ in _abi.BatchSchedule._fields_]}
        return hashlib.sha1(json.dumps(key).encode()).hexdigest()

    def autotune(self, stream=None, reps: int = 3, cache_path: str | None = None,
                 use_cache: bool = True) -> dict:
        """Pick the fastest schedule on the device (or reuse a cached choice)."""
        path = cache_path or default_cache_path()
        key = self.signature()
        if use_cache:
            with _cache_lock:
                try:
                    with open(path) as f:
                        cached = json.load(f).get(key)
                except (OSError, ValueError):
                    cached = None
            if cached:
                self.schedule = {**self.schedule, **cached["schedule"]}
                return {"schedule": self.schedule, "ms": cached["ms"], "cached": True}
        chosen = _abi.BatchSchedule()
        ms = ctypes.c_float()
        _abi.call("fp_hip_batch_autotune", self.handle, _abi._handle(stream), int(reps), ctypes.byref(chosen),
                  ctypes.byref(ms))
        result = {"schedule": chosen.as_dict(), "ms": float(ms.value), "cached": False}
        if use_cache:
            with _cache_lock:
                try:
                    os.makedirs(os.path.dirname(path), exist_ok=True)
                    try:
                        with open(path) as f:
                            data = json.load(f)
                    except (OSError, ValueError):
                        data = {}
                    data[key] = {"schedule": result["schedule"], "ms": result["ms"]}
                    tmp = path + ".tmp"
                    with open(tmp, "w") as f:
                        json.dump(data, f, indent=1, sort_keys=True)
                    os.replace(tmp, path)
                except OSError:
                    pass
        return result

    # -- running -----------------------------------------------------------
    def launch(self, stream=None) -> None:
        _abi.call("fp_hip_batch_launch", self.handle, _abi._handle(stream))

    def destroy(self) -> None:
        if getattr(self, "handle", None):
            _abi.lib().fp_hip_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def uneven_workload(n_ops: int = 256, seed: int = 18, n_range=(8, 18), strings_range=(16, 512),
                    states_range=(1, 64)) -> list[dict]:
    """Operator shapes of the uneven-batch workload: n uniform in n_range, string
    counts log-uniform in strings_range, state counts uniform in states_range."""
    w = syn.words(syn.stream_seed(seed, syn.TAG_MISC), 0, 3 * n_ops).reshape(n_ops, 3)
    lo, hi = np.log(strings_range[0]), np.log(strings_range[1])
    out = []
    for i in range(n_ops):
        n = n_range[0] + int(w[i, 0] % np.uint64(n_range[1] - n_range[0] + 1))
        u = (float(w[i, 1] >> np.uint64(11)) + 0.5) * 2.0 ** -53
        N = int(round(np.exp(lo + u * (hi - lo))))
        S = states_range[0] + int(w[i, 2] % np.uint64(states_range[1] - states_range[0] + 1))
        out.append({"n": n, "strings": N, "states": S, "seed": seed * 1000 + i})
    return out
