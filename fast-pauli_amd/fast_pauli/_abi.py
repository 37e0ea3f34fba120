"""ctypes binding of the C ABI in include/fast_pauli_hip.h (libfastpauli_hip.so).

This is the thin host-side layer the Python tests, the benchmark and the
serving loop use to reach the gfx950 kernels directly with device pointers.
There is no CPU fallback: if the library is missing, loading raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libfastpauli_hip.so"
LIB_PATH = os.path.join(_HERE, LIB_NAME)

FP_OK = 0
FP_ERR_INVALID_ARGUMENT = 1
FP_ERR_RUNTIME = 2
FP_ERR_OUT_OF_MEMORY = 3
FP_ERR_NO_DEVICE = 4
FP_ERR_UNSUPPORTED = 5

FP_C128 = 0
FP_C64 = 1

FP_CONTRACT_FP64 = 0
FP_CONTRACT_FP32 = 1
FP_CONTRACT_BF16X3 = 2
FP_CONTRACT_AUTO = 3

FP_H2D, FP_D2H, FP_D2D = 0, 1, 2

DTYPES = {FP_C128: np.complex128, FP_C64: np.complex64}


class HipLibraryMissing(ImportError):
    """The HIP library was not built; there is deliberately no CPU fallback."""


_lib = None
_lock = threading.Lock()

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_i = ctypes.c_int
_d = ctypes.c_double
_u64 = ctypes.c_uint64

# name -> argtypes (restype is int unless listed in _RESTYPES)
_SIGS = {
    "fp_hip_version": [],
    "fp_hip_init": [_i],
    "fp_hip_device_count": [ctypes.POINTER(_i)],
    "fp_hip_last_error": [ctypes.c_char_p, _sz],
    "fp_hip_malloc": [ctypes.POINTER(_vp), _sz],
    "fp_hip_free": [_vp],
    "fp_hip_memcpy": [_vp, _vp, _sz, _i, _vp],
    "fp_hip_memset": [_vp, _i, _sz, _vp],
    "fp_hip_stream_create": [ctypes.POINTER(_vp)],
    "fp_hip_stream_destroy": [_vp],
    "fp_hip_stream_synchronize": [_vp],
    "fp_hip_device_synchronize": [],
    "fp_hip_is_device_pointer": [_vp, ctypes.POINTER(_i)],
    "fp_hip_mem_info": [ctypes.POINTER(_sz), ctypes.POINTER(_sz)],
    "fp_hip_event_create": [ctypes.POINTER(_vp)],
    "fp_hip_event_destroy": [_vp],
    "fp_hip_event_record": [_vp, _vp],
    "fp_hip_event_synchronize": [_vp],
    "fp_hip_event_query": [_vp, ctypes.POINTER(_i)],
    "fp_hip_event_elapsed_ms": [_vp, _vp, ctypes.POINTER(ctypes.c_float)],
    "fp_hip_fill_states": [_i, _vp, _sz, _sz, _u64, _i, _vp],
    "fp_hip_convert": [_i, _vp, _i, _vp, _sz, _vp],
    "fp_hip_copy": [_vp, _vp, _sz, _vp],
    "fp_hip_ops_create": [ctypes.POINTER(_vp), _sz, _sz, _vp, _vp],
    "fp_hip_ops_destroy": [_vp],
    "fp_hip_ops_assign": [_vp, _sz, _sz, _vp, _vp, _vp],
    "fp_hip_ops_set_coeffs": [_vp, _vp],
    "fp_hip_ops_set_matrix": [_vp, _sz, _vp],
    "fp_hip_ops_info": [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_sz), ctypes.POINTER(_sz)],
    "fp_hip_string_apply_batch": [_vp, _sz, _i, _vp, _vp, _sz, _sz, _d, _d, _i, _vp],
    "fp_hip_string_expectation": [_vp, _sz, _i, _vp, _vp, _sz, _sz, _d, _d, _i, _vp],
    "fp_hip_op_apply": [_vp, _i, _vp, _vp, _sz, _sz, _i, _vp],
    "fp_hip_op_expectation": [_vp, _i, _vp, _vp, _sz, _sz, _i, _vp],
    "fp_hip_op_expectation_plan": [_vp, _i, _sz, _sz, _vp],
    "fp_hip_op_apply_plan": [_vp, _i, _sz, _sz, _i, _vp],
    "fp_hip_summed_apply_plan": [_vp, _i, _sz, _sz, _i, _vp],
    "fp_hip_summed_apply_weighted_plan": [_vp, _i, _sz, _sz, _i, _vp],
    "fp_hip_op_expectation_noalloc": [_vp, _i, _vp, _vp, _sz, _sz, _i, _vp],
    "fp_hip_strings_expectation_matrix": [_vp, _i, _vp, _vp, _sz, _sz, _vp],
    "fp_hip_summed_expectation": [_vp, _i, _i, _vp, _vp, _sz, _sz, _i, _vp],
    "fp_hip_summed_apply": [_vp, _i, _vp, _vp, _sz, _sz, _i, _vp],
    "fp_hip_summed_apply_weighted": [_vp, _i, _vp, _vp, _vp, _sz, _sz, _i, _vp],
    "fp_hip_gemm": [_i, _sz, _sz, _sz, _vp, _vp, _vp, _vp],
    "fp_hip_summed_square": [_vp, _i, _vp, ctypes.POINTER(_sz), _vp],
    "fp_hip_string_sparse_repr": [_vp, _sz, _i, _vp, _vp, _sz, _vp],
    "fp_hip_string_to_tensor": [_vp, _sz, _i, _vp, _sz, _vp],
    "fp_hip_op_to_tensor": [_vp, _i, _vp, _sz, _i, _vp],
    "fp_hip_summed_to_tensor": [_vp, _i, _vp, _sz, _i, _vp],
}


FP_EV_STREAM, FP_EV_TILED = 0, 1


class EvPlan(ctypes.Structure):
    """fp_hip_ev_plan"""
    _fields_ = [("kernel", ctypes.c_int), ("n_groups", ctypes.c_uint32), ("tile_rows_log2", ctypes.c_uint32),
                ("states_per_block", ctypes.c_uint32), ("hbm_bytes", ctypes.c_uint64),
                ("hoisted_groups", ctypes.c_uint32), ("shared_groups", ctypes.c_uint32)]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


def op_expectation_plan(ops, dtype: int, dim: int, n_states: int) -> dict:
    """Which kernel fp_hip_op_expectation runs for these shapes, and its HBM bytes."""
    p = EvPlan()
    call("fp_hip_op_expectation_plan", ops.handle, dtype, dim, n_states, ctypes.byref(p))
    return p.as_dict()


def op_apply_plan(ops, dtype: int, dim: int, n_states: int, accumulate: bool = False) -> dict:
    """Which kernel fp_hip_op_apply runs for these shapes, and its HBM bytes."""
    p = EvPlan()
    call("fp_hip_op_apply_plan", ops.handle, dtype, dim, n_states, int(bool(accumulate)), ctypes.byref(p))
    return p.as_dict()


def summed_apply_weighted_plan(ops, dtype: int, dim: int, n_states: int, accumulate: bool = False) -> dict:
    """Which kernel fp_hip_summed_apply_weighted runs for these shapes, and its HBM bytes."""
    p = EvPlan()
    call("fp_hip_summed_apply_weighted_plan", ops.handle, dtype, dim, n_states, int(bool(accumulate)),
         ctypes.byref(p))
    return p.as_dict()


def summed_apply_plan(ops, dtype: int, dim: int, n_states: int, accumulate: bool = False) -> dict:
    """Which kernel fp_hip_summed_apply runs for these shapes, and its HBM bytes."""
    p = EvPlan()
    call("fp_hip_summed_apply_plan", ops.handle, dtype, dim, n_states, int(bool(accumulate)), ctypes.byref(p))
    return p.as_dict()


class BatchSchedule(ctypes.Structure):
    """fp_hip_batch_schedule"""
    _fields_ = [("tile_vectors", ctypes.c_uint32), ("ev_tile_work", ctypes.c_uint64),
                ("blocks_per_cu", ctypes.c_uint32), ("order", ctypes.c_int),
                ("coset_tiles", ctypes.c_uint32), ("coset_min_kib", ctypes.c_uint32),
                ("tiled_min_kib", ctypes.c_uint32)]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


_SIGS.update({
    "fp_hip_batch_create": [ctypes.POINTER(_vp)],
    "fp_hip_batch_destroy": [_vp],
    "fp_hip_batch_clear": [_vp],
    "fp_hip_batch_add": [_vp, _i, _vp, _i, _vp, _vp, _sz, _sz, _i],
    "fp_hip_batch_set_schedule": [_vp, ctypes.POINTER(BatchSchedule)],
    "fp_hip_batch_get_schedule": [_vp, ctypes.POINTER(BatchSchedule)],
    "fp_hip_batch_info": [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_sz), ctypes.POINTER(_sz)],
    "fp_hip_batch_coset_info": [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_sz), ctypes.POINTER(_sz),
                                ctypes.POINTER(ctypes.c_double)],
    "fp_hip_batch_tiled_info": [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_sz)],
    "fp_hip_batch_check": [_vp],
    "fp_hip_batch_launch": [_vp, _vp],
    "fp_hip_batch_set_graph": [_vp, _i],
    "fp_hip_batch_autotune": [_vp, _vp, _i, ctypes.POINTER(BatchSchedule), ctypes.POINTER(ctypes.c_float)],
})


class ServeConfig(ctypes.Structure):
    """fp_serve_config"""
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("n_slots", ctypes.c_int),
                ("max_batch", ctypes.c_int), ("poll_us", ctypes.c_int)]


class ServeRequest(ctypes.Structure):
    """fp_serve_request"""
    _fields_ = [("kind", ctypes.c_int), ("dtype", ctypes.c_int), ("n_qubits", ctypes.c_size_t),
                ("n_strings", ctypes.c_size_t), ("codes", ctypes.c_void_p), ("coeffs", ctypes.c_void_p),
                ("n_states", ctypes.c_size_t), ("state_seed", ctypes.c_uint64), ("states", ctypes.c_void_p),
                ("keep_output", ctypes.c_int)]


class ServeResult(ctypes.Structure):
    """fp_serve_result"""
    _fields_ = [("done", ctypes.c_int), ("status", ctypes.c_int), ("latency_s", ctypes.c_double),
                ("queue_s", ctypes.c_double), ("batch_index", ctypes.c_uint64)]


class ServeStats(ctypes.Structure):
    """fp_serve_stats"""
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("batched_requests", ctypes.c_uint64), ("queued", ctypes.c_uint64),
                ("peak_bytes_in_use", ctypes.c_size_t), ("bytes_allocated", ctypes.c_size_t),
                ("max_concurrent_batches", ctypes.c_int)]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


_SIGS.update({
    "fp_hip_server_create": [ctypes.POINTER(_vp), ctypes.POINTER(ServeConfig)],
    "fp_hip_server_destroy": [_vp],
    "fp_hip_server_submit": [_vp, ctypes.POINTER(ServeRequest), ctypes.POINTER(_u64)],
    "fp_hip_server_wait": [_vp, _u64, _d, ctypes.POINTER(ServeResult)],
    "fp_hip_server_fetch": [_vp, _u64, _vp, _sz],
    "fp_hip_server_release": [_vp, _u64],
    "fp_hip_server_stats": [_vp, ctypes.POINTER(ServeStats)],
})
_RESTYPES = {"fp_hip_version": ctypes.c_char_p, "fp_hip_last_error": _sz}

FP_BATCH_APPLY, FP_BATCH_EXPECT = 0, 1
FP_BATCH_ORDER_LARGEST_FIRST, FP_BATCH_ORDER_AS_ADDED = 0, 1


def lib() -> ctypes.CDLL:
    """Load libfastpauli_hip.so (raises HipLibraryMissing when it was not built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise HipLibraryMissing(
                    f"{LIB_PATH} is missing: build it first (python -c 'import "
                    "__graft_entry__ as g; g.build()'); there is no CPU fallback")
            L = ctypes.CDLL(LIB_PATH)
            for name, args in _SIGS.items():
                f = getattr(L, name)
                f.argtypes = args
                f.restype = _RESTYPES.get(name, ctypes.c_int)
            _lib = L
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().fp_hip_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int) -> None:
    """Map an fp_status to the reference's exception types."""
    if status == FP_OK:
        return
    msg = last_error()
    if status == FP_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if status == FP_ERR_OUT_OF_MEMORY:
        raise MemoryError(msg)
    if status == FP_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(f"fast_pauli_hip error {status}: {msg}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


_initialised = False


def init(device: int = 0) -> None:
    global _initialised
    if not _initialised:
        call("fp_hip_init", device)
        _initialised = True


def device_count() -> int:
    n = _i(0)
    st = lib().fp_hip_device_count(ctypes.byref(n))
    return n.value if st == FP_OK else 0


def mem_info() -> tuple[int, int]:
    f, t = _sz(0), _sz(0)
    call("fp_hip_mem_info", ctypes.byref(f), ctypes.byref(t))
    return f.value, t.value


# ---------------------------------------------------------------------------
# device memory, streams, events
# ---------------------------------------------------------------------------

class DeviceBuffer:
    """Owning device allocation (hipMalloc through the C ABI)."""

    def __init__(self, nbytes: int):
        init()
        p = _vp()
        call("fp_hip_malloc", ctypes.byref(p), int(nbytes))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a: np.ndarray, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a, stream)
        return b

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        call("fp_hip_memcpy", self.ptr, a.ctypes.data, a.nbytes, FP_H2D, _handle(stream))

    def download(self, a: np.ndarray, stream=None) -> np.ndarray:
        assert a.flags.c_contiguous and a.nbytes <= self.nbytes
        call("fp_hip_memcpy", a.ctypes.data, self.ptr, a.nbytes, FP_D2H, _handle(stream))
        return a

    def to_array(self, dtype, shape, stream=None) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        return self.download(out, stream)

    def zero(self, stream=None) -> None:
        call("fp_hip_memset", self.ptr, 0, self.nbytes, _handle(stream))

    def free(self) -> None:
        if self.ptr:
            lib().fp_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        init()
        p = _vp()
        call("fp_hip_stream_create", ctypes.byref(p))
        self.handle = p.value

    def synchronize(self) -> None:
        call("fp_hip_stream_synchronize", self.handle)

    def destroy(self) -> None:
        if self.handle:
            lib().fp_hip_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _handle(stream):
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


class Event:
    def __init__(self):
        init()
        p = _vp()
        call("fp_hip_event_create", ctypes.byref(p))
        self.handle = p.value

    def record(self, stream=None) -> None:
        call("fp_hip_event_record", self.handle, _handle(stream))

    def synchronize(self) -> None:
        call("fp_hip_event_synchronize", self.handle)

    def query(self) -> bool:
        d = _i(0)
        call("fp_hip_event_query", self.handle, ctypes.byref(d))
        return bool(d.value)

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0)
        call("fp_hip_event_elapsed_ms", self.handle, end.handle, ctypes.byref(ms))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                lib().fp_hip_event_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def synchronize() -> None:
    call("fp_hip_device_synchronize")


# ---------------------------------------------------------------------------
# operator sets
# ---------------------------------------------------------------------------

class Ops:
    """Device-resident Pauli strings (letter codes) with complex128 coefficients."""

    def __init__(self, codes: np.ndarray, coeffs=None, matrix=None):
        init()
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        if codes.ndim != 2:
            raise ValueError("codes must be (n_strings, n_qubits)")
        self.n_strings, self.n_qubits = codes.shape
        self._codes = codes
        c_ptr = None
        if coeffs is not None:
            self._coeffs = np.ascontiguousarray(coeffs, dtype=np.complex128).reshape(-1)
            if self._coeffs.size != self.n_strings:
                raise ValueError("coeffs and pauli_strings must have the same size")
            c_ptr = self._coeffs.ctypes.data
        p = _vp()
        call("fp_hip_ops_create", ctypes.byref(p), self.n_qubits, self.n_strings,
             codes.ctypes.data, c_ptr)
        self.handle = p.value
        self.n_operators = 0
        if matrix is not None:
            self.set_matrix(matrix)

    @property
    def dim(self) -> int:
        return 1 << self.n_qubits

    def assign(self, codes: np.ndarray, coeffs=None, stream=None) -> None:
        """Re-targets this set to new strings in place (fp_hip_ops_assign); the
        device tables are rewritten on `stream` and reallocated only if they grow."""
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        if codes.ndim != 2:
            raise ValueError("codes must be (n_strings, n_qubits)")
        c_ptr = None
        cf = None
        if coeffs is not None:
            cf = np.ascontiguousarray(coeffs, dtype=np.complex128).reshape(-1)
            if cf.size != codes.shape[0]:
                raise ValueError("coeffs and pauli_strings must have the same size")
            c_ptr = cf.ctypes.data
        call("fp_hip_ops_assign", self.handle, codes.shape[1], codes.shape[0], codes.ctypes.data, c_ptr, stream)
        self.n_strings, self.n_qubits = codes.shape
        self._codes, self._coeffs = codes, cf
        self.n_operators = 0

    def set_coeffs(self, coeffs) -> None:
        c = np.ascontiguousarray(coeffs, dtype=np.complex128).reshape(-1)
        if c.size != self.n_strings:
            raise ValueError("coeffs and pauli_strings must have the same size")
        self._coeffs = c
        call("fp_hip_ops_set_coeffs", self.handle, c.ctypes.data)

    def set_matrix(self, matrix_nk) -> None:
        m = np.ascontiguousarray(matrix_nk, dtype=np.complex128)
        if m.ndim != 2 or m.shape[0] != self.n_strings:
            raise ValueError("The number of PauliStrings must match the number of rows in the coeffs matrix")
        self._matrix = m
        self.n_operators = m.shape[1]
        call("fp_hip_ops_set_matrix", self.handle, m.shape[1], m.ctypes.data)

    def destroy(self) -> None:
        if getattr(self, "handle", None):
            lib().fp_hip_ops_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
