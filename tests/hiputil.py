"""Helpers for GPU tests: run the C-ABI entry points on numpy inputs."""

from __future__ import annotations

import sys

import numpy as np

from fast_pauli import _abi

CDT = {_abi.FP_C128: np.complex128, _abi.FP_C64: np.complex64}


def say(msg: str) -> None:
    """Name the case before a GPU step (flushed, so a fault leaves a trace)."""
    sys.stdout.write(f"[gpu] {msg}\n")
    sys.stdout.flush()


def string_apply(ops, idx, psi, c=1.0, dtype=_abi.FP_C128, out0=None):
    psi = np.ascontiguousarray(psi, dtype=CDT[dtype])
    dim, S = psi.shape
    d_in = _abi.DeviceBuffer.from_array(psi)
    if out0 is not None:
        d_out = _abi.DeviceBuffer.from_array(np.ascontiguousarray(out0, dtype=CDT[dtype]))
    else:
        d_out = _abi.DeviceBuffer(psi.nbytes)
    c = complex(c)
    _abi.call("fp_hip_string_apply_batch", ops.handle, idx, dtype, d_in.ptr, d_out.ptr, dim, S,
              c.real, c.imag, int(out0 is not None), None)
    return d_out.to_array(CDT[dtype], psi.shape)


def op_apply(ops, psi, dtype=_abi.FP_C128, out0=None, entry="fp_hip_op_apply"):
    psi = np.ascontiguousarray(psi, dtype=CDT[dtype])
    dim, S = psi.shape
    d_in = _abi.DeviceBuffer.from_array(psi)
    if out0 is not None:
        d_out = _abi.DeviceBuffer.from_array(np.ascontiguousarray(out0, dtype=CDT[dtype]))
    else:
        d_out = _abi.DeviceBuffer(psi.nbytes)
    _abi.call(entry, ops.handle, dtype, d_in.ptr, d_out.ptr, dim, S, int(out0 is not None), None)
    return d_out.to_array(CDT[dtype], psi.shape)


def string_expectation(ops, idx, psi, c=1.0, dtype=_abi.FP_C128, out0=None):
    psi = np.ascontiguousarray(psi, dtype=CDT[dtype])
    dim, S = psi.shape
    d_in = _abi.DeviceBuffer.from_array(psi)
    out = np.zeros(S, CDT[dtype]) if out0 is None else np.ascontiguousarray(out0, CDT[dtype])
    d_out = _abi.DeviceBuffer.from_array(out)
    c = complex(c)
    _abi.call("fp_hip_string_expectation", ops.handle, idx, dtype, d_in.ptr, d_out.ptr, dim, S,
              c.real, c.imag, int(out0 is not None), None)
    return d_out.to_array(CDT[dtype], (S,))


def op_expectation(ops, psi, dtype=_abi.FP_C128, out0=None):
    psi = np.ascontiguousarray(psi, dtype=CDT[dtype])
    dim, S = psi.shape
    d_in = _abi.DeviceBuffer.from_array(psi)
    out = np.zeros(S, CDT[dtype]) if out0 is None else np.ascontiguousarray(out0, CDT[dtype])
    d_out = _abi.DeviceBuffer.from_array(out)
    _abi.call("fp_hip_op_expectation", ops.handle, dtype, d_in.ptr, d_out.ptr, dim, S,
              int(out0 is not None), None)
    return d_out.to_array(CDT[dtype], (S,))


def ev_matrix(ops, psi, dtype=_abi.FP_C128):
    psi = np.ascontiguousarray(psi, dtype=CDT[dtype])
    dim, S = psi.shape
    rdt = np.float64 if dtype == _abi.FP_C128 else np.float32
    d_in = _abi.DeviceBuffer.from_array(psi)
    d_ev = _abi.DeviceBuffer(ops.n_strings * S * np.dtype(rdt).itemsize)
    _abi.call("fp_hip_strings_expectation_matrix", ops.handle, dtype, d_in.ptr, d_ev.ptr, dim, S, None)
    return d_ev.to_array(rdt, (ops.n_strings, S))


def summed_expectation(ops, psi, dtype=_abi.FP_C128, contraction=_abi.FP_CONTRACT_AUTO, out0=None):
    psi = np.ascontiguousarray(psi, dtype=CDT[dtype])
    dim, S = psi.shape
    K = ops.n_operators
    d_in = _abi.DeviceBuffer.from_array(psi)
    out = np.zeros((K, S), CDT[dtype]) if out0 is None else np.ascontiguousarray(out0, CDT[dtype])
    d_out = _abi.DeviceBuffer.from_array(out)
    _abi.call("fp_hip_summed_expectation", ops.handle, dtype, contraction, d_in.ptr, d_out.ptr, dim,
              S, int(out0 is not None), None)
    return d_out.to_array(CDT[dtype], (K, S))


def gemm(mode, a, b):
    dt = np.float64 if mode == _abi.FP_CONTRACT_FP64 else np.float32
    a = np.ascontiguousarray(a, dt)
    b = np.ascontiguousarray(b, dt)
    m, k = a.shape
    k2, n = b.shape
    assert k == k2
    da, db = _abi.DeviceBuffer.from_array(a), _abi.DeviceBuffer.from_array(b)
    dc = _abi.DeviceBuffer(m * n * np.dtype(dt).itemsize)
    _abi.call("fp_hip_gemm", mode, m, n, k, da.ptr, db.ptr, dc.ptr, None)
    return dc.to_array(dt, (m, n))


def rel_err(got, ref) -> float:
    got = np.asarray(got, np.complex128)
    ref = np.asarray(ref, np.complex128)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(got - ref) / den) if den > 0 else float(np.linalg.norm(got))


def ev_scaled_err(got, ref, coeffs, psi) -> float:
    """max_t |dE_t| / (sum_s |c_s| * ||psi_t||^2): error relative to the bound |E_t| can reach."""
    scale = np.abs(np.asarray(coeffs)).sum() * (np.abs(np.asarray(psi, np.complex128)) ** 2).sum(axis=0)
    d = np.abs(np.asarray(got, np.complex128) - np.asarray(ref, np.complex128))
    return float((d / scale).max())
