"""Uneven batches in one persistent, graph-captured launch: every entry's result
equals the same operator run alone (bit-exact for apply, within the stated
tolerance for expectation values), graph replays equal the eager result
bit-exactly, and schedules change apply results not at all and expectation
values only within rounding."""

import numpy as np
import pytest

from fast_pauli import _abi
from fast_pauli import synthetic as syn
from fast_pauli.batch import Batch, uneven_workload
from oracle import oracle as O
from tests import hiputil as H

pytestmark = pytest.mark.gpu

C128, C64 = _abi.FP_C128, _abi.FP_C64


class Entry:
    def __init__(self, kind, n, N, S, dtype, accumulate, seed):
        self.kind, self.n, self.N, self.S, self.dtype, self.accumulate = kind, n, N, S, dtype, accumulate
        cdt = H.CDT[dtype]
        self.codes = syn.random_codes(N, n, seed=seed)
        self.coeffs = syn.random_coeffs(N, seed=seed)
        self.ops = _abi.Ops(self.codes, self.coeffs)
        self.psi = syn.random_states(1 << n, S, seed=seed).astype(cdt)
        self.d_in = _abi.DeviceBuffer.from_array(self.psi)
        out_shape = (1 << n, S) if kind == _abi.FP_BATCH_APPLY else (S,)
        self.out_shape = out_shape
        self.base = (syn.random_coeffs(int(np.prod(out_shape)), seed=seed + 1).astype(cdt).reshape(out_shape)
                     if accumulate else np.zeros(out_shape, cdt))
        self.d_out = _abi.DeviceBuffer.from_array(self.base)

    def reset(self):
        self.d_out.upload(self.base)

    def result(self):
        return self.d_out.to_array(H.CDT[self.dtype], self.out_shape)

    def alone(self):
        """The same operator through the single-operator entry points."""
        if self.kind == _abi.FP_BATCH_APPLY:
            return H.op_apply(self.ops, self.psi, self.dtype, out0=self.base if self.accumulate else None)
        return H.op_expectation(self.ops, self.psi, self.dtype, out0=self.base if self.accumulate else None)


def check(e: Entry, got):
    ref = e.alone()
    if e.kind == _abi.FP_BATCH_APPLY:
        assert np.array_equal(got, ref), (e.n, e.N, e.S, e.dtype)
    else:
        tol = 1e-12 if e.dtype == C128 else 1e-5
        assert H.rel_err(got - e.base, ref - e.base) <= tol, (e.n, e.N, e.S, e.dtype)


def mixed_entries():
    specs = [
        (_abi.FP_BATCH_APPLY, 3, 1, 1, C128, False), (_abi.FP_BATCH_APPLY, 10, 40, 9, C64, False),
        (_abi.FP_BATCH_EXPECT, 8, 17, 5, C64, False), (_abi.FP_BATCH_APPLY, 12, 7, 16, C64, True),
        (_abi.FP_BATCH_EXPECT, 11, 64, 33, C128, True), (_abi.FP_BATCH_APPLY, 9, 3, 3, C128, True),
        (_abi.FP_BATCH_EXPECT, 1, 2, 1, C64, False), (_abi.FP_BATCH_APPLY, 13, 20, 2, C64, False),
        (_abi.FP_BATCH_EXPECT, 14, 5, 64, C64, False), (_abi.FP_BATCH_APPLY, 6, 100, 7, C64, False),
        (_abi.FP_BATCH_EXPECT, 10, 128, 300, C64, False), (_abi.FP_BATCH_APPLY, 7, 9, 600, C128, False),
    ]
    return [Entry(*s, seed=17 + i) for i, s in enumerate(specs)]


def test_mixed_batch_matches_single_operator_runs(hip):
    H.say("batch mixed")
    entries = mixed_entries()
    b = Batch()
    for e in entries:
        b.add(e.kind, e.ops, e.d_in.ptr, e.d_out.ptr, 1 << e.n, e.S, e.dtype, e.accumulate)
    b.launch()
    _abi.synchronize()
    for e in entries:
        check(e, e.result())


def test_graph_replays_match_eager(hip):
    H.say("batch graph replays")
    entries = mixed_entries()
    b = Batch()
    for e in entries:
        b.add(e.kind, e.ops, e.d_in.ptr, e.d_out.ptr, 1 << e.n, e.S, e.dtype, e.accumulate)
    first = None
    for rep in range(4):
        for e in entries:
            e.reset()
        b.launch()
        _abi.synchronize()
        res = [e.result() for e in entries]
        # eager single-operator work between replays
        for e in entries[:3]:
            e.alone()
        if first is None:
            first = res
            for e, r in zip(entries, res):
                check(e, r)
        else:
            for r0, r in zip(first, res):
                assert np.array_equal(r0, r)


def test_schedules_do_not_change_results(hip):
    H.say("batch schedules")
    entries = mixed_entries()
    b = Batch()
    for e in entries:
        b.add(e.kind, e.ops, e.d_in.ptr, e.d_out.ptr, 1 << e.n, e.S, e.dtype, e.accumulate)
    ref = None
    for sched in ({"tile_vectors": 512, "ev_tile_work": 1 << 14, "blocks_per_cu": 1, "order": 1},
                  {"tile_vectors": 8192, "ev_tile_work": 1 << 20, "blocks_per_cu": 8, "order": 0}):
        b.schedule = sched
        assert {k: b.schedule[k] for k in sched} == sched
        for e in entries:
            e.reset()
        b.launch()
        _abi.synchronize()
        res = [e.result() for e in entries]
        if ref is None:
            ref = res
        else:
            for e, r0, r in zip(entries, ref, res):
                if e.kind == _abi.FP_BATCH_APPLY:
                    assert np.array_equal(r0, r)
                else:  # another tile size regroups the FP64 partial sums
                    tol = 1e-12 if e.dtype == C128 else 1e-6
                    assert H.rel_err(r - e.base, r0 - e.base) <= tol
    with pytest.raises(ValueError):
        b.schedule = {"tile_vectors": 0, "ev_tile_work": 1, "blocks_per_cu": 1, "order": 0}
    with pytest.raises(ValueError):
        b.schedule = {"coset_tiles": 0}


def test_overlapping_outputs_rejected(hip):
    H.say("batch overlap rejection")
    e = Entry(_abi.FP_BATCH_APPLY, 4, 2, 2, C64, False, 5)
    b = Batch()
    b.add_apply(e.ops, e.d_in.ptr, e.d_out.ptr, 16, 2, C64)
    with pytest.raises(ValueError):
        b.add_apply(e.ops, e.d_out.ptr, e.d_in.ptr, 16, 2, C64)  # writes the first entry's input
    with pytest.raises(ValueError):
        b.add_apply(e.ops, e.d_in.ptr, e.d_in.ptr, 16, 2, C64)   # in place


def test_uneven_workload_autotuned_bit_exact(hip, tmp_path):
    """The 256-operator uneven workload (n 8..18, 16..512 strings, 1..64 states,
    complex64), autotuned, against each operator run alone."""
    specs = uneven_workload(256, seed=18)
    H.say(f"batch uneven workload, {len(specs)} operators")
    b = Batch()
    items = []
    for sp in specs:
        n, N, S = sp["n"], sp["strings"], sp["states"]
        codes = syn.random_codes(N, n, seed=sp["seed"])
        ops = _abi.Ops(codes, syn.random_coeffs(N, seed=sp["seed"]))
        d_in = _abi.DeviceBuffer((1 << n) * S * 8)
        _abi.call("fp_hip_fill_states", C64, d_in.ptr, 1 << n, S, sp["seed"], 1, None)
        d_out = _abi.DeviceBuffer((1 << n) * S * 8)
        b.add_apply(ops, d_in.ptr, d_out.ptr, 1 << n, S, C64)
        items.append((ops, d_in, d_out, n, S))
    res = b.autotune(reps=1, cache_path=str(tmp_path / "sched.json"))
    assert res["ms"] > 0
    b.launch()
    _abi.synchronize()
    b.check()
    d_ref = _abi.DeviceBuffer(max((1 << n) * S * 8 for _, _, _, n, S in items))
    for ops, d_in, d_out, n, S in items:
        _abi.call("fp_hip_op_apply", ops.handle, C64, d_in.ptr, d_ref.ptr, 1 << n, S, 0, None)
        a = d_out.to_array(np.complex64, ((1 << n), S))
        r = d_ref.to_array(np.complex64, ((1 << n), S))
        assert np.array_equal(a, r), (n, ops.n_strings, S)
    # cached schedule is reused
    again = b.autotune(reps=1, cache_path=str(tmp_path / "sched.json"))
    assert again["cached"] and again["schedule"] == res["schedule"]


@pytest.mark.parametrize("graph", [True, False])
def test_batch_rebuilds_after_ops_reassigned(hip, graph):
    """A batch built over an operator set must not replay stale tables after the
    set is re-targeted to more strings (its device arena is reallocated) or given
    new coefficients: the next launch rebuilds and matches the operator alone."""
    H.say(f"batch after ops.assign graph={graph}")
    n, S = 8, 5
    psi = syn.random_states(1 << n, S, seed=3).astype(np.complex64)
    d_in = _abi.DeviceBuffer.from_array(psi)
    d_out = _abi.DeviceBuffer(psi.nbytes)
    d_ev = _abi.DeviceBuffer(S * 8)
    ops = _abi.Ops(syn.random_codes(16, n, seed=1), syn.random_coeffs(16, seed=1))
    b = Batch()
    if not graph:
        _abi.call("fp_hip_batch_set_graph", b.handle, 0)
    b.add_apply(ops, d_in.ptr, d_out.ptr, 1 << n, S, C64)
    b.add_expectation(ops, d_in.ptr, d_ev.ptr, 1 << n, S, C64)
    b.launch()
    _abi.synchronize()
    assert np.array_equal(d_out.to_array(np.complex64, psi.shape), H.op_apply(ops, psi, C64))
    # grow the set well beyond its arena, then launch the same batch again
    codes2, coeffs2 = syn.random_codes(300, n, seed=2), syn.random_coeffs(300, seed=2)
    ops.assign(codes2, coeffs2)
    b.launch()
    _abi.synchronize()
    fresh = _abi.Ops(codes2, coeffs2)
    assert np.array_equal(d_out.to_array(np.complex64, psi.shape), H.op_apply(fresh, psi, C64))
    assert H.rel_err(d_ev.to_array(np.complex64, (S,)), H.op_expectation(fresh, psi, C64)) <= 1e-5
    # new coefficients only
    coeffs3 = syn.random_coeffs(300, seed=3)
    ops.set_coeffs(coeffs3)
    b.launch()
    _abi.synchronize()
    fresh = _abi.Ops(codes2, coeffs3)
    assert np.array_equal(d_out.to_array(np.complex64, psi.shape), H.op_apply(fresh, psi, C64))
    # a set re-targeted to another qubit count no longer fits the entry: a clear error
    ops.assign(syn.random_codes(4, n + 1, seed=4))
    with pytest.raises(ValueError):
        b.launch()
    b.destroy()


def coset_entries():
    """Apply entries large enough (with a 1 KiB threshold) and with few enough string
    groups to run as coset items: power-of-two and other state counts, one and
    several state chunks, complex64 with one state per slot (odd counts) and two
    (even counts), complex128 (128-byte row slices with 16 and 24 states), overwrite
    and accumulate; plus streaming apply and
    expectation entries in the same 512-thread kernel."""
    A, E = _abi.FP_BATCH_APPLY, _abi.FP_BATCH_EXPECT
    specs = [
        (A, 12, 40, 3, C64, False), (A, 13, 64, 59, C64, False), (A, 12, 48, 64, C64, True),
        (A, 12, 30, 16, C128, False), (A, 12, 33, 5, C128, True), (A, 11, 24, 8, C64, False),
        (E, 10, 20, 7, C64, False), (A, 9, 12, 4, C64, False), (E, 12, 30, 16, C128, True),
        (A, 12, 6, 1, C64, False), (A, 13, 40, 58, C64, False), (A, 12, 40, 24, C128, False),
    ]
    return [Entry(*s, seed=40 + i) for i, s in enumerate(specs)]


@pytest.mark.parametrize("graph", [True, False])
def test_coset_items_bit_exact(hip, graph):
    """Large apply entries run as coset items (string groups sharing 64 KiB LDS
    tiles, groups of one operator in order): IEEE-equal to the operator alone,
    for every coset-item length, over repeated launches."""
    H.say(f"batch coset items graph={graph}")
    entries = coset_entries()
    b = Batch()
    if not graph:
        _abi.call("fp_hip_batch_set_graph", b.handle, 0)
    for e in entries:
        b.add(e.kind, e.ops, e.d_in.ptr, e.d_out.ptr, 1 << e.n, e.S, e.dtype, e.accumulate)
    refs = [e.alone() for e in entries]
    for tiles in (1, 3, 16):
        b.schedule = {"coset_min_kib": 1, "coset_tiles": tiles}
        for rep in range(3):
            for e in entries:
                e.reset()
            b.launch()
            _abi.synchronize()
            info = b.info()
            assert info["coset_entries"] == 8 and info["coset_items"] > 0, info
            for e, r in zip(entries, refs):
                got = e.result()
                if e.kind == _abi.FP_BATCH_APPLY:
                    assert np.array_equal(got, r), (tiles, rep, e.n, e.N, e.S, e.dtype, e.accumulate)
                else:
                    tol = 1e-12 if e.dtype == C128 else 1e-5
                    assert H.rel_err(got - e.base, r - e.base) <= tol
            b.check()
    # threshold 0: no coset items, same results
    b.schedule = {"coset_min_kib": 0}
    for e in entries:
        e.reset()
    b.launch()
    _abi.synchronize()
    assert b.info()["coset_entries"] == 0
    for e, r in zip(entries, refs):
        if e.kind == _abi.FP_BATCH_APPLY:
            assert np.array_equal(e.result(), r)
    b.destroy()


@pytest.mark.parametrize("graph", [True, False])
def test_tiled_launch_entries_bit_exact(hip, graph):
    """Apply entries over the tiled-launch threshold run as coset-tiled launches of their
    own (one per string group, after the persistent kernel, in the same graph): IEEE-equal
    to the operator alone, beside coset items and streaming entries, over repeated
    launches; a batch of such entries alone launches no persistent kernel."""
    H.say(f"batch tiled launches graph={graph}")
    entries = coset_entries()
    b = Batch()
    if not graph:
        _abi.call("fp_hip_batch_set_graph", b.handle, 0)
    for e in entries:
        b.add(e.kind, e.ops, e.d_in.ptr, e.d_out.ptr, 1 << e.n, e.S, e.dtype, e.accumulate)
    refs = [e.alone() for e in entries]
    # 3000 KiB: the two 2^13-row entries with 58 and 59 states launch on their own, the
    # other large entries run as coset items
    for sched in ({"tiled_min_kib": 1, "coset_min_kib": 0}, {"tiled_min_kib": 1, "coset_min_kib": 1},
                  {"tiled_min_kib": 3000, "coset_min_kib": 1}):
        b.schedule = sched
        for rep in range(2):
            for e in entries:
                e.reset()
            b.launch()
            _abi.synchronize()
            info = b.info()
            assert info["tiled_entries"] >= 2, info
            if sched["tiled_min_kib"] == 3000:
                assert info["coset_entries"] >= 1, info
            for e, r in zip(entries, refs):
                got = e.result()
                if e.kind == _abi.FP_BATCH_APPLY:
                    assert np.array_equal(got, r), (sched, rep, e.n, e.N, e.S, e.dtype, e.accumulate)
                else:
                    tol = 1e-12 if e.dtype == C128 else 1e-5
                    assert H.rel_err(got - e.base, r - e.base) <= tol
            b.check()
    b.destroy()
    big = [entries[1], entries[10]]
    b = Batch()
    if not graph:
        _abi.call("fp_hip_batch_set_graph", b.handle, 0)
    for e in big:
        b.add(e.kind, e.ops, e.d_in.ptr, e.d_out.ptr, 1 << e.n, e.S, e.dtype, e.accumulate)
    b.schedule = {"tiled_min_kib": 1}
    for rep in range(2):
        for e in big:
            e.reset()
        b.launch()
        _abi.synchronize()
        info = b.info()
        assert info["tiled_entries"] == 2 and info["tiles"] == 0, info
        for e in big:
            assert np.array_equal(e.result(), e.alone())
    b.destroy()
