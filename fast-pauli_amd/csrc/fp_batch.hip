// fp_batch.hip — uneven batches of operators in one persistent, graph-captured
// launch, with a schedule picked by timing candidates on the device.
//
// A batch lists (operator set, apply | expectation, states, output) entries of
// any qubit count, string count, state count and dtype.  The host cuts each
// entry into tiles, orders the tiles (largest first by default), uploads the
// descriptors, and captures [reset counters, persistent kernel] into a
// hipGraph that is replayed on every launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "fast_pauli_hip.h"
#include "fp_batch_kernels.hpp"
#include "fp_runtime.hpp"

using fprt::fail;

struct BatchEntry
{
    fp_hip_ops const *ops;
    int kind;
    int dtype;
    void const *states;
    void *out;
    size_t dim;
    size_t n_states;
    int accumulate;
    uint64_t generation; // ops->generation the device tables were built from
};

struct fp_hip_batch
{
    std::vector<BatchEntry> entries;
    fp_hip_batch_schedule sched{2048, 1u << 18, 4, FP_BATCH_ORDER_LARGEST_FIRST};
    bool use_graph = true;
    // device tables, sized by capacity and reused across rebuilds
    bool built = false;
    fpb::BatchOp *d_ops = nullptr;
    fpb::WorkItem *d_items = nullptr;
    uint32_t *d_counter = nullptr; // [0] work counter, [1..] per-operator completion counters
    fpk::Dbl2 *d_partials = nullptr;
    size_t cap_ops = 0, cap_items = 0, cap_partials = 0, cap_counter = 0;
    std::vector<fpb::BatchOp> h_ops;
    std::vector<fpb::WorkItem> h_items;
    uint32_t n_items = 0;
    uint32_t n_blocks = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture_stream = nullptr;
    hipEvent_t last_done = nullptr; // recorded after every launch
    double est_cost = 0.0;
};

namespace
{

// Waits for the last launch of this batch (its tables may be rewritten afterwards).
void wait_idle(fp_hip_batch *b)
{
    if (b->last_done)
        (void)hipEventSynchronize(b->last_done);
}

void drop_graph(fp_hip_batch *b)
{
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->built = false;
}

void release(fp_hip_batch *b)
{
    wait_idle(b);
    drop_graph(b);
    for (void *p : {static_cast<void *>(b->d_ops), static_cast<void *>(b->d_items),
                    static_cast<void *>(b->d_counter), static_cast<void *>(b->d_partials)})
        if (p)
            (void)hipFree(p);
    b->d_ops = nullptr;
    b->d_items = nullptr;
    b->d_counter = nullptr;
    b->d_partials = nullptr;
    b->cap_ops = b->cap_items = b->cap_partials = b->cap_counter = 0;
}

template <class T> fp_status ensure_capacity(T **ptr, size_t &cap, size_t need, hipStream_t stream)
{
    if (cap >= need && *ptr)
        return FP_OK;
    if (*ptr)
    {
        FP_HIP(hipStreamSynchronize(stream));
        FP_HIP(hipFree(*ptr));
        *ptr = nullptr;
    }
    size_t const want = std::max<size_t>(need + need / 2, 16);
    FP_HIP(hipMalloc(reinterpret_cast<void **>(ptr), want * sizeof(T)));
    cap = want;
    return FP_OK;
}

bool ranges_overlap(void const *a, size_t na, void const *b, size_t nb)
{
    auto const pa = reinterpret_cast<uintptr_t>(a), pb = reinterpret_cast<uintptr_t>(b);
    return pa < pb + nb && pb < pa + na;
}

size_t entry_in_bytes(BatchEntry const &e)
{
    return e.dim * e.n_states * (e.dtype == FP_C128 ? 16 : 8);
}

size_t entry_out_bytes(BatchEntry const &e)
{
    return (e.kind == FP_BATCH_APPLY ? e.dim : 1) * e.n_states * (e.dtype == FP_C128 ? 16 : 8);
}

int device_cus()
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    return cus;
}

void enqueue_kernels(fp_hip_batch *b, hipStream_t stream)
{
    uint32_t const n_ops = uint32_t(b->h_ops.size());
    hipLaunchKernelGGL(fpb::k_batch_reset, dim3((n_ops + 1 + fpk::kBlock - 1) / fpk::kBlock), dim3(fpk::kBlock), 0,
                       stream, b->d_counter, b->d_counter + 1, n_ops);
    hipLaunchKernelGGL(fpb::k_batch_persistent, dim3(b->n_blocks), dim3(fpk::kBlock), 0, stream, b->d_ops, b->d_items,
                       b->n_items, b->d_counter, b->d_counter + 1, b->d_partials);
}

// Cuts the entries into tiles, uploads the tables on `stream` and (in graph
// mode) captures the launch.  The previous launch must have completed.
fp_status build(fp_hip_batch *b, hipStream_t stream)
{
    wait_idle(b);
    drop_graph(b);
    if (b->entries.empty())
        return fail(FP_ERR_INVALID_ARGUMENT, "batch is empty");
    std::vector<fpb::BatchOp> ops;
    struct Item
    {
        fpb::WorkItem w;
        double cost;
    };
    std::vector<Item> items;
    uint64_t partial_elems = 0;
    fp_hip_batch_schedule const &sc = b->sched;
    for (size_t i = 0; i < b->entries.size(); ++i)
    {
        BatchEntry &e = b->entries[i];
        fp_hip_ops const *o = e.ops;
        // the operator set may have been re-targeted since the entry was added
        if (e.dim != (size_t(1) << o->n_qubits))
            return fail(FP_ERR_INVALID_ARGUMENT,
                        "batch entry %zu: its operator set now acts on %zu qubits, not on the entry's %zu rows", i,
                        o->n_qubits, e.dim);
        e.generation = o->generation;
        fprt::VecGeom const g = fprt::vec_geom(e.dtype, e.dim, e.n_states, e.states, e.out);
        fpb::BatchOp op{};
        op.in = e.states;
        op.out = e.out;
        op.xs = o->d_x;
        op.zs = o->d_z;
        op.cf = e.dtype == FP_C128 ? static_cast<void const *>(o->d_cf128) : static_cast<void const *>(o->d_cf64);
        op.meta = o->d_meta;
        op.scale = o->d_scale;
        op.n_vec = g.n_vec;
        op.vpr = g.vpr;
        op.n_pairs = e.dim / 2;
        op.n_states = e.n_states;
        op.n_strings = uint32_t(o->n_strings);
        op.shift = g.shift;
        op.pow2 = g.pow2 ? 1u : 0u;
        op.nc = uint32_t(g.nc);
        op.dtype = uint32_t(e.dtype);
        op.kind = e.kind == FP_BATCH_APPLY ? fpb::kApply : fpb::kExpect;
        op.accumulate = e.accumulate ? 1u : 0u;
        op.ct = uint32_t(std::min<uint64_t>(g.vpr, fpk::kBlock));
        op.rt = fpk::kBlock / op.ct;
        uint32_t tiles = 0;
        if (e.kind == FP_BATCH_APPLY)
        {
            uint64_t const tv = std::max<uint64_t>(sc.tile_vectors, 64);
            for (uint64_t v0 = 0; v0 < g.n_vec; v0 += tv, ++tiles)
            {
                uint64_t const cnt = std::min<uint64_t>(tv, g.n_vec - v0);
                items.push_back({fpb::WorkItem{uint32_t(i), tiles, v0, cnt}, double(cnt) * op.n_strings});
            }
        }
        else
        {
            // pairs per tile so that strings x pairs x vectors-per-row ~ ev_tile_work
            uint64_t const per_pair = uint64_t(op.n_strings) * std::max<uint64_t>(g.vpr, 1);
            uint64_t const quantum = uint64_t(op.rt) * 4;
            uint64_t pairs = std::max<uint64_t>(sc.ev_tile_work / std::max<uint64_t>(per_pair, 1), quantum);
            pairs = (pairs + quantum - 1) / quantum * quantum;
            pairs = std::min<uint64_t>(pairs, op.n_pairs);
            op.partial_base = partial_elems;
            for (uint64_t p0 = 0; p0 < op.n_pairs; p0 += pairs, ++tiles)
            {
                uint64_t const cnt = std::min<uint64_t>(pairs, op.n_pairs - p0);
                items.push_back({fpb::WorkItem{uint32_t(i), tiles, p0, cnt}, 2.0 * double(cnt) * double(per_pair)});
            }
            partial_elems += uint64_t(tiles) * e.n_states;
        }
        op.n_tiles = tiles;
        ops.push_back(op);
    }
    if (sc.order == FP_BATCH_ORDER_LARGEST_FIRST)
        std::stable_sort(items.begin(), items.end(), [](Item const &a, Item const &c) { return a.cost > c.cost; });
    std::vector<fpb::WorkItem> work(items.size());
    double total = 0.0;
    for (size_t i = 0; i < items.size(); ++i)
    {
        work[i] = items[i].w;
        total += items[i].cost;
    }
    b->est_cost = total;
    if (work.size() > 0xffffffffu)
        return fail(FP_ERR_INVALID_ARGUMENT, "batch has too many tiles");
    b->n_items = uint32_t(work.size());
    b->n_blocks = uint32_t(std::max<int>(1, device_cus() * int(std::max<uint32_t>(sc.blocks_per_cu, 1))));

    fp_status st;
    if ((st = ensure_capacity(&b->d_ops, b->cap_ops, ops.size(), stream)) != FP_OK ||
        (st = ensure_capacity(&b->d_items, b->cap_items, std::max<size_t>(work.size(), 1), stream)) != FP_OK ||
        (st = ensure_capacity(&b->d_partials, b->cap_partials, std::max<uint64_t>(partial_elems, 1), stream)) !=
            FP_OK ||
        (st = ensure_capacity(&b->d_counter, b->cap_counter, ops.size() + 1, stream)) != FP_OK)
        return st;
    b->h_ops = std::move(ops);
    b->h_items = std::move(work);
    FP_HIP(hipMemcpyAsync(b->d_ops, b->h_ops.data(), b->h_ops.size() * sizeof(fpb::BatchOp), hipMemcpyHostToDevice,
                          stream));
    if (!b->h_items.empty())
        FP_HIP(hipMemcpyAsync(b->d_items, b->h_items.data(), b->h_items.size() * sizeof(fpb::WorkItem),
                              hipMemcpyHostToDevice, stream));
    if (b->use_graph)
    {
        // capture [reset, persistent kernel] once; replayed by every launch
        if (!b->capture_stream)
            FP_HIP(hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking));
        FP_HIP(hipStreamSynchronize(stream)); // tables uploaded before the graph can run on any stream
        FP_HIP(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal));
        enqueue_kernels(b, b->capture_stream);
        hipError_t const le = hipGetLastError();
        hipError_t const ce = hipStreamEndCapture(b->capture_stream, &b->graph);
        if (le != hipSuccess)
            return fail(FP_ERR_RUNTIME, "batch kernel capture failed: %s", hipGetErrorString(le));
        if (ce != hipSuccess)
            return fail(FP_ERR_RUNTIME, "hipStreamEndCapture failed: %s", hipGetErrorString(ce));
        FP_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    b->built = true;
    return FP_OK;
}

} // namespace

extern "C"
{

fp_status fp_hip_batch_create(fp_hip_batch **out)
{
    if (!out)
        return fail(FP_ERR_INVALID_ARGUMENT, "batch out-pointer is NULL");
    *out = new (std::nothrow) fp_hip_batch();
    if (!*out)
        return fail(FP_ERR_OUT_OF_MEMORY, "host allocation failed");
    return FP_OK;
}

fp_status fp_hip_batch_destroy(fp_hip_batch *b)
{
    if (!b)
        return FP_OK;
    release(b);
    if (b->capture_stream)
        (void)hipStreamDestroy(b->capture_stream);
    if (b->last_done)
        (void)hipEventDestroy(b->last_done);
    delete b;
    return FP_OK;
}

fp_status fp_hip_batch_clear(fp_hip_batch *b)
{
    if (!b)
        return fail(FP_ERR_INVALID_ARGUMENT, "batch is NULL");
    wait_idle(b);
    drop_graph(b);
    b->entries.clear();
    return FP_OK;
}

fp_status fp_hip_batch_set_graph(fp_hip_batch *b, int enabled)
{
    if (!b)
        return fail(FP_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (b->use_graph != (enabled != 0))
    {
        wait_idle(b);
        drop_graph(b);
        b->use_graph = enabled != 0;
    }
    return FP_OK;
}

fp_status fp_hip_batch_add(fp_hip_batch *b, int kind, fp_hip_ops const *ops, int dtype, void const *states, void *out,
                           size_t dim, size_t n_states, int accumulate)
{
    if (!b || !ops)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL batch or operator");
    if (kind != FP_BATCH_APPLY && kind != FP_BATCH_EXPECT)
        return fail(FP_ERR_INVALID_ARGUMENT, "unknown batch entry kind %d", kind);
    if (dtype != FP_C128 && dtype != FP_C64)
        return fail(FP_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    if (!states || !out)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL data pointer");
    size_t const want = size_t(1) << ops->n_qubits;
    if (dim != want)
        return fail(FP_ERR_INVALID_ARGUMENT, "states shape (%zu) must match the dimension of the operators (%zu)", dim,
                    want);
    if (n_states == 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "n_states must be positive");
    uintptr_t const align = dtype == FP_C128 ? 16 : 8;
    if (((reinterpret_cast<uintptr_t>(states) | reinterpret_cast<uintptr_t>(out)) & (align - 1)) != 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "device buffers must be aligned to the complex element size");
    BatchEntry const e{ops, kind, dtype, states, out, dim, n_states, accumulate, ops->generation};
    // an output may not overlap any other entry's input or output (they run concurrently)
    for (BatchEntry const &o : b->entries)
    {
        if (ranges_overlap(e.out, entry_out_bytes(e), o.states, entry_in_bytes(o)) ||
            ranges_overlap(e.out, entry_out_bytes(e), o.out, entry_out_bytes(o)) ||
            ranges_overlap(o.out, entry_out_bytes(o), e.states, entry_in_bytes(e)))
            return fail(FP_ERR_INVALID_ARGUMENT, "batch outputs must not overlap other entries' buffers");
    }
    if (ranges_overlap(e.out, entry_out_bytes(e), e.states, entry_in_bytes(e)))
        return fail(FP_ERR_INVALID_ARGUMENT, "an entry's output must not overlap its input");
    wait_idle(b);
    drop_graph(b);
    b->entries.push_back(e);
    return FP_OK;
}

fp_status fp_hip_batch_set_schedule(fp_hip_batch *b, fp_hip_batch_schedule const *s)
{
    if (!b || !s)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL batch or schedule");
    if (s->tile_vectors == 0 || s->ev_tile_work == 0 || s->blocks_per_cu == 0 || s->blocks_per_cu > 64 ||
        (s->order != FP_BATCH_ORDER_LARGEST_FIRST && s->order != FP_BATCH_ORDER_AS_ADDED))
        return fail(FP_ERR_INVALID_ARGUMENT, "invalid batch schedule");
    if (std::memcmp(&b->sched, s, sizeof(*s)) != 0)
    {
        wait_idle(b);
        drop_graph(b);
        b->sched = *s;
    }
    return FP_OK;
}

fp_status fp_hip_batch_get_schedule(fp_hip_batch const *b, fp_hip_batch_schedule *s)
{
    if (!b || !s)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL batch or schedule");
    *s = b->sched;
    return FP_OK;
}

fp_status fp_hip_batch_info(fp_hip_batch const *b, size_t *n_entries, size_t *n_tiles, size_t *n_blocks)
{
    if (!b)
        return fail(FP_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (n_entries)
        *n_entries = b->entries.size();
    if (n_tiles)
        *n_tiles = b->n_items;
    if (n_blocks)
        *n_blocks = b->n_blocks;
    return FP_OK;
}

fp_status fp_hip_batch_launch(fp_hip_batch *b, void *stream)
{
    if (!b)
        return fail(FP_ERR_INVALID_ARGUMENT, "batch is NULL");
    hipStream_t const st = static_cast<hipStream_t>(stream);
    // tables and graphs hold the operator sets' device pointers and string counts:
    // rebuild when a set has been re-targeted (or given new coefficients) since
    for (BatchEntry const &e : b->entries)
        if (b->built && e.ops->generation != e.generation)
        {
            wait_idle(b);
            drop_graph(b);
        }
    if (!b->built)
    {
        fp_status const r = build(b, st);
        if (r != FP_OK)
            return r;
    }
    if (b->use_graph)
        FP_HIP(hipGraphLaunch(b->exec, st));
    else
    {
        enqueue_kernels(b, st);
        FP_LAUNCHED();
    }
    if (!b->last_done)
        FP_HIP(hipEventCreateWithFlags(&b->last_done, hipEventDisableTiming));
    FP_HIP(hipEventRecord(b->last_done, st));
    return FP_OK;
}

// Coordinate descent over the schedule: tile size, expectation-tile work,
// blocks per CU, then tile order; each candidate is timed over `reps` graph
// launches (after one warm-up) and the fastest is kept.  Outputs are
// overwritten by the trial launches (accumulating entries accumulate reps+1
// times), so callers tune on scratch outputs or re-initialise them.
fp_status fp_hip_batch_autotune(fp_hip_batch *b, void *stream, int reps, fp_hip_batch_schedule *chosen,
                                float *best_ms)
{
    if (!b)
        return fail(FP_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (reps <= 0)
        reps = 3;
    hipStream_t const st = static_cast<hipStream_t>(stream);
    hipEvent_t e0, e1;
    FP_HIP(hipEventCreate(&e0));
    FP_HIP(hipEventCreate(&e1));
    auto time_sched = [&](fp_hip_batch_schedule const &s, float &ms) -> fp_status {
        fp_status r = fp_hip_batch_set_schedule(b, &s);
        if (r != FP_OK)
            return r;
        if ((r = fp_hip_batch_launch(b, stream)) != FP_OK)
            return r;
        FP_HIP(hipEventRecord(e0, st));
        for (int i = 0; i < reps; ++i)
            if ((r = fp_hip_batch_launch(b, stream)) != FP_OK)
                return r;
        FP_HIP(hipEventRecord(e1, st));
        FP_HIP(hipEventSynchronize(e1));
        FP_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= float(reps);
        return FP_OK;
    };
    fp_hip_batch_schedule best = b->sched;
    float best_t = 0.f;
    fp_status r = time_sched(best, best_t);
    if (r == FP_OK)
    {
        auto try_field = [&](auto setter, std::vector<uint64_t> const &values) -> fp_status {
            for (uint64_t v : values)
            {
                fp_hip_batch_schedule s = best;
                setter(s, v);
                if (std::memcmp(&s, &best, sizeof(s)) == 0)
                    continue;
                float t = 0.f;
                fp_status const q = time_sched(s, t);
                if (q != FP_OK)
                    return q;
                if (t < best_t)
                {
                    best_t = t;
                    best = s;
                }
            }
            return FP_OK;
        };
        if ((r = try_field([](fp_hip_batch_schedule &s, uint64_t v) { s.tile_vectors = uint32_t(v); },
                           {512, 1024, 2048, 4096, 8192})) == FP_OK &&
            (r = try_field([](fp_hip_batch_schedule &s, uint64_t v) { s.ev_tile_work = v; },
                           {uint64_t(1) << 14, uint64_t(1) << 16, uint64_t(1) << 18, uint64_t(1) << 20})) == FP_OK &&
            (r = try_field([](fp_hip_batch_schedule &s, uint64_t v) { s.blocks_per_cu = uint32_t(v); },
                           {1, 2, 4, 8})) == FP_OK)
            r = try_field([](fp_hip_batch_schedule &s, uint64_t v) { s.order = int(v); },
                          {FP_BATCH_ORDER_LARGEST_FIRST, FP_BATCH_ORDER_AS_ADDED});
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r != FP_OK)
        return r;
    if ((r = fp_hip_batch_set_schedule(b, &best)) != FP_OK)
        return r;
    if (chosen)
        *chosen = best;
    if (best_ms)
        *best_ms = best_t;
    return FP_OK;
}

} // extern "C"
