// fp_batch.hip — uneven batches of operators in one persistent, graph-captured
// launch, with a schedule picked by timing candidates on the device.
//
// A batch lists (operator set, apply | expectation, states, output) entries of
// any qubit count, string count, state count and dtype.  The host cuts each
// entry into work items, orders them (largest first by default), uploads the
// descriptors, and captures [reset counters, persistent kernel] into a
// hipGraph that is replayed on every launch.  Large apply entries whose string
// groups share coset tiles run as coset-tiled launches of their own (one per
// string group, after the persistent kernel in the same graph) or become coset
// items (one string group, a run of 64 KiB tiles each): about three state passes
// per group instead of one per string.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
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
    // Coset-tiled launches of their own from 64 MiB of states (the single-operator rule);
    // coset items would apply from 256 MiB (below that the states stay in the Infinity
    // Cache, whose repeated-read rate the streaming items reach) but the launches take
    // those entries first: in a 256-entry batch of complex64 applies, the 15 entries of
    // 64 MiB and more took 33.4 ms as launches of their own, 38.9 ms as streaming items
    // and 43.9 ms as coset items.
    fp_hip_batch_schedule sched{2048, 1u << 18, 4, FP_BATCH_ORDER_LARGEST_FIRST, 8, 256u << 10, 64u << 10};
    bool use_graph = true;
    // device tables, sized by capacity and reused across rebuilds
    bool built = false;
    fpb::BatchOp *d_ops = nullptr;
    fpb::WorkItem *d_items = nullptr;
    // [0] work counter, [1, 1 + n_done) per-operator and per-string-group completion
    // counters (zeroed by every launch), [1 + n_done] error word (zeroed by build)
    uint32_t *d_counter = nullptr;
    fpk::Dbl2 *d_partials = nullptr;
    size_t cap_ops = 0, cap_items = 0, cap_partials = 0, cap_counter = 0;
    std::vector<fpb::BatchOp> h_ops;
    std::vector<fpb::WorkItem> h_items;
    uint32_t n_items = 0;
    uint32_t n_blocks = 0;
    uint32_t n_done = 0;     // completion counters reset by every launch
    bool coset = false;      // some entry runs as coset items: the 512-thread kernel
    size_t lds = 0;          // dynamic LDS per block
    size_t coset_entries = 0, coset_groups = 0, coset_items = 0;
    // entries run as coset-tiled launches of their own (after the persistent kernel)
    struct Split
    {
        uint32_t entry;
        fprt::TileTables *tt;
        fprt::TileGeom geo;
    };
    std::vector<Split> split;
    size_t tiled_groups = 0;
    double plan_bytes = 0.0; // state passes of the plan: coset groups, or one per string
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

fp_status enqueue_kernels(fp_hip_batch *b, hipStream_t stream)
{
    uint32_t const n = b->n_done;
    if (b->n_items > 0)
    {
        hipLaunchKernelGGL(fpb::k_batch_reset, dim3((n + 1 + fpk::kBlock - 1) / fpk::kBlock), dim3(fpk::kBlock), 0,
                           stream, b->d_counter, b->d_counter + 1, n);
        if (b->coset)
            hipLaunchKernelGGL(fpb::k_batch_persistent<fpb::kTiledBlock>, dim3(b->n_blocks), dim3(fpb::kTiledBlock),
                               b->lds, stream, b->d_ops, b->d_items, b->n_items, b->d_counter, b->d_counter + 1,
                               b->d_partials, b->d_counter + 1 + n);
        else
            hipLaunchKernelGGL(fpb::k_batch_persistent<fpk::kBlock>, dim3(b->n_blocks), dim3(fpk::kBlock), b->lds,
                               stream, b->d_ops, b->d_items, b->n_items, b->d_counter, b->d_counter + 1,
                               b->d_partials, b->d_counter + 1 + n);
        FP_LAUNCHED();
    }
    // (a second stream beside the persistent kernel gained nothing: 64.96 against 64.97 ms
    // for the 256-entry complex64 batch)
    for (fp_hip_batch::Split const &sp : b->split)
    {
        BatchEntry const &e = b->entries[sp.entry];
        fp_status const r = fprt::launch_op_apply_tiled(sp.tt, sp.geo, e.dtype, e.states, e.out, e.dim, e.n_states,
                                                        e.accumulate, stream);
        if (r != FP_OK)
            return r;
    }
    return FP_OK;
}

// Persistent blocks per CU that can be resident for the kernel the batch launches.
int resident_blocks(bool coset, size_t lds)
{
    int nb = 0;
    hipError_t const e =
        coset ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fpb::k_batch_persistent<fpb::kTiledBlock>,
                                                             fpb::kTiledBlock, lds)
              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fpb::k_batch_persistent<fpk::kBlock>, fpk::kBlock,
                                                             lds);
    return (e == hipSuccess && nb > 0) ? nb : 1;
}

// Coset items of complex64 entries with an even state count (and 16-byte aligned
// buffers) take two states per slot (16-byte loads, stores and LDS reads).
bool coset_paired(BatchEntry const &e, fprt::TileGeom const &g)
{
    bool const aligned16 = ((reinterpret_cast<uintptr_t>(e.states) | reinterpret_cast<uintptr_t>(e.out)) & 15u) == 0;
    return e.dtype == FP_C64 && e.n_states % 2 == 0 && g.lsb >= 1 && aligned16;
}

// Coset geometry and string-group tables of an apply entry, when its state is at
// least `min_kib` KiB and its groups move fewer bytes than the streaming form:
// (3 G - 1) state passes (2 G with no output read for the first group, plus a
// read of the output per later group) against N + 1.  `tt` null: streaming.
fp_status coset_plan(BatchEntry const &e, uint32_t min_kib, fprt::TileGeom *geo, fprt::TileTables **tt)
{
    *tt = nullptr;
    if (e.kind != FP_BATCH_APPLY || min_kib == 0 || e.ops->n_strings < 8)
        return FP_OK;
    size_t const esz = e.dtype == FP_C128 ? 16 : 8;
    if (e.dim * e.n_states * esz < size_t(min_kib) * 1024 || (uint64_t(e.dim) * e.n_states) >> 32 != 0)
        return FP_OK;
    fprt::TileGeom const g = fprt::tile_geom(e.dtype, e.ops->n_qubits, e.n_states, false);
    // the kernel's slot decomposition needs the low k row bits inside one thread's
    // 2^(9 - lsbx) consecutive local rows (rm >= k); lsbx = lsb - 1 with two states per slot
    uint32_t const log2_block = uint32_t(__builtin_ctz(uint32_t(fpb::kTiledBlock)));
    uint32_t const lsbx = g.lsb - (coset_paired(e, g) ? 1u : 0u);
    if (g.dh == 0 || g.d >= e.ops->n_qubits || lsbx > log2_block || log2_block - lsbx < g.k)
        return FP_OK;
    fprt::TileTables *t = nullptr;
    fp_status s = fprt::tile_tables(e.ops, fprt::kTilesOpApply, g.k, g.dh, &t);
    if (s != FP_OK)
        return s;
    double const passes = 3.0 * t->n_groups - (e.accumulate ? 0.0 : 1.0);
    if (passes > 0.8 * double(e.ops->n_strings + 1))
        return FP_OK;
    if ((s = fprt::check_tile_fit(t, g, e.dtype, fpa::kTileBytes)) != FP_OK)
        return s;
    *geo = g;
    *tt = t;
    return FP_OK;
}

// Cuts the entries into tiles, uploads the tables on `stream` and (in graph
// mode) captures the launch.  The previous launch must have completed.
fp_status build(fp_hip_batch *b, hipStream_t stream)
{
    wait_idle(b);
    drop_graph(b);
    if (b->entries.empty())
        return fail(FP_ERR_INVALID_ARGUMENT, "batch is empty");
    b->coset_entries = b->coset_groups = b->coset_items = 0;
    b->split.clear();
    b->tiled_groups = 0;
    b->plan_bytes = 0.0;
    std::vector<fpb::BatchOp> ops;
    struct Item
    {
        fpb::WorkItem w;
        double cost;  // ordering key of the streaming items
        double bytes; // estimated memory traffic (interleaving of the two lists)
    };
    std::vector<Item> items;
    struct CosetOp
    {
        uint32_t op, groups;
        double bytes;
    };
    std::vector<CosetOp> coset_ops;
    uint64_t partial_elems = 0;
    uint32_t n_deps = 0;
    bool any_expect = false;
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
        fprt::TileGeom geo{};
        fprt::TileTables *tt = nullptr;
        if (e.kind == FP_BATCH_APPLY && sc.tiled_min_kib != 0)
        {
            fp_status ts = fprt::apply_tiled_plan(o, e.dtype, e.n_states, size_t(sc.tiled_min_kib) << 10,
                                                  fprt::kTilesOpApply, &geo, &tt);
            if (ts == FP_OK && tt)
                ts = fprt::check_tile_fit(tt, geo, e.dtype, fpa::kTileBytes);
            if (ts == FP_OK && tt)
                ts = fprt::tile_device(o, tt, stream);
            if (ts == FP_OK && tt)
                ts = fprt::apply_tiled_attributes();
            if (ts != FP_OK)
                return ts;
            if (tt)
            {
                // launches of their own: no items in the persistent kernel
                b->split.push_back(fp_hip_batch::Split{uint32_t(i), tt, geo});
                b->tiled_groups += tt->n_groups;
                b->plan_bytes += double(entry_in_bytes(e)) * (3.0 * tt->n_groups - (e.accumulate ? 0.0 : 1.0));
                op.sshift = 64u;
                op.n_tiles = 0;
                ops.push_back(op);
                continue;
            }
        }
        fp_status const cs = coset_plan(e, sc.coset_min_kib, &geo, &tt);
        if (cs != FP_OK)
            return cs;
        op.sshift = 64u;
        uint32_t tiles = 0;
        if (tt)
        {
            fp_status const us = fprt::tile_device(o, tt, stream);
            if (us != FP_OK)
                return us;
            op.coset = 1;
            op.nc = coset_paired(e, geo) ? 2u : 1u;
            op.groups = tt->groups;
            op.btab = tt->btab;
            op.astrs = static_cast<fpa::ApplyStr const *>(tt->strs);
            op.lsb = geo.lsb;
            op.k = geo.k;
            op.d = geo.d;
            op.n_cosets = uint64_t(e.dim) >> geo.d;
            if ((e.n_states & (e.n_states - 1)) == 0)
                op.sshift = uint32_t(__builtin_ctzll(e.n_states));
            uint32_t const cpi = std::max<uint32_t>(sc.coset_tiles, 1);
            uint64_t const chunks = (e.n_states + (size_t(1) << geo.lsb) - 1) >> geo.lsb;
            uint64_t const per_group = chunks * ((op.n_cosets + cpi - 1) / cpi);
            if (per_group > 0xffffffffu || uint64_t(n_deps) + tt->n_groups > 0xffffffu)
                return fail(FP_ERR_INVALID_ARGUMENT, "batch entry %zu has too many coset items", i);
            op.items_per_group = uint32_t(per_group);
            op.dep_base = 0; // set below, after the per-operator counters
            op.n_tiles = tt->n_groups;
            n_deps += tt->n_groups;
            // first group: state read + output write (+ output read when accumulating);
            // later groups: state read + output read and write
            b->plan_bytes += double(entry_in_bytes(e)) * (3.0 * tt->n_groups - (e.accumulate ? 0.0 : 1.0));
            coset_ops.push_back(CosetOp{uint32_t(i), tt->n_groups, double(entry_in_bytes(e))});
            b->coset_groups += tt->n_groups;
            ++b->coset_entries;
            ops.push_back(op);
            continue;
        }
        op.ct = uint32_t(std::min<uint64_t>(g.vpr, fpk::kBlock));
        op.rt = fpk::kBlock / op.ct;
        b->plan_bytes += double(entry_in_bytes(e)) * (e.kind == FP_BATCH_APPLY
                                                          ? double(o->n_strings) + 1.0 + (e.accumulate ? 1.0 : 0.0)
                                                          : double(o->n_strings));
        if (e.kind == FP_BATCH_APPLY)
        {
            uint64_t const tv = std::max<uint64_t>(sc.tile_vectors, 64);
            for (uint64_t v0 = 0; v0 < g.n_vec; v0 += tv, ++tiles)
            {
                uint64_t const cnt = std::min<uint64_t>(tv, g.n_vec - v0);
                items.push_back({fpb::WorkItem{uint32_t(i), tiles, v0, cnt, 0, 0}, double(cnt) * op.n_strings,
                                 double(cnt) * g.nc * (e.dtype == FP_C128 ? 16 : 8) * (op.n_strings + 1.0)});
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
            any_expect = true;
            for (uint64_t p0 = 0; p0 < op.n_pairs; p0 += pairs, ++tiles)
            {
                uint64_t const cnt = std::min<uint64_t>(pairs, op.n_pairs - p0);
                double const work = 2.0 * double(cnt) * double(per_pair);
                items.push_back({fpb::WorkItem{uint32_t(i), tiles, p0, cnt, 0, 0}, work,
                                 work * g.nc * (e.dtype == FP_C128 ? 16 : 8)});
            }
            partial_elems += uint64_t(tiles) * e.n_states;
        }
        op.n_tiles = tiles;
        ops.push_back(op);
    }
    if (sc.order == FP_BATCH_ORDER_LARGEST_FIRST)
    {
        std::stable_sort(items.begin(), items.end(), [](Item const &a, Item const &c) { return a.cost > c.cost; });
        std::stable_sort(coset_ops.begin(), coset_ops.end(),
                         [](CosetOp const &a, CosetOp const &c) { return a.bytes > c.bytes; });
    }
    // coset items, group level by group level (every item of an operator's group
    // g - 1 precedes its group g: the waits' progress condition)
    std::vector<Item> citems;
    uint32_t const n_ops = uint32_t(ops.size());
    {
        uint32_t dep = n_ops;
        for (CosetOp const &c : coset_ops)
        {
            ops[c.op].dep_base = dep;
            dep += c.groups;
        }
        uint32_t max_groups = 0;
        for (CosetOp const &c : coset_ops)
            max_groups = std::max(max_groups, c.groups);
        uint32_t const cpi = std::max<uint32_t>(sc.coset_tiles, 1);
        for (uint32_t g = 0; g < max_groups; ++g)
            for (CosetOp const &c : coset_ops)
            {
                if (g >= c.groups)
                    continue;
                BatchEntry const &e = b->entries[c.op];
                fpb::BatchOp const &op = ops[c.op];
                uint64_t const chunks = (e.n_states + (size_t(1) << op.lsb) - 1) >> op.lsb;
                double const tile_bytes = double(size_t(1) << (op.d + op.lsb)) * (e.dtype == FP_C128 ? 16 : 8);
                double const passes = (g == 0 && !e.accumulate) ? 2.0 : 3.0;
                for (uint64_t y = 0; y < chunks; ++y)
                    for (uint64_t q0 = 0; q0 < op.n_cosets; q0 += cpi)
                    {
                        uint64_t const cnt = std::min<uint64_t>(cpi, op.n_cosets - q0);
                        citems.push_back({fpb::WorkItem{c.op, g, q0, cnt, uint32_t(y), 0}, 0.0,
                                          double(cnt) * tile_bytes * passes});
                    }
            }
    }
    b->coset_items = citems.size();
    // interleave the two lists by their shares of estimated bytes, so that small
    // streaming items fill in while coset groups complete
    std::vector<fpb::WorkItem> work;
    work.reserve(items.size() + citems.size());
    {
        double ts = 0.0, tc = 0.0, cost = 0.0;
        for (Item const &it : items)
        {
            ts += it.bytes;
            cost += it.cost;
        }
        for (Item const &it : citems)
            tc += it.bytes;
        double ds = 0.0, dc = 0.0;
        size_t is = 0, ic = 0;
        while (is < items.size() || ic < citems.size())
        {
            bool const take_c =
                ic < citems.size() && (is >= items.size() || dc / std::max(tc, 1.0) <= ds / std::max(ts, 1.0));
            if (take_c)
            {
                dc += citems[ic].bytes;
                work.push_back(citems[ic++].w);
            }
            else
            {
                ds += items[is].bytes;
                work.push_back(items[is++].w);
            }
        }
        b->est_cost = cost;
    }
    if (work.size() > 0xffffffffu)
        return fail(FP_ERR_INVALID_ARGUMENT, "batch has too many tiles");
    b->n_items = uint32_t(work.size());
    b->n_done = n_ops + n_deps;
    b->coset = !coset_ops.empty();
    if (b->coset) // expectation items of the 512-thread kernel: twice the rows per column
        for (fpb::BatchOp &op : ops)
            if (op.kind == fpb::kExpect)
            {
                op.ct = uint32_t(std::min<uint64_t>(op.vpr, fpb::kTiledBlock));
                op.rt = fpb::kTiledBlock / op.ct;
            }
    {
        int const B = b->coset ? fpb::kTiledBlock : fpk::kBlock;
        b->lds = std::max<size_t>(b->coset ? fpa::kTileBytes : 0, any_expect ? size_t(4) * B * sizeof(double) : 0);
        if (b->lds > fprt::lds_bytes())
            return fail(FP_ERR_UNSUPPORTED, "batch kernel needs %zu bytes of LDS per workgroup, the device has %zu",
                        b->lds, fprt::lds_bytes());
        static std::once_flag attr_once;
        static hipError_t attr_err = hipSuccess;
        std::call_once(attr_once, [] {
            attr_err = hipFuncSetAttribute(reinterpret_cast<void const *>(&fpb::k_batch_persistent<fpb::kTiledBlock>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, int(fpa::kTileBytes));
        });
        if (attr_err != hipSuccess)
            return fail(FP_ERR_RUNTIME, "hipFuncSetAttribute failed: %s", hipGetErrorString(attr_err));
        int const per_cu = std::min<int>(int(std::max<uint32_t>(sc.blocks_per_cu, 1)), resident_blocks(b->coset, b->lds));
        b->n_blocks = uint32_t(std::max<int>(1, device_cus() * per_cu));
    }

    fp_status st;
    if ((st = ensure_capacity(&b->d_ops, b->cap_ops, ops.size(), stream)) != FP_OK ||
        (st = ensure_capacity(&b->d_items, b->cap_items, std::max<size_t>(work.size(), 1), stream)) != FP_OK ||
        (st = ensure_capacity(&b->d_partials, b->cap_partials, std::max<uint64_t>(partial_elems, 1), stream)) !=
            FP_OK ||
        (st = ensure_capacity(&b->d_counter, b->cap_counter, size_t(b->n_done) + 2, stream)) != FP_OK)
        return st;
    FP_HIP(hipMemsetAsync(b->d_counter + 1 + b->n_done, 0, sizeof(uint32_t), stream)); // error word
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
        fp_status const es = enqueue_kernels(b, b->capture_stream);
        hipError_t const le = hipGetLastError();
        hipError_t const ce = hipStreamEndCapture(b->capture_stream, &b->graph);
        if (es != FP_OK || le != hipSuccess || ce != hipSuccess)
        {
            if (b->graph)
                (void)hipGraphDestroy(b->graph);
            b->graph = nullptr;
        }
        if (es != FP_OK)
            return es;
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
        (s->order != FP_BATCH_ORDER_LARGEST_FIRST && s->order != FP_BATCH_ORDER_AS_ADDED) || s->coset_tiles == 0 ||
        s->coset_tiles > 4096)
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

fp_status fp_hip_batch_coset_info(fp_hip_batch const *b, size_t *n_entries, size_t *n_groups, size_t *n_items,
                                  double *plan_bytes)
{
    if (!b)
        return fail(FP_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (n_entries)
        *n_entries = b->coset_entries;
    if (n_groups)
        *n_groups = b->coset_groups;
    if (n_items)
        *n_items = b->coset_items;
    if (plan_bytes)
        *plan_bytes = b->plan_bytes;
    return FP_OK;
}

fp_status fp_hip_batch_tiled_info(fp_hip_batch const *b, size_t *n_entries, size_t *n_groups)
{
    if (!b)
        return fail(FP_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (n_entries)
        *n_entries = b->split.size();
    if (n_groups)
        *n_groups = b->tiled_groups;
    return FP_OK;
}

fp_status fp_hip_batch_check(fp_hip_batch *b)
{
    if (!b)
        return fail(FP_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (!b->built || !b->d_counter)
        return FP_OK;
    wait_idle(b);
    uint32_t err = 0;
    FP_HIP(hipMemcpy(&err, b->d_counter + 1 + b->n_done, sizeof(err), hipMemcpyDeviceToHost));
    if (err != 0)
        return fail(FP_ERR_RUNTIME, "a coset work item gave up waiting for the previous string group");
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
        fp_status const r = enqueue_kernels(b, st);
        if (r != FP_OK)
            return r;
    }
    if (!b->last_done)
        FP_HIP(hipEventCreateWithFlags(&b->last_done, hipEventDisableTiming));
    FP_HIP(hipEventRecord(b->last_done, st));
    return FP_OK;
}

// Coordinate descent over the schedule: tile size, expectation-tile work,
// blocks per CU, tiled-launch threshold, coset-item threshold and length, then tile order; each candidate is timed over `reps` graph
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
                           {1, 2, 4, 8})) == FP_OK &&
            (r = try_field([](fp_hip_batch_schedule &s, uint64_t v) { s.tiled_min_kib = uint32_t(v); },
                           {0, 16384, 65536, 262144})) == FP_OK &&
            (r = try_field([](fp_hip_batch_schedule &s, uint64_t v) { s.coset_min_kib = uint32_t(v); },
                           {0, 4096, 32768, 262144})) == FP_OK &&
            (r = try_field([](fp_hip_batch_schedule &s, uint64_t v) { s.coset_tiles = uint32_t(v); },
                           {4, 8, 16, 32})) == FP_OK)
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
