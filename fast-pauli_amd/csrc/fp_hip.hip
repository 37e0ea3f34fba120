// fp_hip.hip — runtime and C ABI of libfastpauli_hip.so (gfx950).
//
// Implements include/fast_pauli_hip.h: operator sets, per-stream scratch,
// launch-shape selection and the entry points of the hot path.  Kernels live
// in fp_kernels.hpp (apply / expectation) and fp_mfma.hpp (contraction).
#include "fast_pauli_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "fp_dense.hpp"
#include "fp_kernels.hpp"
#include "fp_mfma.hpp"
#include "fp_runtime.hpp"

namespace fprt
{

// ---------------------------------------------------------------------------
// Per-stream scratch: reduction partials etc.  Grown on demand (after the
// stream has drained), never shrunk.  Calls on one stream are ordered, and the
// entry points that use scratch hold the stream's StreamGuard until their
// kernels are enqueued, so one buffer per stream suffices.
// ---------------------------------------------------------------------------
struct Scratch
{
    void *ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_scratch_mu;
std::map<std::pair<int, hipStream_t>, Scratch> g_scratch;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<std::recursive_mutex>> g_stream_mu;

StreamGuard::StreamGuard(hipStream_t stream)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        auto &m = g_stream_mu[{dev, stream}];
        if (!m)
            m = std::make_unique<std::recursive_mutex>();
        mu_ = m.get();
    }
    mu_->lock();
}

StreamGuard::~StreamGuard()
{
    mu_->unlock();
}

fp_status scratch_get(hipStream_t stream, size_t bytes, void **out)
{
    int dev = 0;
    FP_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    Scratch &s = g_scratch[{dev, stream}];
    if (s.bytes < bytes)
    {
        if (s.ptr)
        {
            // earlier calls on this stream have enqueued all their kernels (the
            // caller holds the stream's guard): once they finish, the buffer is free
            FP_HIP(hipStreamSynchronize(stream));
            FP_HIP(hipFree(s.ptr));
            s.ptr = nullptr;
            s.bytes = 0;
        }
        size_t const want = std::max<size_t>(bytes, size_t(1) << 20);
        FP_HIP(hipMalloc(&s.ptr, want));
        s.bytes = want;
    }
    *out = s.ptr;
    return FP_OK;
}

size_t lds_bytes()
{
    static std::mutex mu;
    static std::map<int, size_t> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lock(mu);
    auto const it = cache.find(dev);
    if (it != cache.end())
        return it->second;
    int v = 0;
    size_t bytes = 64 * 1024; // the smallest workgroup LDS of the CDNA parts, should the query fail
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
        bytes = size_t(v);
    else
        (void)hipGetLastError();
    cache[dev] = bytes;
    return bytes;
}

VecGeom vec_geom(int dtype, size_t dim, size_t n_states, void const *a, void const *b)
{
    VecGeom g{};
    bool const aligned16 = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15u) == 0;
    g.nc = (dtype == FP_C64 && (n_states % 2 == 0) && aligned16) ? 2 : 1;
    g.vpr = n_states / g.nc;
    g.n_vec = uint64_t(dim) * g.vpr;
    g.pow2 = (g.vpr & (g.vpr - 1)) == 0;
    g.shift = 0;
    while ((uint64_t(1) << g.shift) < g.vpr)
        ++g.shift;
    g.inv_vpr = 1.0 / double(g.vpr);
    return g;
}

} // namespace fprt

using namespace fprt;


namespace
{

void free_device(fp_hip_ops *o)
{
    fprt::release_tiles(o);
    void *ptrs[] = {o->d_arena, o->d_rowsum128, o->d_rowsum64, o->d_wr64, o->d_wi64,
                    o->d_wr32,  o->d_wi32,      o->d_h128,     o->d_h64};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    o->d_arena = nullptr;
    o->arena_bytes = 0;
    o->d_x = o->d_z = nullptr;
    o->d_meta = nullptr;
    o->d_ny = nullptr;
    o->d_cf128 = o->d_scale = o->d_kappa_c = o->d_kappa = o->d_rowsum128 = o->d_wr64 = o->d_wi64 = nullptr;
    o->d_cf64 = o->d_rowsum64 = o->d_wr32 = o->d_wi32 = nullptr;
    o->d_h128 = nullptr;
    o->d_h64 = nullptr;
}

// (-i)^k as (re, im)
void minus_i_pow(uint32_t k, double &re, double &im)
{
    switch (k & 3u)
    {
    case 0:
        re = 1, im = 0;
        break;
    case 1:
        re = 0, im = -1;
        break;
    case 2:
        re = -1, im = 0;
        break;
    default:
        re = 0, im = 1;
        break;
    }
}

// exact product of c by a unit on an axis (a swap and/or negation)
void mul_axis_unit(double cr, double ci, double ur, double ui, double &orr, double &oi)
{
    if (ui == 0)
    {
        orr = ur > 0 ? cr : -cr;
        oi = ur > 0 ? ci : -ci;
    }
    else
    {
        // (cr + i ci) * (i ui) = -ci ui + i cr ui
        orr = ui > 0 ? -ci : ci;
        oi = ui > 0 ? cr : -cr;
    }
}

double kappa_of(uint64_t x, uint32_t ny)
{
    if (x == 0)
        return 1.0;
    if ((ny & 1u) == 0)
        return (ny & 2u) ? -2.0 : 2.0;
    return ((ny & 3u) == 1u) ? 2.0 : -2.0;
}

template <class T> fp_status upload(T **dptr, std::vector<T> const &h)
{
    if (*dptr == nullptr)
        FP_HIP(hipMalloc(reinterpret_cast<void **>(dptr), std::max<size_t>(h.size(), 1) * sizeof(T)));
    if (!h.empty())
        FP_HIP(hipMemcpy(*dptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return FP_OK;
}

// Arena layout: x | z | ny | meta | cf128 | cf64 | scale | kappa_c | kappa,
// each section 256-byte aligned.
struct ArenaLayout
{
    size_t x, z, ny, meta, cf128, cf64, scale, kappa_c, kappa, ev_chunks, ev_strs, total;
};

ArenaLayout arena_layout(size_t N)
{
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    ArenaLayout L{};
    size_t off = 0;
    L.x = off, off += up(N * sizeof(uint64_t));
    L.z = off, off += up(N * sizeof(uint64_t));
    L.ny = off, off += up(N * sizeof(uint32_t));
    L.meta = off, off += up(N * sizeof(fpk::StringMeta));
    L.cf128 = off, off += up(2 * N * sizeof(double));
    L.cf64 = off, off += up(2 * N * sizeof(float));
    L.scale = off, off += up(2 * N * sizeof(double));
    L.kappa_c = off, off += up(2 * N * sizeof(double));
    L.kappa = off, off += up(N * sizeof(double));
    L.ev_chunks = off, off += up(N * sizeof(fpl::EvChunk));
    L.ev_strs = off, off += up(N * sizeof(fpl::EvStr));
    L.total = off;
    return L;
}

template <class T> T *at(void *base, size_t off)
{
    return reinterpret_cast<T *>(static_cast<unsigned char *>(base) + off);
}

// Coefficient-dependent sections of the host mirror.
void fill_coeff_tables(fp_hip_ops *o, ArenaLayout const &L)
{
    void *h = o->h_arena.data();
    size_t const N = o->n_strings;
    o->cf128.resize(2 * N);
    o->cf64.resize(2 * N);
    for (size_t s = 0; s < N; ++s)
    {
        double ur, ui;
        minus_i_pow(o->ny[s], ur, ui);
        double const cr = o->coeffs[2 * s], ci = o->coeffs[2 * s + 1];
        double *cf128 = at<double>(h, L.cf128);
        mul_axis_unit(cr, ci, ur, ui, cf128[2 * s], cf128[2 * s + 1]);
        o->cf128[2 * s] = cf128[2 * s];
        o->cf128[2 * s + 1] = cf128[2 * s + 1];
        double fr, fi;
        mul_axis_unit(double(float(cr)), double(float(ci)), ur, ui, fr, fi);
        at<float>(h, L.cf64)[2 * s] = float(fr);
        o->cf64[2 * s] = float(fr);
        o->cf64[2 * s + 1] = float(fi);
        at<float>(h, L.cf64)[2 * s + 1] = float(fi);
        double const k = kappa_of(o->x[s], o->ny[s]);
        at<double>(h, L.kappa)[s] = k;
        at<double>(h, L.kappa_c)[2 * s] = k;
        at<double>(h, L.kappa_c)[2 * s + 1] = 0.0;
        at<double>(h, L.scale)[2 * s] = cr * k;
        at<double>(h, L.scale)[2 * s + 1] = ci * k;
    }
}

// Groups strings with equal x and equal half of w (the parity of nY) into
// chunks of at most fpl::kChunk for the LDS expectation kernel, largest work
// first.  Classes of at least fpl::kWhtMin strings (columns of 2^kWhtMinRowsLog2
// rows and more) become transform classes of at most fpl::kWhtMax strings, listed
// first (*n_wht of them).  Returns the number of chunks, transform classes included.
uint32_t build_ev_chunks(fp_hip_ops const *o, fpl::EvChunk *chunks, fpl::EvStr *strs, uint32_t *n_wht)
{
    size_t const N = o->n_strings;
    std::vector<uint32_t> order(N);
    for (size_t i = 0; i < N; ++i)
        order[i] = uint32_t(i);
    auto key = [&](uint32_t a) { return std::make_pair(o->x[a], o->ny[a] & 1u); };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
    struct Pending
    {
        fpl::EvChunk c;
        std::vector<fpl::EvStr> s;
        double cost;
        bool wht;
    };
    std::vector<Pending> pend;
    size_t i = 0;
    while (i < N)
    {
        size_t j = i;
        while (j < N && key(order[j]) == key(order[i]))
            ++j;
        uint64_t const x = o->x[order[i]];
        uint32_t const h = x ? 63u - uint32_t(__builtin_clzll(x)) : 0u;
        uint32_t const mode = fpl::ev_mode(x, h, uint64_t(1) << o->n_qubits);
        bool const wht = j - i >= fprt::wht_min_strings() && o->n_qubits >= fpl::kWhtMinRowsLog2;
        size_t const per_chunk = wht ? fpl::kWhtMax : size_t(fpl::kChunk);
        for (size_t k = i; k < j; k += per_chunk)
        {
            Pending p{};
            p.wht = wht;
            p.c.x = uint32_t(x);
            p.c.h = h;
            p.c.mode = mode;
            p.c.imag = o->ny[order[i]] & 1u;
            p.c.count = uint32_t(std::min<size_t>(per_chunk, j - k));
            for (size_t m = k; m < k + p.c.count; ++m)
            {
                uint32_t const sidx = order[m];
                uint32_t const ny = o->ny[sidx];
                double const eps = ((ny & 3u) <= 1u) ? 1.0 : -1.0;
                p.s.push_back(fpl::EvStr{uint32_t(o->z[sidx]), ny & 1u, sidx, 0u,
                                         eps * (mode == fpl::kAllRows ? 1.0 : 2.0)});
            }
            p.cost = double(p.c.count + 2) * (mode == fpl::kAllRows ? 2.0 : 1.0);
            pend.push_back(std::move(p));
        }
        i = j;
    }
    // transform classes first, then per-string chunks, each costliest first
    std::stable_sort(pend.begin(), pend.end(), [](Pending const &a, Pending const &b) {
        return a.wht != b.wht ? a.wht : a.cost > b.cost;
    });
    *n_wht = 0;
    for (auto const &p : pend)
        *n_wht += p.wht ? 1u : 0u;
    uint32_t first = 0;
    for (size_t c = 0; c < pend.size(); ++c)
    {
        pend[c].c.first = first;
        chunks[c] = pend[c].c;
        for (auto const &e : pend[c].s)
            strs[first++] = e;
    }
    return uint32_t(pend.size());
}

// (Re)builds the string tables of `o` from letter codes and uploads them on
// `stream`.  Allocates only when the arena must grow.
fp_status ops_assign(fp_hip_ops *o, size_t n_qubits, size_t n_strings, uint8_t const *codes, double const *coeffs,
                     hipStream_t stream, bool sync)
{
    if (n_qubits == 0 || n_qubits > 62)
        return fail(FP_ERR_INVALID_ARGUMENT, "n_qubits must be in 1..62 (got %zu)", n_qubits);
    if (n_strings == 0 || n_strings > 0xffffffffu)
        return fail(FP_ERR_INVALID_ARGUMENT, "n_strings must be in 1..2^32-1 (got %zu)", n_strings);
    if (!codes)
        return fail(FP_ERR_INVALID_ARGUMENT, "codes is NULL");
    for (size_t i = 0; i < n_qubits * n_strings; ++i)
        if (codes[i] > 3)
            return fail(FP_ERR_INVALID_ARGUMENT, "Pauli code must be 0, 1, 2, or 3 (got %u)", unsigned(codes[i]));
    fprt::invalidate_tiles(o); // string groups belong to the previous strings
    ++o->generation;
    ArenaLayout const L = arena_layout(n_strings);
    if (o->upload_done)
        FP_HIP(hipEventSynchronize(o->upload_done)); // the previous upload may still read h_arena
    else
        FP_HIP(hipEventCreateWithFlags(&o->upload_done, hipEventDisableTiming));
    o->n_qubits = n_qubits;
    o->n_strings = n_strings;
    o->n_operators = 0;
    o->x.resize(n_strings);
    o->z.resize(n_strings);
    o->ny.resize(n_strings);
    o->coeffs.assign(2 * n_strings, 0.0);
    o->h_arena.assign(L.total, 0);
    void *h = o->h_arena.data();
    for (size_t s = 0; s < n_strings; ++s)
    {
        uint64_t x = 0, z = 0;
        for (size_t i = 0; i < n_qubits; ++i)
        {
            uint8_t const c = codes[s * n_qubits + i];
            uint64_t const bit = uint64_t(1) << (n_qubits - 1 - i);
            if (c == 1 || c == 2)
                x |= bit;
            if (c == 2 || c == 3)
                z |= bit;
        }
        uint32_t const ny = uint32_t(__builtin_popcountll(x & z));
        o->x[s] = x;
        o->z[s] = z;
        o->ny[s] = ny;
        uint32_t const hb = x ? 63u - uint32_t(__builtin_clzll(x)) : 0u;
        uint32_t const mode = x == 0 ? fpk::kModeSelf : ((ny & 1u) ? fpk::kModeIm : fpk::kModeRe);
        at<uint64_t>(h, L.x)[s] = x;
        at<uint64_t>(h, L.z)[s] = z;
        at<uint32_t>(h, L.ny)[s] = ny;
        at<fpk::StringMeta>(h, L.meta)[s] = fpk::StringMeta{x, z, hb, mode};
        o->coeffs[2 * s] = coeffs ? coeffs[2 * s] : 1.0;
        o->coeffs[2 * s + 1] = coeffs ? coeffs[2 * s + 1] : 0.0;
    }
    fill_coeff_tables(o, L);
    o->n_ev_chunks = build_ev_chunks(o, at<fpl::EvChunk>(h, L.ev_chunks), at<fpl::EvStr>(h, L.ev_strs), &o->n_ev_wht);
    if (o->arena_bytes < L.total)
    {
        if (o->d_arena)
        {
            FP_HIP(hipStreamSynchronize(stream));
            FP_HIP(hipFree(o->d_arena));
            o->d_arena = nullptr;
            o->arena_bytes = 0;
        }
        FP_HIP(hipMalloc(&o->d_arena, L.total));
        o->arena_bytes = L.total;
    }
    void *d = o->d_arena;
    o->d_x = at<uint64_t>(d, L.x);
    o->d_z = at<uint64_t>(d, L.z);
    o->d_ny = at<uint32_t>(d, L.ny);
    o->d_meta = at<fpk::StringMeta>(d, L.meta);
    o->d_cf128 = at<double>(d, L.cf128);
    o->d_cf64 = at<float>(d, L.cf64);
    o->d_scale = at<double>(d, L.scale);
    o->d_kappa_c = at<double>(d, L.kappa_c);
    o->d_kappa = at<double>(d, L.kappa);
    o->d_ev_chunks = at<fpl::EvChunk>(d, L.ev_chunks);
    o->d_ev_strs = at<fpl::EvStr>(d, L.ev_strs);
    FP_HIP(hipMemcpyAsync(o->d_arena, o->h_arena.data(), L.total, hipMemcpyHostToDevice, stream));
    FP_HIP(hipEventRecord(o->upload_done, stream));
    if (sync)
        FP_HIP(hipStreamSynchronize(stream));
    return FP_OK;
}

} // namespace

namespace fprt
{

fp_status check_ops(fp_hip_ops const *ops)
{
    if (!ops)
        return fail(FP_ERR_INVALID_ARGUMENT, "ops is NULL");
    return FP_OK;
}

fp_status check_shape(fp_hip_ops const *ops, int dtype, void const *a, void const *b, size_t dim, size_t n_states)
{
    fp_status st = check_ops(ops);
    if (st != FP_OK)
        return st;
    if (dtype != FP_C128 && dtype != FP_C64)
        return fail(FP_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    if (ops->n_strings == 0 || ops->n_qubits == 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "operator has no Pauli strings");
    size_t const want = size_t(1) << ops->n_qubits;
    if (dim != want)
        return fail(FP_ERR_INVALID_ARGUMENT, "states shape (%zu) must match the dimension of the operators (%zu)", dim,
                    want);
    if (n_states == 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "n_states must be positive");
    if (!a || !b)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL data pointer");
    uintptr_t const align = dtype == FP_C128 ? 16 : 8;
    if (((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & (align - 1)) != 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "device buffers must be aligned to the complex element size (%zu bytes)",
                    size_t(align));
    return FP_OK;
}

} // namespace fprt

namespace
{

bool ranges_overlap(void const *a, size_t na, void const *b, size_t nb)
{
    auto const pa = reinterpret_cast<uintptr_t>(a), pb = reinterpret_cast<uintptr_t>(b);
    return pa < pb + nb && pb < pa + na;
}

// An apply writes `out` while other threads still read `states`: the byte ranges
// of the two (dim x n_states) batches must be disjoint.
fp_status check_no_alias(int dtype, void const *states, void const *out, size_t dim, size_t n_states)
{
    size_t const bytes = dim * n_states * (dtype == FP_C128 ? 16 : 8);
    if (ranges_overlap(states, bytes, out, bytes))
        return fail(FP_ERR_INVALID_ARGUMENT, "apply cannot run in place (out overlaps states)");
    return FP_OK;
}

template <class T, int NC, int U, bool POW2, bool ACC>
void launch_string_apply(VecGeom const &g, void const *in, void *out, uint64_t x, uint64_t z, double cr, double ci,
                         hipStream_t st)
{
    typedef typename fpk::VecOf<T, NC>::type V;
    uint64_t const per_block = uint64_t(fpk::kBlock) * U;
    for (uint64_t v0 = 0; v0 < g.n_vec; v0 += kMaxLaunchBlocks * per_block)
    {
        uint64_t const rest = g.n_vec - v0;
        uint64_t const blocks = std::min<uint64_t>((rest + per_block - 1) / per_block, kMaxLaunchBlocks);
        hipLaunchKernelGGL((fpk::k_string_apply<T, NC, U, POW2, ACC>), dim3(uint32_t(blocks)), dim3(fpk::kBlock), 0,
                           st, static_cast<V const *>(in), static_cast<V *>(out), v0, g.n_vec, g.shift, g.vpr,
                           g.inv_vpr, x, z, T(cr), T(ci));
    }
}

template <class T, int NC>
void dispatch_string_apply(VecGeom const &g, void const *in, void *out, uint64_t x, uint64_t z, double cr, double ci,
                           bool acc, hipStream_t st)
{
    // one 16-byte vector per thread: at n = 24 with 16 states this moves 6.47 TB/s against
    // 5.87 TB/s with four vectors per thread (the same as a plain copy of each form)
    constexpr int U = 1;
    if (g.pow2)
    {
        if (acc)
            launch_string_apply<T, NC, U, true, true>(g, in, out, x, z, cr, ci, st);
        else
            launch_string_apply<T, NC, U, true, false>(g, in, out, x, z, cr, ci, st);
    }
    else
    {
        if (acc)
            launch_string_apply<T, NC, U, false, true>(g, in, out, x, z, cr, ci, st);
        else
            launch_string_apply<T, NC, U, false, false>(g, in, out, x, z, cr, ci, st);
    }
}

template <class T, int NC, int U, bool POW2, bool ACC>
void launch_op_apply(VecGeom const &g, void const *in, void *out, uint64_t const *xs, uint64_t const *zs, T const *cf,
                     uint32_t n_strings, hipStream_t st)
{
    typedef typename fpk::VecOf<T, NC>::type V;
    uint64_t const per_block = uint64_t(fpk::kBlock) * U;
    for (uint64_t v0 = 0; v0 < g.n_vec; v0 += kMaxLaunchBlocks * per_block)
    {
        uint64_t const rest = g.n_vec - v0;
        uint64_t const blocks = std::min<uint64_t>((rest + per_block - 1) / per_block, kMaxLaunchBlocks);
        hipLaunchKernelGGL((fpk::k_op_apply<T, NC, U, POW2, ACC>), dim3(uint32_t(blocks)), dim3(fpk::kBlock), 0, st,
                           static_cast<V const *>(in), static_cast<V *>(out), v0, g.n_vec, g.shift, g.vpr, g.inv_vpr,
                           xs, zs, cf, n_strings);
    }
}

template <class T, int NC>
void dispatch_op_apply(VecGeom const &g, void const *in, void *out, uint64_t const *xs, uint64_t const *zs,
                       T const *cf, uint32_t n_strings, bool acc, hipStream_t st)
{
    // fewer vectors per thread when the grid would not fill the chip
    bool const small = g.n_vec < uint64_t(fpk::kBlock) * 2048 * 2;
    if (small)
    {
        if (g.pow2)
            acc ? launch_op_apply<T, NC, 1, true, true>(g, in, out, xs, zs, cf, n_strings, st)
                : launch_op_apply<T, NC, 1, true, false>(g, in, out, xs, zs, cf, n_strings, st);
        else
            acc ? launch_op_apply<T, NC, 1, false, true>(g, in, out, xs, zs, cf, n_strings, st)
                : launch_op_apply<T, NC, 1, false, false>(g, in, out, xs, zs, cf, n_strings, st);
    }
    else
    {
        if (g.pow2)
            acc ? launch_op_apply<T, NC, 2, true, true>(g, in, out, xs, zs, cf, n_strings, st)
                : launch_op_apply<T, NC, 2, true, false>(g, in, out, xs, zs, cf, n_strings, st);
        else
            acc ? launch_op_apply<T, NC, 2, false, true>(g, in, out, xs, zs, cf, n_strings, st)
                : launch_op_apply<T, NC, 2, false, false>(g, in, out, xs, zs, cf, n_strings, st);
    }
}

// ---- expectation launch shape ----
struct EvGeom
{
    uint32_t ct, rt;  // column threads, row threads
    uint32_t gx, gy, gz;
    uint64_t n_pairs, ppb;
    uint32_t spb;
};

constexpr int kEvUnroll = 4;

EvGeom ev_geom(VecGeom const &vg, size_t dim, size_t n_strings, bool per_string, size_t n_states)
{
    EvGeom e{};
    e.ct = uint32_t(std::min<uint64_t>(vg.vpr, fpk::kBlock));
    e.rt = fpk::kBlock / e.ct;
    e.gy = uint32_t((vg.vpr + e.ct - 1) / e.ct);
    e.n_pairs = dim / 2;
    uint64_t const min_ppb = uint64_t(e.rt) * kEvUnroll * 8;
    uint64_t gx = (e.n_pairs + min_ppb - 1) / min_ppb;
    uint64_t cap = std::max<uint64_t>(1, 2048 / e.gy);
    if (per_string)
    {
        // keep the (strings x gx x states) FP64 partials modest
        uint64_t const budget = uint64_t(256) << 20;
        uint64_t const per_gx = uint64_t(n_strings) * n_states * sizeof(double);
        cap = std::min<uint64_t>(cap, std::max<uint64_t>(1, budget / std::max<uint64_t>(per_gx, 1)));
    }
    gx = std::max<uint64_t>(1, std::min(gx, cap));
    e.ppb = (e.n_pairs + gx - 1) / gx;
    e.gx = uint32_t((e.n_pairs + e.ppb - 1) / e.ppb);
    uint64_t const blocks_xy = uint64_t(e.gx) * e.gy;
    uint64_t gz = std::min<uint64_t>(n_strings, std::max<uint64_t>(1, (4096 + blocks_xy - 1) / blocks_xy));
    e.spb = uint32_t((n_strings + gz - 1) / gz);
    e.gz = uint32_t((n_strings + e.spb - 1) / e.spb);
    return e;
}

template <class T, int NC>
void launch_op_expect(VecGeom const &vg, EvGeom const &e, void const *states, fpk::StringMeta const *meta,
                      double const *scale, uint32_t n_strings, uint64_t n_states, fpk::Dbl2 *partial, hipStream_t st)
{
    typedef typename fpk::VecOf<T, NC>::type V;
    hipLaunchKernelGGL((fpk::k_op_expect<T, NC, kEvUnroll>), dim3(e.gx, e.gy, e.gz), dim3(fpk::kBlock), 0, st,
                       static_cast<V const *>(states), vg.vpr, e.ct, e.rt, e.n_pairs, e.ppb, meta, scale, n_strings,
                       e.spb, n_states, partial);
}

template <class T, int NC>
void launch_strings_expect(VecGeom const &vg, EvGeom const &e, void const *states, fpk::StringMeta const *meta,
                           uint32_t n_strings, uint64_t n_states, double *partial, hipStream_t st)
{
    typedef typename fpk::VecOf<T, NC>::type V;
    hipLaunchKernelGGL((fpk::k_strings_expect<T, NC, kEvUnroll>), dim3(e.gx, e.gy, e.gz), dim3(fpk::kBlock), 0, st,
                       static_cast<V const *>(states), vg.vpr, e.ct, e.rt, e.n_pairs, e.ppb, meta, n_strings, e.spb,
                       n_states, partial);
}

// Shrinks an expectation grid to at most `max_parts` partials per state: fewer
// string blocks first, then fewer row blocks.  Every (pair, string) stays covered.
void fit_ev_geom(EvGeom &e, size_t count, uint64_t max_parts)
{
    if (uint64_t(e.gx) * e.gz <= max_parts)
        return;
    uint64_t gz = std::max<uint64_t>(1, max_parts / e.gx);
    e.spb = uint32_t((count + gz - 1) / gz);
    e.gz = uint32_t((count + e.spb - 1) / e.spb);
    if (uint64_t(e.gx) * e.gz > max_parts)
    {
        uint64_t const gx = std::max<uint64_t>(1, max_parts / e.gz);
        e.ppb = (e.n_pairs + gx - 1) / gx;
        e.gx = uint32_t((e.n_pairs + e.ppb - 1) / e.ppb);
    }
}

} // namespace

namespace fprt
{

// Expectation values of strings [s_begin, s_begin + count) combined with
// per-string complex scales, reduced to out[t] (+)= post * sum.
fp_status run_op_expect(fp_hip_ops const *ops, int dtype, void const *states, void *out, size_t dim, size_t n_states,
                        size_t s_begin, size_t count, double const *scale_table, double post_re, double post_im,
                        int accumulate, hipStream_t st, Workspace const *ws)
{
    VecGeom const vg = vec_geom(dtype, dim, n_states, states, states);
    EvGeom e = ev_geom(vg, dim, count, false, n_states);
    void *scratch = nullptr;
    if (ws)
    {
        uint64_t const per_part = uint64_t(n_states) * sizeof(fpk::Dbl2);
        if (per_part > ws->bytes)
            return fail(FP_ERR_INVALID_ARGUMENT,
                        "%zu states need %llu bytes of expectation partials, more than the %zu-byte workspace",
                        n_states, static_cast<unsigned long long>(per_part), ws->bytes);
        fit_ev_geom(e, count, ws->bytes / per_part);
        scratch = ws->base;
    }
    uint64_t const n_parts = uint64_t(e.gx) * e.gz;
    if (!ws)
    {
        fp_status const s = scratch_get(st, n_parts * n_states * sizeof(fpk::Dbl2), &scratch);
        if (s != FP_OK)
            return s;
    }
    auto *partial = static_cast<fpk::Dbl2 *>(scratch);
    fpk::StringMeta const *meta = ops->d_meta + s_begin;
    double const *scale = scale_table + 2 * s_begin;
    if (dtype == FP_C128)
        launch_op_expect<double, 1>(vg, e, states, meta, scale, uint32_t(count), n_states, partial, st);
    else if (vg.nc == 2)
        launch_op_expect<float, 2>(vg, e, states, meta, scale, uint32_t(count), n_states, partial, st);
    else
        launch_op_expect<float, 1>(vg, e, states, meta, scale, uint32_t(count), n_states, partial, st);
    FP_LAUNCHED();
    if (dtype == FP_C128)
        hipLaunchKernelGGL(fpk::k_reduce_partials<double>, dim3(uint32_t(n_states)), dim3(fpk::kBlock), 0, st,
                           partial, n_parts, uint64_t(n_states), static_cast<double *>(out), accumulate, post_re,
                           post_im);
    else
        hipLaunchKernelGGL(fpk::k_reduce_partials<float>, dim3(uint32_t(n_states)), dim3(fpk::kBlock), 0, st, partial,
                           n_parts, uint64_t(n_states), static_cast<float *>(out), accumulate, post_re, post_im);
    FP_LAUNCHED();
    return FP_OK;
}

} // namespace fprt

namespace
{

// The state column fits in a workgroup's LDS: the x-grouped LDS kernel applies.
bool lds_eligible(fp_hip_ops const *ops, int dtype, size_t dim)
{
    size_t const esz = dtype == FP_C128 ? 16 : 8;
    size_t const extra = ops->n_ev_wht ? fpl::kWhtLdsBytes : 0;
    return ops->n_qubits >= 6 && dim * esz + extra <= column_lds_bytes() && ops->n_ev_chunks > 0;
}

// Scratch bytes of stage 1 (run_strings_expect): the transposed states for the
// LDS kernel, otherwise the streaming kernel's partials.
size_t strings_expect_partial_bytes(fp_hip_ops const *ops, int dtype, size_t dim, size_t n_states, void const *states)
{
    if (lds_eligible(ops, dtype, dim))
        return dim * n_states * (dtype == FP_C128 ? 16 : 8);
    if (fprt::ev_coset_eligible(ops, dtype, dim, n_states))
        return fprt::ev_coset_scratch_bytes(ops, dtype, dim, n_states);
    VecGeom const vg = vec_geom(dtype, dim, n_states, states, states);
    EvGeom const e = ev_geom(vg, dim, ops->n_strings, true, n_states);
    return size_t(e.gx) * ops->n_strings * n_states * sizeof(double);
}

template <class T>
fp_status launch_ev_lds(fp_hip_ops const *ops, void const *states, void *ev, size_t dim, size_t n_states,
                        void *transposed, hipStream_t st)
{
    typedef typename fpk::VecOf<T, 1>::type V;
    dim3 const tg(uint32_t((n_states + 31) / 32), uint32_t((dim + 31) / 32));
    hipLaunchKernelGGL(fpl::k_transpose<V>, tg, dim3(256), 0, st, static_cast<V const *>(states),
                       static_cast<V *>(transposed), uint64_t(dim), uint64_t(n_states));
    FP_LAUNCHED();
    size_t const lds = dim * sizeof(V) + (ops->n_ev_wht ? fpl::kWhtLdsBytes : 0);
    static bool attr_set[2] = {false, false};
    int const which = sizeof(T) == 8 ? 1 : 0;
    if (!attr_set[which])
    {
        FP_HIP(hipFuncSetAttribute(reinterpret_cast<void const *>(&fpl::k_ev_lds<T>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(column_lds_bytes())));
        attr_set[which] = true;
    }
    uint32_t const gy = ev_matrix_split(n_states, 1, ops->n_ev_chunks);
    hipLaunchKernelGGL(fpl::k_ev_lds<T>, dim3(uint32_t(n_states), gy), dim3(fpl::kThreads), lds, st,
                       static_cast<V const *>(transposed), uint32_t(dim), ops->d_ev_chunks, ops->n_ev_chunks,
                       ops->n_ev_wht, ops->d_ev_strs, uint64_t(n_states), static_cast<T *>(ev));
    FP_LAUNCHED();
    return FP_OK;
}

// Real (n_strings x n_states) matrix of per-string expectation values.
// `partial` must hold strings_expect_partial_bytes(); NULL = use the stream's scratch.
fp_status run_strings_expect(fp_hip_ops const *ops, int dtype, void const *states, void *ev, size_t dim,
                             size_t n_states, hipStream_t st, double *partial)
{
    size_t const N = ops->n_strings;
    if (!partial)
    {
        void *scratch = nullptr;
        fp_status s = scratch_get(st, strings_expect_partial_bytes(ops, dtype, dim, n_states, states), &scratch);
        if (s != FP_OK)
            return s;
        partial = static_cast<double *>(scratch);
    }
    if (lds_eligible(ops, dtype, dim))
    {
        if (dtype == FP_C128)
            return launch_ev_lds<double>(ops, states, ev, dim, n_states, partial, st);
        return launch_ev_lds<float>(ops, states, ev, dim, n_states, partial, st);
    }
    if (fprt::ev_coset_eligible(ops, dtype, dim, n_states))
        return fprt::run_ev_coset(ops, dtype, states, ev, dim, n_states, partial, st);
    VecGeom const vg = vec_geom(dtype, dim, n_states, states, states);
    EvGeom const e = ev_geom(vg, dim, N, true, n_states);
    if (dtype == FP_C128)
        launch_strings_expect<double, 1>(vg, e, states, ops->d_meta, uint32_t(N), n_states, partial, st);
    else if (vg.nc == 2)
        launch_strings_expect<float, 2>(vg, e, states, ops->d_meta, uint32_t(N), n_states, partial, st);
    else
        launch_strings_expect<float, 1>(vg, e, states, ops->d_meta, uint32_t(N), n_states, partial, st);
    FP_LAUNCHED();
    uint64_t const total = uint64_t(N) * n_states;
    uint32_t const blocks = uint32_t((total + fpk::kBlock - 1) / fpk::kBlock);
    if (dtype == FP_C128)
        hipLaunchKernelGGL(fpk::k_reduce_string_partials<double>, dim3(blocks), dim3(fpk::kBlock), 0, st, partial,
                           e.gx, uint64_t(n_states), uint64_t(N), ops->d_kappa, static_cast<double *>(ev));
    else
        hipLaunchKernelGGL(fpk::k_reduce_string_partials<float>, dim3(blocks), dim3(fpk::kBlock), 0, st, partial,
                           e.gx, uint64_t(n_states), uint64_t(N), ops->d_kappa, static_cast<float *>(ev));
    FP_LAUNCHED();
    return FP_OK;
}

size_t round256(size_t b)
{
    return (b + 255) / 256 * 256;
}

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C"
{

const char *fp_hip_version(void)
{
    return "fast_pauli_hip 0.1 (gfx950)";
}

fp_status fp_hip_device_count(int *count)
{
    if (!count)
        return fail(FP_ERR_INVALID_ARGUMENT, "count is NULL");
    int n = 0;
    hipError_t const e = hipGetDeviceCount(&n);
    if (e != hipSuccess)
    {
        *count = 0;
        return fail(FP_ERR_NO_DEVICE, "hipGetDeviceCount failed: %s", hipGetErrorString(e));
    }
    *count = n;
    return FP_OK;
}

fp_status fp_hip_init(int device)
{
    int n = 0;
    hipError_t const e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(FP_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= n)
        return fail(FP_ERR_INVALID_ARGUMENT, "device %d out of range (%d devices)", device, n);
    FP_HIP(hipSetDevice(device));
    FP_HIP(hipFree(nullptr)); // force context creation
    return noalloc_setup();
}

fp_status fp_hip_malloc(void **ptr, size_t bytes)
{
    if (!ptr)
        return fail(FP_ERR_INVALID_ARGUMENT, "ptr is NULL");
    *ptr = nullptr;
    FP_HIP(hipMalloc(ptr, std::max<size_t>(bytes, 1)));
    return FP_OK;
}

fp_status fp_hip_free(void *ptr)
{
    if (ptr)
        FP_HIP(hipFree(ptr));
    return FP_OK;
}

fp_status fp_hip_memcpy(void *dst, void const *src, size_t bytes, int kind, void *stream)
{
    if (bytes == 0)
        return FP_OK;
    if (!dst || !src)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL pointer in memcpy");
    hipMemcpyKind k = kind == FP_H2D ? hipMemcpyHostToDevice
                      : kind == FP_D2H ? hipMemcpyDeviceToHost
                                       : hipMemcpyDeviceToDevice;
    FP_HIP(hipMemcpyAsync(dst, src, bytes, k, static_cast<hipStream_t>(stream)));
    if (kind != FP_D2D)
        FP_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return FP_OK;
}

fp_status fp_hip_memset(void *ptr, int value, size_t bytes, void *stream)
{
    if (bytes == 0)
        return FP_OK;
    FP_HIP(hipMemsetAsync(ptr, value, bytes, static_cast<hipStream_t>(stream)));
    return FP_OK;
}

fp_status fp_hip_stream_create(void **stream)
{
    if (!stream)
        return fail(FP_ERR_INVALID_ARGUMENT, "stream is NULL");
    hipStream_t s = nullptr;
    FP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return FP_OK;
}

fp_status fp_hip_stream_destroy(void *stream)
{
    if (!stream)
        return FP_OK;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    FP_HIP(hipStreamSynchronize(s));
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        for (auto it = g_scratch.begin(); it != g_scratch.end();)
        {
            if (it->first.second == s)
            {
                (void)hipFree(it->second.ptr);
                it = g_scratch.erase(it);
            }
            else
                ++it;
        }
    }
    FP_HIP(hipStreamDestroy(s));
    return FP_OK;
}

fp_status fp_hip_stream_synchronize(void *stream)
{
    FP_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return FP_OK;
}

fp_status fp_hip_device_synchronize(void)
{
    FP_HIP(hipDeviceSynchronize());
    return FP_OK;
}

fp_status fp_hip_is_device_pointer(void const *ptr, int *is_device)
{
    if (!is_device)
        return fail(FP_ERR_INVALID_ARGUMENT, "is_device is NULL");
    *is_device = 0;
    if (!ptr)
        return FP_OK;
    hipPointerAttribute_t attr;
    hipError_t const e = hipPointerGetAttributes(&attr, ptr);
    if (e != hipSuccess)
    {
        (void)hipGetLastError(); // plain host memory: clear the sticky error
        return FP_OK;
    }
    *is_device = attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
    return FP_OK;
}

fp_status fp_hip_mem_info(size_t *free_bytes, size_t *total_bytes)
{
    size_t f = 0, t = 0;
    FP_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes)
        *free_bytes = f;
    if (total_bytes)
        *total_bytes = t;
    return FP_OK;
}

fp_status fp_hip_event_create(void **event)
{
    if (!event)
        return fail(FP_ERR_INVALID_ARGUMENT, "event is NULL");
    hipEvent_t e = nullptr;
    FP_HIP(hipEventCreate(&e));
    *event = e;
    return FP_OK;
}

fp_status fp_hip_event_destroy(void *event)
{
    if (event)
        FP_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return FP_OK;
}

fp_status fp_hip_event_record(void *event, void *stream)
{
    if (!event)
        return fail(FP_ERR_INVALID_ARGUMENT, "event is NULL");
    FP_HIP(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
    return FP_OK;
}

fp_status fp_hip_event_synchronize(void *event)
{
    if (!event)
        return fail(FP_ERR_INVALID_ARGUMENT, "event is NULL");
    FP_HIP(hipEventSynchronize(static_cast<hipEvent_t>(event)));
    return FP_OK;
}

fp_status fp_hip_event_query(void *event, int *done)
{
    if (!event || !done)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL event or out-pointer");
    hipError_t const e = hipEventQuery(static_cast<hipEvent_t>(event));
    if (e == hipSuccess)
    {
        *done = 1;
        return FP_OK;
    }
    if (e == hipErrorNotReady)
    {
        (void)hipGetLastError();
        *done = 0;
        return FP_OK;
    }
    return fail(FP_ERR_RUNTIME, "hipEventQuery failed: %s", hipGetErrorString(e));
}

fp_status fp_hip_event_elapsed_ms(void *start, void *stop, float *ms)
{
    if (!start || !stop || !ms)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL event or out-pointer");
    FP_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return FP_OK;
}

fp_status fp_hip_fill_states(int dtype, void *states, size_t dim, size_t n_states, uint64_t seed, int normalise,
                             void *stream)
{
    if (!states || dim == 0 || n_states == 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "empty or NULL states");
    if (dtype != FP_C128 && dtype != FP_C64)
        return fail(FP_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    hipStream_t const st = static_cast<hipStream_t>(stream);
    StreamGuard guard(st);
    uint64_t const n_elem = uint64_t(dim) * n_states;
    uint64_t const n_real = 2 * n_elem;
    uint32_t const blocks = uint32_t((n_real + fpk::kBlock - 1) / fpk::kBlock);
    if (dtype == FP_C128)
        hipLaunchKernelGGL(fpk::k_fill_normal<double>, dim3(blocks), dim3(fpk::kBlock), 0, st,
                           static_cast<double *>(states), n_real, seed);
    else
        hipLaunchKernelGGL(fpk::k_fill_normal<float>, dim3(blocks), dim3(fpk::kBlock), 0, st,
                           static_cast<float *>(states), n_real, seed);
    FP_LAUNCHED();
    if (!normalise)
        return FP_OK;
    uint64_t const rows_per_block = 1024;
    uint64_t const n_parts = (dim + rows_per_block - 1) / rows_per_block;
    void *scratch = nullptr;
    fp_status s = scratch_get(st, (n_parts + 1) * n_states * sizeof(double), &scratch);
    if (s != FP_OK)
        return s;
    double *partial = static_cast<double *>(scratch);
    double *inv_norm = partial + n_parts * n_states;
    dim3 const g1(uint32_t(n_parts), uint32_t((n_states + fpk::kBlock - 1) / fpk::kBlock));
    if (dtype == FP_C128)
        hipLaunchKernelGGL(fpk::k_colnorm_partial<double>, g1, dim3(fpk::kBlock), 0, st,
                           static_cast<double const *>(states), uint64_t(dim), uint64_t(n_states), rows_per_block,
                           partial);
    else
        hipLaunchKernelGGL(fpk::k_colnorm_partial<float>, g1, dim3(fpk::kBlock), 0, st,
                           static_cast<float const *>(states), uint64_t(dim), uint64_t(n_states), rows_per_block,
                           partial);
    FP_LAUNCHED();
    hipLaunchKernelGGL(fpk::k_colnorm_final<double>, dim3(uint32_t((n_states + fpk::kBlock - 1) / fpk::kBlock)),
                       dim3(fpk::kBlock), 0, st, partial, n_parts, uint64_t(n_states), inv_norm);
    FP_LAUNCHED();
    uint32_t const b2 = uint32_t((n_elem + fpk::kBlock - 1) / fpk::kBlock);
    if (dtype == FP_C128)
        hipLaunchKernelGGL(fpk::k_scale_columns<double>, dim3(b2), dim3(fpk::kBlock), 0, st,
                           static_cast<double *>(states), n_elem, uint64_t(n_states), inv_norm);
    else
        hipLaunchKernelGGL(fpk::k_scale_columns<float>, dim3(b2), dim3(fpk::kBlock), 0, st,
                           static_cast<float *>(states), n_elem, uint64_t(n_states), inv_norm);
    FP_LAUNCHED();
    return FP_OK;
}

fp_status fp_hip_copy(void *dst, void const *src, size_t bytes, void *stream)
{
    if (bytes == 0)
        return FP_OK;
    if (!dst || !src || (bytes & 15u) || ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15u))
        return fail(FP_ERR_INVALID_ARGUMENT, "copy needs non-NULL, 16-byte aligned buffers and a multiple of 16 bytes");
    hipStream_t const st = static_cast<hipStream_t>(stream);
    typedef fpk::VecOf<double, 1>::type V;
    uint64_t const n_vec = bytes / 16;
    constexpr int U = 1; // one vector per thread: the fastest form of the sizes measured (1, 2, 4)
    uint64_t const per_block = uint64_t(fpk::kBlock) * U;
    for (uint64_t v0 = 0; v0 < n_vec; v0 += kMaxLaunchBlocks * per_block)
    {
        uint64_t const blocks = std::min<uint64_t>((n_vec - v0 + per_block - 1) / per_block, kMaxLaunchBlocks);
        hipLaunchKernelGGL(fpk::k_copy<U>, dim3(uint32_t(blocks)), dim3(fpk::kBlock), 0, st,
                           static_cast<V const *>(src), static_cast<V *>(dst), v0, n_vec);
        FP_LAUNCHED();
    }
    return FP_OK;
}

fp_status fp_hip_convert(int src_dtype, void const *src, int dst_dtype, void *dst, size_t n_elements, void *stream)
{
    if ((src_dtype != FP_C128 && src_dtype != FP_C64) || (dst_dtype != FP_C128 && dst_dtype != FP_C64))
        return fail(FP_ERR_INVALID_ARGUMENT, "unknown dtype");
    if (n_elements == 0)
        return FP_OK;
    if (!src || !dst)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL pointer in convert");
    hipStream_t const st = static_cast<hipStream_t>(stream);
    uint64_t const n_real = 2 * uint64_t(n_elements);
    if (src_dtype == dst_dtype)
    {
        FP_HIP(hipMemcpyAsync(dst, src, n_real * (src_dtype == FP_C128 ? 8 : 4), hipMemcpyDeviceToDevice, st));
        return FP_OK;
    }
    if (src == dst)
        return fail(FP_ERR_INVALID_ARGUMENT, "convert cannot run in place between dtypes");
    for (uint64_t b = 0; b < n_real; b += kMaxLaunchBlocks * fpk::kBlock)
    {
        uint64_t const cnt = std::min<uint64_t>(n_real - b, kMaxLaunchBlocks * fpk::kBlock);
        uint32_t const blocks = uint32_t((cnt + fpk::kBlock - 1) / fpk::kBlock);
        if (src_dtype == FP_C128)
            hipLaunchKernelGGL((fpk::k_convert<double, float>), dim3(blocks), dim3(fpk::kBlock), 0, st,
                               static_cast<double const *>(src) + b, static_cast<float *>(dst) + b, cnt);
        else
            hipLaunchKernelGGL((fpk::k_convert<float, double>), dim3(blocks), dim3(fpk::kBlock), 0, st,
                               static_cast<float const *>(src) + b, static_cast<double *>(dst) + b, cnt);
        FP_LAUNCHED();
    }
    return FP_OK;
}

fp_status fp_hip_ops_create(fp_hip_ops **out, size_t n_qubits, size_t n_strings, uint8_t const *codes,
                            double const *coeffs)
{
    if (!out)
        return fail(FP_ERR_INVALID_ARGUMENT, "ops out-pointer is NULL");
    *out = nullptr;
    auto *o = new (std::nothrow) fp_hip_ops();
    if (!o)
        return fail(FP_ERR_OUT_OF_MEMORY, "host allocation failed");
    (void)hipGetDevice(&o->device);
    fp_status const st = ops_assign(o, n_qubits, n_strings, codes, coeffs, nullptr, true);
    if (st != FP_OK)
    {
        free_device(o);
        if (o->upload_done)
            (void)hipEventDestroy(o->upload_done);
        delete o;
        return st;
    }
    *out = o;
    return FP_OK;
}

fp_status fp_hip_ops_assign(fp_hip_ops *ops, size_t n_qubits, size_t n_strings, uint8_t const *codes,
                            double const *coeffs, void *stream)
{
    fp_status st = check_ops(ops);
    if (st != FP_OK)
        return st;
    // a SummedPauliOp matrix belongs to the previous strings
    for (void **p : {reinterpret_cast<void **>(&ops->d_wr64), reinterpret_cast<void **>(&ops->d_wi64),
                     reinterpret_cast<void **>(&ops->d_wr32), reinterpret_cast<void **>(&ops->d_wi32),
                     reinterpret_cast<void **>(&ops->d_rowsum128), reinterpret_cast<void **>(&ops->d_rowsum64),
                     reinterpret_cast<void **>(&ops->d_h128), reinterpret_cast<void **>(&ops->d_h64)})
        if (*p)
        {
            FP_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
            FP_HIP(hipFree(*p));
            *p = nullptr;
        }
    ops->matrix.clear();
    return ops_assign(ops, n_qubits, n_strings, codes, coeffs, static_cast<hipStream_t>(stream), false);
}

fp_status fp_hip_ops_destroy(fp_hip_ops *ops)
{
    if (!ops)
        return FP_OK;
    (void)hipDeviceSynchronize();
    free_device(ops);
    if (ops->upload_done)
        (void)hipEventDestroy(ops->upload_done);
    delete ops;
    return FP_OK;
}

fp_status fp_hip_ops_set_coeffs(fp_hip_ops *ops, double const *coeffs)
{
    fp_status st = check_ops(ops);
    if (st != FP_OK)
        return st;
    if (!coeffs)
        return fail(FP_ERR_INVALID_ARGUMENT, "coeffs is NULL");
    (void)hipDeviceSynchronize();
    fprt::invalidate_tiles(ops); // the tiled expectation tables hold the scaled coefficients
    ++ops->generation;
    std::memcpy(ops->coeffs.data(), coeffs, 2 * ops->n_strings * sizeof(double));
    ArenaLayout const L = arena_layout(ops->n_strings);
    fill_coeff_tables(ops, L);
    FP_HIP(hipMemcpy(ops->d_arena, ops->h_arena.data(), L.total, hipMemcpyHostToDevice));
    return FP_OK;
}

fp_status fp_hip_ops_set_matrix(fp_hip_ops *ops, size_t n_operators, double const *coeffs_nk)
{
    fp_status st = check_ops(ops);
    if (st != FP_OK)
        return st;
    if (n_operators == 0 || !coeffs_nk)
        return fail(FP_ERR_INVALID_ARGUMENT, "empty coefficient matrix");
    (void)hipDeviceSynchronize();
    ++ops->generation;
    fprt::invalidate_tiles(ops); // the tiled SummedPauliOp::apply tables hold the row sums
    size_t const N = ops->n_strings, K = n_operators;
    ops->n_operators = K;
    ops->matrix.assign(coeffs_nk, coeffs_nk + 2 * N * K);
    // transposed real / imaginary planes (K x N) for the contraction, and the
    // per-string row sums used by SummedPauliOp::apply
    std::vector<double> wr(K * N), wi(K * N), rowsum(2 * N, 0.0), rowsum_cf(2 * N);
    std::vector<float> wr32(K * N), wi32(K * N), rowsum_cf32(2 * N);
    bool imag_zero = true;
    for (size_t j = 0; j < N; ++j)
    {
        double sr = 0.0, si = 0.0;
        for (size_t k = 0; k < K; ++k)
        {
            double const re = coeffs_nk[2 * (j * K + k)], im = coeffs_nk[2 * (j * K + k) + 1];
            wr[k * N + j] = re;
            wi[k * N + j] = im;
            wr32[k * N + j] = float(re);
            wi32[k * N + j] = float(im);
            imag_zero = imag_zero && im == 0.0;
            sr += re;
            si += im;
        }
        double ur, ui;
        minus_i_pow(ops->ny[j], ur, ui);
        mul_axis_unit(sr, si, ur, ui, rowsum_cf[2 * j], rowsum_cf[2 * j + 1]);
        // complex64: the reference sums the float coefficients in float
        float fr = 0.f, fi = 0.f;
        for (size_t k = 0; k < K; ++k)
        {
            fr += float(coeffs_nk[2 * (j * K + k)]);
            fi += float(coeffs_nk[2 * (j * K + k) + 1]);
        }
        double gr, gi;
        mul_axis_unit(double(fr), double(fi), ur, ui, gr, gi);
        rowsum_cf32[2 * j] = float(gr);
        rowsum_cf32[2 * j + 1] = float(gi);
    }
    ops->matrix_imag_zero = imag_zero;
    std::vector<double> h128(coeffs_nk, coeffs_nk + 2 * N * K);
    std::vector<float> h64(2 * N * K);
    for (size_t i = 0; i < 2 * N * K; ++i)
        h64[i] = float(coeffs_nk[i]);
    for (void **p : {reinterpret_cast<void **>(&ops->d_wr64), reinterpret_cast<void **>(&ops->d_wi64),
                     reinterpret_cast<void **>(&ops->d_wr32), reinterpret_cast<void **>(&ops->d_wi32),
                     reinterpret_cast<void **>(&ops->d_rowsum128), reinterpret_cast<void **>(&ops->d_rowsum64),
                     reinterpret_cast<void **>(&ops->d_h128), reinterpret_cast<void **>(&ops->d_h64)})
    {
        if (*p)
        {
            (void)hipFree(*p);
            *p = nullptr;
        }
    }
    ops->rowsum128 = rowsum_cf;
    ops->rowsum64 = rowsum_cf32;
    if ((st = upload(&ops->d_wr64, wr)) != FP_OK || (st = upload(&ops->d_wi64, wi)) != FP_OK ||
        (st = upload(&ops->d_wr32, wr32)) != FP_OK || (st = upload(&ops->d_wi32, wi32)) != FP_OK ||
        (st = upload(&ops->d_rowsum128, rowsum_cf)) != FP_OK || (st = upload(&ops->d_rowsum64, rowsum_cf32)) != FP_OK ||
        (st = upload(&ops->d_h128, h128)) != FP_OK || (st = upload(&ops->d_h64, h64)) != FP_OK)
        return st;
    return FP_OK;
}

fp_status fp_hip_ops_info(fp_hip_ops const *ops, size_t *n_qubits, size_t *n_strings, size_t *n_operators)
{
    fp_status st = check_ops(ops);
    if (st != FP_OK)
        return st;
    if (n_qubits)
        *n_qubits = ops->n_qubits;
    if (n_strings)
        *n_strings = ops->n_strings;
    if (n_operators)
        *n_operators = ops->n_operators;
    return FP_OK;
}

fp_status fp_hip_string_apply_batch(fp_hip_ops const *ops, size_t string_index, int dtype, void const *states,
                                    void *out, size_t dim, size_t n_states, double c_re, double c_im, int accumulate,
                                    void *stream)
{
    fp_status st = check_shape(ops, dtype, states, out, dim, n_states);
    if (st != FP_OK)
        return st;
    if (string_index >= ops->n_strings)
        return fail(FP_ERR_INVALID_ARGUMENT, "string index %zu out of range (%zu strings)", string_index,
                    ops->n_strings);
    if ((st = check_no_alias(dtype, states, out, dim, n_states)) != FP_OK)
        return st;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    VecGeom const g = vec_geom(dtype, dim, n_states, states, out);
    // fold (-i)^{nY} into the coefficient; exact
    double ur, ui, cr, ci;
    minus_i_pow(ops->ny[string_index], ur, ui);
    if (dtype == FP_C64)
    {
        c_re = double(float(c_re));
        c_im = double(float(c_im));
    }
    mul_axis_unit(c_re, c_im, ur, ui, cr, ci);
    uint64_t const x = ops->x[string_index], z = ops->z[string_index];
    if (dtype == FP_C128)
        dispatch_string_apply<double, 1>(g, states, out, x, z, cr, ci, accumulate != 0, s);
    else if (g.nc == 2)
        dispatch_string_apply<float, 2>(g, states, out, x, z, cr, ci, accumulate != 0, s);
    else
        dispatch_string_apply<float, 1>(g, states, out, x, z, cr, ci, accumulate != 0, s);
    FP_LAUNCHED();
    return FP_OK;
}

fp_status fp_hip_string_expectation(fp_hip_ops const *ops, size_t string_index, int dtype, void const *states,
                                    void *out, size_t dim, size_t n_states, double c_re, double c_im, int accumulate,
                                    void *stream)
{
    fp_status st = check_shape(ops, dtype, states, out, dim, n_states);
    if (st != FP_OK)
        return st;
    if (string_index >= ops->n_strings)
        return fail(FP_ERR_INVALID_ARGUMENT, "string index %zu out of range (%zu strings)", string_index,
                    ops->n_strings);
    StreamGuard guard(static_cast<hipStream_t>(stream));
    return run_op_expect(ops, dtype, states, out, dim, n_states, string_index, 1, ops->d_kappa_c, c_re, c_im,
                         accumulate, static_cast<hipStream_t>(stream));
}

fp_status fp_hip_op_apply(fp_hip_ops const *ops, int dtype, void const *states, void *out, size_t dim,
                          size_t n_states, int accumulate, void *stream)
{
    fp_status st = check_shape(ops, dtype, states, out, dim, n_states);
    if (st != FP_OK)
        return st;
    if ((st = check_no_alias(dtype, states, out, dim, n_states)) != FP_OK)
        return st;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    bool tiled = false;
    if ((st = fprt::choose_op_apply_tiled(ops, dtype, n_states, &tiled)) != FP_OK)
        return st;
    if (tiled)
        return fprt::run_op_apply_tiled(ops, dtype, states, out, dim, n_states, accumulate, s);
    VecGeom const g = vec_geom(dtype, dim, n_states, states, out);
    uint32_t const N = uint32_t(ops->n_strings);
    if (dtype == FP_C128)
        dispatch_op_apply<double, 1>(g, states, out, ops->d_x, ops->d_z, ops->d_cf128, N, accumulate != 0, s);
    else if (g.nc == 2)
        dispatch_op_apply<float, 2>(g, states, out, ops->d_x, ops->d_z, ops->d_cf64, N, accumulate != 0, s);
    else
        dispatch_op_apply<float, 1>(g, states, out, ops->d_x, ops->d_z, ops->d_cf64, N, accumulate != 0, s);
    FP_LAUNCHED();
    return FP_OK;
}

fp_status fp_hip_op_expectation(fp_hip_ops const *ops, int dtype, void const *states, void *out, size_t dim,
                                size_t n_states, int accumulate, void *stream)
{
    fp_status st = check_shape(ops, dtype, states, out, dim, n_states);
    if (st != FP_OK)
        return st;
    StreamGuard guard(static_cast<hipStream_t>(stream));
    bool tiled = false;
    if ((st = fprt::choose_op_expect_tiled(ops, dtype, n_states, &tiled)) != FP_OK)
        return st;
    if (tiled)
        return fprt::run_op_expect_tiled(ops, dtype, states, out, dim, n_states, accumulate,
                                         static_cast<hipStream_t>(stream));
    return run_op_expect(ops, dtype, states, out, dim, n_states, 0, ops->n_strings, ops->d_scale, 1.0, 0.0,
                         accumulate, static_cast<hipStream_t>(stream));
}

fp_status fp_hip_strings_expectation_matrix(fp_hip_ops const *ops, int dtype, void const *states, void *ev,
                                            size_t dim, size_t n_states, void *stream)
{
    fp_status st = check_shape(ops, dtype, states, ev, dim, n_states);
    if (st != FP_OK)
        return st;
    StreamGuard guard(static_cast<hipStream_t>(stream));
    return run_strings_expect(ops, dtype, states, ev, dim, n_states, static_cast<hipStream_t>(stream), nullptr);
}

fp_status fp_hip_summed_expectation(fp_hip_ops const *ops, int dtype, int contraction, void const *states, void *out,
                                    size_t dim, size_t n_states, int accumulate, void *stream)
{
    fp_status st = check_shape(ops, dtype, states, out, dim, n_states);
    if (st != FP_OK)
        return st;
    if (ops->n_operators == 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "SummedPauliOp has no coefficient matrix");
    if (contraction == FP_CONTRACT_AUTO)
        contraction = dtype == FP_C128 ? FP_CONTRACT_FP64 : FP_CONTRACT_FP32;
    if (contraction < FP_CONTRACT_FP64 || contraction > FP_CONTRACT_BF16X3)
        return fail(FP_ERR_INVALID_ARGUMENT, "unknown contraction mode %d", contraction);
    hipStream_t const s = static_cast<hipStream_t>(stream);
    StreamGuard guard(s);
    size_t const N = ops->n_strings, K = ops->n_operators, S = n_states;
    bool const fp64 = contraction == FP_CONTRACT_FP64;
    size_t const esz = fp64 ? sizeof(double) : sizeof(float);
    size_t const nat_esz = dtype == FP_C128 ? sizeof(double) : sizeof(float);
    bool const convert = esz != nat_esz;
    // scratch: [ev at the states' precision | ev converted (if needed) | C (planes x K x S) | stage-1 partials]
    size_t const ev_nat_bytes = round256(N * S * nat_esz);
    size_t const ev_conv_bytes = convert ? round256(N * S * esz) : 0;
    size_t const c_bytes = round256(2 * K * S * esz);
    size_t const part_bytes = round256(strings_expect_partial_bytes(ops, dtype, dim, S, states));
    size_t const split_bytes = fpm::gemm_scratch_bytes(contraction, K, S, N);
    void *scratch = nullptr;
    if ((st = scratch_get(s, ev_nat_bytes + ev_conv_bytes + c_bytes + part_bytes + split_bytes, &scratch)) != FP_OK)
        return st;
    char *base = static_cast<char *>(scratch);
    void *ev_nat = base;
    void *ev = convert ? static_cast<void *>(base + ev_nat_bytes) : ev_nat;
    void *cmat = base + ev_nat_bytes + ev_conv_bytes;
    auto *partial = reinterpret_cast<double *>(base + ev_nat_bytes + ev_conv_bytes + c_bytes);
    void *split = split_bytes ? base + ev_nat_bytes + ev_conv_bytes + c_bytes + part_bytes : nullptr;
    if ((st = run_strings_expect(ops, dtype, states, ev_nat, dim, S, s, partial)) != FP_OK)
        return st;
    if (convert)
    {
        uint32_t const cb = uint32_t((N * S + 255) / 256);
        if (dtype == FP_C128)
            hipLaunchKernelGGL((fpk::k_convert<double, float>), dim3(cb), dim3(256), 0, s,
                               static_cast<double const *>(ev_nat), static_cast<float *>(ev), uint64_t(N * S));
        else
            hipLaunchKernelGGL((fpk::k_convert<float, double>), dim3(cb), dim3(256), 0, s,
                               static_cast<float const *>(ev_nat), static_cast<double *>(ev), uint64_t(N * S));
        FP_LAUNCHED();
    }
    // stage 2: C_re = Wr^T ev, C_im = Wi^T ev  (ev is real: Pauli strings are Hermitian)
    size_t const planes = ops->matrix_imag_zero ? 1 : 2;
    for (size_t p = 0; p < planes; ++p)
    {
        void const *w = fp64 ? static_cast<void const *>(p == 0 ? ops->d_wr64 : ops->d_wi64)
                             : static_cast<void const *>(p == 0 ? ops->d_wr32 : ops->d_wi32);
        void *c = static_cast<char *>(cmat) + p * K * S * esz;
        if ((st = fpm::gemm(contraction, K, S, N, w, ev, c, split, s)) != FP_OK)
            return fail(st, "contraction failed: %s", fpm::last_error());
    }
    uint64_t const total = uint64_t(K) * S;
    uint32_t const blocks = uint32_t((total + 255) / 256);
    int const imag_zero = planes == 1;
    if (fp64)
    {
        if (dtype == FP_C128)
            hipLaunchKernelGGL((fpk::k_pack_complex<double, double>), dim3(blocks), dim3(256), 0, s,
                               static_cast<double const *>(cmat), static_cast<double *>(out), total, imag_zero,
                               accumulate);
        else
            hipLaunchKernelGGL((fpk::k_pack_complex<double, float>), dim3(blocks), dim3(256), 0, s,
                               static_cast<double const *>(cmat), static_cast<float *>(out), total, imag_zero,
                               accumulate);
    }
    else
    {
        if (dtype == FP_C128)
            hipLaunchKernelGGL((fpk::k_pack_complex<float, double>), dim3(blocks), dim3(256), 0, s,
                               static_cast<float const *>(cmat), static_cast<double *>(out), total, imag_zero,
                               accumulate);
        else
            hipLaunchKernelGGL((fpk::k_pack_complex<float, float>), dim3(blocks), dim3(256), 0, s,
                               static_cast<float const *>(cmat), static_cast<float *>(out), total, imag_zero,
                               accumulate);
    }
    FP_LAUNCHED();
    return FP_OK;
}

fp_status fp_hip_summed_apply(fp_hip_ops const *ops, int dtype, void const *states, void *out, size_t dim,
                              size_t n_states, int accumulate, void *stream)
{
    fp_status st = check_shape(ops, dtype, states, out, dim, n_states);
    if (st != FP_OK)
        return st;
    if (ops->n_operators == 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "SummedPauliOp has no coefficient matrix");
    if ((st = check_no_alias(dtype, states, out, dim, n_states)) != FP_OK)
        return st;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    // large states whose strings share coset tiles: the tiled kernel (same operations, same order)
    bool use_tiled = false;
    if ((st = fprt::choose_op_apply_tiled(ops, dtype, n_states, &use_tiled, fprt::kTilesSummedApply)) != FP_OK)
        return st;
    if (use_tiled)
        return fprt::run_op_apply_tiled(ops, dtype, states, out, dim, n_states, accumulate, s,
                                        fprt::kTilesSummedApply);
    VecGeom const g = vec_geom(dtype, dim, n_states, states, out);
    uint32_t const N = uint32_t(ops->n_strings);
    if (dtype == FP_C128)
        dispatch_op_apply<double, 1>(g, states, out, ops->d_x, ops->d_z, ops->d_rowsum128, N, accumulate != 0, s);
    else if (g.nc == 2)
        dispatch_op_apply<float, 2>(g, states, out, ops->d_x, ops->d_z, ops->d_rowsum64, N, accumulate != 0, s);
    else
        dispatch_op_apply<float, 1>(g, states, out, ops->d_x, ops->d_z, ops->d_rowsum64, N, accumulate != 0, s);
    FP_LAUNCHED();
    return FP_OK;
}

fp_status fp_hip_summed_apply_weighted(fp_hip_ops const *ops, int dtype, void const *states, void const *data,
                                       void *out, size_t dim, size_t n_states, int accumulate, void *stream)
{
    fp_status st = check_shape(ops, dtype, states, out, dim, n_states);
    if (st != FP_OK)
        return st;
    if (ops->n_operators == 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "SummedPauliOp has no coefficient matrix");
    if (!data)
        return fail(FP_ERR_INVALID_ARGUMENT, "data is NULL");
    if ((st = check_no_alias(dtype, states, out, dim, n_states)) != FP_OK)
        return st;
    if (ranges_overlap(data, ops->n_operators * n_states * (dtype == FP_C128 ? 8 : 4), out,
                       dim * n_states * (dtype == FP_C128 ? 16 : 8)))
        return fail(FP_ERR_INVALID_ARGUMENT, "out must not overlap data");
    hipStream_t const s = static_cast<hipStream_t>(stream);
    StreamGuard guard(s);
    size_t const N = ops->n_strings, K = ops->n_operators, S = n_states;
    size_t const esz = dtype == FP_C128 ? sizeof(double) : sizeof(float);
    void *scratch = nullptr;
    if ((st = scratch_get(s, 2 * N * S * esz, &scratch)) != FP_OK)
        return st;
    uint32_t const wb = uint32_t((N * S + fpk::kBlock - 1) / fpk::kBlock);
    if (dtype == FP_C128)
        hipLaunchKernelGGL(fpk::k_weighted_coeffs<double>, dim3(wb), dim3(fpk::kBlock), 0, s, ops->d_h128,
                           static_cast<double const *>(data), static_cast<double *>(scratch), uint64_t(N),
                           uint64_t(K), uint64_t(S));
    else
        hipLaunchKernelGGL(fpk::k_weighted_coeffs<float>, dim3(wb), dim3(fpk::kBlock), 0, s, ops->d_h64,
                           static_cast<float const *>(data), static_cast<float *>(scratch), uint64_t(N), uint64_t(K),
                           uint64_t(S));
    FP_LAUNCHED();
    // large states whose string groups pay: the coset-tiled kernel with per-state weights
    bool use_tiled = false;
    if ((st = fprt::choose_op_apply_tiled(ops, dtype, n_states, &use_tiled, fprt::kTilesWeightedApply)) != FP_OK)
        return st;
    if (use_tiled)
        return fprt::run_op_apply_tiled(ops, dtype, states, out, dim, n_states, accumulate, s,
                                        fprt::kTilesWeightedApply, scratch);
    VecGeom const g = vec_geom(dtype, dim, n_states, states, out);
    bool const acc = accumulate != 0;
    // four rows per thread (one for a 1-qubit operator, whose 2 rows do not split)
    auto run = [&](auto rows) -> fp_status {
        constexpr int R = decltype(rows)::value;
        uint64_t const n_slot = g.n_vec / R;
        for (uint64_t v0 = 0; v0 < n_slot; v0 += kMaxLaunchBlocks * fpk::kBlock)
        {
            uint32_t const blocks =
                uint32_t(std::min<uint64_t>((n_slot - v0 + fpk::kBlock - 1) / fpk::kBlock, kMaxLaunchBlocks));
            auto launch = [&](auto kernel, auto const *in_p, auto *out_p, auto const *w_p) {
                hipLaunchKernelGGL(kernel, dim3(blocks), dim3(fpk::kBlock), 0, s, in_p, out_p, v0, g.n_vec, g.shift,
                                   g.vpr, g.inv_vpr, ops->d_x, ops->d_z, ops->d_ny, w_p, uint32_t(N), uint64_t(S));
            };
            if (dtype == FP_C128)
            {
                using V = fpk::VecOf<double, 1>::type;
                auto const *ip = static_cast<V const *>(states);
                auto *op = static_cast<V *>(out);
                auto const *wp = static_cast<double const *>(scratch);
                if (g.pow2)
                    acc ? launch(fpk::k_op_apply_weighted<double, 1, true, true, R>, ip, op, wp)
                        : launch(fpk::k_op_apply_weighted<double, 1, true, false, R>, ip, op, wp);
                else
                    acc ? launch(fpk::k_op_apply_weighted<double, 1, false, true, R>, ip, op, wp)
                        : launch(fpk::k_op_apply_weighted<double, 1, false, false, R>, ip, op, wp);
            }
            else if (g.nc == 2)
            {
                using V = fpk::VecOf<float, 2>::type;
                auto const *ip = static_cast<V const *>(states);
                auto *op = static_cast<V *>(out);
                auto const *wp = static_cast<float const *>(scratch);
                if (g.pow2)
                    acc ? launch(fpk::k_op_apply_weighted<float, 2, true, true, R>, ip, op, wp)
                        : launch(fpk::k_op_apply_weighted<float, 2, true, false, R>, ip, op, wp);
                else
                    acc ? launch(fpk::k_op_apply_weighted<float, 2, false, true, R>, ip, op, wp)
                        : launch(fpk::k_op_apply_weighted<float, 2, false, false, R>, ip, op, wp);
            }
            else
            {
                using V = fpk::VecOf<float, 1>::type;
                auto const *ip = static_cast<V const *>(states);
                auto *op = static_cast<V *>(out);
                auto const *wp = static_cast<float const *>(scratch);
                if (g.pow2)
                    acc ? launch(fpk::k_op_apply_weighted<float, 1, true, true, R>, ip, op, wp)
                        : launch(fpk::k_op_apply_weighted<float, 1, true, false, R>, ip, op, wp);
                else
                    acc ? launch(fpk::k_op_apply_weighted<float, 1, false, true, R>, ip, op, wp)
                        : launch(fpk::k_op_apply_weighted<float, 1, false, false, R>, ip, op, wp);
            }
            FP_LAUNCHED();
        }
        return FP_OK;
    };
    return dim % 4 == 0 ? run(std::integral_constant<int, 4>{}) : run(std::integral_constant<int, 1>{});
}

fp_status fp_hip_gemm(int contraction, size_t m, size_t n, size_t k, void const *a, void const *b, void *c,
                      void *stream)
{
    if (!a || !b || !c || m == 0 || n == 0 || k == 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "empty or NULL GEMM operand");
    hipStream_t const s = static_cast<hipStream_t>(stream);
    StreamGuard guard(s);
    size_t const split_bytes = fpm::gemm_scratch_bytes(contraction, m, n, k);
    void *split = nullptr;
    fp_status st = FP_OK;
    if (split_bytes && (st = scratch_get(s, split_bytes, &split)) != FP_OK)
        return st;
    st = fpm::gemm(contraction, m, n, k, a, b, c, split, s);
    if (st != FP_OK)
        return fail(st, "%s", fpm::last_error());
    return FP_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Dense and sparse materialisation (debugging, interop)
// ---------------------------------------------------------------------------

namespace
{

fp_status check_dense(fp_hip_ops const *ops, int dtype, void const *out, size_t dim, size_t n_mats)
{
    fp_status st = check_ops(ops);
    if (st != FP_OK)
        return st;
    if (dtype != FP_C128 && dtype != FP_C64)
        return fail(FP_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    if (ops->n_strings == 0 || ops->n_qubits == 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "operator has no Pauli strings");
    if (ops->n_qubits > 30)
        return fail(FP_ERR_INVALID_ARGUMENT, "a dense matrix of %zu qubits is not addressable", ops->n_qubits);
    if (ops->n_strings > 0xffffffffu)
        return fail(FP_ERR_INVALID_ARGUMENT, "too many strings for the dense form (%zu)", ops->n_strings);
    size_t const want = size_t(1) << ops->n_qubits;
    if (dim != want)
        return fail(FP_ERR_INVALID_ARGUMENT, "tensor dimension (%zu) must match the dimension of the operators (%zu)",
                    dim, want);
    size_t bytes = 0;
    if (__builtin_mul_overflow(dim * dim, n_mats, &bytes) ||
        __builtin_mul_overflow(bytes, size_t(dtype == FP_C128 ? 16 : 8), &bytes))
        return fail(FP_ERR_INVALID_ARGUMENT, "dense tensor size overflows");
    if (!out)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL data pointer");
    uintptr_t const align = dtype == FP_C128 ? 16 : 8;
    if ((reinterpret_cast<uintptr_t>(out) & (align - 1)) != 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "device buffers must be aligned to the complex element size (%zu bytes)",
                    size_t(align));
    return FP_OK;
}

uint32_t grid_for(uint64_t n)
{
    return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>((n + fpk::kBlock - 1) / fpk::kBlock, 65536)));
}

// k_to_tensor over n_mats (K) matrices, after clearing them unless accumulating
fp_status run_to_tensor(fp_hip_ops const *ops, int dtype, void *out, size_t dim, size_t n_mats, bool assign,
                        size_t s0, void const *coef, uint32_t cstride, int folded, bool clear, hipStream_t s)
{
    uint64_t const n_rows = uint64_t(n_mats) * dim;
    uint32_t const n = uint32_t(ops->n_qubits), N = uint32_t(ops->n_strings);
    if (clear)
    {
        uint64_t const n_el = n_rows * dim;
        if (dtype == FP_C128)
            hipLaunchKernelGGL(fpd::k_fill_zero<double>, dim3(grid_for(n_el)), dim3(fpk::kBlock), 0, s,
                               static_cast<fpk::VecOf<double, 1>::type *>(out), n_el);
        else
            hipLaunchKernelGGL(fpd::k_fill_zero<float>, dim3(grid_for(n_el)), dim3(fpk::kBlock), 0, s,
                               static_cast<fpk::VecOf<float, 1>::type *>(out), n_el);
        FP_LAUNCHED();
    }
    auto launch = [&](auto T_tag) {
        typedef decltype(T_tag) T;
        auto *o = static_cast<typename fpk::VecOf<T, 1>::type *>(out);
        auto const *c = static_cast<T const *>(coef);
        if (assign)
            hipLaunchKernelGGL((fpd::k_to_tensor<T, true>), dim3(grid_for(n_rows)), dim3(fpk::kBlock), 0, s, o, n,
                               n_rows, ops->d_x, ops->d_z, ops->d_ny, uint32_t(s0), N, c, cstride, folded);
        else
            hipLaunchKernelGGL((fpd::k_to_tensor<T, false>), dim3(grid_for(n_rows)), dim3(fpk::kBlock), 0, s, o, n,
                               n_rows, ops->d_x, ops->d_z, ops->d_ny, uint32_t(s0), N, c, cstride, folded);
    };
    if (dtype == FP_C128)
        launch(double{});
    else
        launch(float{});
    FP_LAUNCHED();
    return FP_OK;
}

} // namespace

extern "C"
{

fp_status fp_hip_string_sparse_repr(fp_hip_ops const *ops, size_t string_index, int dtype, uint64_t *cols,
                                    void *vals, size_t dim, void *stream)
{
    fp_status st = check_shape(ops, dtype, cols, vals, dim, 1);
    if (st != FP_OK)
        return st;
    if (string_index >= ops->n_strings)
        return fail(FP_ERR_INVALID_ARGUMENT, "string index %zu out of range (%zu strings)", string_index,
                    ops->n_strings);
    if ((reinterpret_cast<uintptr_t>(cols) & 7u) != 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "cols must be aligned to 8 bytes");
    hipStream_t const s = static_cast<hipStream_t>(stream);
    uint64_t const x = ops->x[string_index], z = ops->z[string_index];
    uint32_t const ny = ops->ny[string_index];
    if (dtype == FP_C128)
        hipLaunchKernelGGL(fpd::k_sparse_repr<double>, dim3(grid_for(dim)), dim3(fpk::kBlock), 0, s, cols,
                           static_cast<fpk::VecOf<double, 1>::type *>(vals), uint64_t(dim), x, z, ny);
    else
        hipLaunchKernelGGL(fpd::k_sparse_repr<float>, dim3(grid_for(dim)), dim3(fpk::kBlock), 0, s, cols,
                           static_cast<fpk::VecOf<float, 1>::type *>(vals), uint64_t(dim), x, z, ny);
    FP_LAUNCHED();
    return FP_OK;
}

fp_status fp_hip_string_to_tensor(fp_hip_ops const *ops, size_t string_index, int dtype, void *out, size_t dim,
                                  void *stream)
{
    fp_status st = check_dense(ops, dtype, out, dim, 1);
    if (st != FP_OK)
        return st;
    if (string_index >= ops->n_strings)
        return fail(FP_ERR_INVALID_ARGUMENT, "string index %zu out of range (%zu strings)", string_index,
                    ops->n_strings);
    return run_to_tensor(ops, dtype, out, dim, 1, true, string_index, nullptr, 0, 0, false,
                         static_cast<hipStream_t>(stream));
}

fp_status fp_hip_op_to_tensor(fp_hip_ops const *ops, int dtype, void *out, size_t dim, int accumulate, void *stream)
{
    fp_status st = check_dense(ops, dtype, out, dim, 1);
    if (st != FP_OK)
        return st;
    void const *coef = dtype == FP_C128 ? static_cast<void const *>(ops->d_cf128) : static_cast<void const *>(ops->d_cf64);
    return run_to_tensor(ops, dtype, out, dim, 1, false, 0, coef, 1, 1, accumulate == 0,
                         static_cast<hipStream_t>(stream));
}

fp_status fp_hip_summed_to_tensor(fp_hip_ops const *ops, int dtype, void *out, size_t dim, int accumulate,
                                  void *stream)
{
    fp_status st = check_ops(ops);
    if (st != FP_OK)
        return st;
    if (ops->n_operators == 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "SummedPauliOp has no coefficient matrix");
    if ((st = check_dense(ops, dtype, out, dim, ops->n_operators)) != FP_OK)
        return st;
    if (ops->n_operators > 0xffffffffu)
        return fail(FP_ERR_INVALID_ARGUMENT, "too many operators");
    void const *coef = dtype == FP_C128 ? static_cast<void const *>(ops->d_h128) : static_cast<void const *>(ops->d_h64);
    return run_to_tensor(ops, dtype, out, dim, ops->n_operators, false, 0, coef, uint32_t(ops->n_operators), 0,
                         accumulate == 0, static_cast<hipStream_t>(stream));
}

} // extern "C"
