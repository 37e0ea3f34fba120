// fp_tiled.hip — host side of the coset-tiled PauliOp expectation kernel
// (fp_ev_tiled.hpp): string grouping, the per-operator table cache, the
// path choice and the launch.
//
// Grouping is greedy in x order: a string joins the open group if its x mask
// (above the low k bits) lies in the group's span or the span can still grow;
// otherwise the group is closed and padded with unit vectors to dimension dh.
// The reduced basis keeps every pivot bit in exactly one basis vector, so a
// string's local mask is read off its pivot bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "fast_pauli_hip.h"
#include "fp_ev_tiled.hpp"
#include "fp_runtime.hpp"

using fprt::fail;

namespace
{

uint32_t log2u(uint64_t v)
{
    uint32_t r = 0;
    while ((uint64_t(1) << (r + 1)) <= v)
        ++r;
    return r;
}

uint32_t parity(uint64_t v)
{
    return uint32_t(__builtin_popcountll(v) & 1);
}

uint32_t env_u32(char const *name, uint32_t dflt)
{
    char const *v = std::getenv(name);
    if (!v || !*v)
        return dflt;
    long const x = std::strtol(v, nullptr, 10);
    return x > 0 ? uint32_t(x) : dflt;
}

struct Span
{
    std::vector<uint64_t> b;
    std::vector<uint32_t> piv;

    uint64_t reduce(uint64_t r) const
    {
        for (size_t i = 0; i < b.size(); ++i)
            if ((r >> piv[i]) & 1u)
                r ^= b[i];
        return r;
    }
    void add(uint64_t r) // r already reduced, non-zero
    {
        uint32_t const p = 63u - uint32_t(__builtin_clzll(r));
        for (auto &v : b)
            if ((v >> p) & 1u)
                v ^= r;
        b.push_back(r);
        piv.push_back(p);
    }
};

} // namespace

namespace fprt
{

TileGeom tile_geom(int dtype, size_t n_qubits, size_t n_states)
{
    uint32_t const esz = dtype == FP_C128 ? 16u : 8u;
    uint32_t const row_target = env_u32("FAST_PAULI_TILE_ROW_BYTES", 256u);
    uint32_t const elems = fpt::kTileBytes / esz;
    uint32_t sblk = 1;
    while (sblk < n_states && sblk * esz < row_target)
        sblk *= 2;
    TileGeom g{};
    g.lsb = log2u(sblk);
    g.d = std::min<uint32_t>(log2u(elems / sblk), uint32_t(n_qubits));
    g.k = 0;
    if (sblk >= n_states) // one block covers every state: 2^k consecutive rows are contiguous
        while ((uint64_t(1) << g.k) * n_states * esz < 128 && g.k < g.d)
            ++g.k;
    g.dh = g.d - g.k;
    if (g.dh > uint32_t(fpt::kMaxDh))
    {
        g.dh = fpt::kMaxDh;
        g.k = g.d - g.dh;
    }
    return g;
}

// Builds (or returns the cached) tile tables of `ops` for geometry (k, dh).
fp_status tile_tables(fp_hip_ops const *ops, TileGeom const &geo, TileTables const **out)
{
    std::lock_guard<std::mutex> lock(ops->tiles_mu);
    for (auto const &t : ops->tiles)
        if (t.k == geo.k && t.dh == geo.dh)
        {
            *out = &t;
            return FP_OK;
        }
    size_t const N = ops->n_strings, n = ops->n_qubits;
    uint32_t const k = geo.k, dh = geo.dh;
    uint64_t const lowmask = (uint64_t(1) << k) - 1;
    std::vector<uint32_t> order(N);
    for (size_t i = 0; i < N; ++i)
        order[i] = uint32_t(i);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return ops->x[a] < ops->x[b]; });

    std::vector<fpt::TileGroup> groups;
    std::vector<uint64_t> btab;
    std::vector<fpt::TileChunk> chunks;
    std::vector<fpt::TileStr> strs;
    Span span;
    std::vector<uint32_t> members;
    auto close = [&]() {
        if (members.empty())
            return;
        for (uint32_t i = k; i < n && span.b.size() < dh; ++i)
        {
            uint64_t const r = span.reduce(uint64_t(1) << i);
            if (r)
                span.add(r);
        }
        // basis in ascending pivot order: bit i of a_hi <-> i-th smallest pivot
        std::vector<size_t> idx(span.b.size());
        for (size_t i = 0; i < idx.size(); ++i)
            idx[i] = i;
        std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return span.piv[a] < span.piv[b]; });
        fpt::TileGroup g{};
        g.n_piv = uint32_t(idx.size());
        g.btab_off = uint32_t(btab.size());
        for (size_t i = 0; i < idx.size(); ++i)
            (i < 8 ? g.piv_lo : g.piv_hi) |= uint64_t(span.piv[idx[i]]) << (8 * (i & 7));
        size_t const n_bt = size_t(1) << g.n_piv;
        size_t const base = btab.size();
        btab.resize(base + n_bt, 0);
        for (size_t a = 1; a < n_bt; ++a)
            btab[base + a] = btab[base + (a & (a - 1))] ^ span.b[idx[__builtin_ctzll(a)]];
        struct Loc
        {
            uint32_t A, Z, s;
        };
        std::vector<Loc> loc;
        for (uint32_t s : members)
        {
            uint64_t const xh = ops->x[s] & ~lowmask;
            uint32_t alpha = 0, zeta = 0;
            for (size_t i = 0; i < idx.size(); ++i)
            {
                alpha |= uint32_t((xh >> span.piv[idx[i]]) & 1u) << i;
                zeta |= parity(span.b[idx[i]] & ops->z[s]) << i;
            }
            loc.push_back(Loc{(alpha << k) | uint32_t(ops->x[s] & lowmask), (zeta << k) | uint32_t(ops->z[s] & lowmask),
                              s});
        }
        std::stable_sort(loc.begin(), loc.end(), [](Loc const &a, Loc const &b) { return a.A < b.A; });
        g.chunk_first = uint32_t(chunks.size());
        for (size_t i = 0; i < loc.size();)
        {
            size_t j = i;
            while (j < loc.size() && j - i < size_t(fpt::kChunk) && loc[j].A == loc[i].A)
                ++j;
            fpt::TileChunk c{};
            c.A = loc[i].A;
            c.h = c.A ? 31u - uint32_t(__builtin_clz(c.A)) : 0u;
            c.count = uint32_t(j - i);
            c.first = uint32_t(strs.size());
            for (size_t m = i; m < j; ++m)
            {
                fpt::TileStr st{};
                st.z = ops->z[loc[m].s];
                st.zl = loc[m].Z;
                st.sidx = loc[m].s;
                st.imag = ops->ny[loc[m].s] & 1u;
                strs.push_back(st);
            }
            chunks.push_back(c);
            i = j;
        }
        g.n_chunks = uint32_t(chunks.size()) - g.chunk_first;
        groups.push_back(g);
        span = Span{};
        members.clear();
    };
    if (dh == 0)
        for (uint32_t s : order)
            if (ops->x[s] & ~lowmask)
                return fail(FP_ERR_RUNTIME, "tile geometry has no room for x mask bits above bit %u", k);
    for (uint32_t s : order)
    {
        uint64_t const r = span.reduce(ops->x[s] & ~lowmask);
        if (r && span.b.size() >= dh)
        {
            close();
            uint64_t const r2 = ops->x[s] & ~lowmask;
            if (r2)
                span.add(r2);
        }
        else if (r)
            span.add(r);
        members.push_back(s);
    }
    close();

    TileTables t{};
    t.k = k;
    t.dh = dh;
    t.n_groups = uint32_t(groups.size());
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    size_t const o_groups = 0;
    size_t const o_btab = o_groups + up(groups.size() * sizeof(fpt::TileGroup));
    size_t const o_chunks = o_btab + up(btab.size() * sizeof(uint64_t));
    size_t const o_strs = o_chunks + up(chunks.size() * sizeof(fpt::TileChunk));
    size_t const total = o_strs + up(strs.size() * sizeof(fpt::TileStr));
    std::vector<unsigned char> h(total, 0);
    std::memcpy(h.data() + o_groups, groups.data(), groups.size() * sizeof(fpt::TileGroup));
    std::memcpy(h.data() + o_btab, btab.data(), btab.size() * sizeof(uint64_t));
    std::memcpy(h.data() + o_chunks, chunks.data(), chunks.size() * sizeof(fpt::TileChunk));
    std::memcpy(h.data() + o_strs, strs.data(), strs.size() * sizeof(fpt::TileStr));
    FP_HIP(hipMalloc(&t.d_mem, total));
    hipError_t const e = hipMemcpy(t.d_mem, h.data(), total, hipMemcpyHostToDevice);
    if (e != hipSuccess)
    {
        (void)hipFree(t.d_mem);
        return fail(FP_ERR_RUNTIME, "tile table upload failed: %s", hipGetErrorString(e));
    }
    auto *base = static_cast<unsigned char *>(t.d_mem);
    t.groups = reinterpret_cast<fpt::TileGroup *>(base + o_groups);
    t.btab = reinterpret_cast<uint64_t *>(base + o_btab);
    t.chunks = reinterpret_cast<fpt::TileChunk *>(base + o_chunks);
    t.strs = reinterpret_cast<fpt::TileStr *>(base + o_strs);
    ops->tiles.push_back(t);
    *out = &ops->tiles.back();
    return FP_OK;
}

void release_tiles(fp_hip_ops *ops)
{
    std::lock_guard<std::mutex> lock(ops->tiles_mu);
    for (auto &t : ops->tiles)
        if (t.d_mem)
            (void)hipFree(t.d_mem);
    ops->tiles.clear();
}

// 0: streaming kernel, 1: tiled kernel, per FAST_PAULI_EV_KERNEL ("stream",
// "tiled") or, by default, by estimated HBM traffic (groups vs strings).
fp_status choose_op_expect_tiled(fp_hip_ops const *ops, int dtype, size_t n_states, bool *use_tiled)
{
    *use_tiled = false;
    char const *v = std::getenv("FAST_PAULI_EV_KERNEL");
    if (v && std::strcmp(v, "stream") == 0)
        return FP_OK;
    bool const forced = v && std::strcmp(v, "tiled") == 0;
    if (!forced && ops->n_strings < 2)
        return FP_OK;
    TileGeom const geo = tile_geom(dtype, ops->n_qubits, n_states);
    TileTables const *tt = nullptr;
    fp_status const st = tile_tables(ops, geo, &tt);
    if (st != FP_OK)
        return st;
    *use_tiled = forced || double(tt->n_groups) <= 0.8 * double(ops->n_strings);
    return FP_OK;
}

fp_status run_op_expect_tiled(fp_hip_ops const *ops, int dtype, void const *states, void *out, size_t dim,
                              size_t n_states, int accumulate, hipStream_t st)
{
    TileGeom const geo = tile_geom(dtype, ops->n_qubits, n_states);
    TileTables const *tt = nullptr;
    fp_status s = tile_tables(ops, geo, &tt);
    if (s != FP_OK)
        return s;
    uint64_t const n_cosets = uint64_t(dim) >> geo.d;
    uint32_t const chunks_y = uint32_t((n_states + (size_t(1) << geo.lsb) - 1) >> geo.lsb);
    if (chunks_y > 65535u)
        return fail(FP_ERR_INVALID_ARGUMENT, "too many states for the tiled expectation kernel");
    // about 2048 blocks in total, each over a contiguous range of cosets
    uint64_t const target = 2048;
    uint64_t want = (target + uint64_t(tt->n_groups) * chunks_y - 1) / (uint64_t(tt->n_groups) * chunks_y);
    want = std::max<uint64_t>(1, std::min<uint64_t>(want, n_cosets));
    uint64_t const per_range = (n_cosets + want - 1) / want;
    uint32_t const ranges = uint32_t((n_cosets + per_range - 1) / per_range);
    uint64_t const n_parts = uint64_t(tt->n_groups) * ranges;
    if (n_parts > 0x7fffffffu)
        return fail(FP_ERR_INVALID_ARGUMENT, "too many tile groups");
    void *scratch = nullptr;
    if ((s = scratch_get(st, n_parts * n_states * sizeof(fpk::Dbl2), &scratch)) != FP_OK)
        return s;
    auto *partial = static_cast<fpk::Dbl2 *>(scratch);
    size_t const lds = fpt::kTileBytes + (size_t(1) << geo.dh) * sizeof(uint64_t);
    static std::once_flag attr_once;
    static hipError_t attr_err = hipSuccess;
    std::call_once(attr_once, [] {
        hipError_t e1 = hipFuncSetAttribute(reinterpret_cast<void const *>(&fpt::k_op_expect_tiled<double>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            int(fpt::kTileBytes + (1u << fpt::kMaxDh) * sizeof(uint64_t)));
        hipError_t e2 = hipFuncSetAttribute(reinterpret_cast<void const *>(&fpt::k_op_expect_tiled<float>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            int(fpt::kTileBytes + (1u << fpt::kMaxDh) * sizeof(uint64_t)));
        attr_err = e1 != hipSuccess ? e1 : e2;
    });
    if (attr_err != hipSuccess)
        return fail(FP_ERR_RUNTIME, "hipFuncSetAttribute failed: %s", hipGetErrorString(attr_err));
    dim3 const grid(uint32_t(n_parts), chunks_y);
    if (dtype == FP_C128)
    {
        typedef fpk::VecOf<double, 1>::type V;
        hipLaunchKernelGGL(fpt::k_op_expect_tiled<double>, grid, dim3(fpt::kThreads), lds, st,
                           static_cast<V const *>(states), uint64_t(n_states), geo.lsb, geo.k, geo.d, ranges,
                           per_range, n_cosets, tt->groups, tt->btab, tt->chunks, tt->strs, ops->d_scale, partial);
        FP_LAUNCHED();
        hipLaunchKernelGGL(fpk::k_reduce_partials<double>, dim3(uint32_t(n_states)), dim3(fpk::kBlock), 0, st,
                           partial, n_parts, uint64_t(n_states), static_cast<double *>(out), accumulate, 1.0, 0.0);
    }
    else
    {
        typedef fpk::VecOf<float, 1>::type V;
        hipLaunchKernelGGL(fpt::k_op_expect_tiled<float>, grid, dim3(fpt::kThreads), lds, st,
                           static_cast<V const *>(states), uint64_t(n_states), geo.lsb, geo.k, geo.d, ranges,
                           per_range, n_cosets, tt->groups, tt->btab, tt->chunks, tt->strs, ops->d_scale, partial);
        FP_LAUNCHED();
        hipLaunchKernelGGL(fpk::k_reduce_partials<float>, dim3(uint32_t(n_states)), dim3(fpk::kBlock), 0, st,
                           partial, n_parts, uint64_t(n_states), static_cast<float *>(out), accumulate, 1.0, 0.0);
    }
    FP_LAUNCHED();
    return FP_OK;
}

} // namespace fprt
