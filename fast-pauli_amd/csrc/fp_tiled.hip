// fp_tiled.hip — host side of the coset-tiled kernels (fp_ev_tiled.hpp,
// fp_apply_tiled.hpp, the per-string EV matrix on cosets): the per-operator
// cache of string-group tables (built by fp_tiles_host.cpp) and their device
// copies, the path choices and the launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>
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

uint32_t env_u32(char const *name, uint32_t dflt)
{
    char const *v = std::getenv(name);
    if (!v || !*v)
        return dflt;
    long const x = std::strtol(v, nullptr, 10);
    return x > 0 ? uint32_t(x) : dflt;
}

} // namespace

namespace fprt
{


// Builds (or returns the cached) tables of `ops` of one kind for geometry (k, dh).
fp_status tile_tables(fp_hip_ops const *ops, int kind, uint32_t k, uint32_t dh, TileTables **out, uint32_t wht_min)
{
    std::lock_guard<std::mutex> lock(ops->tiles_mu);
    for (auto &t : ops->tiles)
        if (t.valid && t.kind == kind && t.k == k && t.dh == dh && t.wht_min == wht_min)
        {
            *out = &t;
            return FP_OK;
        }
    bool const rows = kind == kTilesSummedApply;
    if (rows && (ops->rowsum128.size() != 2 * ops->n_strings || ops->rowsum64.size() != 2 * ops->n_strings))
        return fail(FP_ERR_INVALID_ARGUMENT, "SummedPauliOp has no coefficient matrix");
    // weighted apply: the records carry (-i)^{nY} only
    std::vector<double> unit128;
    std::vector<float> unit64;
    if (kind == kTilesWeightedApply)
    {
        unit128.resize(2 * ops->n_strings);
        unit64.resize(2 * ops->n_strings);
        for (size_t i = 0; i < ops->n_strings; ++i)
        {
            static double const re[4] = {1.0, 0.0, -1.0, 0.0}, im[4] = {0.0, -1.0, 0.0, 1.0};
            uint32_t const ny = ops->ny[i] & 3u;
            unit128[2 * i] = re[ny];
            unit128[2 * i + 1] = im[ny];
            unit64[2 * i] = float(re[ny]);
            unit64[2 * i + 1] = float(im[ny]);
        }
    }
    double const *c128 = rows ? ops->rowsum128.data() : (unit128.empty() ? ops->cf128.data() : unit128.data());
    float const *c64 = rows ? ops->rowsum64.data() : (unit64.empty() ? ops->cf64.data() : unit64.data());
    StringTable const st{ops->n_qubits, ops->n_strings, ops->x.data(), ops->z.data(), ops->ny.data(),
                         ops->coeffs.data(), c128, c64};
    TileImage img;
    fp_status const bs = build_tile_image(st, kind, k, dh, wht_min, img);
    if (bs != FP_OK)
        return bs;

    // reuse an invalidated entry (its host image and device memory) when there is one
    TileTables *t = nullptr;
    for (auto &e : ops->tiles)
        if (!e.valid)
        {
            t = &e;
            break;
        }
    if (!t)
    {
        ops->tiles.emplace_back();
        t = &ops->tiles.back();
        hipError_t const e = hipEventCreateWithFlags(&t->uploaded, hipEventDisableTiming);
        if (e != hipSuccess)
            return fail(FP_ERR_RUNTIME, "tile table event creation failed: %s", hipGetErrorString(e));
    }
    if (t->uploaded)
    {
        // the previous upload may still be reading the host image
        hipError_t const e = hipEventSynchronize(t->uploaded);
        if (e != hipSuccess)
            return fail(FP_ERR_RUNTIME, "tile table event wait failed: %s", hipGetErrorString(e));
    }
    t->kind = kind;
    t->k = k;
    t->dh = dh;
    t->wht_min = wht_min;
    t->n_groups = img.n_groups;
    t->max_chunks = img.max_chunks;
    t->o_btab = img.o_btab;
    t->o_chunks = img.o_chunks;
    t->o_strs = img.o_strs;
    t->o_sh = img.o_sh;
    t->h = std::move(img.h);
    t->on_device = false;
    t->valid = true;
    *out = t;
    return FP_OK;
}

// Makes the device copy of `t` current for kernels on `stream`: reuses its
// device memory when large enough, uploads asynchronously on `stream` and
// records an event that guards the host image.
fp_status tile_device(fp_hip_ops const *ops, TileTables *t, hipStream_t stream)
{
    std::lock_guard<std::mutex> lock(ops->tiles_mu);
    if (t->on_device)
    {
        // the upload may have been issued on another stream: order this one after it
        if (t->uploaded)
            FP_HIP(hipStreamWaitEvent(stream, t->uploaded, 0));
        return FP_OK;
    }
    size_t const total = t->h.size();
    if (t->cap < total)
    {
        if (t->d_mem)
        {
            FP_HIP(hipFree(t->d_mem));
            t->d_mem = nullptr;
            t->cap = 0;
        }
        FP_HIP(hipMalloc(&t->d_mem, total));
        t->cap = total;
    }
    if (!t->uploaded)
        FP_HIP(hipEventCreateWithFlags(&t->uploaded, hipEventDisableTiming));
    FP_HIP(hipMemcpyAsync(t->d_mem, t->h.data(), total, hipMemcpyHostToDevice, stream));
    FP_HIP(hipEventRecord(t->uploaded, stream));
    auto *base = static_cast<unsigned char *>(t->d_mem);
    t->groups = reinterpret_cast<fpt::TileGroup *>(base);
    t->btab = reinterpret_cast<uint64_t *>(base + t->o_btab);
    t->chunks = base + t->o_chunks;
    t->strs = base + t->o_strs;
    t->sh = reinterpret_cast<fpt::ShStr *>(base + t->o_sh);
    t->on_device = true;
    return FP_OK;
}

void invalidate_tiles(fp_hip_ops *ops)
{
    std::lock_guard<std::mutex> lock(ops->tiles_mu);
    for (auto &t : ops->tiles)
    {
        t.valid = false;
        t.on_device = false;
    }
}

void release_tiles(fp_hip_ops *ops)
{
    std::lock_guard<std::mutex> lock(ops->tiles_mu);
    for (auto &t : ops->tiles)
    {
        if (t.uploaded)
        {
            (void)hipEventSynchronize(t.uploaded);
            (void)hipEventDestroy(t.uploaded);
        }
        if (t.d_mem)
            (void)hipFree(t.d_mem);
    }
    ops->tiles.clear();
}

// The string-group tables `tt` were cut for `geo`, and one tile of that geometry
// (2^d rows x 2^lsb states) fits the LDS tile of the kernel instance about to run.
fp_status check_tile_fit(TileTables const *tt, TileGeom const &geo, int dtype, size_t tile_bytes)
{
    size_t const esz = dtype == FP_C128 ? 16 : 8;
    if (tt->k != geo.k || tt->dh != geo.dh)
        return fail(FP_ERR_RUNTIME, "string-group tables built for k=%u dh=%u, launch geometry k=%u dh=%u", tt->k,
                    tt->dh, geo.k, geo.dh);
    if (geo.d != geo.k + geo.dh || (size_t(1) << (geo.d + geo.lsb)) * esz > tile_bytes)
        return fail(FP_ERR_RUNTIME, "a tile of 2^%u rows x 2^%u states does not fit the kernel's %zu-byte LDS tile",
                    geo.d, geo.lsb, tile_bytes);
    if (tile_bytes > lds_bytes())
        return fail(FP_ERR_UNSUPPORTED, "%zu-byte LDS tiles exceed the device's %zu bytes per workgroup", tile_bytes,
                    lds_bytes());
    return FP_OK;
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
    TileGeom const geo = tile_geom(dtype, ops->n_qubits, n_states, true);
    TileTables *tt = nullptr;
    fp_status const st = tile_tables(ops, kTilesOpExpect, geo.k, geo.dh, &tt);
    if (st != FP_OK)
        return st;
    *use_tiled = forced || double(tt->n_groups) <= 0.8 * double(ops->n_strings);
    return FP_OK;
}

// Instance of k_op_expect_tiled that runs group g at this launch shape (2 shared
// leaders, 1 set up once, 0 general), by the kernel's own rule: modes 1 and 2
// run 2^r leader rows per pass, r = 6 - lsb, or r + 1 for complex64 (two states
// per lane, which needs lsb >= 1).
int ev_group_mode(fpt::TileGroup const &g, int dtype, TileGeom const &geo, size_t n_states)
{
    bool const c64 = dtype != FP_C128;
    uint32_t const W = geo.tile_kb == 128 ? 16u : 8u;
    uint32_t const rrx = (c64 ? 7u : 6u) - geo.lsb;
    bool const can_pair = !c64 || geo.lsb >= 1;
    // complex64 shared leaders need the paired 16-byte tile loads: full tiles and a
    // power-of-two state count (the kernel's own rule)
    bool const pow2 = (n_states & (n_states - 1)) == 0;
    bool const pairs_ok = !c64 || (pow2 && geo.lsb >= 1 && geo.d + geo.lsb == log2u(geo.tile_kb * 1024u / 8u));
    if (can_pair && pairs_ok && ((g.shared >> fpt::shared_bit(W, c64, rrx)) & 1u) != 0)
        return 2;
    if (can_pair && ((g.hoist >> fpt::hoist_bit(W, 4u, rrx)) & 1u) != 0)
        return 1;
    return 0;
}

fp_status run_op_expect_tiled(fp_hip_ops const *ops, int dtype, void const *states, void *out, size_t dim,
                              size_t n_states, int accumulate, hipStream_t st, Workspace const *ws)
{
    TileGeom const geo = tile_geom(dtype, ops->n_qubits, n_states, true);
    TileTables *tt = nullptr;
    fp_status s = tile_tables(ops, kTilesOpExpect, geo.k, geo.dh, &tt);
    if (s != FP_OK)
        return s;
    if ((s = check_tile_fit(tt, geo, dtype, size_t(geo.tile_kb) << 10)) != FP_OK)
        return s;
    // device addresses of the tables: their own device copy, or the workspace's head
    fpt::TileGroup const *d_groups;
    uint64_t const *d_btab;
    void const *d_chunks, *d_strs;
    fpt::ShStr const *d_sh;
    size_t const table_bytes = (tt->h.size() + 255) / 256 * 256;
    if (ws)
    {
        if (table_bytes > ws->bytes)
            return fail(FP_ERR_INVALID_ARGUMENT, "the string-group tables (%zu bytes) exceed the %zu-byte workspace",
                        table_bytes, ws->bytes);
        std::lock_guard<std::mutex> lock(ops->tiles_mu);
        FP_HIP(hipMemcpyAsync(ws->base, tt->h.data(), tt->h.size(), hipMemcpyHostToDevice, st));
        // the host image may be rebuilt only once this copy has read it
        FP_HIP(hipEventRecord(tt->uploaded, st));
        auto *base = static_cast<unsigned char *>(ws->base);
        d_groups = reinterpret_cast<fpt::TileGroup const *>(base);
        d_btab = reinterpret_cast<uint64_t const *>(base + tt->o_btab);
        d_chunks = base + tt->o_chunks;
        d_strs = base + tt->o_strs;
        d_sh = reinterpret_cast<fpt::ShStr const *>(base + tt->o_sh);
    }
    else
    {
        if ((s = tile_device(ops, tt, st)) != FP_OK)
            return s;
        d_groups = tt->groups;
        d_btab = tt->btab;
        d_chunks = tt->chunks;
        d_strs = tt->strs;
        d_sh = tt->sh;
    }
    uint64_t const n_cosets = uint64_t(dim) >> geo.d;
    uint32_t const chunks_y = uint32_t((n_states + (size_t(1) << geo.lsb) - 1) >> geo.lsb);
    if (chunks_y > 65535u)
        return fail(FP_ERR_INVALID_ARGUMENT, "too many states for the tiled expectation kernel");
    // up to 16384 blocks (FAST_PAULI_TILE_BLOCKS) with at least ~16 tiles each; the
    // `ranges` blocks of a group take every ranges-th coset
    uint64_t const tiles = uint64_t(tt->n_groups) * chunks_y * n_cosets;
    uint64_t const target = std::max<uint64_t>(256, std::min<uint64_t>(env_u32("FAST_PAULI_TILE_BLOCKS", 16384), tiles / 16));
    uint64_t want = (target + uint64_t(tt->n_groups) * chunks_y - 1) / (uint64_t(tt->n_groups) * chunks_y);
    want = std::max<uint64_t>(1, std::min<uint64_t>(want, n_cosets));
    uint64_t const part_bytes = uint64_t(n_states) * sizeof(fpk::Dbl2); // per (group, range)
    if (ws)
    {
        // fewer blocks per group when the workspace cannot hold their partials
        uint64_t const room = (ws->bytes - table_bytes) / part_bytes;
        if (room < tt->n_groups)
            return fail(FP_ERR_INVALID_ARGUMENT,
                        "%u string groups x %zu states need %llu bytes of partials beside %zu bytes of tables, "
                        "more than the %zu-byte workspace",
                        tt->n_groups, n_states, static_cast<unsigned long long>(uint64_t(tt->n_groups) * part_bytes),
                        table_bytes, ws->bytes);
        want = std::min<uint64_t>(want, room / tt->n_groups);
    }
    uint32_t const ranges = uint32_t(std::min<uint64_t>(want, 0x7fffffffu));
    uint64_t const n_parts = uint64_t(tt->n_groups) * ranges;
    if (n_parts > 0x7fffffffu)
        return fail(FP_ERR_INVALID_ARGUMENT, "too many tile groups");
    void *scratch = nullptr;
    if (ws)
        scratch = static_cast<unsigned char *>(ws->base) + table_bytes;
    else if ((s = scratch_get(st, n_parts * part_bytes, &scratch)) != FP_OK)
        return s;
    auto *partial = static_cast<fpk::Dbl2 *>(scratch);
    size_t const lds = size_t(geo.tile_kb) * 1024u;
    uint32_t const sshift = (n_states & (n_states - 1)) == 0 ? log2u(n_states) : 64u;
    // 64-bit element offsets when the state has 2^32 elements or more (FAST_PAULI_TILE_WIDE=1
    // forces them, so that tests cover both forms at small sizes)
    bool const wide = (uint64_t(dim) * n_states) >> 32 != 0 || env_u32("FAST_PAULI_TILE_WIDE", 0) == 1;
    static std::once_flag attr_once;
    static hipError_t attr_err = hipSuccess;
    std::call_once(attr_once, [] {
        auto set = [](void const *fn, int bytes) {
            hipError_t const e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
            if (e != hipSuccess && attr_err == hipSuccess)
                attr_err = e;
        };
        auto set4 = [&](auto T_tag, auto tkb_tag) {
            typedef decltype(T_tag) T;
            constexpr int TKB = decltype(tkb_tag)::value;
            set(reinterpret_cast<void const *>(&fpt::k_op_expect_tiled<T, false, TKB, 0>), TKB << 10);
            set(reinterpret_cast<void const *>(&fpt::k_op_expect_tiled<T, true, TKB, 0>), TKB << 10);
            set(reinterpret_cast<void const *>(&fpt::k_op_expect_tiled<T, false, TKB, 1>), TKB << 10);
            set(reinterpret_cast<void const *>(&fpt::k_op_expect_tiled<T, true, TKB, 1>), TKB << 10);
            set(reinterpret_cast<void const *>(&fpt::k_op_expect_tiled<T, false, TKB, 2>), TKB << 10);
            set(reinterpret_cast<void const *>(&fpt::k_op_expect_tiled<T, true, TKB, 2>), TKB << 10);
        };
        set4(double{}, std::integral_constant<int, 64>{});
        set4(float{}, std::integral_constant<int, 64>{});
        set(reinterpret_cast<void const *>(&fpt::k_op_expect_tiled<double, false, 64, 3>), 64 << 10);
        set(reinterpret_cast<void const *>(&fpt::k_op_expect_tiled<float, false, 64, 3>), 64 << 10);
        set4(double{}, std::integral_constant<int, 128>{});
        set4(float{}, std::integral_constant<int, 128>{});
    });
    if (attr_err != hipSuccess)
        return fail(FP_ERR_RUNTIME, "hipFuncSetAttribute failed: %s", hipGetErrorString(attr_err));
    uint64_t const n_ids = (n_parts + 7) / 8 * 8 * chunks_y;
    if (n_ids > 0x7fffffffu)
        return fail(FP_ERR_INVALID_ARGUMENT, "too many tile blocks");
    dim3 const grid = dim3(static_cast<uint32_t>(n_ids));
    // Groups whose waves each own at most one fast-path unit run in the HOIST
    // instance, the others in the general one; blocks of the other kind exit at
    // once, so a launch with both kinds runs both instances over the same grid.
    bool run_mode[3] = {false, false, false};
    auto const *hgroups = reinterpret_cast<fpt::TileGroup const *>(tt->h.data());
    for (uint32_t g = 0; g < tt->n_groups; ++g)
        run_mode[ev_group_mode(hgroups[g], dtype, geo, n_states)] = true;
    auto launch = [&](auto T_tag, auto wide_tag, auto tkb_tag, void *typed_out_v) -> fp_status {
        typedef decltype(T_tag) T;
        typedef typename fpk::VecOf<T, 1>::type V;
        constexpr bool WIDE = decltype(wide_tag)::value;
        constexpr int TKB = decltype(tkb_tag)::value;
        auto const *typed_states = static_cast<V const *>(states);
        // shared leaders with each string's partner offset and coset sign mask in one word
        // when every coset index is below 2^16 (64 KiB tiles: offsets below 2^16 too);
        // FAST_PAULI_EV_PACK=2 keeps them apart
        bool const pack = n_cosets <= 65536u && env_u32("FAST_PAULI_EV_PACK", 1) == 1;
        for (int mode = 2; mode >= 0; --mode)
        {
            if (!run_mode[mode])
                continue;
            auto kfn = mode == 2   ? &fpt::k_op_expect_tiled<T, WIDE, TKB, 2>
                       : mode == 1 ? &fpt::k_op_expect_tiled<T, WIDE, TKB, 1>
                                   : &fpt::k_op_expect_tiled<T, WIDE, TKB, 0>;
            if constexpr (!WIDE && TKB == 64)
                if (mode == 2 && pack)
                    kfn = &fpt::k_op_expect_tiled<T, false, 64, 3>;
            hipLaunchKernelGGL(kfn, grid, dim3(TKB * 8), lds, st, typed_states, uint64_t(n_states), sshift, geo.lsb,
                               geo.k, geo.d, ranges, uint32_t(n_parts), chunks_y, n_cosets, d_groups, d_btab,
                               static_cast<fpt::TileChunk const *>(d_chunks),
                               static_cast<fpt::TileStr const *>(d_strs), d_sh, partial);
            FP_LAUNCHED();
        }
        hipLaunchKernelGGL(fpk::k_reduce_partials<T>, dim3(uint32_t(n_states)), dim3(fpk::kBlock), 0, st, partial,
                           n_parts, uint64_t(n_states), static_cast<T *>(typed_out_v), accumulate, 1.0, 0.0);
        FP_LAUNCHED();
        return FP_OK;
    };
    typedef std::integral_constant<int, 64> K64;
    typedef std::integral_constant<int, 128> K128;
    if (dtype == FP_C128)
    {
        if (geo.tile_kb == 128)
            return wide ? launch(double{}, std::true_type{}, K128{}, out) : launch(double{}, std::false_type{}, K128{}, out);
        return wide ? launch(double{}, std::true_type{}, K64{}, out) : launch(double{}, std::false_type{}, K64{}, out);
    }
    if (geo.tile_kb == 128)
        return wide ? launch(float{}, std::true_type{}, K128{}, out) : launch(float{}, std::false_type{}, K128{}, out);
    return wide ? launch(float{}, std::true_type{}, K64{}, out) : launch(float{}, std::false_type{}, K64{}, out);
}

// ---------------------------------------------------------------------------
// PauliOp::apply on coset tiles (groups of consecutive strings, one launch each)
// ---------------------------------------------------------------------------

fp_status apply_tiled_plan(fp_hip_ops const *ops, int dtype, size_t n_states, size_t min_bytes, int kind,
                           TileGeom *geo, TileTables **out)
{
    *out = nullptr;
    size_t const state_bytes = (size_t(1) << ops->n_qubits) * n_states * (dtype == FP_C128 ? 16 : 8);
    if (ops->n_strings < 8 || state_bytes < min_bytes)
        return FP_OK;
    TileGeom const g = tile_geom(dtype, ops->n_qubits, n_states, false);
    TileTables *tt = nullptr;
    fp_status const st = tile_tables(ops, kind, g.k, g.dh, &tt);
    if (st != FP_OK)
        return st;
    if (3.0 * double(tt->n_groups) > 0.8 * double(ops->n_strings))
        return FP_OK;
    *geo = g;
    *out = tt;
    return FP_OK;
}

// 0: streaming kernel, 1: tiled kernel, per FAST_PAULI_APPLY_KERNEL ("stream",
// "tiled") or, by default, by apply_tiled_plan with kApplyTiledMinBytes.
fp_status choose_op_apply_tiled(fp_hip_ops const *ops, int dtype, size_t n_states, bool *use_tiled, int kind)
{
    *use_tiled = false;
    char const *v = std::getenv("FAST_PAULI_APPLY_KERNEL");
    if (v && std::strcmp(v, "stream") == 0)
        return FP_OK;
    if (v && std::strcmp(v, "tiled") == 0)
    {
        *use_tiled = true;
        return FP_OK;
    }
    TileGeom geo{};
    TileTables *tt = nullptr;
    fp_status const st = apply_tiled_plan(ops, dtype, n_states, kApplyTiledMinBytes, kind, &geo, &tt);
    *use_tiled = tt != nullptr;
    return st;
}

fp_status apply_tiled_attributes()
{
    static std::once_flag attr_once;
    static hipError_t attr_err = hipSuccess;
    std::call_once(attr_once, [] {
        void const *fns[] = {reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<double, false, false, false>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<double, true, false, false>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<float, false, false, false>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<float, true, false, false>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<float, false, true, false>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<float, true, true, false>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<double, false, false, true>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<double, true, false, true>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<float, false, false, true>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<float, true, false, true>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<float, false, true, true>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<float, true, true, true>)};
        for (void const *f : fns)
        {
            hipError_t const e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(fpa::kTileBytes));
            if (e != hipSuccess && attr_err == hipSuccess)
                attr_err = e;
        }
    });
    if (attr_err != hipSuccess)
        return fail(FP_ERR_RUNTIME, "hipFuncSetAttribute failed: %s", hipGetErrorString(attr_err));
    return FP_OK;
}

fp_status launch_op_apply_tiled(TileTables const *tt, TileGeom const &geo, int dtype, void const *states, void *out,
                                size_t dim, size_t n_states, int accumulate, hipStream_t st, void const *weights)
{
    uint64_t const n_cosets = uint64_t(dim) >> geo.d;
    uint32_t const chunks_y = uint32_t((n_states + (size_t(1) << geo.lsb) - 1) >> geo.lsb);
    if (chunks_y > 65535u)
        return fail(FP_ERR_INVALID_ARGUMENT, "too many states for the tiled apply kernel");
    // About four tiles per block (a block with one tile has nothing to overlap its loads
    // with), between 1024 and 8192 blocks; but complex128 launches of at most 8192 tiles
    // take one tile per block, which the hardware deals to CUs as they free up
    // (SummedPauliOp::apply at n = 14, 1000 states, 4000 tiles per group launch: 28.2 ms
    // with 1125 blocks of 3-4 tiles, 25.2 ms with 500 blocks of 8, 24.4 ms with one each;
    // the complex64 group launches of a batch's 64-118 MiB entries, about 2048 tiles each,
    // lose 2 ms with one tile per block).  FAST_PAULI_TILE_BLOCKS overrides.
    uint64_t const tiles = n_cosets * chunks_y;
    uint64_t const dflt = dtype == FP_C128 && tiles <= 8192 ? tiles : std::clamp<uint64_t>(tiles / 4, 1024, 8192);
    uint64_t const target = env_u32("FAST_PAULI_TILE_BLOCKS", uint32_t(dflt));
    uint64_t const blocks_x = std::max<uint64_t>(1, std::min<uint64_t>(n_cosets, (target + chunks_y - 1) / chunks_y));
    uint32_t const sshift = (n_states & (n_states - 1)) == 0 ? log2u(n_states) : 64u;
    bool const wide = (uint64_t(dim) * n_states) >> 32 != 0 || env_u32("FAST_PAULI_TILE_WIDE", 0) == 1;
    // complex64 full tiles with an even state count: two states per slot (16-byte
    // loads, stores and LDS reads); a slot's two states are both live or both not
    bool const aligned16 = ((reinterpret_cast<uintptr_t>(states) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0;
    bool const paired = dtype == FP_C64 && geo.lsb >= 1 && n_states % 2 == 0 && aligned16 &&
                        geo.d + geo.lsb == log2u(fpa::kTileBytes / 8u);
    fp_status const as = apply_tiled_attributes();
    if (as != FP_OK)
        return as;
    dim3 const grid(uint32_t(blocks_x), chunks_y);
    auto const *strs = static_cast<fpa::ApplyStr const *>(tt->strs);
    bool const wtd = weights != nullptr;
    for (uint32_t g = 0; g < tt->n_groups; ++g)
    {
        int const init_zero = (g == 0 && !accumulate) ? 1 : 0;
        if (dtype == FP_C128)
        {
            typedef fpk::VecOf<double, 1>::type V;
            auto const kfn = wtd ? (wide ? &fpa::k_op_apply_tiled<double, true, false, true>
                                         : &fpa::k_op_apply_tiled<double, false, false, true>)
                                 : (wide ? &fpa::k_op_apply_tiled<double, true, false, false>
                                         : &fpa::k_op_apply_tiled<double, false, false, false>);
            hipLaunchKernelGGL(kfn, grid, dim3(fpa::kThreads), fpa::kTileBytes, st, static_cast<V const *>(states),
                               static_cast<V *>(out), uint64_t(n_states), sshift, geo.lsb, geo.k, geo.d, n_cosets,
                               tt->groups + g, tt->btab, strs, init_zero, static_cast<double const *>(weights));
        }
        else
        {
            typedef fpk::VecOf<float, 1>::type V;
            auto const kfn =
                wtd ? (paired ? (wide ? &fpa::k_op_apply_tiled<float, true, true, true>
                                      : &fpa::k_op_apply_tiled<float, false, true, true>)
                              : (wide ? &fpa::k_op_apply_tiled<float, true, false, true>
                                      : &fpa::k_op_apply_tiled<float, false, false, true>))
                    : (paired ? (wide ? &fpa::k_op_apply_tiled<float, true, true, false>
                                      : &fpa::k_op_apply_tiled<float, false, true, false>)
                              : (wide ? &fpa::k_op_apply_tiled<float, true, false, false>
                                      : &fpa::k_op_apply_tiled<float, false, false, false>));
            hipLaunchKernelGGL(kfn, grid, dim3(fpa::kThreads), fpa::kTileBytes, st, static_cast<V const *>(states),
                               static_cast<V *>(out), uint64_t(n_states), sshift, geo.lsb, geo.k, geo.d, n_cosets,
                               tt->groups + g, tt->btab, strs, init_zero, static_cast<float const *>(weights));
        }
        FP_LAUNCHED();
    }
    return FP_OK;
}

fp_status run_op_apply_tiled(fp_hip_ops const *ops, int dtype, void const *states, void *out, size_t dim,
                             size_t n_states, int accumulate, hipStream_t st, int kind, void const *weights)
{
    if ((kind == kTilesWeightedApply) != (weights != nullptr))
        return fail(FP_ERR_INVALID_ARGUMENT, "weights go with the weighted tables only");
    TileGeom const geo = tile_geom(dtype, ops->n_qubits, n_states, false);
    TileTables *tt = nullptr;
    fp_status s = tile_tables(ops, kind, geo.k, geo.dh, &tt);
    if (s != FP_OK)
        return s;
    if ((s = check_tile_fit(tt, geo, dtype, fpa::kTileBytes)) != FP_OK)
        return s;
    if ((s = tile_device(ops, tt, st)) != FP_OK)
        return s;
    return launch_op_apply_tiled(tt, geo, dtype, states, out, dim, n_states, accumulate, st, weights);
}

// ---------------------------------------------------------------------------
// Per-string EV matrix on cosets (columns larger than LDS)
// ---------------------------------------------------------------------------
namespace
{

// Smallest transform class of the complex128 coset kernel (see wht_min_strings).
// Measured at n = 14, 2000 weight-3 strings, 1000 states (stage 1, one box): 2.21 ms
// with 24, 2.00 with 28, 2.01 with 32, 2.05 with 40, 2.10 with 48, 2.14 with 64; all
// per-string: 2.35 ms.
constexpr uint32_t kWhtMinCosetC128 = 32;

struct CosetGeom
{
    uint32_t d, k, n_cos;
};

CosetGeom coset_geom(int dtype, size_t n_qubits)
{
    size_t const esz = dtype == FP_C128 ? 16 : 8;
    CosetGeom g{};
    g.d = log2u(column_lds_bytes() / esz);     // 13 (complex128) or 14 (complex64) with 160 KiB
    g.k = dtype == FP_C128 ? 3u : 4u;          // 128-byte runs of consecutive rows
    g.n_cos = n_qubits > g.d && n_qubits - g.d < 31 ? 1u << (n_qubits - g.d) : 0u;
    return g;
}

size_t round256(size_t v)
{
    return (v + 255) / 256 * 256;
}

} // namespace

bool ev_coset_eligible(fp_hip_ops const *ops, int dtype, size_t dim, size_t n_states)
{
    char const *v = std::getenv("FAST_PAULI_EV_MATRIX_KERNEL");
    if (v && std::strcmp(v, "stream") == 0)
        return false;
    CosetGeom const g = coset_geom(dtype, ops->n_qubits);
    if (g.n_cos == 0 || g.n_cos > 16 || (size_t(1) << ops->n_qubits) != dim)
        return false;
    return double(g.n_cos) * double(ops->n_strings) * double(n_states) * 8.0 <= double(size_t(1) << 32) &&
           n_states <= 0x7fffffffu;
}

size_t ev_coset_scratch_bytes(fp_hip_ops const *ops, int dtype, size_t dim, size_t n_states)
{
    CosetGeom const g = coset_geom(dtype, ops->n_qubits);
    return round256(dim * n_states * (dtype == FP_C128 ? 16 : 8)) +
           size_t(g.n_cos) * ops->n_strings * n_states * sizeof(double);
}

template <class T>
fp_status launch_ev_coset(fp_hip_ops const *ops, TileTables *tt, CosetGeom const &g, void const *states,
                          void *ev, size_t dim, size_t n_states, void *scratch, hipStream_t st)
{
    typedef typename fpk::VecOf<T, 1>::type V;
    V *transposed = static_cast<V *>(scratch);
    auto *partial =
        reinterpret_cast<double *>(static_cast<unsigned char *>(scratch) + round256(dim * n_states * sizeof(V)));
    dim3 const tg(uint32_t((n_states + 31) / 32), uint32_t((dim + 31) / 32));
    hipLaunchKernelGGL(fpl::k_transpose<V>, tg, dim3(256), 0, st, static_cast<V const *>(states), transposed,
                       uint64_t(dim), uint64_t(n_states));
    FP_LAUNCHED();
    static std::once_flag once;
    static hipError_t attr = hipSuccess;
    std::call_once(once, [] {
        attr = hipFuncSetAttribute(reinterpret_cast<void const *>(&fpl::k_ev_coset<T>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(column_lds_bytes()));
    });
    if (attr != hipSuccess)
        return fail(FP_ERR_RUNTIME, "hipFuncSetAttribute failed: %s", hipGetErrorString(attr));
    uint32_t const gy = ev_matrix_split(n_states, g.n_cos, tt->max_chunks);
    bool any_wht = false;
    auto const *hgroups = reinterpret_cast<fpt::TileGroup const *>(tt->h.data());
    for (uint32_t i = 0; i < tt->n_groups; ++i)
        any_wht = any_wht || hgroups[i].n_wht > 0;
    size_t const lds = (size_t(1) << g.d) * sizeof(V) + (any_wht ? fpl::kWhtLdsBytes : 0);
    if (lds > column_lds_bytes())
        return fail(FP_ERR_RUNTIME, "coset column and transform buffers (%zu bytes) exceed the LDS", lds);
    dim3 const grid(uint32_t(n_states), gy, tt->n_groups * g.n_cos);
    hipLaunchKernelGGL(fpl::k_ev_coset<T>, grid, dim3(fpl::kThreads), lds, st, transposed, uint64_t(dim), g.d, g.k,
                       g.n_cos, tt->groups, tt->btab, static_cast<fpl::EvChunk const *>(tt->chunks),
                       static_cast<fpl::EvStr const *>(tt->strs), uint64_t(n_states), uint64_t(ops->n_strings),
                       partial);
    FP_LAUNCHED();
    uint64_t const total = uint64_t(ops->n_strings) * n_states;
    hipLaunchKernelGGL(fpl::k_sum_cosets<T>, dim3(uint32_t((total + 255) / 256)), dim3(256), 0, st, partial, g.n_cos,
                       total, static_cast<T *>(ev));
    FP_LAUNCHED();
    return FP_OK;
}

fp_status run_ev_coset(fp_hip_ops const *ops, int dtype, void const *states, void *ev, size_t dim, size_t n_states,
                       void *scratch, hipStream_t st)
{
    CosetGeom const g = coset_geom(dtype, ops->n_qubits);
    TileTables *tt = nullptr;
    fp_status s = tile_tables(ops, kTilesEvMatrix, g.k, g.d - g.k, &tt,
                              wht_min_strings(dtype == FP_C128 ? kWhtMinCosetC128 : fpl::kWhtMin));
    if (s != FP_OK)
        return s;
    if ((s = tile_device(ops, tt, st)) != FP_OK)
        return s;
    if (tt->n_groups * g.n_cos > 65535u)
        return fail(FP_ERR_INVALID_ARGUMENT, "too many string groups for the coset kernel");
    if (dtype == FP_C128)
        return launch_ev_coset<double>(ops, tt, g, states, ev, dim, n_states, scratch, st);
    return launch_ev_coset<float>(ops, tt, g, states, ev, dim, n_states, scratch, st);
}

} // namespace fprt

extern "C" fp_status fp_hip_op_expectation_plan(fp_hip_ops const *ops, int dtype, size_t dim, size_t n_states,
                                                fp_hip_ev_plan *plan)
{
    if (!ops || !plan)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL ops or plan");
    if (dtype != FP_C128 && dtype != FP_C64)
        return fail(FP_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    if (n_states == 0 || dim != (size_t(1) << ops->n_qubits))
        return fail(FP_ERR_INVALID_ARGUMENT, "shape (%zu, %zu) does not match %zu qubits", dim, n_states,
                    ops->n_qubits);
    bool tiled = false;
    fp_status st = fprt::choose_op_expect_tiled(ops, dtype, n_states, &tiled);
    if (st != FP_OK)
        return st;
    uint64_t const state_bytes = uint64_t(dim) * n_states * (dtype == FP_C128 ? 16u : 8u);
    *plan = fp_hip_ev_plan{};
    if (!tiled)
    {
        plan->kernel = FP_EV_STREAM;
        plan->n_groups = uint32_t(ops->n_strings);
        plan->hbm_bytes = state_bytes * ops->n_strings;
        return FP_OK;
    }
    fprt::TileGeom const geo = fprt::tile_geom(dtype, ops->n_qubits, n_states, true);
    fprt::TileTables *tt = nullptr;
    if ((st = fprt::tile_tables(ops, fprt::kTilesOpExpect, geo.k, geo.dh, &tt)) != FP_OK)
        return st;
    plan->kernel = FP_EV_TILED;
    plan->n_groups = tt->n_groups;
    plan->tile_rows_log2 = geo.d;
    plan->states_per_block = 1u << geo.lsb;
    plan->hbm_bytes = state_bytes * tt->n_groups;
    auto const *groups = reinterpret_cast<fpt::TileGroup const *>(tt->h.data());
    for (uint32_t g = 0; g < tt->n_groups; ++g)
    {
        int const mode = fprt::ev_group_mode(groups[g], dtype, geo, n_states);
        plan->hoisted_groups += mode == 1 ? 1u : 0u;
        plan->shared_groups += mode == 2 ? 1u : 0u;
    }
    return FP_OK;
}

namespace
{

fp_status apply_plan(fp_hip_ops const *ops, int dtype, size_t dim, size_t n_states, int accumulate, int kind,
                     fp_hip_ev_plan *plan)
{
    if (!ops || !plan)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL ops or plan");
    if (dtype != FP_C128 && dtype != FP_C64)
        return fail(FP_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    if (n_states == 0 || dim != (size_t(1) << ops->n_qubits))
        return fail(FP_ERR_INVALID_ARGUMENT, "shape (%zu, %zu) does not match %zu qubits", dim, n_states,
                    ops->n_qubits);
    if (kind != fprt::kTilesOpApply && ops->n_operators == 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "SummedPauliOp has no coefficient matrix");
    bool tiled = false;
    fp_status st = fprt::choose_op_apply_tiled(ops, dtype, n_states, &tiled, kind);
    if (st != FP_OK)
        return st;
    uint64_t const state_bytes = uint64_t(dim) * n_states * (dtype == FP_C128 ? 16u : 8u);
    uint64_t const out_reads = accumulate ? 1u : 0u;
    *plan = fp_hip_ev_plan{};
    if (!tiled)
    {
        plan->kernel = FP_EV_STREAM;
        plan->n_groups = uint32_t(ops->n_strings);
        plan->hbm_bytes = state_bytes * (ops->n_strings + 1u + out_reads);
        return FP_OK;
    }
    fprt::TileGeom const geo = fprt::tile_geom(dtype, ops->n_qubits, n_states, false);
    fprt::TileTables *tt = nullptr;
    if ((st = fprt::tile_tables(ops, kind, geo.k, geo.dh, &tt)) != FP_OK)
        return st;
    uint64_t const G = tt->n_groups;
    plan->kernel = FP_EV_TILED;
    plan->n_groups = uint32_t(G);
    plan->tile_rows_log2 = geo.d;
    plan->states_per_block = 1u << geo.lsb;
    plan->hbm_bytes = state_bytes * (3u * G - 1u + out_reads);
    return FP_OK;
}

} // namespace

extern "C" fp_status fp_hip_op_apply_plan(fp_hip_ops const *ops, int dtype, size_t dim, size_t n_states, int accumulate,
                                          fp_hip_ev_plan *plan)
{
    return apply_plan(ops, dtype, dim, n_states, accumulate, fprt::kTilesOpApply, plan);
}

extern "C" fp_status fp_hip_summed_apply_plan(fp_hip_ops const *ops, int dtype, size_t dim, size_t n_states,
                                              int accumulate, fp_hip_ev_plan *plan)
{
    return apply_plan(ops, dtype, dim, n_states, accumulate, fprt::kTilesSummedApply, plan);
}

extern "C" fp_status fp_hip_summed_apply_weighted_plan(fp_hip_ops const *ops, int dtype, size_t dim,
                                                       size_t n_states, int accumulate, fp_hip_ev_plan *plan)
{
    return apply_plan(ops, dtype, dim, n_states, accumulate, fprt::kTilesWeightedApply, plan);
}
