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

uint32_t parity(uint64_t v)
{
    return uint32_t(__builtin_popcountll(v) & 1);
}

// kappa of the pair identity: 1 for x = 0, else 2 eps, eps = +1 for nY = 0, 1 (mod 4), -1 otherwise
double kappa_of(uint64_t x, uint32_t ny)
{
    if (x == 0)
        return 1.0;
    return ((ny & 3u) <= 1u) ? 2.0 : -2.0;
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

TileGeom tile_geom(int dtype, size_t n_qubits, size_t n_states, bool expect)
{
    uint32_t const esz = dtype == FP_C128 ? 16u : 8u;
    // complex128 tiles with a power-of-two state count take 128-byte row slices
    // (8 states): 2^9-row tiles, so groups of 9 strings (in the shared-leader form
    // for expectation values), against 8 strings with 256-byte rows; the slices of
    // a row are fetched by blocks of one XCD.  Apply gains the same way (256 random
    // strings: 29 groups instead of 32).  With other state counts the slices would
    // not align with rows.
    bool const pow2 = n_states != 0 && (n_states & (n_states - 1)) == 0;
    uint32_t const row_target = env_u32("FAST_PAULI_TILE_ROW_BYTES", dtype == FP_C128 && pow2 ? 128u : 256u);
    // 128 KiB tiles (FAST_PAULI_TILE_KB=128) exist for the expectation kernel only; the
    // tiled apply kernel's tiles are 64 KiB
    uint32_t const tile_kb = expect && env_u32("FAST_PAULI_TILE_KB", 64u) >= 128u ? 128u : 64u;
    uint32_t const elems = tile_kb * 1024u / esz;
    uint32_t sblk = 1;
    while (sblk < n_states && sblk * esz < row_target)
        sblk *= 2;
    TileGeom g{};
    g.tile_kb = tile_kb;
    g.lsb = log2u(sblk);
    g.d = std::min<uint32_t>(log2u(elems / sblk), uint32_t(n_qubits));
    g.k = 0;
    if (sblk >= n_states) // one block covers every state: 2^k consecutive rows are contiguous
        while ((uint64_t(1) << g.k) * n_states * esz < 256 && g.k < g.d)
            ++g.k;
    // FAST_PAULI_TILE_K: at least this many low row bits in every tile (longer contiguous segments)
    uint32_t const kmin = std::min<uint32_t>(env_u32("FAST_PAULI_TILE_K", 0), g.d);
    if (sblk >= n_states)
        g.k = std::max(g.k, kmin);
    g.dh = g.d - g.k;
    if (g.dh > uint32_t(fpt::kMaxDh))
    {
        g.dh = fpt::kMaxDh;
        g.k = g.d - g.dh;
    }
    return g;
}

namespace
{

struct LocalStr
{
    uint32_t s;  // string index
    uint32_t A;  // local x mask
    uint32_t Z;  // local z mask
    uint64_t zc; // z compressed onto the coset-index bits
};

struct HostGroup
{
    std::vector<uint64_t> b;   // basis, ascending pivot order
    std::vector<uint32_t> piv; // pivots, ascending
    std::vector<LocalStr> mem; // members, sorted by local x mask
};

// Greedy grouping of the strings of `ops` into subspaces of dimension k + dh
// (the low k bits plus dh reduced basis vectors), in x order.
fp_status group_strings(fp_hip_ops const *ops, uint32_t k, uint32_t dh, bool ordered, std::vector<HostGroup> &out)
{
    size_t const N = ops->n_strings, n = ops->n_qubits;
    uint64_t const lowmask = (uint64_t(1) << k) - 1;
    if (k + dh > n)
        return fail(FP_ERR_RUNTIME, "tile geometry k=%u dh=%u exceeds %zu qubits", k, dh, n);
    std::vector<uint32_t> order(N);
    for (size_t i = 0; i < N; ++i)
        order[i] = uint32_t(i);
    if (!ordered) // x order groups related strings; apply keeps the operator's order
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return ops->x[a] < ops->x[b]; });
    if (dh == 0)
        for (uint32_t s : order)
            if (ops->x[s] & ~lowmask)
                return fail(FP_ERR_RUNTIME, "tile geometry has no room for x mask bits above bit %u", k);
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
        std::vector<size_t> idx(span.b.size());
        for (size_t i = 0; i < idx.size(); ++i)
            idx[i] = i;
        std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return span.piv[a] < span.piv[b]; });
        HostGroup g;
        uint64_t pivmask = 0;
        for (size_t i : idx)
        {
            g.b.push_back(span.b[i]);
            g.piv.push_back(span.piv[i]);
            pivmask |= uint64_t(1) << span.piv[i];
        }
        // coset-index bits: positions >= k that are not pivots, ascending
        std::vector<uint32_t> cpos;
        for (uint32_t i = k; i < n; ++i)
            if (!((pivmask >> i) & 1u))
                cpos.push_back(i);
        for (uint32_t s : members)
        {
            uint64_t const xh = ops->x[s] & ~lowmask, z = ops->z[s];
            uint32_t alpha = 0, zeta = 0;
            uint64_t zc = 0;
            for (size_t i = 0; i < g.b.size(); ++i)
            {
                alpha |= uint32_t((xh >> g.piv[i]) & 1u) << i;
                zeta |= parity(g.b[i] & z) << i;
            }
            for (size_t i = 0; i < cpos.size(); ++i)
                zc |= ((z >> cpos[i]) & 1u) << i;
            g.mem.push_back(LocalStr{s, (alpha << k) | uint32_t(ops->x[s] & lowmask),
                                     (zeta << k) | uint32_t(z & lowmask), zc});
        }
        if (!ordered)
            std::stable_sort(g.mem.begin(), g.mem.end(), [](LocalStr const &a, LocalStr const &b) { return a.A < b.A; });
        out.push_back(std::move(g));
        span = Span{};
        members.clear();
    };
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
    for (auto const &g : out)
        if (g.b.size() != dh)
            return fail(FP_ERR_RUNTIME, "string group has dimension %zu, expected %u", g.b.size(), dh);
    return FP_OK;
}

// TileGroup::hoist of an expectation group of dimension d = k + dh: for each
// block shape (W waves, U1 passes per fast-path step, 2^r leader rows per pass),
// whether every wave owns at most one unit and that unit takes the one-string
// fast path (the unit split of k_op_expect_tiled, restated).
uint32_t hoist_bits(fpt::TileChunk const *ch, uint32_t n_chunks, uint32_t d)
{
    uint32_t bits = 0;
    for (uint32_t W : {8u, 16u})
        for (uint32_t U1 : {4u, 8u})
            for (uint32_t r = 0; r <= 6; ++r)
            {
                uint32_t split_log2 = 0;
                while ((n_chunks << (split_log2 + 1)) <= W && split_log2 < 6)
                    ++split_log2;
                bool ok = (n_chunks << split_log2) <= W;
                uint32_t const split = 1u << split_log2;
                for (uint32_t c = 0; ok && c < n_chunks; ++c)
                {
                    uint32_t const n_rows = ch[c].A ? (1u << (d - 1)) : (1u << d);
                    uint32_t const n_it = std::max(1u, n_rows >> r);
                    ok = ch[c].count == 1 && n_it % (split * U1) == 0;
                }
                if (ok)
                    bits |= 1u << fpt::hoist_bit(W, U1, r);
            }
    return bits;
}

// Shared leaders for an expectation group.  Local coordinates may use any basis
// of V (the coset structure depends on V only): with generators g_0 .. g_m, the
// strings' x masks (above the low k bits) that first raise the span, the basis
// g_0, g_1 ^ g_0, ..., g_m ^ g_0 (padded with vectors of the former basis) gives
// every string whose x is the sum of an odd number of generators local bit k
// set.  When that holds for every string of the group (at most kMaxShared of
// them), the group switches to this basis and returns true: the rows with local
// bit k clear are then leaders for all of its strings.
bool shared_basis(HostGroup &g, fp_hip_ops const *ops, uint32_t k, uint32_t dh)
{
    if (g.mem.empty() || g.mem.size() > size_t(fpt::kMaxShared) || dh == 0)
        return false;
    uint64_t const lowmask = (uint64_t(1) << k) - 1;
    // elimination over the new basis, tracking which basis vectors make up each row
    std::vector<uint64_t> nb, red, comb;
    std::vector<uint32_t> rpiv;
    auto reduce = [&](uint64_t v, uint64_t &c) {
        c = 0;
        for (size_t i = 0; i < red.size(); ++i)
            if ((v >> rpiv[i]) & 1u)
            {
                v ^= red[i];
                c ^= comb[i];
            }
        return v;
    };
    auto push = [&](uint64_t v) { // v independent of nb
        uint64_t c = 0;
        uint64_t const r = reduce(v, c);
        nb.push_back(v);
        red.push_back(r);
        comb.push_back(c ^ (uint64_t(1) << (nb.size() - 1)));
        rpiv.push_back(63u - uint32_t(__builtin_clzll(r)));
    };
    uint64_t g0 = 0;
    for (LocalStr const &m : g.mem)
    {
        uint64_t const xh = ops->x[m.s] & ~lowmask;
        uint64_t c = 0;
        if (!xh || !reduce(xh, c))
            continue;
        if (nb.empty())
        {
            g0 = xh;
            push(xh);
        }
        else
            push(xh ^ g0);
    }
    if (nb.empty())
        return false;
    for (uint64_t b : g.b) // pad with the former basis
    {
        uint64_t c = 0;
        if (nb.size() < dh && reduce(b, c))
            push(b);
    }
    if (nb.size() != dh)
        return false;
    // coordinates of every string in the new basis; bit 0 must be set
    std::vector<uint32_t> alpha(g.mem.size());
    for (size_t i = 0; i < g.mem.size(); ++i)
    {
        uint64_t c = 0;
        uint64_t const rest = reduce(ops->x[g.mem[i].s] & ~lowmask, c);
        if (rest != 0 || !(c & 1u))
            return false;
        alpha[i] = uint32_t(c);
    }
    g.b = nb;
    for (size_t i = 0; i < g.mem.size(); ++i)
    {
        LocalStr &m = g.mem[i];
        uint64_t const z = ops->z[m.s];
        uint32_t zeta = 0;
        for (size_t j = 0; j < nb.size(); ++j)
            zeta |= parity(nb[j] & z) << j;
        m.A = (alpha[i] << k) | uint32_t(ops->x[m.s] & lowmask);
        m.Z = (zeta << k) | uint32_t(z & lowmask);
    }
    std::stable_sort(g.mem.begin(), g.mem.end(), [](LocalStr const &a, LocalStr const &b) { return a.A < b.A; });
    return true;
}

// TileGroup::shared of a shared-leader group of dimension d: the shared form
// needs full 64 or 128 KiB tiles (2^(d-1) leader rows = 4 passes per wave of
// 2^rr rows), the leader bit k inside the lane part (k < rr) and at most
// max_shared(W) strings (the instance's per-string registers).  It is used
// for complex128 only: complex64 tiles already evaluate at about their load
// time in the set-up-once form (the shared form measured no faster there).
uint32_t shared_bits(uint32_t d, uint32_t k, size_t n_str)
{
    uint32_t bits = 0;
    for (uint32_t W : {8u, 16u})
        for (uint32_t rr = 0; rr <= 6; ++rr)
            if (k < rr && d >= rr + 1 && (uint64_t(1) << (d - 1 - rr)) == 4u * W &&
                n_str <= size_t(fpt::max_shared(W)))
                bits |= 1u << fpt::shared_bit(W, false, rr);
    return bits;
}

fpt::TileGroup group_header(HostGroup const &hg, std::vector<uint64_t> &btab)
{
    fpt::TileGroup g{};
    g.n_piv = uint32_t(hg.b.size());
    g.btab_off = uint32_t(btab.size());
    for (size_t i = 0; i < hg.piv.size(); ++i)
        (i < 8 ? g.piv_lo : g.piv_hi) |= uint64_t(hg.piv[i]) << (8 * (i & 7));
    size_t const n_bt = size_t(1) << g.n_piv;
    size_t const base = btab.size();
    btab.resize(base + n_bt, 0);
    for (size_t a = 1; a < n_bt; ++a)
        btab[base + a] = btab[base + (a & (a - 1))] ^ hg.b[__builtin_ctzll(a)];
    return g;
}

} // namespace

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
    std::vector<HostGroup> hgs;
    fp_status const st = group_strings(ops, k, dh, kind == kTilesOpApply, hgs);
    if (st != FP_OK)
        return st;
    std::vector<fpt::TileGroup> groups;
    std::vector<uint64_t> btab;
    std::vector<unsigned char> chunks, strs; // raw bytes of the kind's chunk / string records
    std::vector<fpt::ShStr> sh;              // shared-leader records (expectation tables)
    size_t n_chunks = 0;
    for (auto &hg : hgs)
    {
        bool const shared = kind == kTilesOpExpect && shared_basis(hg, ops, k, dh);
        fpt::TileGroup g = group_header(hg, btab);
        g.chunk_first = uint32_t(n_chunks);
        if (kind == kTilesOpExpect)
        {
            // the shared-leader instance keeps coset indices in 32 bits
            g.shared = shared && ops->n_qubits - (k + dh) <= 32 ? shared_bits(k + dh, k, hg.mem.size()) : 0u;
            g.str_first = uint32_t(sh.size());
            g.n_str = uint32_t(hg.mem.size());
            for (LocalStr const &m : hg.mem)
            {
                double const kap = kappa_of(ops->x[m.s], ops->ny[m.s]);
                sh.push_back(fpt::ShStr{m.zc, m.A, m.Z, ops->ny[m.s] & 1u, 0u, ops->coeffs[2 * m.s] * kap,
                                        ops->coeffs[2 * m.s + 1] * kap});
            }
        }
        if (kind == kTilesOpApply)
        {
            // one record per string, in the operator's order; n_chunks counts them
            for (LocalStr const &m : hg.mem)
            {
                fpa::ApplyStr r{};
                r.zc = m.zc;
                r.A = m.A;
                r.Z = m.Z;
                r.cr = ops->cf128[2 * m.s];
                r.ci = ops->cf128[2 * m.s + 1];
                r.fr = ops->cf64[2 * m.s];
                r.fi = ops->cf64[2 * m.s + 1];
                strs.insert(strs.end(), reinterpret_cast<unsigned char const *>(&r),
                            reinterpret_cast<unsigned char const *>(&r) + sizeof(r));
                ++n_chunks;
            }
        }
        else if (kind == kTilesOpExpect)
        {
            for (size_t i = 0; i < hg.mem.size();)
            {
                size_t j = i;
                while (j < hg.mem.size() && j - i < size_t(fpt::kChunk) && hg.mem[j].A == hg.mem[i].A)
                    ++j;
                fpt::TileChunk c{};
                c.A = hg.mem[i].A;
                c.h = c.A ? 31u - uint32_t(__builtin_clz(c.A)) : 0u;
                c.count = uint32_t(j - i);
                c.first = uint32_t(strs.size() / sizeof(fpt::TileStr));
                for (size_t m = i; m < j; ++m)
                {
                    uint32_t const sidx = hg.mem[m].s;
                    uint32_t const Z = hg.mem[m].Z;
                    fpt::TileStr r{};
                    r.zc = hg.mem[m].zc;
                    // leader p maps to a = insert_zero_bit(p, h): drop bit h of Z
                    r.zt = c.A ? ((Z & ((1u << c.h) - 1u)) | ((Z >> (c.h + 1)) << c.h)) : Z;
                    r.imag = ops->ny[sidx] & 1u;
                    c.need |= r.imag ? 2u : 1u;
                    double const kap = kappa_of(ops->x[sidx], ops->ny[sidx]);
                    r.sre = ops->coeffs[2 * sidx] * kap;
                    r.sim = ops->coeffs[2 * sidx + 1] * kap;
                    strs.insert(strs.end(), reinterpret_cast<unsigned char const *>(&r),
                                reinterpret_cast<unsigned char const *>(&r) + sizeof(r));
                }
                chunks.insert(chunks.end(), reinterpret_cast<unsigned char const *>(&c),
                              reinterpret_cast<unsigned char const *>(&c) + sizeof(c));
                ++n_chunks;
                i = j;
            }
        }
        else
        {
            // per-string EV matrix: chunks of <= fpl::kChunk strings with equal local x
            // and equal half of w, costliest first (the LDS kernel's chunk records)
            struct Pending
            {
                fpl::EvChunk c;
                std::vector<fpl::EvStr> s;
                double cost;
                bool wht;
            };
            std::vector<Pending> pend;
            std::vector<LocalStr> mem = hg.mem;
            auto key = [&](LocalStr const &m) { return std::make_pair(m.A, ops->ny[m.s] & 1u); };
            std::stable_sort(mem.begin(), mem.end(),
                             [&](LocalStr const &a, LocalStr const &b) { return key(a) < key(b); });
            for (size_t i = 0; i < mem.size();)
            {
                // classes of at least kWhtMin strings become transform classes
                size_t cls = i;
                while (cls < mem.size() && key(mem[cls]) == key(mem[i]))
                    ++cls;
                bool const wht = cls - i >= wht_min && k + dh >= fpl::kWhtMinRowsLog2;
                size_t const per_chunk = wht ? fpl::kWhtMax : size_t(fpl::kChunk);
                size_t j = i;
                while (j < mem.size() && j - i < per_chunk && key(mem[j]) == key(mem[i]))
                    ++j;
                Pending p{};
                p.wht = wht;
                p.c.x = mem[i].A;
                p.c.h = p.c.x ? 31u - uint32_t(__builtin_clz(p.c.x)) : 0u;
                p.c.mode = fpl::ev_mode(p.c.x, p.c.h, uint64_t(1) << (k + dh));
                p.c.imag = ops->ny[mem[i].s] & 1u;
                p.c.count = uint32_t(j - i);
                for (size_t m = i; m < j; ++m)
                {
                    uint32_t const sidx = mem[m].s;
                    uint32_t const ny = ops->ny[sidx];
                    double const eps = ((ny & 3u) <= 1u) ? 1.0 : -1.0;
                    p.s.push_back(fpl::EvStr{mem[m].Z, ny & 1u, sidx, uint32_t(mem[m].zc),
                                             eps * (p.c.mode == fpl::kAllRows ? 1.0 : 2.0)});
                }
                p.cost = double(p.c.count + 2) * (p.c.mode == fpl::kAllRows ? 2.0 : 1.0);
                pend.push_back(std::move(p));
                i = j;
            }
            // transform classes first, then per-string chunks, each costliest first
            std::stable_sort(pend.begin(), pend.end(), [](Pending const &a, Pending const &b) {
                return a.wht != b.wht ? a.wht : a.cost > b.cost;
            });
            for (auto &p : pend)
            {
                g.n_wht += p.wht ? 1u : 0u;
                p.c.first = uint32_t(strs.size() / sizeof(fpl::EvStr));
                for (auto const &e : p.s)
                    strs.insert(strs.end(), reinterpret_cast<unsigned char const *>(&e),
                                reinterpret_cast<unsigned char const *>(&e) + sizeof(e));
                chunks.insert(chunks.end(), reinterpret_cast<unsigned char const *>(&p.c),
                              reinterpret_cast<unsigned char const *>(&p.c) + sizeof(p.c));
                ++n_chunks;
            }
        }
        g.n_chunks = uint32_t(n_chunks) - g.chunk_first;
        if (kind == kTilesOpExpect)
            g.hoist = hoist_bits(reinterpret_cast<fpt::TileChunk const *>(chunks.data()) + g.chunk_first, g.n_chunks,
                                 k + dh);
        groups.push_back(g);
    }
    uint32_t max_chunks = 0;
    for (auto const &g : groups)
        max_chunks = std::max(max_chunks, g.n_chunks);

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
    t->n_groups = uint32_t(groups.size());
    t->max_chunks = max_chunks;
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    t->o_btab = up(groups.size() * sizeof(fpt::TileGroup));
    t->o_chunks = t->o_btab + up(btab.size() * sizeof(uint64_t));
    t->o_strs = t->o_chunks + up(chunks.size());
    t->o_sh = t->o_strs + up(strs.size());
    size_t const total = t->o_sh + up(sh.size() * sizeof(fpt::ShStr));
    t->h.assign(total, 0);
    std::memcpy(t->h.data(), groups.data(), groups.size() * sizeof(fpt::TileGroup));
    std::memcpy(t->h.data() + t->o_btab, btab.data(), btab.size() * sizeof(uint64_t));
    std::memcpy(t->h.data() + t->o_chunks, chunks.data(), chunks.size());
    std::memcpy(t->h.data() + t->o_strs, strs.data(), strs.size());
    if (!sh.empty())
        std::memcpy(t->h.data() + t->o_sh, sh.data(), sh.size() * sizeof(fpt::ShStr));
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
int ev_group_mode(fpt::TileGroup const &g, int dtype, TileGeom const &geo)
{
    bool const c64 = dtype != FP_C128;
    uint32_t const W = geo.tile_kb == 128 ? 16u : 8u;
    uint32_t const rrx = (c64 ? 7u : 6u) - geo.lsb;
    bool const can_pair = !c64 || geo.lsb >= 1;
    if (can_pair && ((g.shared >> fpt::shared_bit(W, c64, rrx)) & 1u) != 0)
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
            if constexpr (sizeof(T) == 8)
            {
                set(reinterpret_cast<void const *>(&fpt::k_op_expect_tiled<T, false, TKB, 2>), TKB << 10);
                set(reinterpret_cast<void const *>(&fpt::k_op_expect_tiled<T, true, TKB, 2>), TKB << 10);
            }
        };
        set4(double{}, std::integral_constant<int, 64>{});
        set4(float{}, std::integral_constant<int, 64>{});
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
        run_mode[ev_group_mode(hgroups[g], dtype, geo)] = true;
    auto launch = [&](auto T_tag, auto wide_tag, auto tkb_tag, void *typed_out_v) -> fp_status {
        typedef decltype(T_tag) T;
        typedef typename fpk::VecOf<T, 1>::type V;
        constexpr bool WIDE = decltype(wide_tag)::value;
        constexpr int TKB = decltype(tkb_tag)::value;
        auto const *typed_states = static_cast<V const *>(states);
        // the shared-leader instance exists for complex128 only
        constexpr int kTop = sizeof(T) == 8 ? 2 : 1;
        for (int mode = 2; mode >= 0; --mode)
        {
            if (!run_mode[mode])
                continue;
            if (mode > kTop)
                return fail(FP_ERR_RUNTIME, "no shared-leader instance for this precision");
            auto const kfn = mode == 2   ? &fpt::k_op_expect_tiled<T, WIDE, TKB, kTop>
                             : mode == 1 ? &fpt::k_op_expect_tiled<T, WIDE, TKB, 1>
                                         : &fpt::k_op_expect_tiled<T, WIDE, TKB, 0>;
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

// 0: streaming kernel, 1: tiled kernel, per FAST_PAULI_APPLY_KERNEL ("stream",
// "tiled") or, by default, when the state is large (64 MiB and up: beyond the
// caches that serve the streaming kernel's repeated row reads) and the groups'
// traffic (about three state passes each) is below the strings' (one each).
fp_status choose_op_apply_tiled(fp_hip_ops const *ops, int dtype, size_t n_states, bool *use_tiled)
{
    *use_tiled = false;
    char const *v = std::getenv("FAST_PAULI_APPLY_KERNEL");
    if (v && std::strcmp(v, "stream") == 0)
        return FP_OK;
    bool const forced = v && std::strcmp(v, "tiled") == 0;
    size_t const state_bytes = (size_t(1) << ops->n_qubits) * n_states * (dtype == FP_C128 ? 16 : 8);
    if (!forced && (ops->n_strings < 8 || state_bytes < (size_t(64) << 20)))
        return FP_OK;
    TileGeom const geo = tile_geom(dtype, ops->n_qubits, n_states, false);
    TileTables *tt = nullptr;
    fp_status const st = tile_tables(ops, kTilesOpApply, geo.k, geo.dh, &tt);
    if (st != FP_OK)
        return st;
    *use_tiled = forced || 3.0 * double(tt->n_groups) <= 0.8 * double(ops->n_strings);
    return FP_OK;
}

fp_status run_op_apply_tiled(fp_hip_ops const *ops, int dtype, void const *states, void *out, size_t dim,
                             size_t n_states, int accumulate, hipStream_t st)
{
    TileGeom const geo = tile_geom(dtype, ops->n_qubits, n_states, false);
    TileTables *tt = nullptr;
    fp_status s = tile_tables(ops, kTilesOpApply, geo.k, geo.dh, &tt);
    if (s != FP_OK)
        return s;
    if ((s = check_tile_fit(tt, geo, dtype, fpa::kTileBytes)) != FP_OK)
        return s;
    if ((s = tile_device(ops, tt, st)) != FP_OK)
        return s;
    uint64_t const n_cosets = uint64_t(dim) >> geo.d;
    uint32_t const chunks_y = uint32_t((n_states + (size_t(1) << geo.lsb) - 1) >> geo.lsb);
    if (chunks_y > 65535u)
        return fail(FP_ERR_INVALID_ARGUMENT, "too many states for the tiled apply kernel");
    // about four tiles per block (a block with one tile has nothing to overlap its
    // loads with), between 1024 and 8192 blocks; FAST_PAULI_TILE_BLOCKS overrides
    uint64_t const tiles = n_cosets * chunks_y;
    uint64_t const target = env_u32("FAST_PAULI_TILE_BLOCKS", uint32_t(std::clamp<uint64_t>(tiles / 4, 1024, 8192)));
    uint64_t const blocks_x = std::max<uint64_t>(1, std::min<uint64_t>(n_cosets, (target + chunks_y - 1) / chunks_y));
    uint32_t const sshift = (n_states & (n_states - 1)) == 0 ? log2u(n_states) : 64u;
    bool const wide = (uint64_t(dim) * n_states) >> 32 != 0 || env_u32("FAST_PAULI_TILE_WIDE", 0) == 1;
    // complex64 full tiles with a power-of-two state count that one block's
    // states divide: two states per slot (16-byte loads, stores and LDS reads)
    bool const paired = dtype == FP_C64 && geo.lsb >= 1 && sshift < 64u && n_states % (size_t(1) << geo.lsb) == 0 &&
                        geo.d + geo.lsb == log2u(fpa::kTileBytes / 8u);
    static std::once_flag attr_once;
    static hipError_t attr_err = hipSuccess;
    std::call_once(attr_once, [] {
        void const *fns[] = {reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<double, false, false>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<double, true, false>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<float, false, false>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<float, true, false>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<float, false, true>),
                             reinterpret_cast<void const *>(&fpa::k_op_apply_tiled<float, true, true>)};
        for (void const *f : fns)
        {
            hipError_t const e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(fpa::kTileBytes));
            if (e != hipSuccess && attr_err == hipSuccess)
                attr_err = e;
        }
    });
    if (attr_err != hipSuccess)
        return fail(FP_ERR_RUNTIME, "hipFuncSetAttribute failed: %s", hipGetErrorString(attr_err));
    dim3 const grid(uint32_t(blocks_x), chunks_y);
    auto const *strs = static_cast<fpa::ApplyStr const *>(tt->strs);
    for (uint32_t g = 0; g < tt->n_groups; ++g)
    {
        int const init_zero = (g == 0 && !accumulate) ? 1 : 0;
        if (dtype == FP_C128)
        {
            typedef fpk::VecOf<double, 1>::type V;
            auto const kfn =
                wide ? &fpa::k_op_apply_tiled<double, true, false> : &fpa::k_op_apply_tiled<double, false, false>;
            hipLaunchKernelGGL(kfn, grid, dim3(fpa::kThreads), fpa::kTileBytes, st, static_cast<V const *>(states),
                               static_cast<V *>(out), uint64_t(n_states), sshift, geo.lsb, geo.k, geo.d, n_cosets,
                               tt->groups + g, tt->btab, strs, init_zero);
        }
        else
        {
            typedef fpk::VecOf<float, 1>::type V;
            auto const kfn = paired ? (wide ? &fpa::k_op_apply_tiled<float, true, true>
                                            : &fpa::k_op_apply_tiled<float, false, true>)
                                    : (wide ? &fpa::k_op_apply_tiled<float, true, false>
                                            : &fpa::k_op_apply_tiled<float, false, false>);
            hipLaunchKernelGGL(kfn, grid, dim3(fpa::kThreads), fpa::kTileBytes, st, static_cast<V const *>(states),
                               static_cast<V *>(out), uint64_t(n_states), sshift, geo.lsb, geo.k, geo.d, n_cosets,
                               tt->groups + g, tt->btab, strs, init_zero);
        }
        FP_LAUNCHED();
    }
    return FP_OK;
}

// ---------------------------------------------------------------------------
// Per-string EV matrix on cosets (columns larger than LDS)
// ---------------------------------------------------------------------------
namespace
{

// Smallest transform class of the complex128 coset kernel (see wht_min_strings).
constexpr uint32_t kWhtMinCosetC128 = 48;

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
    uint32_t gy = 1;
    if (n_states < 1024)
        gy = uint32_t(std::max<size_t>(1, std::min<size_t>(tt->max_chunks, (1024 + n_states - 1) / n_states)));
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
        int const mode = fprt::ev_group_mode(groups[g], dtype, geo);
        plan->hoisted_groups += mode == 1 ? 1u : 0u;
        plan->shared_groups += mode == 2 ? 1u : 0u;
    }
    return FP_OK;
}

extern "C" fp_status fp_hip_op_apply_plan(fp_hip_ops const *ops, int dtype, size_t dim, size_t n_states, int accumulate,
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
    fp_status st = fprt::choose_op_apply_tiled(ops, dtype, n_states, &tiled);
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
    if ((st = fprt::tile_tables(ops, fprt::kTilesOpApply, geo.k, geo.dh, &tt)) != FP_OK)
        return st;
    uint64_t const G = tt->n_groups;
    plan->kernel = FP_EV_TILED;
    plan->n_groups = uint32_t(G);
    plan->tile_rows_log2 = geo.d;
    plan->states_per_block = 1u << geo.lsb;
    plan->hbm_bytes = state_bytes * (3u * G - 1u + out_reads);
    return FP_OK;
}
