// fp_tiles_host.cpp — the host-side builder of the coset-tile tables: greedy
// grouping of strings into subspaces over GF(2), local coordinates, shared
// leaders, chunking, and the packed image the kernels read.  Plain C++ (no
// HIP): the library compiles it with the host compiler, and its unit test
// builds it under the address and undefined-behaviour sanitizers.
//
// Grouping is greedy: a string joins the open group if its x mask (above the
// low k bits) lies in the group's span or the span can still grow; otherwise
// the group is closed and padded with unit vectors to dimension dh.  The reduced
// basis keeps every pivot bit in exactly one basis vector, so a string's local
// mask is read off its pivot bits.
#include "fp_tiles_host.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

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
    // strings: 29 groups instead of 32), and also with any multiple of 8 states,
    // whose slices stay 128-byte aligned (SummedPauliOp::apply at n = 14, 2000
    // weight-3 strings, 1000 states: 149 groups instead of 172, 31.2 -> 28.2 ms;
    // 64-byte slices: 123 groups but 36.0 ms).  With other state counts the slices
    // would not align with rows.
    bool const pow2 = n_states != 0 && (n_states & (n_states - 1)) == 0;
    bool const slices128 = dtype == FP_C128 && (pow2 || (!expect && n_states % 8 == 0));
    uint32_t const row_target = env_u32("FAST_PAULI_TILE_ROW_BYTES", slices128 ? 128u : 256u);
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
    // one block covers every state: 2^k consecutive rows are contiguous, at least 256
    // bytes; complex64 expectation tiles take 128 bytes like complex128's row slices
    // (S = 16: one fewer low row bit, so groups of 9 strings instead of 8)
    uint32_t const seg = expect && dtype != FP_C128 ? 128u : 256u;
    if (sblk >= n_states)
        while ((uint64_t(1) << g.k) * n_states * esz < seg && g.k < g.d)
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
fp_status group_strings(StringTable const *ops, uint32_t k, uint32_t dh, bool ordered, std::vector<HostGroup> &out)
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
bool shared_basis(HostGroup &g, StringTable const *ops, uint32_t k, uint32_t dh)
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
// max_shared(W) strings (the instance's per-string registers).  Complex64 also
// needs full tiles and a power-of-two state count (the paired 16-byte loads),
// which the launch checks.
uint32_t shared_bits(uint32_t d, uint32_t k, size_t n_str)
{
    // the same rule for complex128 (2^rr leader rows per pass, one state per lane) and
    // complex64 (the paired lanes: two states per lane, 2^rr = 2^(7 - lsb) rows per pass)
    uint32_t bits = 0;
    for (uint32_t W : {8u, 16u})
        for (uint32_t rr = 0; rr <= 6; ++rr)
            if (k < rr && d >= rr + 1 && (uint64_t(1) << (d - 1 - rr)) == 4u * W &&
                n_str <= size_t(fpt::max_shared(W)))
                bits |= (1u << fpt::shared_bit(W, false, rr)) | (1u << fpt::shared_bit(W, true, rr));
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

fp_status build_tile_image(StringTable const &table, int kind, uint32_t k, uint32_t dh, uint32_t wht_min,
                           TileImage &out)
{
    StringTable const *ops = &table;
    std::vector<HostGroup> hgs;
    fp_status const st = group_strings(ops, k, dh, apply_kind(kind), hgs);
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
        if (apply_kind(kind))
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
                r.s = m.s;
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
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    out.n_groups = uint32_t(groups.size());
    out.max_chunks = max_chunks;
    out.o_btab = up(groups.size() * sizeof(fpt::TileGroup));
    out.o_chunks = out.o_btab + up(btab.size() * sizeof(uint64_t));
    out.o_strs = out.o_chunks + up(chunks.size());
    out.o_sh = out.o_strs + up(strs.size());
    size_t const total = out.o_sh + up(sh.size() * sizeof(fpt::ShStr));
    out.h.assign(total, 0);
    std::memcpy(out.h.data(), groups.data(), groups.size() * sizeof(fpt::TileGroup));
    if (!btab.empty())
        std::memcpy(out.h.data() + out.o_btab, btab.data(), btab.size() * sizeof(uint64_t));
    if (!chunks.empty())
        std::memcpy(out.h.data() + out.o_chunks, chunks.data(), chunks.size());
    if (!strs.empty())
        std::memcpy(out.h.data() + out.o_strs, strs.data(), strs.size());
    if (!sh.empty())
        std::memcpy(out.h.data() + out.o_sh, sh.data(), sh.size() * sizeof(fpt::ShStr));
    return FP_OK;
}

} // namespace fprt
