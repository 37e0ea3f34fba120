// Unit test of the host-side coset-tile table builder (fast-pauli_amd/csrc/
// fp_tiles_host.cpp), built with the address and undefined-behaviour
// sanitizers.  For random, weight-limited, diagonal, shared-x and duplicated
// string sets and the geometries the kernels use, it checks every image against
// the properties the kernels rely on:
//   * layout: every part inside the image, groups' tables in bounds;
//   * each group's basis: dh vectors in reduced echelon form above the low k
//     bits, the B table linear over XOR;
//   * every string in exactly one record, with its coefficient scaled as the
//     kernels expect (apply: c (-i)^nY bitwise; expectation: c * kappa);
//   * local coordinates: a string's local x mask maps back to its x, and for
//     every coset q and local row a the row sign factors as
//     parity(row & z) = parity(a & Z) ^ parity(q & zc)  (the kernels' split);
//   * apply groups are runs of consecutive strings that the greedy rule closes
//     only when the next string's x leaves the span;
//   * chunking limits, transform classes first, chunk modes;
// and that impossible geometries are reported as errors.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "fp_tiles_host.hpp"

namespace
{

int g_failures = 0;
long g_checks = 0;

#define CHECK(cond, ...)                                                                                               \
    do                                                                                                                 \
    {                                                                                                                  \
        ++g_checks;                                                                                                    \
        if (!(cond))                                                                                                   \
        {                                                                                                              \
            if (g_failures < 20)                                                                                       \
            {                                                                                                          \
                std::fprintf(stderr, "FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond);                                  \
                std::fprintf(stderr, __VA_ARGS__);                                                                     \
                std::fprintf(stderr, "\n");                                                                            \
            }                                                                                                          \
            ++g_failures;                                                                                              \
        }                                                                                                              \
    } while (0)

struct Rng
{
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    uint32_t below(uint32_t n) { return uint32_t(next() % n); }
    double normal() // sum of uniforms: any spread of values will do here
    {
        double a = 0;
        for (int i = 0; i < 4; ++i)
            a += double(next() >> 11) * 0x1.0p-53;
        return a - 2.0;
    }
};

uint32_t parity(uint64_t v)
{
    return uint32_t(__builtin_popcountll(v) & 1);
}

struct Strings
{
    size_t n = 0;
    std::vector<uint64_t> x, z;
    std::vector<uint32_t> ny;
    std::vector<double> coeffs, cf128;
    std::vector<float> cf64;

    void add(std::vector<int> const &letters, Rng &r) // 0 I, 1 X, 2 Y, 3 Z on character c -> bit n-1-c
    {
        uint64_t xs = 0, zs = 0;
        uint32_t y = 0;
        for (size_t c = 0; c < n; ++c)
        {
            uint64_t const bit = uint64_t(1) << (n - 1 - c);
            int const l = letters[c];
            if (l == 1 || l == 2)
                xs |= bit;
            if (l == 2 || l == 3)
                zs |= bit;
            y += l == 2 ? 1u : 0u;
        }
        x.push_back(xs);
        z.push_back(zs);
        ny.push_back(y);
        double const cr = r.normal(), ci = r.normal();
        coeffs.push_back(cr);
        coeffs.push_back(ci);
        // c (-i)^nY: multiply by -i nY times (exact: swaps and negations)
        double fr = cr, fi = ci;
        for (uint32_t k = 0; k < (y & 3u); ++k)
        {
            double const t = fr;
            fr = fi;
            fi = -t;
        }
        cf128.push_back(fr);
        cf128.push_back(fi);
        cf64.push_back(float(fr));
        cf64.push_back(float(fi));
    }
    fprt::StringTable table() const
    {
        return fprt::StringTable{n, x.size(), x.data(), z.data(), ny.data(), coeffs.data(), cf128.data(), cf64.data()};
    }
};

Strings random_strings(size_t n, size_t N, int kind, Rng &r)
{
    Strings s;
    s.n = n;
    std::vector<int> L(n);
    std::vector<std::vector<int>> made;
    for (size_t i = 0; i < N; ++i)
    {
        std::fill(L.begin(), L.end(), 0);
        if (kind == 0) // every site uniform
            for (auto &l : L)
                l = int(r.below(4));
        else if (kind == 1) // weight <= 3
            for (uint32_t w = 1 + r.below(3), t = 0; t < w; ++t)
                L[r.below(uint32_t(n))] = 1 + int(r.below(3));
        else if (kind == 2) // diagonal
            for (auto &l : L)
                l = r.below(2) ? 3 : 0;
        else if (kind == 3) // a few distinct x masks, random z
        {
            uint32_t const pick = r.below(3);
            for (size_t c = 0; c < n; ++c)
            {
                bool const xb = ((c * 7 + pick * 3) % 5) == 0;
                bool const zb = r.below(2) != 0;
                L[c] = xb ? (zb ? 2 : 1) : (zb ? 3 : 0);
            }
        }
        else // duplicates of earlier strings
        {
            if (!made.empty() && r.below(2))
                L = made[r.below(uint32_t(made.size()))];
            else
                for (auto &l : L)
                    l = int(r.below(4));
        }
        made.push_back(L);
        s.add(L, r);
    }
    return s;
}

struct View
{
    fprt::TileImage const &im;
    template <class T> T const *at(size_t off, size_t count) const
    {
        CHECK(off + count * sizeof(T) <= im.h.size(), "record range %zu + %zu x %zu beyond image %zu", off, count,
              sizeof(T), im.h.size());
        return reinterpret_cast<T const *>(im.h.data() + off);
    }
};

// Group geometry decoded from a TileGroup: pivots, basis, coset bases.
struct Geo
{
    uint32_t n, k, dh;
    std::vector<uint32_t> piv;
    std::vector<uint64_t> basis;
    uint64_t const *bt;

    uint64_t coset_base(uint64_t q) const // q << k with zeros inserted at the pivots, ascending
    {
        uint64_t v = q << k;
        for (uint32_t p : piv)
            v = ((v >> p) << (p + 1)) | (v & ((uint64_t(1) << p) - 1));
        return v;
    }
    uint64_t row(uint64_t q, uint32_t a) const
    {
        uint64_t const low = (uint64_t(1) << k) - 1;
        return coset_base(q) ^ bt[a >> k] ^ (uint64_t(a) & low);
    }
    uint64_t global_x(uint32_t A) const
    {
        return bt[A >> k] ^ (uint64_t(A) & ((uint64_t(1) << k) - 1));
    }
    // z compressed onto the coset-index bits (the non-pivot bits >= k, ascending)
    uint64_t compress(uint64_t z) const
    {
        uint64_t out = 0;
        uint32_t j = 0;
        for (uint32_t b = k; b < n; ++b)
        {
            bool pv = false;
            for (uint32_t p : piv)
                pv = pv || p == b;
            if (!pv)
                out |= ((z >> b) & 1u) << j++;
        }
        return out;
    }
};

// `reduced`: the basis is in reduced echelon form at the pivots (apply and EV-matrix
// tables); expectation tables may switch a group to another basis of the same
// span (shared leaders), for which the pivot coordinates must still be a
// bijection (every row decomposes uniquely over the cosets).
Geo check_group(View const &v, fpt::TileGroup const &g, uint32_t n, uint32_t k, uint32_t dh, bool reduced = true)
{
    Geo geo{n, k, dh, {}, {}, nullptr};
    CHECK(g.n_piv == dh, "group has %u pivots, want %u", g.n_piv, dh);
    for (uint32_t i = 0; i < g.n_piv && i < 16; ++i)
        geo.piv.push_back(uint32_t((i < 8 ? g.piv_lo : g.piv_hi) >> (8 * (i & 7))) & 0xffu);
    for (size_t i = 0; i < geo.piv.size(); ++i)
    {
        CHECK(geo.piv[i] >= k && geo.piv[i] < n, "pivot %u outside [%u, %u)", geo.piv[i], k, n);
        if (i)
            CHECK(geo.piv[i] > geo.piv[i - 1], "pivots not ascending");
    }
    size_t const n_bt = size_t(1) << g.n_piv;
    uint64_t const *bt = v.at<uint64_t>(v.im.o_btab + size_t(g.btab_off) * 8, n_bt);
    geo.bt = bt;
    CHECK(bt[0] == 0, "B(0) = %" PRIx64, bt[0]);
    for (size_t a = 1; a < n_bt; ++a)
        CHECK(bt[a] == (bt[a & (a - 1)] ^ bt[a & (~a + 1)]), "B table not linear at %zu", a);
    uint64_t const low = (uint64_t(1) << k) - 1;
    for (uint32_t i = 0; i < g.n_piv; ++i)
    {
        uint64_t const b = bt[size_t(1) << i];
        geo.basis.push_back(b);
        CHECK((b & low) == 0, "basis vector %u has low bits", i);
        CHECK(b >> n == 0, "basis vector %u beyond %u qubits", i, n);
        if (reduced)
            for (uint32_t j = 0; j < g.n_piv; ++j)
                CHECK(((b >> geo.piv[j]) & 1u) == (i == j ? 1u : 0u), "basis not reduced at pivots (%u, %u)", i, j);
    }
    // the pivot-coordinate matrix M[i][j] = bit piv_j of b_i is invertible over GF(2)
    std::vector<uint32_t> rows_m;
    for (uint64_t b : geo.basis)
    {
        uint32_t m = 0;
        for (size_t j = 0; j < geo.piv.size(); ++j)
            m |= uint32_t((b >> geo.piv[j]) & 1u) << j;
        rows_m.push_back(m);
    }
    uint32_t rank = 0;
    for (uint32_t col = 0; col < uint32_t(rows_m.size()); ++col)
    {
        size_t sel = rank;
        while (sel < rows_m.size() && !((rows_m[sel] >> col) & 1u))
            ++sel;
        if (sel == rows_m.size())
            continue;
        std::swap(rows_m[sel], rows_m[rank]);
        for (size_t i = 0; i < rows_m.size(); ++i)
            if (i != rank && ((rows_m[i] >> col) & 1u))
                rows_m[i] ^= rows_m[rank];
        ++rank;
    }
    CHECK(rank == g.n_piv, "pivot coordinates of the basis have rank %u of %u", rank, g.n_piv);
    return geo;
}

// The kernels' sign split at sampled cosets and local rows.
void check_signs(Geo const &geo, uint64_t z, uint32_t Z, uint64_t zc, Rng &r, char const *what, size_t s)
{
    uint64_t const n_cos = uint64_t(1) << (geo.n - geo.k - geo.dh);
    uint32_t const n_loc = 1u << (geo.k + geo.dh);
    CHECK(zc == geo.compress(z), "%s %zu: zc %" PRIx64 " want %" PRIx64, what, s, zc, geo.compress(z));
    for (int t = 0; t < 24; ++t)
    {
        uint64_t const q = r.next() % n_cos;
        uint32_t const a = r.below(n_loc);
        uint64_t const row = geo.row(q, a);
        CHECK(parity(row & z) == (parity(uint64_t(a) & Z) ^ parity(q & zc)), "%s %zu: sign split fails at q=%" PRIu64
              " a=%u", what, s, q, a);
    }
}

void check_apply(Strings const &S, uint32_t k, uint32_t dh, Rng &r)
{
    fprt::TileImage im;
    fp_status const st = fprt::build_tile_image(S.table(), fprt::kTilesOpApply, k, dh, 0, im);
    CHECK(st == FP_OK, "apply tables failed (n=%zu k=%u dh=%u)", S.n, k, dh);
    if (st != FP_OK)
        return;
    View v{im};
    auto const *groups = v.at<fpt::TileGroup>(0, im.n_groups);
    auto const *recs = v.at<fpa::ApplyStr>(im.o_strs, S.x.size());
    uint32_t next = 0;
    for (uint32_t gi = 0; gi < im.n_groups; ++gi)
    {
        fpt::TileGroup const &g = groups[gi];
        Geo const geo = check_group(v, g, uint32_t(S.n), k, dh);
        CHECK(g.chunk_first == next, "apply group %u starts at %u, want %u", gi, g.chunk_first, next);
        CHECK(g.n_chunks >= 1, "empty apply group");
        CHECK(im.max_chunks >= g.n_chunks, "max_chunks %u < %u", im.max_chunks, g.n_chunks);
        for (uint32_t j = 0; j < g.n_chunks && next + j < S.x.size(); ++j)
        {
            size_t const s = next + j;
            fpa::ApplyStr const &a = recs[s];
            CHECK(geo.global_x(a.A) == S.x[s], "apply string %zu: local x maps to %" PRIx64 ", x = %" PRIx64, s,
                  geo.global_x(a.A), S.x[s]);
            CHECK(std::memcmp(&a.cr, &S.cf128[2 * s], 8) == 0 && std::memcmp(&a.ci, &S.cf128[2 * s + 1], 8) == 0,
                  "apply string %zu: complex128 coefficient", s);
            CHECK(std::memcmp(&a.fr, &S.cf64[2 * s], 4) == 0 && std::memcmp(&a.fi, &S.cf64[2 * s + 1], 4) == 0,
                  "apply string %zu: complex64 coefficient", s);
            check_signs(geo, S.z[s], a.Z, a.zc, r, "apply string", s);
        }
        next += g.n_chunks;
        // greedy: the group closed because the next string's x leaves its (full) span
        if (next < S.x.size())
        {
            uint64_t const low = (uint64_t(1) << k) - 1;
            uint64_t v2 = S.x[next] & ~low;
            for (size_t i = 0; i < geo.basis.size(); ++i)
                if ((v2 >> geo.piv[i]) & 1u)
                    v2 ^= geo.basis[i];
            CHECK(v2 != 0, "apply group %u closed although string %u lies in its span", gi, next);
        }
    }
    CHECK(next == S.x.size(), "apply groups cover %u of %zu strings", next, S.x.size());
}

double kappa(uint64_t x, uint32_t ny)
{
    return x == 0 ? 1.0 : (((ny & 3u) <= 1u) ? 2.0 : -2.0);
}

void check_expect(Strings const &S, uint32_t k, uint32_t dh, Rng &r)
{
    fprt::TileImage im;
    fp_status const st = fprt::build_tile_image(S.table(), fprt::kTilesOpExpect, k, dh, 0, im);
    CHECK(st == FP_OK, "expectation tables failed (n=%zu k=%u dh=%u)", S.n, k, dh);
    if (st != FP_OK)
        return;
    View v{im};
    auto const *groups = v.at<fpt::TileGroup>(0, im.n_groups);
    // every string once: multiset of (x, imag, scale) against the strings'
    std::map<std::tuple<uint64_t, uint32_t, double, double>, int> want;
    for (size_t s = 0; s < S.x.size(); ++s)
    {
        double const kap = kappa(S.x[s], S.ny[s]);
        ++want[{S.x[s], S.ny[s] & 1u, S.coeffs[2 * s] * kap, S.coeffs[2 * s + 1] * kap}];
    }
    auto find_string = [&](uint64_t x, uint32_t imag, double sre, double sim) -> long {
        for (size_t s = 0; s < S.x.size(); ++s)
        {
            double const kap = kappa(S.x[s], S.ny[s]);
            if (S.x[s] == x && (S.ny[s] & 1u) == imag && S.coeffs[2 * s] * kap == sre &&
                S.coeffs[2 * s + 1] * kap == sim)
                return long(s);
        }
        return -1;
    };
    uint32_t total = 0;
    for (uint32_t gi = 0; gi < im.n_groups; ++gi)
    {
        fpt::TileGroup const &g = groups[gi];
        Geo const geo = check_group(v, g, uint32_t(S.n), k, dh, false);
        auto const *ch = v.at<fpt::TileChunk>(im.o_chunks + size_t(g.chunk_first) * sizeof(fpt::TileChunk), g.n_chunks);
        uint32_t in_group = 0;
        for (uint32_t c = 0; c < g.n_chunks; ++c)
        {
            fpt::TileChunk const &t = ch[c];
            CHECK(t.count >= 1 && t.count <= uint32_t(fpt::kChunk), "chunk count %u", t.count);
            CHECK(t.A < (1u << (k + dh)), "chunk local x %u beyond the tile", t.A);
            if (t.A)
                CHECK(t.h == 31u - uint32_t(__builtin_clz(t.A)), "chunk h %u for A %u", t.h, t.A);
            auto const *ts = v.at<fpt::TileStr>(im.o_strs + size_t(t.first) * sizeof(fpt::TileStr), t.count);
            uint32_t need = 0;
            for (uint32_t j = 0; j < t.count; ++j)
            {
                fpt::TileStr const &e = ts[j];
                uint64_t const x = geo.global_x(t.A);
                auto it = want.find({x, e.imag, e.sre, e.sim});
                CHECK(it != want.end() && it->second > 0, "expectation record (x %" PRIx64 ") matches no string", x);
                if (it != want.end())
                    --it->second;
                need |= e.imag ? 2u : 1u;
                long const s = find_string(x, e.imag, e.sre, e.sim);
                if (s >= 0)
                {
                    // leader rows p map to a = p with a zero inserted at bit h
                    uint32_t const Zs = [&] {
                        uint32_t zz = 0;
                        for (size_t i = 0; i < geo.basis.size(); ++i)
                            zz |= parity(geo.basis[i] & S.z[s]) << i;
                        return (zz << k) | uint32_t(S.z[s] & ((uint64_t(1) << k) - 1));
                    }();
                    uint32_t const zt = t.A ? ((Zs & ((1u << t.h) - 1u)) | ((Zs >> (t.h + 1)) << t.h)) : Zs;
                    CHECK(e.zt == zt, "expectation record: zt %u want %u", e.zt, zt);
                    check_signs(geo, S.z[s], Zs, e.zc, r, "expectation string", size_t(s));
                }
            }
            CHECK(t.need == need, "chunk need %u want %u", t.need, need);
            in_group += t.count;
        }
        CHECK(in_group == g.n_str, "group lists %u strings, chunks hold %u", g.n_str, in_group);
        if (g.shared)
        {
            auto const *sh = v.at<fpt::ShStr>(im.o_sh + size_t(g.str_first) * sizeof(fpt::ShStr), g.n_str);
            CHECK(g.n_str <= uint32_t(fpt::kMaxShared), "shared group of %u strings", g.n_str);
            for (uint32_t j = 0; j < g.n_str; ++j)
            {
                CHECK((sh[j].A >> k) & 1u, "shared-leader string without local bit k");
                long const s = find_string(geo.global_x(sh[j].A), sh[j].imag, sh[j].sre, sh[j].sim);
                CHECK(s >= 0, "shared-leader record matches no string");
                if (s >= 0)
                    check_signs(geo, S.z[s], sh[j].Z, sh[j].zc, r, "shared-leader string", size_t(s));
            }
        }
        total += in_group;
    }
    CHECK(total == S.x.size(), "expectation groups hold %u of %zu strings", total, S.x.size());
    for (auto const &kv : want)
        CHECK(kv.second == 0, "a string is missing from (or repeated in) the expectation tables");
}

void check_ev_matrix(Strings const &S, uint32_t k, uint32_t dh, uint32_t wht_min, Rng &r)
{
    fprt::TileImage im;
    fp_status const st = fprt::build_tile_image(S.table(), fprt::kTilesEvMatrix, k, dh, wht_min, im);
    CHECK(st == FP_OK, "EV-matrix tables failed (n=%zu k=%u dh=%u)", S.n, k, dh);
    if (st != FP_OK)
        return;
    View v{im};
    auto const *groups = v.at<fpt::TileGroup>(0, im.n_groups);
    std::vector<int> seen(S.x.size(), 0);
    uint32_t const rows = 1u << (k + dh);
    for (uint32_t gi = 0; gi < im.n_groups; ++gi)
    {
        fpt::TileGroup const &g = groups[gi];
        Geo const geo = check_group(v, g, uint32_t(S.n), k, dh);
        auto const *ch = v.at<fpl::EvChunk>(im.o_chunks + size_t(g.chunk_first) * sizeof(fpl::EvChunk), g.n_chunks);
        CHECK(g.n_wht <= g.n_chunks, "n_wht %u > chunks %u", g.n_wht, g.n_chunks);
        for (uint32_t c = 0; c < g.n_chunks; ++c)
        {
            fpl::EvChunk const &t = ch[c];
            bool const wht = c < g.n_wht;
            CHECK(t.count >= 1 && t.count <= (wht ? fpl::kWhtMax : uint32_t(fpl::kChunk)), "EV chunk count %u", t.count);
            if (wht)
                CHECK(k + dh >= fpl::kWhtMinRowsLog2, "transform class on a %u-row tile", rows);
            if (t.x)
                CHECK(t.h == 31u - uint32_t(__builtin_clz(t.x)), "EV chunk h");
            CHECK(t.mode == fpl::ev_mode(t.x, t.h, rows), "EV chunk mode %u", t.mode);
            auto const *es = v.at<fpl::EvStr>(im.o_strs + size_t(t.first) * sizeof(fpl::EvStr), t.count);
            for (uint32_t j = 0; j < t.count; ++j)
            {
                fpl::EvStr const &e = es[j];
                CHECK(e.out < S.x.size(), "EV record output row %u", e.out);
                if (e.out >= S.x.size())
                    continue;
                size_t const s = e.out;
                ++seen[s];
                CHECK(geo.global_x(t.x) == S.x[s], "EV string %zu: local x maps to %" PRIx64, s, geo.global_x(t.x));
                CHECK(e.imag == (S.ny[s] & 1u) && t.imag == e.imag, "EV string %zu: half of w", s);
                double const eps = ((S.ny[s] & 3u) <= 1u) ? 1.0 : -1.0;
                CHECK(e.scale == eps * (t.mode == fpl::kAllRows ? 1.0 : 2.0), "EV string %zu: scale %g", s, e.scale);
                check_signs(geo, S.z[s], e.z, e.zc, r, "EV string", s);
            }
        }
    }
    for (size_t s = 0; s < seen.size(); ++s)
        CHECK(seen[s] == 1, "EV string %zu listed %d times", s, seen[s]);
}

void check_errors()
{
    Rng r{7};
    Strings S = random_strings(6, 10, 0, r);
    fprt::TileImage im;
    CHECK(fprt::build_tile_image(S.table(), fprt::kTilesOpApply, 4, 3, 0, im) != FP_OK,
          "k + dh beyond the qubit count accepted");
    // dh = 0 leaves no room for x bits above the low k bits
    bool high = false;
    for (uint64_t x : S.x)
        high = high || (x >> 2) != 0;
    if (high)
        CHECK(fprt::build_tile_image(S.table(), fprt::kTilesOpExpect, 2, 0, 0, im) != FP_OK,
              "dh = 0 accepted with x bits above k");
}

// The tile geometries the kernels are tuned for: complex128 takes 128-byte row slices
// (8 states) for power-of-two state counts and, for apply, any multiple of 8; complex64
// expectation tiles keep 128-byte runs of consecutive rows (k = 0 at 16 states), the
// other tiles 256-byte runs.
void check_geometry_rules()
{
    struct G
    {
        int dtype;
        size_t n, S;
        bool expect;
        uint32_t lsb, d, k;
    };
    G const cases[] = {
        {FP_C128, 24, 16, true, 3, 9, 0},    {FP_C128, 24, 16, false, 3, 9, 0},
        {FP_C128, 14, 1000, false, 3, 9, 0}, {FP_C128, 14, 1000, true, 4, 8, 0},
        {FP_C128, 14, 24, false, 3, 9, 0},   {FP_C128, 14, 24, true, 4, 8, 0},
        {FP_C64, 24, 16, true, 4, 9, 0},     {FP_C64, 24, 16, false, 4, 9, 1},
        {FP_C64, 24, 1, true, 0, 13, 4},     {FP_C64, 24, 1, false, 0, 13, 5},
    };
    for (G const &g : cases)
    {
        fprt::TileGeom const t = fprt::tile_geom(g.dtype, g.n, g.S, g.expect);
        CHECK(t.lsb == g.lsb && t.d == g.d && t.k == g.k && t.dh == t.d - t.k,
              "tile_geom(dtype %d, n %zu, S %zu, expect %d): lsb %u d %u k %u, want %u %u %u", g.dtype, g.n, g.S,
              int(g.expect), t.lsb, t.d, t.k, g.lsb, g.d, g.k);
    }
}

} // namespace

int main()
{
    check_geometry_rules();
    Rng r{2024};
    struct Case
    {
        size_t n, N;
        int kind;
    };
    Case const cases[] = {{6, 40, 0},  {10, 64, 0}, {12, 200, 1}, {13, 150, 2}, {14, 300, 3}, {16, 256, 4},
                          {18, 120, 0}, {20, 64, 1}, {24, 96, 0},  {11, 33, 3},  {9, 1, 0},    {12, 2000, 1}};
    int n_images = 0;
    for (Case const &c : cases)
    {
        Strings const S = random_strings(c.n, c.N, c.kind, r);
        // the kernels' geometries: apply and expectation tiles for several state counts
        for (int dtype : {FP_C128, FP_C64})
            for (size_t n_states : {size_t(1), size_t(3), size_t(16), size_t(24), size_t(59), size_t(64)})
            {
                fprt::TileGeom const ga = fprt::tile_geom(dtype, c.n, n_states, false);
                if (ga.dh > 0 && ga.k + ga.dh <= c.n)
                {
                    check_apply(S, ga.k, ga.dh, r);
                    ++n_images;
                }
                fprt::TileGeom const ge = fprt::tile_geom(dtype, c.n, n_states, true);
                if (ge.k + ge.dh <= c.n)
                {
                    check_expect(S, ge.k, ge.dh, r);
                    ++n_images;
                }
            }
        // per-string EV matrix on cosets: k = 3 / 4, one dimension below n
        for (uint32_t k : {3u, 4u})
            if (c.n > k + 1 && c.n <= 20)
                for (uint32_t wht_min : {0xffffffffu, 2u, 12u})
                {
                    check_ev_matrix(S, k, uint32_t(c.n) - 1 - k, wht_min, r);
                    ++n_images;
                }
    }
    check_errors();
    std::printf("tile tables: %d images, %ld checks, %d failures\n", n_images, g_checks, g_failures);
    return g_failures == 0 ? 0 : 1;
}
