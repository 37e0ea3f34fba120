// fp_ev_tiled.hpp — PauliOp expectation values with coset tiles in LDS.
//
// The streaming kernel reads the whole state once per string.  Here strings
// are grouped so that one read of the state serves a whole group:
//
//  * The host picks, per group, a subspace V of row indices (over GF(2)) that
//    contains the x mask of every string in the group and the low k bits (so
//    that 2^k consecutive rows form one contiguous segment in HBM).  V has
//    dimension d = k + dh; it is the span of the group's x masks, padded with
//    unit vectors to a fixed d.
//  * The rows split into 2^(n-d) cosets  c + V.  Row j ^ x of a string in the
//    group lies in the same coset as row j, so each coset is a closed 2^d-row
//    sub-problem: a tile of 2^d rows x S_blk states is loaded into LDS once and
//    every string of the group is evaluated on it.
//  * In local coordinates a = (a_hi, a_lo) of the tile (a_hi: coefficients in
//    the reduced basis b_i of V, a_lo: the low k bits), row
//        r(c, a) = c ^ B(a_hi) ^ a_lo,   B(a_hi) = XOR of b_i over the set bits of a_hi,
//    the string (x, z) acts as the local string (A, Z) with
//        A = (alpha << k) | (x & lowmask),  alpha_i = bit piv_i of x,
//        Z = (zeta  << k) | (z & lowmask),  zeta_i  = parity(b_i & z),
//    and the sign of row r is (-1)^{parity(c & z)} (-1)^{parity(a & Z)}.
//  * Pairs (a, a ^ A) are visited once from the leader a (bit h of a clear,
//    h the top bit of A); the pair contributes
//        sign * (Re w or Im w),  w = conj(psi_a) psi_{a^A},
//    times the string's scale c_s * kappa_s (kappa = 2 eps, the same pair
//    identity as the streaming kernel).  A = 0 (x = 0) walks every row with
//    kappa = 1.
//
// HBM traffic is (number of groups) x (state bytes) instead of
// (number of strings) x (state bytes).  The tile is double-buffered through
// registers: the next coset is in flight while the current one is evaluated.
// Each launch shape has three instances: shared leaders, for groups whose
// strings all pair rows with local bit k clear (one leader read serves every
// string); set-up-once, for groups whose waves each own at most one
// single-string unit; and the general one, which sets units up per tile.
// Modes 1 and 2 give complex64 lanes two states each.
// Each block accumulates FP64 sums for its state and its range of cosets and
// writes one partial per state; partials are reduced in a fixed order, so
// results are identical run to run.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "fp_kernels.hpp"
#include "fp_tile_types.hpp"

namespace fpt
{

// Tiles of TKB KiB (64: two 512-thread blocks per CU; 128: one 1024-thread block).
template <int TKB> struct TileCfg
{
    static constexpr uint32_t kBytes = uint32_t(TKB) * 1024u;
    static constexpr int kThreads = TKB * 8;
};
__device__ __forceinline__ uint32_t insert_zero_bit32(uint32_t v, uint32_t b)
{
    uint32_t const lo = v & ((1u << b) - 1u);
    return ((v >> b) << (b + 1)) | lo;
}

// v with its sign bit flipped when bit31 = 0x80000000 (exact negation)
__device__ __forceinline__ float xor_sign_bit(float v, uint32_t bit31)
{
    return __uint_as_float(__float_as_uint(v) ^ bit31);
}
__device__ __forceinline__ double xor_sign_bit(double v, uint32_t bit31)
{
    return __longlong_as_double(__double_as_longlong(v) ^ (static_cast<long long>(bit31) << 32));
}

__device__ __forceinline__ uint64_t coset_base(uint64_t q, uint32_t k, TileGroup const &g)
{
    uint64_t v = q << k;
    for (uint32_t i = 0; i < g.n_piv; ++i)
        v = fpk::insert_zero_bit(v, uint32_t((i < 8 ? g.piv_lo : g.piv_hi) >> (8 * (i & 7))) & 0xffu);
    return v;
}

// Leader rows of a chunk: p in [0, n_rows), a = insert_zero_bit(p, h) (or p
// itself when A = 0).  A wave evaluates a chunk (or a part of its rows): lane
// (p_first, t) visits p = p_first | (it << r), with 2^r = 64 / S_blk leader rows
// per pass.  Since insert_zero_bit is linear over GF(2), a = a_first ^ a_it
// with a_it wave-uniform, and the row sign parity(p & zt) = parity(p_first & zt)
// ^ parity(it & (zt >> r)): a per-lane bit applied once per chunk and a
// wave-uniform bit per pass.
//
// Per-block set-up that does not depend on the coset is done once: a wave that
// owns a single one-string unit keeps its addresses, signs and scale in
// registers for all of the block's cosets, and coset bases advance by one
// deposited addition per tile (see below) instead of a pivot-insertion loop.
// MODE 3 is MODE 2 with each string's partner offset and coset sign mask packed in one
// 32-bit word (fewer scalar registers, fewer spills in the string loop); the host takes it
// when every coset index and every byte offset in the tile is below 2^16.
template <class T, bool WIDE, int TKB, int MODE_>
__global__ __launch_bounds__(TKB * 8, WIDE ? 2 : 4) void k_op_expect_tiled(typename fpk::VecOf<T, 1>::type const *__restrict__ states,
                                                              uint64_t n_states, uint32_t sshift, uint32_t lsb,
                                                              uint32_t k, uint32_t d, uint32_t ranges,
                                                              uint32_t n_parts, uint32_t n_chunks_y,
                                                              uint64_t n_cosets,
                                                              TileGroup const *__restrict__ groups,
                                                              uint64_t const *__restrict__ btab,
                                                              TileChunk const *__restrict__ chunks,
                                                              TileStr const *__restrict__ strs,
                                                              ShStr const *__restrict__ sh,
                                                              fpk::Dbl2 *__restrict__ partial)
{
    constexpr int MODE = MODE_ == 3 ? 2 : MODE_; // the evaluation form
    constexpr bool PK = MODE_ == 3;
    static_assert(!PK || TKB * 1024 <= 65536, "packed offsets: tile offsets below 2^16");
    typedef typename fpk::VecOf<T, 1>::type V;
    typedef typename fpk::RowAcc<T>::type Acc;
    constexpr int kThreads = TileCfg<TKB>::kThreads;
    constexpr uint32_t kTileBytes = TileCfg<TKB>::kBytes;
    constexpr uint32_t kElems = kTileBytes / sizeof(V);
    constexpr uint32_t EPT = kElems / kThreads;
    constexpr uint32_t kEsz = sizeof(V);
    constexpr int U = 4;
    // passes per step of the one-string fast path in the per-tile instance (the
    // set-up-once instance uses 4)
    constexpr int U1 = sizeof(V) == 16 ? 4 : 8;
    extern __shared__ __align__(16) unsigned char smem[];
    V *tile = reinterpret_cast<V *>(smem);

    uint32_t const tid = threadIdx.x;
    // Block id -> (part, state chunk): the state chunks of one part get ids
    // 8 apart within a run of 8 * n_chunks_y ids, i.e. the same XCD (workgroups
    // are dealt round-robin over the 8 XCDs) at about the same time, so their
    // slices of each row share that XCD's L2 / the Infinity Cache.
    uint32_t const id = blockIdx.x;
    uint32_t const part = (id / 8u / n_chunks_y) * 8u + (id % 8u);
    uint32_t const chunk_y = (id / 8u) % n_chunks_y;
    if (part >= n_parts)
        return; // padding of the part count to a multiple of 8 (whole block, before any barrier)
    uint32_t const g = part / ranges, rg = part % ranges;
    TileGroup const grp = groups[g];
    uint64_t const *bt = btab + grp.btab_off;
    constexpr uint32_t W = kThreads / 64;
    uint32_t const r = 6u - lsb;
    // Instances (MODE): 2 shared leaders (TileGroup::shared), 1 set up once
    // (TileGroup::hoist), 0 general; a group runs in the first that applies, and a
    // launch may need several (each block runs in one).  Modes 1 and 2 run 4
    // passes per step, complex64 in the paired lane form (2^(r+1) leader rows per
    // pass), which needs S_blk >= 2.
    constexpr bool HOIST = MODE == 1;
    constexpr bool PAIRED = MODE >= 1 && sizeof(V) == 8;
    uint32_t const rrx = sizeof(V) == 8 ? r + 1u : r;
    bool const can_pair = sizeof(V) != 8 || lsb >= 1;
    // complex64 takes the shared-leader form only with the paired 16-byte tile loads (full
    // tiles, a power-of-two state count), which its instance assumes; the host applies the
    // same rule when it picks the instances to launch
    bool const pairs_ok = sizeof(V) != 8 || (sshift < 64u && lsb >= 1u && ((1u << d) << lsb) == kElems);
    bool const g_shared =
        can_pair && pairs_ok && ((grp.shared >> shared_bit(W, sizeof(V) == 8, rrx)) & 1u) != 0;
    bool const g_hoist = !g_shared && can_pair && ((grp.hoist >> hoist_bit(W, 4u, rrx)) & 1u) != 0;
    int const gmode = g_shared ? 2 : (g_hoist ? 1 : 0);
    if (gmode != MODE)
        return; // whole block, before any barrier

    uint32_t const sblk = 1u << lsb;
    uint32_t const t = tid & (sblk - 1u);
    uint64_t const s0 = uint64_t(chunk_y) << lsb;
    bool const t_live = s0 + t < n_states;
    uint32_t const tile_elems = (1u << d) << lsb;
    uint32_t const lowmask = (1u << k) - 1u;
    // every element of every tile of this block is live: loads and stores need no guard
    bool const full = tile_elems == kElems && s0 + sblk <= n_states;
    // cosets rg, rg + ranges, rg + 2 ranges, ...: at any moment the blocks of a
    // group work on consecutive cosets, whose rows are neighbours in HBM (the
    // lowest coset-index bits are the lowest non-pivot row bits)
    uint64_t const q0 = rg;
    uint64_t const q_step = ranges;
    // Coset bases: base(q) = coset_base(q) deposits the bits of q << k into the
    // non-pivot row bits.  Addition carries over into that deposited form with
    // the other bits set to 1, which pass carries through:
    //     base(q + q_step) = ((base(q) | ~cmask) + base(q_step)) & cmask,
    // cmask = base(n_cosets - 1) (n_cosets is a power of two: every coset bit).
    uint64_t const cmask = coset_base(n_cosets - 1u, k, grp);
    uint64_t const cstep = coset_base(q_step, k, grp);

    // this thread's tile elements: element tid + m * kThreads is local row
    // a_m = p_first + m * 2^r, global row coset_base ^ roff_m.  With a
    // power-of-two state count (sshift = log2 S) the element index is
    // (coset_base << sshift) ^ eoff[m]; otherwise eoff[m] = roff_m.
    // WIDE = false when 2^n * S < 2^32: offsets fit 32 bits (half the registers).
    typedef typename std::conditional<WIDE, uint64_t, uint32_t>::type Idx;
    bool const s_pow2 = sshift < 64u;
    // complex64 full tiles with a power-of-two state count and S_blk >= 2 move two
    // neighbouring states per 16-byte load and LDS store: the thread's m-th pair is
    // elements 2 (tid + m * kThreads) and the one after (eoff[m]: its 16-byte index)
    typedef typename fpk::VecOf<float, 2>::type V4;
    constexpr uint32_t EPT2 = EPT / 2;
    // the complex64 shared-leader instance runs only the paired form (see pairs_ok), so
    // it keeps offsets for the 16-byte slots only
    constexpr bool V2 = MODE == 2 && sizeof(V) == 8;
    constexpr uint32_t NOFF = V2 ? EPT2 : EPT;
    bool const vec2 = V2 || (sizeof(V) == 8 && full && s_pow2 && lsb >= 1);
    Idx eoff[NOFF];
    if (vec2)
    {
#pragma unroll
        for (uint32_t m = 0; m < EPT2; ++m)
        {
            uint32_t const e = 2u * (tid + m * kThreads);
            uint32_t const a = e >> lsb;
            uint64_t const roff = bt[a >> k] ^ uint64_t(a & lowmask);
            eoff[m] = Idx(((roff << sshift) | (s0 + (e & (sblk - 1u)))) >> 1);
            if constexpr (!V2)
                eoff[EPT2 + m] = 0;
        }
    }
    else
    {
        if constexpr (!V2)
        {
#pragma unroll
            for (uint32_t m = 0; m < EPT; ++m)
            {
                uint32_t const e = tid + m * kThreads;
                uint32_t const a = e >> lsb;
                bool const live = e < tile_elems && t_live;
                uint64_t const roff = live ? (bt[a >> k] ^ uint64_t(a & lowmask)) : 0;
                eoff[m] = Idx(s_pow2 ? ((roff << sshift) | (s0 + t)) : roff);
            }
        }
    }

    V reg[EPT];
    auto fetch = [&](V(&rg)[EPT], uint64_t cb) {
        Idx const cbs = Idx(s_pow2 ? (cb << sshift) : cb);
        auto index = [&](uint32_t m) {
            return s_pow2 ? Idx(cbs ^ eoff[m]) : Idx((cbs ^ eoff[m]) * Idx(n_states) + Idx(s0 + t));
        };
        if constexpr (sizeof(V) == 8)
        {
            if (vec2)
            {
                auto const *st4 = reinterpret_cast<V4 const *>(states);
                Idx const cbs4 = Idx(cb << (sshift - 1u));
#pragma unroll
                for (uint32_t m = 0; m < EPT2; ++m)
                {
                    V4 const v = __builtin_nontemporal_load(&st4[cbs4 ^ eoff[m]]);
                    rg[2 * m] = V{v[0], v[1]};
                    rg[2 * m + 1] = V{v[2], v[3]};
                }
                return;
            }
        }
        if constexpr (V2)
            return;
        else if (full)
        {
#pragma unroll
            for (uint32_t m = 0; m < EPT; ++m)
                rg[m] = __builtin_nontemporal_load(&states[index(m)]);
        }
        else
        {
#pragma unroll
            for (uint32_t m = 0; m < EPT; ++m)
            {
                rg[m] = V(0);
                if (tid + m * kThreads < tile_elems && t_live)
                    rg[m] = __builtin_nontemporal_load(&states[index(m)]);
            }
        }
    };

    // Waves own work units: unit u = (chunk u / split, row part u % split),
    // u = wave, wave + W, ...; split is the largest power of two with
    // n_chunks * split <= W (1 when a group has W chunks or more), so that a group
    // of at most W chunks gives each wave at most one unit.  Within a wave, lane =
    // (p_first, t): 2^r = 64 / S_blk leader rows per pass.
    uint32_t const wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63u;
    uint32_t const p_first = lane >> lsb;
    uint32_t split_log2 = 0;
    while ((grp.n_chunks << (split_log2 + 1)) <= W && split_log2 < 6)
        ++split_log2;
    uint32_t const split = 1u << split_log2;
    uint32_t const n_units = grp.n_chunks << split_log2;

    // Lane geometry of the fast path: lane = (pf, tx) with 2^rr leader rows per
    // pass; the lane reads LDS at byte ((a << lsbx) | tx) * ESZ for row a.  The
    // plain form has one state per lane (lsbx = lsb, tx = t, rr = r, ESZ = the
    // element size).  The paired form (complex64, S_blk >= 2) gives a lane the two
    // states 2 tx, 2 tx + 1 with one 16-byte read (lsbx = lsb - 1, rr = r + 1):
    // half the LDS instructions per (pair, state).
    struct Lanes
    {
        uint32_t lsbx, tx, rr, pf;
    };
    Lanes const plain{lsb, t, r, p_first};

    struct Unit
    {
        uint32_t A, h, count, first;
        uint32_t it_begin, it_end; // passes of this unit
        uint32_t hmask;            // insert_zero_bit(v, h) = v + (v & hmask) (0 when A = 0)
        uint32_t a_first;          // this lane's first leader row
        bool lane_live;            // dead lanes (fewer leader rows than a pass) drop their sums
        bool fast;                 // one string, whole groups of UU passes
    };
    auto make_unit = [&](uint32_t u, Lanes const &lg, uint32_t UU) {
        Unit un;
        TileChunk const ch = chunks[grp.chunk_first + (u >> split_log2)];
        un.A = ch.A;
        un.h = ch.h;
        un.count = ch.count;
        un.first = ch.first;
        uint32_t const n_rows = un.A ? (1u << (d - 1)) : (1u << d);
        uint32_t const n_it = (n_rows >> lg.rr) ? (n_rows >> lg.rr) : 1u;
        uint32_t const it_per = (n_it + split - 1) >> split_log2;
        un.it_begin = (u & (split - 1u)) * it_per;
        uint32_t const e = un.it_begin + it_per < n_it ? un.it_begin + it_per : n_it;
        un.it_end = e > un.it_begin ? e : un.it_begin;
        // Dead lanes (only when a tile has fewer leader rows than a pass) read a
        // valid row and drop their sums; the loops stay wave-uniform, so the pass
        // index, addresses' uniform part and signs live in scalar registers.
        un.lane_live = lg.pf < n_rows;
        uint32_t const p_lane = un.lane_live ? lg.pf : 0u;
        un.hmask = un.A ? ~((1u << un.h) - 1u) : 0u;
        un.a_first = p_lane + (p_lane & un.hmask);
        un.fast = un.count == 1 && (un.it_begin % UU) == 0 && ((un.it_end - un.it_begin) % UU) == 0;
        return un;
    };

    // One string, whole groups of UU passes: pass i0 + uu (i0 a multiple of UU)
    // has uniform address term uoff(i0) ^ uoff(uu) and row sign parity(i0 & zr) ^
    // parity(uu & zr).  Per-uu lane bases and per-uu accumulators carry the uu
    // parts, so each group of passes costs one scalar offset and one scalar sign.
    struct Fast
    {
        uint32_t it_begin, it_end, zr, hmask, A_hx, A_hy;
        uint32_t pa0, pb0; // LDS byte offsets of this lane's first leader / partner
        bool im, lane_neg, lane_live;
        uint64_t zc;
        double sre, sim;
    };
    auto make_fast = [&](Unit const &un, Lanes const &lg, uint32_t esz, uint32_t UU) {
        Fast f;
        TileStr const st = strs[un.first];
        uint32_t const A = un.A, hmask = un.hmask, rr = lg.rr;
        f.it_begin = un.it_begin;
        f.it_end = un.it_end;
        f.zr = st.zt >> rr;
        f.hmask = hmask;
        f.im = st.imag != 0;
        f.zc = st.zc;
        f.sre = st.sre;
        f.sim = st.sim;
        f.lane_live = un.lane_live;
        f.lane_neg = (__popc(lg.pf & st.zt) & 1) != 0;
        // Leader a = a_first + a_it and partner a ^ A = (a_first ^ A_lo) +
        // (a_it ^ A_hi) + 2^h, with A_lo / A_hi the bits of A in the lane / pass
        // part of the row index (disjoint bit ranges, so + is ^).  Both byte
        // addresses are a per-lane offset plus a wave-uniform one; the pass part
        // splits again into uu (per lane) and i0 (uniform).
        uint32_t const lr = ((1u << rr) - 1u) + (((1u << rr) - 1u) & hmask);
        uint32_t const A_lo = A & lr;
        uint32_t const A_hi = A ? (A & ~lr & ~(1u << un.h)) : 0u;
        uint32_t const ym0 = (UU - 1u) << rr;
        uint32_t const ymask = ym0 + (ym0 & hmask);
        f.A_hy = A_hi & ymask;
        f.A_hx = A_hi & ~ymask;
        uint32_t const bh = A ? (1u << un.h) : 0u;
        f.pa0 = ((un.a_first << lg.lsbx) | lg.tx) * esz;
        f.pb0 = ((((un.a_first ^ A_lo) << lg.lsbx) | lg.tx) * esz) + ((bh << lg.lsbx) * esz);
        return f;
    };

    // acc_re/acc_im: this thread's state (plain form) or the even state of its
    // pair (paired form); acc_re1/acc_im1: the odd state of the pair
    double acc_re = 0.0, acc_im = 0.0, acc_re1 = 0.0, acc_im1 = 0.0;
    auto run_fast = [&](Fast const &f, uint64_t q, Lanes const &lg, auto im_tag, auto paired_tag) {
        constexpr bool IM = decltype(im_tag)::value;
        constexpr bool PAIRED = decltype(paired_tag)::value;
        typedef typename std::conditional<PAIRED, typename fpk::VecOf<float, 2>::type, V>::type LV;
        constexpr int UU = PAIRED ? 4 : U1;
        constexpr uint32_t ESZ = sizeof(LV);
        constexpr int NS = PAIRED ? 2 : 1; // states per lane
        uint32_t const rr = lg.rr, lsbx = lg.lsbx;
        Acc pu[NS][UU];
        uint32_t pa[UU], pb[UU]; // byte offsets of the leader / partner of pass i0 + uu, less the i0 part
#pragma unroll
        for (int uu = 0; uu < UU; ++uu)
        {
#pragma unroll
            for (int j = 0; j < NS; ++j)
                pu[j][uu] = Acc(0);
            uint32_t const p = uint32_t(uu) << rr;
            uint32_t const ai = p + (p & f.hmask);
            pa[uu] = f.pa0 + ((ai << lsbx) * ESZ);
            pb[uu] = f.pb0 + (((ai ^ f.A_hy) << lsbx) * ESZ);
        }
        for (uint32_t i0 = f.it_begin; i0 < f.it_end; i0 += UU)
        {
            uint32_t const p = i0 << rr;
            uint32_t const ai = p + (p & f.hmask);
            uint32_t const sa = (ai << lsbx) * ESZ;
            uint32_t const sb = ((ai ^ f.A_hx) << lsbx) * ESZ;
            Acc const sg = (__popc(i0 & f.zr) & 1) ? Acc(-1) : Acc(1);
            LV va[UU], vb[UU];
#pragma unroll
            for (int uu = 0; uu < UU; ++uu)
            {
                va[uu] = *reinterpret_cast<LV const *>(smem + (pa[uu] + sa));
                vb[uu] = *reinterpret_cast<LV const *>(smem + (pb[uu] + sb));
            }
#pragma unroll
            for (int uu = 0; uu < UU; ++uu)
            {
#pragma unroll
                for (int j = 0; j < NS; ++j)
                {
                    Acc const ar = Acc(va[uu][2 * j]), ai_ = Acc(va[uu][2 * j + 1]);
                    Acc const br = Acc(vb[uu][2 * j]), bi = Acc(vb[uu][2 * j + 1]);
                    Acc const gv = IM ? fma(ar, bi, -(ai_ * br)) : fma(ar, br, ai_ * bi);
                    pu[j][uu] = fma(sg, gv, pu[j][uu]);
                }
            }
        }
        bool const neg = f.lane_neg != ((__popcll(q & f.zc) & 1) != 0);
#pragma unroll
        for (int j = 0; j < NS; ++j)
        {
            Acc p0 = Acc(0);
#pragma unroll
            for (int uu = 0; uu < UU; ++uu)
                p0 += (__popc(uint32_t(uu) & f.zr) & 1) ? -pu[j][uu] : pu[j][uu];
            if (!f.lane_live)
                p0 = Acc(0);
            double const v = neg ? -double(p0) : double(p0);
            if (j == 0)
            {
                acc_re += f.sre * v;
                acc_im += f.sim * v;
            }
            else
            {
                acc_re1 += f.sre * v;
                acc_im1 += f.sim * v;
            }
        }
    };
    auto run_fast_any = [&](Fast const &f, uint64_t q, Lanes const &lg, auto paired_tag) {
        if (f.im)
            run_fast(f, q, lg, std::true_type{}, paired_tag);
        else
            run_fast(f, q, lg, std::false_type{}, paired_tag);
    };

    // a unit on the general paths (ragged pass range, or several strings with one A)
    auto run_general = [&](Unit const &un, uint64_t q) {
        uint32_t const A = un.A;
        uint32_t const hmask = un.hmask;
        uint32_t const it_begin = un.it_begin, it_end = un.it_end;
        uint32_t const off0 = ((un.a_first << lsb) | t) * kEsz;
        uint32_t const pxor = (A << lsb) * kEsz;
        uint32_t zt[kChunk];
        bool imag[kChunk];
        Acc part[kChunk];
#pragma unroll
        for (int s = 0; s < kChunk; ++s)
        {
            part[s] = Acc(0);
            zt[s] = 0;
            imag[s] = false;
            if (s < int(un.count))
            {
                zt[s] = strs[un.first + s].zt;
                imag[s] = strs[un.first + s].imag != 0;
            }
        }
        auto uniform_off = [&](uint32_t it) {
            uint32_t const pit = it << r;
            return ((pit + (pit & hmask)) << lsb) * kEsz;
        };
        if (un.count == 1)
        {
            // one string, ragged pass range: only the needed half of w; scalar +-1 sign
            auto run = [&](auto im_tag) {
                constexpr bool IM = decltype(im_tag)::value;
                uint32_t const zr = zt[0] >> r;
                Acc p0 = Acc(0);
                for (uint32_t i0 = it_begin; i0 < it_end; i0 += U)
                {
                    V va[U], vb[U];
#pragma unroll
                    for (int uu = 0; uu < U; ++uu)
                    {
                        uint32_t const it = i0 + uu < it_end ? i0 + uu : i0;
                        uint32_t const addr = off0 ^ uniform_off(it);
                        va[uu] = *reinterpret_cast<V const *>(smem + addr);
                        vb[uu] = *reinterpret_cast<V const *>(smem + (addr ^ pxor));
                    }
#pragma unroll
                    for (int uu = 0; uu < U; ++uu)
                    {
                        if (i0 + uu < it_end)
                        {
                            Acc const sg = (__popc((i0 + uu) & zr) & 1) ? Acc(-1) : Acc(1);
                            Acc const ar = Acc(va[uu][0]), ai = Acc(va[uu][1]);
                            Acc const br = Acc(vb[uu][0]), bi = Acc(vb[uu][1]);
                            Acc const gv = IM ? fma(ar, bi, -(ai * br)) : fma(ar, br, ai * bi);
                            p0 = fma(sg, gv, p0);
                        }
                    }
                }
                part[0] = p0;
            };
            if (imag[0])
                run(std::true_type{});
            else
                run(std::false_type{});
        }
        else
        {
            for (uint32_t i0 = it_begin; i0 < it_end; i0 += U)
            {
                V va[U], vb[U];
#pragma unroll
                for (int uu = 0; uu < U; ++uu)
                {
                    uint32_t const it = i0 + uu < it_end ? i0 + uu : i0;
                    uint32_t const addr = off0 ^ uniform_off(it);
                    va[uu] = *reinterpret_cast<V const *>(smem + addr);
                    vb[uu] = *reinterpret_cast<V const *>(smem + (addr ^ pxor));
                }
#pragma unroll
                for (int uu = 0; uu < U; ++uu)
                {
                    if (i0 + uu < it_end)
                    {
                        uint32_t const it = i0 + uu;
                        Acc const ar = Acc(va[uu][0]), ai = Acc(va[uu][1]);
                        Acc const br = Acc(vb[uu][0]), bi = Acc(vb[uu][1]);
                        Acc const gre = fma(ar, br, ai * bi);
                        Acc const gim = fma(ar, bi, -(ai * br));
#pragma unroll
                        for (int s = 0; s < kChunk; ++s)
                        {
                            if (s < int(un.count))
                            {
                                Acc const sg = (__popc(it & (zt[s] >> r)) & 1) ? Acc(-1) : Acc(1);
                                part[s] = fma(sg, imag[s] ? gim : gre, part[s]);
                            }
                        }
                    }
                }
            }
        }
        if (!un.lane_live)
        {
#pragma unroll
            for (int s = 0; s < kChunk; ++s)
                part[s] = Acc(0);
        }
#pragma unroll
        for (int s = 0; s < kChunk; ++s)
        {
            if (s < int(un.count))
            {
                TileStr const st = strs[un.first + s];
                bool const neg = ((__popc(p_first & zt[s]) ^ __popcll(q & st.zc)) & 1) != 0;
                double const v = neg ? -double(part[s]) : double(part[s]);
                acc_re += st.sre * v;
                acc_im += st.sim * v;
            }
        }
    };

    auto store_tile = [&](V const(&rg)[EPT]) {
        if constexpr (sizeof(V) == 8)
        {
            if (vec2)
            {
                // 16-byte slot tid + m * kThreads holds row (slot >> (lsb - 1)); MODE 2 stores
                // row a at a ^ ((a >> 1) & 1) as below: row bit 1 is slot bit lsb
                auto *tile4 = reinterpret_cast<V4 *>(smem);
                uint32_t const tsw2 = MODE == 2 ? tid ^ (((tid >> lsb) & 1u) << (lsb - 1u)) : tid;
#pragma unroll
                for (uint32_t m = 0; m < EPT2; ++m)
                    tile4[tsw2 + m * kThreads] = V4{rg[2 * m][0], rg[2 * m][1], rg[2 * m + 1][0], rg[2 * m + 1][1]};
                return;
            }
        }
        // MODE 2 swizzles rows: local row a is stored at a ^ ((a >> 1) & 1), so that the two
        // rows of a 16-lane LDS access group (rows differing in bit 0 or in bit 1 only) always
        // fall on different halves of the 64 banks when a row slice is 128 bytes.  Bit 1 of
        // the row of element tid + m * kThreads is a bit of tid (kThreads >= 2^(lsb + 2)).
        // (tid + m * kThreads) ^ swz = (tid ^ swz) + m * kThreads: one base, immediate offsets
        uint32_t const tsw = MODE == 2 ? tid ^ (((tid >> (lsb + 1u)) & 1u) << lsb) : tid;
        if (full)
        {
#pragma unroll
            for (uint32_t m = 0; m < EPT; ++m)
                tile[tsw + m * kThreads] = rg[m];
        }
        else
        {
#pragma unroll
            for (uint32_t m = 0; m < EPT; ++m)
            {
                if (tid + m * kThreads < tile_elems)
                    tile[tsw + m * kThreads] = rg[m];
            }
        }
    };
    // store_tile(rg) followed by fetch(rg, cb), element by element: each register is
    // reloaded for the next tile as soon as it has been written to LDS, so the next
    // tile's loads are issued during the stores instead of after them (complex128, and
    // the complex64 shared-leader instance; the other complex64 instances keep the
    // separate passes)
    auto store_fetch = [&](V(&rg)[EPT], uint64_t cb) {
        if constexpr (V2)
        {
            // the complex64 shared-leader instance: 16-byte slots (two states), rows swizzled
            auto *tile4 = reinterpret_cast<V4 *>(smem);
            auto const *st4 = reinterpret_cast<V4 const *>(states);
            Idx const cbs4 = Idx(cb << (sshift - 1u));
            uint32_t const tsw2 = tid ^ (((tid >> lsb) & 1u) << (lsb - 1u));
#pragma unroll
            for (uint32_t m = 0; m < EPT2; ++m)
            {
                tile4[tsw2 + m * kThreads] = V4{rg[2 * m][0], rg[2 * m][1], rg[2 * m + 1][0], rg[2 * m + 1][1]};
                V4 const v = __builtin_nontemporal_load(&st4[cbs4 ^ eoff[m]]);
                rg[2 * m] = V{v[0], v[1]};
                rg[2 * m + 1] = V{v[2], v[3]};
            }
        }
        else if constexpr (sizeof(V) == 16)
        {
            Idx const cbs = Idx(s_pow2 ? (cb << sshift) : cb);
            auto index = [&](uint32_t m) {
                return s_pow2 ? Idx(cbs ^ eoff[m]) : Idx((cbs ^ eoff[m]) * Idx(n_states) + Idx(s0 + t));
            };
            uint32_t const tsw = MODE == 2 ? tid ^ (((tid >> (lsb + 1u)) & 1u) << lsb) : tid;
#pragma unroll
            for (uint32_t m = 0; m < EPT; ++m)
            {
                bool const live = full || tid + m * kThreads < tile_elems;
                if (live)
                    tile[tsw + m * kThreads] = rg[m];
                if (full)
                    rg[m] = __builtin_nontemporal_load(&states[index(m)]);
                else
                {
                    rg[m] = V(0);
                    if (live && t_live)
                        rg[m] = __builtin_nontemporal_load(&states[index(m)]);
                }
            }
        }
    };
    auto next_base = [&](uint64_t b) { return ((b | ~cmask) + cstep) & cmask; };
    uint64_t cb = 0;
    if (q0 < n_cosets)
    {
        cb = coset_base(q0, k, grp);
        fetch(reg, cb);
    }
    // tiles q0, q0 + q_step, ... with the next tile in flight during `compute`
    auto tile_loop = [&](auto &&compute) {
        for (uint64_t q = q0; q < n_cosets; q += q_step)
        {
            __syncthreads(); // the previous tile is no longer read
            uint64_t const cbn = next_base(cb);
            bool const more = q + q_step < n_cosets;
            // complex128: the next tile's loads go out during the stores (n = 24, 16 states,
            // 256 strings: 19.10 -> 18.64 ms); complex64 lost with it in the set-up-once
            // instance (10.24 -> 10.62 ms)
            if constexpr (sizeof(V) == 16 || V2)
            {
                if (more)
                    store_fetch(reg, cbn);
                else
                    store_tile(reg);
                __syncthreads();
            }
            else
            {
                store_tile(reg);
                __syncthreads();
                if (more)
                    fetch(reg, cbn);
            }
            compute(q);
            cb = cbn;
        }
    };
    if constexpr (HOIST)
    {
        // every wave has at most one unit, on the fast path: set it up once
        bool const idle = wave >= n_units;
        uint32_t const sb2 = PAIRED ? (sblk >> 1) : sblk;
        Lanes const lg = PAIRED ? Lanes{lsb - 1u, lane & (sb2 - 1u), r + 1u, lane >> (lsb - 1u)} : plain;
        Fast fh{};
        if (!idle)
            fh = make_fast(make_unit(wave, lg, 4u), lg, PAIRED ? 16u : kEsz, 4u);
        auto compute = [&](uint64_t q) {
            if (!idle)
                run_fast_any(fh, q, lg, std::integral_constant<bool, PAIRED>{});
        };
        // complex128 with 32-bit offsets and 64 KiB tiles has the registers for a second tile in
        // flight: tiles j + 1 and j + 2 load while tile j is evaluated
        constexpr bool DEEP = sizeof(V) == 16 && !WIDE && TKB == 64;
        if constexpr (DEEP)
        {
            V reg2[EPT];
            uint64_t nb = next_base(cb); // base of the tile after the two in flight
            if (q0 + q_step < n_cosets)
                fetch(reg2, nb);
            nb = next_base(nb);
            for (uint64_t q = q0; q < n_cosets; q += 2 * q_step)
            {
                __syncthreads(); // the previous tile is no longer read
                store_tile(reg);
                __syncthreads();
                if (q + 2 * q_step < n_cosets)
                    fetch(reg, nb);
                compute(q);
                if (q + q_step >= n_cosets)
                    break;
                uint64_t const nb3 = next_base(nb);
                __syncthreads();
                store_tile(reg2);
                __syncthreads();
                if (q + 3 * q_step < n_cosets)
                    fetch(reg2, nb3);
                compute(q + q_step);
                nb = next_base(nb3);
            }
        }
        else
            tile_loop(compute);
    }
    else if constexpr (MODE == 2)
    {
        // Shared leaders: every string of the group has local bit k set, so the
        // rows with bit k clear lead the pairs of all of them.  Wave w takes passes
        // it = 4w + i (i < 4) of the 2^(d-1) leader rows, 2^rr rows per pass, for
        // every string: one LDS read per leader row and one per (string, leader row).
        //   leader  a_i = af + ((4w + i) << (rr + 1)),  af = insert_zero_bit(pf, k)
        //   partner a_i ^ A = (af ^ A_l) + ((4w << (rr + 1)) ^ A_w) + ((i ^ ci) << (rr + 1))
        // with A_l, ci, A_w the bits of A in the lane, pass and wave parts (disjoint),
        // so both byte addresses are a per-lane base plus (i or i ^ ci) * 128 * ESZ,
        // and the sign parity(a_i & Z) is a per-lane bit (applied once per block)
        // XOR parity(i & zi), zi = the pass part of Z.  k < rr (host condition).
        typedef typename std::conditional<PAIRED, typename fpk::VecOf<float, 2>::type, V>::type LV;
        constexpr uint32_t ESZ = sizeof(LV);
        constexpr int NS = PAIRED ? 2 : 1;    // states per lane
        constexpr uint32_t STRIDE = 128u * ESZ; // one pass: 2^(rr + 1 + lsbx) * ESZ bytes, rr + lsbx = 6
        uint32_t const sb2 = PAIRED ? (sblk >> 1) : sblk;
        Lanes const lg = PAIRED ? Lanes{lsb - 1u, lane & (sb2 - 1u), r + 1u, lane >> (lsb - 1u)} : plain;
        uint32_t const rr = lg.rr;
        constexpr int KMS = max_shared(W);
        uint32_t const n_sh = grp.n_str; // <= KMS (host condition)
        ShStr const *ss = sh + grp.str_first;
        uint32_t const af = insert_zero_bit32(lg.pf, k);
        uint32_t const wpart = (wave * 4u) << (rr + 1u);
        // rows are stored swizzled (see store_tile): row a at a ^ ((a >> 1) & 1), a map that
        // is linear over XOR and touches only row bits 0 and 1, below the pass bits
        auto swz_row = [](uint32_t a) { return a ^ ((a >> 1) & 1u); };
        uint32_t const la0 = ((swz_row(af + wpart) << lg.lsbx) | lg.tx) * ESZ;
        // partner base of string s: row (af | wpart) ^ (A without its pass bits), so
        // byte address la0 ^ px[s] (ESZ a power of two, tx below the row bits)
        uint32_t const passmask = 3u << (rr + 1u);
        uint32_t px[KMS];
        uint32_t neg_bits = 0; // bit s: parity((af + wpart) & Z_s)
        // per string: case (Im flag * 4 + ci) in 3 bits of `cases`, zi in 2 bits of `zis`
        uint32_t cases = 0, zis = 0;
        uint32_t zcs[KMS]; // coset sign masks (coset indices < 2^32: host condition)
        // PK: partner offset in bits 0-15, coset sign mask in bits 16-31 (coset indices and
        // tile byte offsets below 2^16: host condition)
        uint32_t pz[KMS];
        Acc acc_s[KMS][NS];
#pragma unroll
        for (int s = 0; s < KMS; ++s)
        {
            px[s] = 0;
            zcs[s] = 0;
#pragma unroll
            for (int n = 0; n < NS; ++n)
                acc_s[s][n] = Acc(0);
            if (s < int(n_sh))
            {
                zcs[s] = uint32_t(ss[s].zc);
                uint32_t const A = ss[s].A, Z = ss[s].Z;
                px[s] = (swz_row(A & ~passmask) << lg.lsbx) * ESZ;
                neg_bits |= uint32_t(__popc((af + wpart) & Z) & 1) << s;
                cases |= (((ss[s].imag ? 4u : 0u) + ((A >> (rr + 1u)) & 3u)) << (3 * s));
                zis |= ((Z >> (rr + 1u)) & 3u) << (2 * s);
            }
            pz[s] = 0;
            if constexpr (PK)
                pz[s] = __builtin_amdgcn_readfirstlane(px[s] | (zcs[s] << 16));
        }
        cases = __builtin_amdgcn_readfirstlane(cases);
        zis = __builtin_amdgcn_readfirstlane(zis);
        auto compute = [&](uint64_t q) {
            // Per-tile copies the compiler must re-read, so that what is derived from them
            // is recomputed each tile in scalar registers instead of being held across the
            // loop, where it spilled to vector lanes.  Complex64: the per-string sign factors
            // and string-count tests (its string loop had pushed vector registers into
            // scratch; n = 24, 16 states: 9.21 -> 8.87 ms).  Complex128: the string-count
            // tests only (18.67 -> 18.60 ms; re-reading the sign factors too: 18.58 against
            // 18.50 ms without either).
            uint32_t zis_t = zis, nsh_t = n_sh;
            if constexpr (sizeof(V) == 8)
                asm volatile("" : "+s"(zis_t), "+s"(nsh_t));
            else
                asm volatile("" : "+s"(nsh_t));
            // coset signs of every string
            uint32_t cneg = 0;
            if constexpr (PK)
            {
                uint32_t const qq = uint32_t(q) << 16; // low half clear: masks out the offsets
#pragma unroll
                for (int s = 0; s < KMS; ++s)
                    cneg |= uint32_t(__popc(qq & pz[s]) & 1) << s;
            }
            else
            {
#pragma unroll
                for (int s = 0; s < KMS; ++s)
                    cneg |= uint32_t(__popc(uint32_t(q) & zcs[s]) & 1) << s;
            }
            LV L[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                L[i] = *reinterpret_cast<LV const *>(smem + (la0 + uint32_t(i) * STRIDE));
            // partner values of string s
            auto load_R = [&](int s, LV(&R)[4]) {
                uint32_t const base = la0 ^ (PK ? (pz[s] & 0xffffu) : px[s]);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    R[j] = *reinterpret_cast<LV const *>(smem + (base + uint32_t(j) * STRIDE));
            };
            auto eval = [&](int s, LV const(&R)[4]) {
                uint32_t const cse = (cases >> (3 * s)) & 7u;
                uint32_t const zi = (zis_t >> (2 * s)) & 3u;
                Acc const s1 = (zi & 1u) ? Acc(-1) : Acc(1);
                Acc const s2 = (zi & 2u) ? Acc(-1) : Acc(1);
                Acc const cs = ((cneg >> s) & 1u) ? Acc(-1) : Acc(1);
                // sum_i (-1)^{parity(i & zi)} G(L[i], R[i ^ ci]) per state
                auto body = [&](auto im_tag, auto ci_tag) {
                    constexpr bool IM = decltype(im_tag)::value;
                    constexpr int CI = decltype(ci_tag)::value;
#pragma unroll
                    for (int n = 0; n < NS; ++n)
                    {
                        Acc G[4];
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                        {
                            Acc const lr = Acc(L[i][2 * n]), li = Acc(L[i][2 * n + 1]);
                            Acc const pr = Acc(R[i ^ CI][2 * n]), pi = Acc(R[i ^ CI][2 * n + 1]);
                            G[i] = IM ? fma(lr, pi, -(li * pr)) : fma(lr, pr, li * pi);
                        }
                        Acc const t0 = fma(s1, G[1], G[0]);
                        Acc const t1 = fma(s1, G[3], G[2]);
                        acc_s[s][n] = fma(cs, fma(s2, t1, t0), acc_s[s][n]);
                    }
                };
                typedef std::true_type Im;
                typedef std::false_type Re;
                switch (cse)
                {
                case 0: body(Re{}, std::integral_constant<int, 0>{}); break;
                case 1: body(Re{}, std::integral_constant<int, 1>{}); break;
                case 2: body(Re{}, std::integral_constant<int, 2>{}); break;
                case 3: body(Re{}, std::integral_constant<int, 3>{}); break;
                case 4: body(Im{}, std::integral_constant<int, 0>{}); break;
                case 5: body(Im{}, std::integral_constant<int, 1>{}); break;
                case 6: body(Im{}, std::integral_constant<int, 2>{}); break;
                default: body(Im{}, std::integral_constant<int, 3>{}); break;
                }
            };
#pragma unroll
            for (int s = 0; s < KMS; ++s)
            {
                if (s < int(nsh_t))
                {
                    LV R[4];
                    load_R(s, R);
                    eval(s, R);
                }
            }
        };
        tile_loop(compute);
        // per-lane sign and string scale, once per block
#pragma unroll
        for (int s = 0; s < KMS; ++s)
        {
            if (s < int(n_sh))
            {
                double const sre = ss[s].sre, sim = ss[s].sim;
                bool const neg = ((neg_bits >> s) & 1u) != 0;
#pragma unroll
                for (int n = 0; n < NS; ++n)
                {
                    double const v = neg ? -double(acc_s[s][n]) : double(acc_s[s][n]);
                    if (n == 0)
                    {
                        acc_re += sre * v;
                        acc_im += sim * v;
                    }
                    else
                    {
                        acc_re1 += sre * v;
                        acc_im1 += sim * v;
                    }
                }
            }
        }
    }
    else
    {
        tile_loop([&](uint64_t q) {
            for (uint32_t u = wave; u < n_units; u += W)
            {
                Unit const un = make_unit(u, plain, U1);
                if (un.fast)
                    run_fast_any(make_fast(un, plain, kEsz, U1), q, plain, std::false_type{});
                else
                    run_general(un, q);
            }
        });
    }
    // block sum per state, in a fixed order
    __syncthreads();
    double *red = reinterpret_cast<double *>(smem);
    red[tid] = acc_re;
    red[kThreads + tid] = acc_im;
    if constexpr (PAIRED)
    {
        red[2 * kThreads + tid] = acc_re1;
        red[3 * kThreads + tid] = acc_im1;
    }
    __syncthreads();
    if (tid < sblk && s0 + tid < n_states)
    {
        // plain: threads tid + j sblk hold state tid; paired: threads t2 + j (sblk / 2)
        // hold state 2 t2 (first pair) and 2 t2 + 1 (second pair)
        uint32_t const stride = PAIRED ? (sblk >> 1) : sblk;
        uint32_t const first = PAIRED ? (tid >> 1) : tid;
        uint32_t const base = PAIRED ? (tid & 1u) * 2u * uint32_t(kThreads) : 0u;
        double sr = 0.0, si = 0.0;
        for (uint32_t j = first; j < uint32_t(kThreads); j += stride)
        {
            sr += red[base + j];
            si += red[base + kThreads + j];
        }
        partial[uint64_t(part) * n_states + s0 + tid] = fpk::Dbl2{sr, si};
    }
}

} // namespace fpt
