// fp_apply_tiled.hpp — PauliOp::apply with coset tiles in LDS, bit-exact with
// the sequential string order.
//
// The streaming kernel (k_op_apply) reads a partner row of the state for every
// string of every output row: N passes over the state.  Here the strings are
// cut into groups of CONSECUTIVE strings (original order) whose x masks span,
// with the low k row bits, a subspace V of dimension d (the same coset tiles as
// the tiled expectation kernel).  One launch per group, in order: every block
// loads coset tiles of the input state into LDS, and each thread adds the
// group's strings, in order, to the output elements it owns:
//     out(r, t) = out(r, t) + cf_s (-1)^{parity(r & z_s)} psi(r ^ x_s, t)
// with the partner element read from the tile (local row a ^ A_s).  The first
// group starts from 0 (or from out when accumulating), so every output element
// sees exactly the operations of the streaming kernel (cmul_exact, add_exact,
// no fused multiply-add) in the same order: results are IEEE-equal to it.
// HBM traffic per group is one read of the state plus one read and one write of
// the output, instead of one state read per string.
//
// A thread owns "slots" of the tile: one element, or (PAIRED: complex64 full
// tiles with an even state count, S_blk >= 2) the two neighbouring states of one
// row, moved by 16-byte loads, stores and LDS reads.  Coset bases
// advance by one deposited addition per tile.  complex128 (32-bit offsets)
// requests the next coset's output values together with its state values, in a
// second register set; other forms request a coset's output values just before
// its tile store, so their latency overlaps the store.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "fp_ev_tiled.hpp"
#include "fp_kernels.hpp"

namespace fpa
{

constexpr int kThreads = 512;
constexpr uint32_t kTileBytes = 64u * 1024u;

// One group: launch over cosets [0, n_cosets) of group g, blocks striding by
// gridDim.x; state chunk blockIdx.y.  PAIRED requires complex64, S_blk >= 2, an
// even state count (so a slot's two states are both live or both not) and
// 2^d rows x S_blk = one full tile.
// WEIGHTED (SummedPauliOp::apply_weighted): string s of state t has the coefficient
// (-i)^{nY_s} w(s, t), with w the (strings x states) complex table `wts` and the
// records' coefficient the unit (-i)^{nY_s}.  The unit is applied to w by exact
// negation and swap, and the product with the partner value uses the same four
// products and two sums as (-i)^{nY} psi times w, so the terms are IEEE-equal to
// the sequential loop's.
template <class T, bool WIDE, bool PAIRED, bool WEIGHTED>
__global__ __launch_bounds__(kThreads, 4) void k_op_apply_tiled(typename fpk::VecOf<T, 1>::type const *__restrict__ states,
                                                             typename fpk::VecOf<T, 1>::type *__restrict__ out,
                                                             uint64_t n_states, uint32_t sshift, uint32_t lsb,
                                                             uint32_t k, uint32_t d, uint64_t n_cosets,
                                                             fpt::TileGroup const *__restrict__ group,
                                                             uint64_t const *__restrict__ btab,
                                                             ApplyStr const *__restrict__ strs, int init_zero,
                                                             T const *__restrict__ wts)
{
    static_assert(!PAIRED || sizeof(T) == 4, "two states per slot: complex64 only");
    typedef typename fpk::VecOf<T, 1>::type V;
    typedef typename std::conditional<PAIRED, typename fpk::VecOf<float, 2>::type, V>::type LV; // one slot
    constexpr uint32_t NS = PAIRED ? 2 : 1;                                                   // states per slot
    constexpr uint32_t kSlots = kTileBytes / sizeof(LV);
    constexpr uint32_t EPT = kSlots / kThreads;
    constexpr uint32_t kLsz = sizeof(LV);
    extern __shared__ __align__(16) unsigned char smem[];
    LV *tile = reinterpret_cast<LV *>(smem);
    LV const *src_l = reinterpret_cast<LV const *>(states);
    LV *out_l = reinterpret_cast<LV *>(out);

    uint32_t const tid = threadIdx.x;
    fpt::TileGroup const grp = *group;
    uint64_t const *bt = btab + grp.btab_off;
    // for apply groups, n_chunks / chunk_first count and locate the group's string
    // records (in the operator's order)
    uint32_t const n_str = grp.n_chunks;
    ApplyStr const *gs = strs + grp.chunk_first;

    uint32_t const sblk = 1u << lsb;
    uint32_t const lsbx = PAIRED ? lsb - 1u : lsb; // slots per row: 2^lsbx
    uint32_t const tx = tid & ((1u << lsbx) - 1u);
    uint32_t const t0 = PAIRED ? 2u * tx : tx; // first state of this thread's slots within the block
    uint64_t const s0 = uint64_t(blockIdx.y) << lsb;
    bool const t_live = s0 + t0 < n_states;
    uint32_t const tile_slots = (1u << d) << lsbx;
    uint32_t const lowmask = (1u << k) - 1u;
    // every slot of every tile of this block is live: loads and stores need no guard
    bool const full = tile_slots == kSlots && s0 + sblk <= n_states;

    typedef typename std::conditional<WIDE, uint64_t, uint32_t>::type Idx;
    bool const s_pow2 = sshift < 64u;
    // slot m is tile slot tid + m * kThreads: local row a_m = (tid >> lsbx) + (m << rm)
    Idx eoff[EPT];
#pragma unroll
    for (uint32_t m = 0; m < EPT; ++m)
    {
        uint32_t const e = tid + m * kThreads;
        uint32_t const a = e >> lsbx;
        bool const live = e < tile_slots && t_live;
        uint64_t const roff = live ? (bt[a >> k] ^ uint64_t(a & lowmask)) : 0;
        uint64_t const el = s_pow2 ? ((roff << sshift) | (s0 + t0)) : 0; // element index, power-of-two S
        eoff[m] = Idx(s_pow2 ? (PAIRED ? (el >> 1) : el) : roff);
    }
    // slot index of slot m in coset base cb (elements; pairs of elements when PAIRED)
    auto index_of = [&](uint64_t cb, uint32_t m) -> Idx {
        if constexpr (PAIRED)
            return s_pow2 ? Idx(Idx(cb << (sshift - 1u)) ^ eoff[m])
                          : Idx((Idx(cb) ^ eoff[m]) * Idx(n_states >> 1) + Idx((s0 + t0) >> 1));
        else
        {
            Idx const cbs = Idx(s_pow2 ? (cb << sshift) : cb);
            return s_pow2 ? Idx(cbs ^ eoff[m]) : Idx((cbs ^ eoff[m]) * Idx(n_states) + Idx(s0 + t0));
        }
    };
    auto slot_live = [&](uint32_t m) { return full || (tid + m * kThreads < tile_slots && t_live); };

    LV reg[EPT];
    auto fetch = [&](uint64_t cb) {
#pragma unroll
        for (uint32_t m = 0; m < EPT; ++m)
        {
            reg[m] = LV(0);
            if (slot_live(m))
                reg[m] = __builtin_nontemporal_load(&src_l[index_of(cb, m)]);
        }
    };

    // cosets blockIdx.x, + gridDim.x, ...; base(q + step) = ((base(q) | ~cmask) +
    // base(step)) & cmask: the pivot bits, set to 1, pass carries through
    uint64_t const q_step = gridDim.x;
    uint64_t const cmask = fpt::coset_base(n_cosets - 1u, k, grp);
    uint64_t const cstep = fpt::coset_base(q_step, k, grp);
    uint64_t cb = 0;
    if (blockIdx.x < n_cosets)
    {
        cb = fpt::coset_base(blockIdx.x, k, grp);
        fetch(cb);
    }
    // Slot m holds local row a_m = a_lane + (m << rm), a_lane = tid >> lsbx < 2^rm
    // (disjoint bits).  For a string (A, Z): the partner row a_m ^ A = (a_lane ^
    // A_lane) + ((m << rm) ^ A_m) and the sign bit parity(a_m & Z) =
    // parity(a_lane & Z) ^ parity((m << rm) & Z): a per-lane pointer and sign per
    // string, plus a wave-uniform offset and bit per slot m.
    uint32_t const rm = 9u - lsbx; // log2(kThreads) - lsbx
    uint32_t const a_lane = tid >> lsbx;
    uint32_t const lane_mask = (1u << rm) - 1u;
    // this coset's output values (zero for the first group of an overwrite)
    auto read_out = [&](LV(&acc)[EPT], uint64_t cbq) {
#pragma unroll
        for (uint32_t m = 0; m < EPT; ++m)
        {
            acc[m] = LV(0);
            if (!init_zero && slot_live(m))
                acc[m] = out_l[index_of(cbq, m)];
        }
    };
    auto write_out = [&](LV const(&acc)[EPT], uint64_t cbq) {
#pragma unroll
        for (uint32_t m = 0; m < EPT; ++m)
            if (slot_live(m))
                __builtin_nontemporal_store(acc[m], &out_l[index_of(cbq, m)]);
    };
    auto store_tile = [&]() {
#pragma unroll
        for (uint32_t m = 0; m < EPT; ++m)
        {
            uint32_t const e = tid + m * kThreads;
            if (full || e < tile_slots)
                tile[e] = reg[m];
        }
    };
    // WEIGHTED: this thread's weights w(s, t) of string s (zero for dead states)
    auto load_w = [&](uint32_t s, T (&wr)[NS], T (&wi)[NS]) {
#pragma unroll
        for (uint32_t j = 0; j < NS; ++j)
        {
            uint64_t const t = s0 + t0 + j;
            wr[j] = T(0);
            wi[j] = T(0);
            if (t < n_states)
            {
                wr[j] = wts[2 * (uint64_t(gs[s].s) * n_states + t)];
                wi[j] = wts[2 * (uint64_t(gs[s].s) * n_states + t) + 1];
            }
        }
    };
    // the group's strings, in order, on the tile in LDS
    auto compute = [&](uint64_t q, LV(&acc)[EPT]) {
        T nwr[NS], nwi[NS]; // WEIGHTED: the next string's weights, requested one string ahead
        if constexpr (WEIGHTED)
            load_w(0, nwr, nwi);
        for (uint32_t s = 0; s < n_str; ++s)
        {
            ApplyStr const st = gs[s];
            uint32_t const cs = uint32_t(__popcll(q & st.zc) & 1u);
            T cr, ci;
            if constexpr (sizeof(T) == 8)
            {
                cr = st.cr;
                ci = st.ci;
            }
            else
            {
                cr = st.fr;
                ci = st.fi;
            }
            uint32_t const A_lane = st.A & lane_mask, A_m = st.A & ~lane_mask;
            unsigned char const *pl = smem + ((((a_lane ^ A_lane) << lsbx) | tx) * kLsz);
            // lane sign folded into the coefficient (negation is exact)
            bool const lneg = ((uint32_t(__popc(a_lane & st.Z)) ^ cs) & 1u) != 0;
            T lcr[NS], lci[NS];
            if constexpr (WEIGHTED)
            {
                T wr[NS], wi[NS];
#pragma unroll
                for (uint32_t j = 0; j < NS; ++j)
                {
                    wr[j] = nwr[j];
                    wi[j] = nwi[j];
                }
                if (s + 1 < n_str)
                    load_w(s + 1, nwr, nwi);
                // (cr, ci) = (-i)^{nY}: (1, 0), (0, -1), (-1, 0) or (0, 1)
#pragma unroll
                for (uint32_t j = 0; j < NS; ++j)
                {
                    T ur, ui;
                    if (cr != T(0))
                    {
                        ur = cr > T(0) ? wr[j] : -wr[j];
                        ui = cr > T(0) ? wi[j] : -wi[j];
                    }
                    else
                    {
                        ur = ci < T(0) ? wi[j] : -wi[j];
                        ui = ci < T(0) ? -wr[j] : wr[j];
                    }
                    lcr[j] = lneg ? -ur : ur;
                    lci[j] = lneg ? -ui : ui;
                }
            }
            else
            {
#pragma unroll
                for (uint32_t j = 0; j < NS; ++j)
                {
                    lcr[j] = lneg ? -cr : cr;
                    lci[j] = lneg ? -ci : ci;
                }
            }
#pragma unroll
            for (uint32_t m = 0; m < EPT; ++m)
            {
                uint32_t const mrow = m << rm;
                LV const src = *reinterpret_cast<LV const *>(pl + (((mrow ^ A_m) << lsbx) * kLsz));
                uint32_t const mbit = uint32_t(__popc(mrow & st.Z) & 1u) << 31;
#pragma unroll
                for (uint32_t j = 0; j < NS; ++j)
                {
                    T const a = fpt::xor_sign_bit(lcr[j], mbit);
                    T const b = fpt::xor_sign_bit(lci[j], mbit);
                    T orr, oi;
                    fpk::cmul_exact<T>(a, b, src[2 * j], src[2 * j + 1], orr, oi);
                    acc[m][2 * j] = fpk::add_exact<T>(acc[m][2 * j], orr);
                    acc[m][2 * j + 1] = fpk::add_exact<T>(acc[m][2 * j + 1], oi);
                }
            }
        }
    };
    // complex128 with 32-bit offsets has the registers to request the next coset's
    // output values with its state values (two output sets, alternating); other
    // forms (and the weighted one, whose weights take the registers) request a
    // coset's output values just before its tile store
    constexpr bool PF = sizeof(T) == 8 && !WIDE && !WEIGHTED;
    if constexpr (PF)
    {
        LV acc0[EPT], acc1[EPT];
        if (blockIdx.x < n_cosets)
            read_out(acc0, cb);
        for (uint64_t q = blockIdx.x; q < n_cosets; q += 2 * q_step)
        {
            __syncthreads(); // the previous tile is no longer read
            store_tile();
            __syncthreads();
            uint64_t const cb1 = ((cb | ~cmask) + cstep) & cmask;
            if (q + q_step < n_cosets)
            {
                fetch(cb1);
                read_out(acc1, cb1);
            }
            compute(q, acc0);
            write_out(acc0, cb);
            if (q + q_step >= n_cosets)
                break;
            __syncthreads();
            store_tile();
            __syncthreads();
            uint64_t const cb2 = ((cb1 | ~cmask) + cstep) & cmask;
            if (q + 2 * q_step < n_cosets)
            {
                fetch(cb2);
                read_out(acc0, cb2);
            }
            compute(q + q_step, acc1);
            write_out(acc1, cb1);
            cb = cb2;
        }
    }
    else
    {
        for (uint64_t q = blockIdx.x; q < n_cosets; q += q_step)
        {
            LV acc[EPT];
            read_out(acc, cb); // requested before the tile store
            __syncthreads();   // the previous tile is no longer read
            store_tile();
            __syncthreads();
            uint64_t const cbn = ((cb | ~cmask) + cstep) & cmask;
            if (q + q_step < n_cosets)
                fetch(cbn);
            compute(q, acc);
            write_out(acc, cb);
            cb = cbn;
        }
    }
}

} // namespace fpa
