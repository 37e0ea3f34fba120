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

struct ApplyStr
{
    uint64_t zc;  // z compressed onto the coset-index bits: coset sign = parity(q & zc)
    uint32_t A;   // local x mask
    uint32_t Z;   // local z mask
    double cr;    // folded coefficient c_s (-i)^{nY_s}, complex128
    double ci;
    float fr;     // the same, complex64 (rounded as the streaming kernel's table)
    float fi;
    uint32_t pad[2];
};

// One group: launch over cosets [0, n_cosets) of group g, blocks striding by
// gridDim.x; state chunk blockIdx.y.
template <class T, bool WIDE>
__global__ __launch_bounds__(kThreads, 4) void k_op_apply_tiled(typename fpk::VecOf<T, 1>::type const *__restrict__ states,
                                                             typename fpk::VecOf<T, 1>::type *__restrict__ out,
                                                             uint64_t n_states, uint32_t sshift, uint32_t lsb,
                                                             uint32_t k, uint32_t d, uint64_t n_cosets,
                                                             fpt::TileGroup const *__restrict__ group,
                                                             uint64_t const *__restrict__ btab,
                                                             ApplyStr const *__restrict__ strs, int init_zero)
{
    typedef typename fpk::VecOf<T, 1>::type V;
    constexpr uint32_t kElems = kTileBytes / sizeof(V);
    constexpr uint32_t EPT = kElems / kThreads;
    constexpr uint32_t kEsz = sizeof(V);
    extern __shared__ __align__(16) unsigned char smem[];
    V *tile = reinterpret_cast<V *>(smem);

    uint32_t const tid = threadIdx.x;
    fpt::TileGroup const grp = *group;
    uint64_t const *bt = btab + grp.btab_off;
    // for apply groups, n_chunks / chunk_first count and locate the group's string
    // records (in the operator's order)
    uint32_t const n_str = grp.n_chunks;
    ApplyStr const *gs = strs + grp.chunk_first;

    uint32_t const sblk = 1u << lsb;
    uint32_t const t = tid & (sblk - 1u);
    uint64_t const s0 = uint64_t(blockIdx.y) << lsb;
    bool const t_live = s0 + t < n_states;
    uint32_t const tile_elems = (1u << d) << lsb;
    uint32_t const lowmask = (1u << k) - 1u;

    typedef typename std::conditional<WIDE, uint64_t, uint32_t>::type Idx;
    bool const s_pow2 = sshift < 64u;
    Idx eoff[EPT];
#pragma unroll
    for (uint32_t m = 0; m < EPT; ++m)
    {
        uint32_t const e = tid + m * kThreads;
        uint32_t const a = e >> lsb;
        bool const live = e < tile_elems && t_live;
        uint64_t const roff = live ? (bt[a >> k] ^ uint64_t(a & lowmask)) : 0;
        eoff[m] = Idx(s_pow2 ? ((roff << sshift) | (s0 + t)) : roff);
    }

    auto index_of = [&](uint64_t cb, uint32_t m) -> Idx {
        Idx const cbs = Idx(s_pow2 ? (cb << sshift) : cb);
        return s_pow2 ? Idx(cbs ^ eoff[m]) : Idx((cbs ^ eoff[m]) * Idx(n_states) + Idx(s0 + t));
    };

    V reg[EPT];
    auto fetch = [&](uint64_t q) {
        uint64_t const cb = fpt::coset_base(q, k, grp);
#pragma unroll
        for (uint32_t m = 0; m < EPT; ++m)
        {
            reg[m] = V(0);
            if (tid + m * kThreads < tile_elems && t_live)
                reg[m] = __builtin_nontemporal_load(&states[index_of(cb, m)]);
        }
    };

    uint64_t const q_step = gridDim.x;
    if (blockIdx.x < n_cosets)
        fetch(blockIdx.x);
    for (uint64_t q = blockIdx.x; q < n_cosets; q += q_step)
    {
        __syncthreads(); // the previous tile is no longer read
#pragma unroll
        for (uint32_t m = 0; m < EPT; ++m)
        {
            uint32_t const e = tid + m * kThreads;
            if (e < tile_elems)
                tile[e] = reg[m];
        }
        __syncthreads();
        uint64_t const cb = fpt::coset_base(q, k, grp);
        if (q + q_step < n_cosets)
            fetch(q + q_step);
        V acc[EPT];
#pragma unroll
        for (uint32_t m = 0; m < EPT; ++m)
        {
            acc[m] = V(0);
            if (!init_zero && tid + m * kThreads < tile_elems && t_live)
                acc[m] = out[index_of(cb, m)];
        }
        // Element m of this thread is local row a_m = a_lane + (m << rm), a_lane =
        // tid >> lsb < 2^rm (disjoint bits).  For a string (A, Z): the partner row
        // a_m ^ A = (a_lane ^ A_lane) + ((m << rm) ^ A_m) and the sign bit
        // parity(a_m & Z) = parity(a_lane & Z) ^ parity((m << rm) & Z): a per-lane
        // pointer and sign per string, plus a wave-uniform offset and bit per slot m.
        uint32_t const rm = 9u - lsb; // log2(kThreads) - lsb
        uint32_t const a_lane = tid >> lsb;
        uint32_t const lane_mask = (1u << rm) - 1u;
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
            unsigned char const *pl = smem + ((((a_lane ^ A_lane) << lsb) | t) * kEsz);
            // lane sign folded into the coefficient (negation is exact)
            bool const lneg = ((uint32_t(__popc(a_lane & st.Z)) ^ cs) & 1u) != 0;
            T const lcr = lneg ? -cr : cr;
            T const lci = lneg ? -ci : ci;
#pragma unroll
            for (uint32_t m = 0; m < EPT; ++m)
            {
                uint32_t const mrow = m << rm;
                V const src = *reinterpret_cast<V const *>(pl + (((mrow ^ A_m) << lsb) * kEsz));
                uint32_t const mbit = uint32_t(__popc(mrow & st.Z) & 1u) << 31;
                T const a = fpt::xor_sign_bit(lcr, mbit);
                T const b = fpt::xor_sign_bit(lci, mbit);
                T orr, oi;
                fpk::cmul_exact<T>(a, b, src[0], src[1], orr, oi);
                acc[m][0] = fpk::add_exact<T>(acc[m][0], orr);
                acc[m][1] = fpk::add_exact<T>(acc[m][1], oi);
            }
        }
#pragma unroll
        for (uint32_t m = 0; m < EPT; ++m)
            if (tid + m * kThreads < tile_elems && t_live)
                __builtin_nontemporal_store(acc[m], &out[index_of(cb, m)]);
    }
}

} // namespace fpa
