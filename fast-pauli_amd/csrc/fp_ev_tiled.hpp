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
// Each block accumulates FP64 sums for its state and its range of cosets and
// writes one partial per state; partials are reduced in a fixed order, so
// results are identical run to run.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fp_kernels.hpp"

namespace fpt
{

constexpr int kThreads = 512;
constexpr uint32_t kTileBytes = 64u * 1024u;
constexpr int kMaxDh = 9; // B(a_hi) table: at most 2^9 entries (4 KiB of LDS)

struct TileGroup
{
    uint32_t n_chunks;
    uint32_t chunk_first;
    uint32_t n_piv;    // dh: pivots (high dimensions of V)
    uint32_t btab_off; // first entry of this group's B table (2^dh entries)
    uint64_t piv_lo;   // pivot bit positions 0..7, ascending, 8 bits each
    uint64_t piv_hi;   // pivot bit positions 8..15
};

struct TileChunk
{
    uint32_t A;     // local x mask; 0: every row, kappa 1
    uint32_t h;     // top bit of A
    uint32_t count; // strings in this chunk (1..kChunk), all with local mask A
    uint32_t first; // first TileStr
};

struct TileStr
{
    uint64_t z;    // global z mask (coset sign)
    uint32_t zl;   // local z mask
    uint32_t sidx; // string index (scale table row)
    uint32_t imag; // 1: Im w (odd nY), 0: Re w
    uint32_t pad[3];
};

constexpr int kChunk = 4;

__device__ __forceinline__ uint32_t insert_zero_bit32(uint32_t v, uint32_t b)
{
    uint32_t const lo = v & ((1u << b) - 1u);
    return ((v >> b) << (b + 1)) | lo;
}

__device__ __forceinline__ uint64_t coset_base(uint64_t q, uint32_t k, TileGroup const &g)
{
    uint64_t v = q << k;
    for (uint32_t i = 0; i < g.n_piv; ++i)
        v = fpk::insert_zero_bit(v, uint32_t((i < 8 ? g.piv_lo : g.piv_hi) >> (8 * (i & 7))) & 0xffu);
    return v;
}

template <class T>
__global__ __launch_bounds__(kThreads, 2) void k_op_expect_tiled(typename fpk::VecOf<T, 1>::type const *__restrict__ states,
                                                              uint64_t n_states, uint32_t lsb, uint32_t k, uint32_t d,
                                                              uint32_t ranges, uint64_t per_range, uint64_t n_cosets,
                                                              TileGroup const *__restrict__ groups,
                                                              uint64_t const *__restrict__ btab,
                                                              TileChunk const *__restrict__ chunks,
                                                              TileStr const *__restrict__ strs,
                                                              double const *__restrict__ scale,
                                                              fpk::Dbl2 *__restrict__ partial)
{
    typedef typename fpk::VecOf<T, 1>::type V;
    typedef typename fpk::RowAcc<T>::type Acc;
    constexpr uint32_t kElems = kTileBytes / sizeof(V);
    constexpr uint32_t EPT = kElems / kThreads;
    constexpr int U = 4;
    extern __shared__ __align__(16) unsigned char smem[];
    V *tile = reinterpret_cast<V *>(smem);
    uint64_t *bt = reinterpret_cast<uint64_t *>(smem + kTileBytes);

    uint32_t const tid = threadIdx.x;
    uint32_t const g = blockIdx.x / ranges, rg = blockIdx.x % ranges;
    TileGroup const grp = groups[g];
    uint32_t const n_bt = 1u << grp.n_piv;
    for (uint32_t i = tid; i < n_bt; i += kThreads)
        bt[i] = btab[grp.btab_off + i];

    uint32_t const sblk = 1u << lsb;
    uint32_t const t = tid & (sblk - 1u);
    uint64_t const s0 = uint64_t(blockIdx.y) << lsb;
    bool const t_live = s0 + t < n_states;
    uint32_t const tile_elems = (1u << d) << lsb;
    uint32_t const lowmask = (1u << k) - 1u;
    uint64_t const q0 = uint64_t(rg) * per_range;
    uint64_t const q1 = q0 + per_range < n_cosets ? q0 + per_range : n_cosets;
    __syncthreads(); // bt ready

    V reg[EPT];
    auto fetch = [&](uint64_t q) {
        uint64_t const cb = coset_base(q, k, grp);
#pragma unroll
        for (uint32_t m = 0; m < EPT; ++m)
        {
            uint32_t const e = tid + m * kThreads;
            uint32_t const a = e >> lsb;
            reg[m] = V(0);
            if (e < tile_elems && t_live)
            {
                uint64_t const row = cb ^ bt[a >> k] ^ uint64_t(a & lowmask);
                reg[m] = __builtin_nontemporal_load(&states[row * n_states + s0 + t]);
            }
        }
    };

    double acc_re = 0.0, acc_im = 0.0;
    if (q0 < q1)
        fetch(q0);
    uint32_t const rows_step = kThreads >> lsb;
    uint32_t const p_first = tid >> lsb;
    for (uint64_t q = q0; q < q1; ++q)
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
        if (q + 1 < q1)
            fetch(q + 1);
        uint64_t const cb = coset_base(q, k, grp);
        for (uint32_t c = 0; c < grp.n_chunks; ++c)
        {
            TileChunk const ch = chunks[grp.chunk_first + c];
            uint32_t const n_rows = ch.A ? (1u << (d - 1)) : (1u << d);
            uint32_t zl[kChunk];
            bool imag[kChunk];
            Acc part[kChunk];
#pragma unroll
            for (int s = 0; s < kChunk; ++s)
            {
                part[s] = Acc(0);
                zl[s] = 0;
                imag[s] = false;
                if (s < int(ch.count))
                {
                    zl[s] = strs[ch.first + s].zl;
                    imag[s] = strs[ch.first + s].imag != 0;
                }
            }
            for (uint32_t p0 = p_first; p0 < n_rows; p0 += rows_step * U)
            {
                V va[U], vb[U];
                uint32_t aa[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    uint32_t const p = p0 + uint32_t(u) * rows_step;
                    uint32_t const pc = p < n_rows ? p : p_first;
                    uint32_t const a = ch.A ? insert_zero_bit32(pc, ch.h) : pc;
                    aa[u] = a;
                    va[u] = tile[(a << lsb) + t];
                    vb[u] = tile[((a ^ ch.A) << lsb) + t];
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    if (p0 + uint32_t(u) * rows_step < n_rows)
                    {
                        Acc const ar = Acc(va[u][0]), ai = Acc(va[u][1]), br = Acc(vb[u][0]), bi = Acc(vb[u][1]);
                        Acc const gre = ar * br + ai * bi;
                        Acc const gim = ar * bi - ai * br;
#pragma unroll
                        for (int s = 0; s < kChunk; ++s)
                        {
                            if (s < int(ch.count))
                            {
                                Acc const gv = imag[s] ? gim : gre;
                                part[s] += (__popc(aa[u] & zl[s]) & 1) ? -gv : gv;
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int s = 0; s < kChunk; ++s)
            {
                if (s < int(ch.count))
                {
                    TileStr const st = strs[ch.first + s];
                    double sr = scale[2 * st.sidx], si = scale[2 * st.sidx + 1];
                    if (fpk::parity64(cb & st.z))
                    {
                        sr = -sr;
                        si = -si;
                    }
                    acc_re += sr * double(part[s]);
                    acc_im += si * double(part[s]);
                }
            }
        }
    }
    // block sum per state, in a fixed order
    __syncthreads();
    double *red = reinterpret_cast<double *>(smem);
    red[tid] = acc_re;
    red[kThreads + tid] = acc_im;
    __syncthreads();
    if (tid < sblk && s0 + tid < n_states)
    {
        double sr = 0.0, si = 0.0;
        for (uint32_t j = tid; j < uint32_t(kThreads); j += sblk)
        {
            sr += red[j];
            si += red[kThreads + j];
        }
        partial[uint64_t(blockIdx.x) * n_states + s0 + tid] = fpk::Dbl2{sr, si};
    }
}

} // namespace fpt
