// fp_ev_lds.hpp — per-string expectation values with the state held in LDS.
//
// For n <= 14 (complex64) or n <= 13 (complex128) one state column fits in a
// workgroup's LDS (160 KiB on gfx950).  A block loads column t once (from a
// (states x rows) transposed copy, so the load is contiguous) and evaluates
// every string against it; the state is read from HBM once per block instead
// of once per string.
//
// Strings are grouped on the host into chunks of up to kChunk strings sharing
// the same x mask and the same half of w (Re w for even nY, Im w for odd nY),
// so the LDS reads and the product w = conj(psi_j) psi_{j^x} of a row serve every
// string of the chunk, and only the needed half of w is computed.  Rows are
// enumerated as j = u * 64 + lane with u uniform across the wave, which splits
// the sign (-1)^{popcount(j & z)} into a per-lane bit (fixed for the chunk) and a
// wave-uniform bit (popcount(u & z >> 6), scalar ALU).
//   mode kAllRows:   every row j contributes eps sign G(w_j); with x = 0 that is
//                    sign * |psi_j|^2 (no partner read); used for x = 0, and for
//                    h < 6 when there is only one 64-row block.
//   mode kPairs:     h >= 6 (h: highest set bit of x): rows with bit h of j clear
//                    contribute 2 eps sign G(w_j) (the partner row's term is the
//                    same), G = Re w or Im w, eps = +1 for nY = 0, 1 mod 4, -1
//                    otherwise; the leader rows are whole 64-row blocks u with
//                    bit h-6 clear.
//   mode kPairsLane: 0 < h < 6 (partner in the same 64-row block): iteration v
//                    covers blocks 2v and 2v + 1, and lane l takes row
//                    (2v + bit_h(l)) * 64 + l, the member of its pair whose bit h
//                    equals bit 6 of the row; each pair is visited once and every
//                    lane has a pair in every iteration.
// Per-lane row sums use RowAcc<T> (FP32 for complex64); the wave reduction and
// the final scale are done in the same precision and written to ev(s, t).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fp_ev_tiled.hpp"
#include "fp_tile_types.hpp"
#include "fp_kernels.hpp"

namespace fpl
{

constexpr int kThreads = 1024;

// Row offset of lane `lane` within a pair of 64-row blocks (kPairsLane), or
// within one block (other modes); the lane's sign bit is parity(offset & z & 127).
__device__ __forceinline__ uint32_t lane_row(uint32_t mode, uint32_t h, uint32_t lane)
{
    return mode == kPairsLane ? lane + (((lane >> h) & 1u) << 6) : lane;
}


// v with bit 0 of `bits` inserted as a zero at position h
__device__ __forceinline__ uint32_t insert_zero(uint32_t v, uint32_t h)
{
    return ((v >> h) << (h + 1)) | (v & ((1u << h) - 1u));
}

// Orders this wave's LDS accesses: those before complete before those after, for
// every lane (the wave's LDS instructions execute in order; this keeps the
// compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Walsh-Hadamard evaluation of one class (see kWhtMin) by one wave, on a column
// of 2^d rows in LDS, with `scr` its kWhtPartBytes slice.  Leaders: every row
// when x = 0, else the rows with bit h clear; G = |psi|^2, Re w or Im w as in the
// per-string loop.  Calls emit(s, value) on the lane of string s (s = lane + 64 j)
// with value = sum_leaders (-1)^{parity(a & Z_s)} G(a) (the caller scales it).
template <class T, class Emit>
__device__ void wht_class(typename fpk::VecOf<T, 1>::type const *col, uint32_t d, uint32_t lane, EvChunk const &ch,
                          EvStr const *__restrict__ strs, typename fpk::RowAcc<T>::type *scr, Emit emit)
{
    typedef typename fpk::RowAcc<T>::type A;
    typedef typename fpk::VecOf<T, 1>::type V;
    constexpr uint32_t kPart = kWhtPartBytes / sizeof(A); // 256 (FP64) or 512 (FP32) values
    constexpr uint32_t kPartLog2 = kPart == 512 ? 9u : 8u;
    constexpr uint32_t kVals = kPart / 64;                // per lane: 4 or 8
    constexpr uint32_t kRb = kVals == 8 ? 3u : 2u;        // index bits per register round
    constexpr int kRes = kWhtMax / 64;
    uint32_t const x = ch.x, h = ch.h;
    uint32_t const L = x ? d - 1 : d;
    uint32_t const gb = L < kPartLog2 ? L : kPartLog2;
    uint32_t const gsize = 1u << gb, gmask = gsize - 1u, pb = L - gb;
    // FP32 slice layout: index i at i ^ (h << 3) ^ h, h = i >> 6 (a bijection of
    // [0, 512) that keeps bits 6-8).  In every butterfly round, and in the stores of
    // G, the 64 lanes then touch 64 different banks for each of their 8 values
    // (with the plain layout 8 lanes share a bank in the rounds over bits 0-2 and 3-5).
    // The map is linear over XOR, so sw(a | b) = sw(a) ^ sw(b) for disjoint bits: a
    // lane part once per round and a wave-uniform part per value.
    auto sw = [](uint32_t i) -> uint32_t {
        if constexpr (kVals == 8)
            return i ^ (((i >> 6) & 7u) * 9u);
        else
            return i;
    };
    A res[kRes];
    uint32_t zc[kRes], zs[kRes]; // compressed z; its in-part index in the slice layout
#pragma unroll
    for (int j = 0; j < kRes; ++j)
    {
        res[j] = A(0);
        uint32_t const s = lane + 64u * uint32_t(j);
        uint32_t const z = s < ch.count ? strs[ch.first + s].z : 0u;
        zc[j] = x ? ((z & ((1u << h) - 1u)) | ((z >> (h + 1)) << h)) : z;
        zs[j] = sw(zc[j] & gmask);
    }
    uint32_t const sw_lane = sw(lane);
    for (uint32_t part = 0; part < (1u << pb); ++part)
    {
        // G of this part's leaders into the slice
#pragma unroll
        for (uint32_t m = 0; m < kVals; ++m)
        {
            uint32_t const i = lane + 64u * m;
            if (i < gsize)
            {
                uint32_t const ap = (part << gb) | i;
                uint32_t const a = x ? insert_zero(ap, h) : ap;
                V const v = col[a];
                A const ar = A(v[0]), ai = A(v[1]);
                A g;
                if (x == 0)
                    g = fma(ar, ar, ai * ai);
                else
                {
                    V const w = col[a ^ x];
                    A const br = A(w[0]), bi = A(w[1]);
                    g = ch.imag ? fma(ar, bi, -(ai * br)) : fma(ar, br, ai * bi);
                }
                scr[sw_lane ^ sw(64u * m)] = g;
            }
        }
        // W[z] = sum_i (-1)^{parity(i & z)} G[i]: rounds of kRb index bits in registers
        for (uint32_t b0 = 0; b0 < gb; b0 += kRb)
        {
            uint32_t const rb = gb - b0 < kRb ? gb - b0 : kRb;
            uint32_t const n_groups = gsize >> rb;
            wave_lds_sync();
            for (uint32_t grp = lane; grp < n_groups; grp += 64)
            {
                // group base: the group index with rb zero bits inserted at b0
                uint32_t const base = ((grp >> b0) << (b0 + rb)) | (grp & ((1u << b0) - 1u));
                uint32_t const sbase = sw(base);
                A v[kVals];
#pragma unroll
                for (uint32_t m = 0; m < kVals; ++m)
                    v[m] = m < (1u << rb) ? scr[sbase ^ sw(m << b0)] : A(0);
#pragma unroll
                for (uint32_t lvl = 0; lvl < kRb; ++lvl)
#pragma unroll
                    for (uint32_t m = 0; m < kVals; ++m)
                        if (!((m >> lvl) & 1u))
                        {
                            A const u = v[m], w = v[m | (1u << lvl)];
                            v[m] = u + w;
                            v[m | (1u << lvl)] = u - w;
                        }
#pragma unroll
                for (uint32_t m = 0; m < kVals; ++m)
                    if (m < (1u << rb))
                        scr[sbase ^ sw(m << b0)] = v[m];
            }
        }
        wave_lds_sync();
#pragma unroll
        for (int j = 0; j < kRes; ++j)
        {
            uint32_t const s = lane + 64u * uint32_t(j);
            if (s < ch.count)
            {
                A const t = scr[zs[j]];
                res[j] += (__popc((zc[j] >> gb) & part) & 1) ? -t : t;
            }
        }
        wave_lds_sync(); // the next part overwrites the slice
    }
#pragma unroll
    for (int j = 0; j < kRes; ++j)
    {
        uint32_t const s = lane + 64u * uint32_t(j);
        if (s < ch.count)
            emit(s, res[j]);
    }
}

__device__ __forceinline__ float flip_sign(float v, uint32_t bits)
{
    return __uint_as_float(__float_as_uint(v) ^ bits);
}

__device__ __forceinline__ double flip_sign(double v, uint32_t bits)
{
    return __longlong_as_double(__double_as_longlong(v) ^ (static_cast<long long>(bits) << 32));
}

template <class T> __device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// (rows x states) row-major -> (states x rows): 32 x 32 tiles through LDS.
template <class V>
__global__ __launch_bounds__(256) void k_transpose(V const *__restrict__ in, V *__restrict__ out, uint64_t rows,
                                                   uint64_t cols)
{
    __shared__ V tile[32][33];
    uint64_t const r0 = uint64_t(blockIdx.y) * 32, c0 = uint64_t(blockIdx.x) * 32;
    uint32_t const tx = threadIdx.x & 31, ty = threadIdx.x >> 5; // 32 x 8
    for (uint32_t k = ty; k < 32; k += 8)
    {
        uint64_t const r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols)
            tile[k][tx] = in[r * cols + c];
    }
    __syncthreads();
    for (uint32_t k = ty; k < 32; k += 8)
    {
        uint64_t const c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols)
            out[c * rows + r] = tile[tx][k];
    }
}

// v with its sign bit flipped when bit31 = 0x80000000 (exact negation)
template <class T> __device__ __forceinline__ T xor_sign(T v, uint32_t bit31)
{
    if constexpr (sizeof(T) == 4)
        return __uint_as_float(__float_as_uint(v) ^ bit31);
    else
        return __longlong_as_double(__double_as_longlong(v) ^ (static_cast<long long>(bit31) << 32));
}

// Rows of one chunk with exactly CNT strings, all taking G = |psi|^2 (KIND 0,
// x = 0), Re w (KIND 1) or Im w (KIND 2).  Iterations q = q0 | k (q0 a multiple
// of U, k < U) visit row blocks u(q) = insert(q) = insert(q0) | insert(k)
// (insert_zero_bit at hh is linear over GF(2)), so the wave-uniform sign bit
// parity(u & zhi) splits into a per-chunk constant per (string, k), held as a
// +-1 factor, and one flip bit per (string, q0).  The per-lane bit is applied by
// the caller once per chunk.
template <class T, int U, int CNT, int KIND>
__device__ __forceinline__ void chunk_rows_n(typename fpk::VecOf<T, 1>::type const *col, uint32_t lane,
                                             EvChunk const &ch, uint32_t n_blocks64, uint32_t const (&zhi)[kChunk],
                                             typename fpk::RowAcc<T>::type (&acc)[kChunk])
{
    typedef typename fpk::RowAcc<T>::type A;
    typedef typename fpk::VecOf<T, 1>::type V;
    uint32_t const x = ch.x;
    bool const pairs = ch.mode != kAllRows;
    uint32_t const hh = ch.mode == kPairs ? ch.h - 6 : 0;
    uint32_t const lo_mask = (1u << hh) - 1;
    uint32_t const n_iter = pairs ? (n_blocks64 >> 1) : n_blocks64;
    uint32_t const lr = lane_row(ch.mode, ch.h, lane);
    // a zero bit inserted at hh (pairs) or q itself: q + (q & hmask), two scalar ops
    uint32_t const hmask = pairs ? ~lo_mask : 0u;
    auto ins = [&](uint32_t q) { return q + (q & hmask); };
    auto g_of = [&](V const &a, V const &b) {
        A const ar = A(a[0]), ai = A(a[1]);
        if constexpr (KIND == 0)
            return fma(ar, ar, ai * ai);
        else
        {
            A const br = A(b[0]), bi = A(b[1]);
            if constexpr (KIND == 1)
                return fma(ar, br, ai * bi);
            else
                return fma(ar, bi, -(ai * br));
        }
    };
    A sgk[CNT][U];
#pragma unroll
    for (int s = 0; s < CNT; ++s)
#pragma unroll
        for (int k = 0; k < U; ++k)
            sgk[s][k] = (__popc(ins(uint32_t(k)) & zhi[s]) & 1) ? A(-1) : A(1);
    uint32_t const n_full = n_iter & ~uint32_t(U - 1);
    for (uint32_t q0 = 0; q0 < n_full; q0 += U)
    {
        uint32_t const uq0 = ins(q0);
        V a[U], b[U];
#pragma unroll
        for (int k = 0; k < U; ++k)
        {
            uint32_t const j = ((uq0 | ins(uint32_t(k))) << 6) + lr;
            a[k] = col[j];
            if constexpr (KIND != 0)
                b[k] = col[j ^ x];
            else
                b[k] = a[k];
        }
        // all U row blocks' LDS reads in flight before the first use
        __builtin_amdgcn_sched_barrier(0);
        uint32_t flip[CNT];
#pragma unroll
        for (int s = 0; s < CNT; ++s)
            flip[s] = uint32_t(__popc(uq0 & zhi[s]) & 1) << 31;
        // the U terms of a string first, with its per-chunk signs, then this step's
        // sign once: one scalar sign per string instead of one per term
        A tsum[CNT];
#pragma unroll
        for (int s = 0; s < CNT; ++s)
            tsum[s] = A(0);
#pragma unroll
        for (int k = 0; k < U; ++k)
        {
            A const g = g_of(a[k], b[k]);
#pragma unroll
            for (int s = 0; s < CNT; ++s)
                tsum[s] = fma(sgk[s][k], g, tsum[s]);
        }
#pragma unroll
        for (int s = 0; s < CNT; ++s)
            acc[s] = fma(xor_sign(A(1), flip[s]), tsum[s], acc[s]);
    }
    for (uint32_t q = n_full; q < n_iter; ++q) // fewer than U row blocks left
    {
        uint32_t const u = ins(q);
        uint32_t const j = (u << 6) + lr;
        V const av = col[j];
        V const bv = KIND != 0 ? col[j ^ x] : av;
        A const g = g_of(av, bv);
#pragma unroll
        for (int s = 0; s < CNT; ++s)
        {
            A const sg = (__popc(u & zhi[s]) & 1) ? A(-1) : A(1);
            acc[s] = fma(sg, g, acc[s]);
        }
    }
}

template <class T, int U, int KIND>
__device__ __forceinline__ void chunk_rows_k(typename fpk::VecOf<T, 1>::type const *col, uint32_t lane,
                                             EvChunk const &ch, uint32_t n_blocks64, uint32_t const (&zhi)[kChunk],
                                             typename fpk::RowAcc<T>::type (&acc)[kChunk])
{
    switch (ch.count)
    {
    // one- and two-string chunks (most of them for weight-limited sets) keep 8 row
    // blocks of loads in flight per step instead of 4 (n = 14, 2000 weight-3 strings,
    // 1000 states: complex64 1.007 -> 0.986 ms, complex128 unchanged)
    case 1: chunk_rows_n<T, 2 * U, 1, KIND>(col, lane, ch, n_blocks64, zhi, acc); break;
    case 2: chunk_rows_n<T, 2 * U, 2, KIND>(col, lane, ch, n_blocks64, zhi, acc); break;
    case 3: chunk_rows_n<T, U, 3, KIND>(col, lane, ch, n_blocks64, zhi, acc); break;
    case 4: chunk_rows_n<T, U, 4, KIND>(col, lane, ch, n_blocks64, zhi, acc); break;
    case 5: chunk_rows_n<T, U, 5, KIND>(col, lane, ch, n_blocks64, zhi, acc); break;
    case 6: chunk_rows_n<T, U, 6, KIND>(col, lane, ch, n_blocks64, zhi, acc); break;
    case 7: chunk_rows_n<T, U, 7, KIND>(col, lane, ch, n_blocks64, zhi, acc); break;
    default: chunk_rows_n<T, U, kChunk, KIND>(col, lane, ch, n_blocks64, zhi, acc); break;
    }
}

template <class T, int U>
__device__ __forceinline__ void chunk_rows(typename fpk::VecOf<T, 1>::type const *col, uint32_t lane, EvChunk const &ch,
                                           uint32_t n_blocks64, uint32_t const (&zhi)[kChunk],
                                           typename fpk::RowAcc<T>::type (&acc)[kChunk])
{
    if (ch.x == 0)
        chunk_rows_k<T, U, 0>(col, lane, ch, n_blocks64, zhi, acc);
    else if (ch.imag)
        chunk_rows_k<T, U, 2>(col, lane, ch, n_blocks64, zhi, acc);
    else
        chunk_rows_k<T, U, 1>(col, lane, ch, n_blocks64, zhi, acc);
}

// chunks[0, n_wht) are transform classes (wht_class), the rest per-string chunks.
// The LDS holds the column, then (with n_wht > 0) the transform buffers.
template <class T>
__global__ __launch_bounds__(kThreads) void k_ev_lds(typename fpk::VecOf<T, 1>::type const *__restrict__ states_T,
                                                     uint32_t dim, EvChunk const *__restrict__ chunks,
                                                     uint32_t n_chunks, uint32_t n_wht,
                                                     EvStr const *__restrict__ strs, uint64_t n_states,
                                                     T *__restrict__ ev)
{
    typedef typename fpk::VecOf<T, 1>::type V;
    typedef typename fpk::RowAcc<T>::type A;
    extern __shared__ __align__(16) unsigned char smem[];
    V *col = reinterpret_cast<V *>(smem);
    __shared__ uint32_t next_chunk;
    uint64_t const t = blockIdx.x;
    V const *src = states_T + t * dim;
    for (uint32_t i = threadIdx.x; i < dim; i += kThreads)
        col[i] = src[i];
    // per-string chunks [c_begin, c_end) of this block (gridDim.y splits the list)
    uint32_t const n_reg = n_chunks - n_wht;
    uint32_t const per = (n_reg + gridDim.y - 1) / gridDim.y;
    uint32_t const c_begin = n_wht + blockIdx.y * per;
    uint32_t const c_end = c_begin + per < n_chunks ? c_begin + per : n_chunks;
    __shared__ uint32_t next_wht;
    if (threadIdx.x == 0)
    {
        next_wht = blockIdx.y;
        next_chunk = c_begin;
    }
    __syncthreads();
    uint32_t const lane = threadIdx.x & 63u;
    uint32_t const wave = threadIdx.x >> 6;
    if (n_wht && wave < kWhtWaves)
    {
        // transform classes c = y, y + gy, ... of this block, one per wave at a time;
        // a wave then moves on to the per-string chunks without waiting for the others
        A *scr = reinterpret_cast<A *>(smem + size_t(dim) * sizeof(V) + wave * kWhtPartBytes);
        uint32_t const d = 31u - uint32_t(__builtin_clz(dim));
        while (true)
        {
            uint32_t c = 0;
            if (lane == 0)
                c = atomicAdd(&next_wht, gridDim.y);
            c = __builtin_amdgcn_readfirstlane(__shfl(c, 0, 64));
            if (c >= n_wht)
                break;
            EvChunk const ch = chunks[c];
            wht_class<T>(col, d, lane, ch, strs, scr, [&](uint32_t s, A v) {
                EvStr const st = strs[ch.first + s];
                ev[uint64_t(st.out) * n_states + t] = T(double(v) * st.scale);
            });
        }
    }
    uint32_t const n_blocks64 = dim >> 6;
    while (true)
    {
        uint32_t c = 0;
        if (lane == 0)
            c = atomicAdd(&next_chunk, 1u);
        c = __builtin_amdgcn_readfirstlane(__shfl(c, 0, 64));
        if (c >= c_end)
            break;
        EvChunk const ch = chunks[c];
        uint32_t zhi[kChunk], lsign[kChunk];
        A acc[kChunk];
        uint32_t const lr = lane_row(ch.mode, ch.h, lane);
#pragma unroll
        for (int s = 0; s < kChunk; ++s)
        {
            acc[s] = A(0);
            zhi[s] = 0;
            lsign[s] = 0;
            if (s < int(ch.count))
            {
                EvStr const st = strs[ch.first + s];
                zhi[s] = st.z >> 6;
                lsign[s] = uint32_t(__popc(lr & st.z & 127u) & 1) << 31;
            }
        }
        chunk_rows<T, 4>(col, lane, ch, n_blocks64, zhi, acc);
#pragma unroll
        for (int s = 0; s < kChunk; ++s)
        {
            if (s < int(ch.count))
            {
                A const sum = wave_sum(flip_sign(acc[s], lsign[s]));
                if (lane == 0)
                {
                    EvStr const st = strs[ch.first + s];
                    ev[uint64_t(st.out) * n_states + t] = T(double(sum) * st.scale);
                }
            }
        }
    }
}

// The same evaluation on one coset of a subspace V of row indices, for columns
// that do not fit in LDS.  The host groups strings so that every x mask of a
// group lies in its V (dimension d, low k bits included); block (t, y, z) loads
// coset q = z % n_cos of group g = z / n_cos of column t,
//     local row a  ->  global row  c_q ^ B(a >> k) ^ (a & lowmask),
// and runs the group's chunks in local coordinates (local x, local z).  The
// coset's sign (-1)^{parity(c_q & z)} equals (-1)^{parity(q & zc)}, zc being z
// compressed onto the coset-index bits (EvStr::zc).  Each (coset, string,
// state) sum is written to `partial`; k_sum_cosets adds the cosets in order.
template <class T>
__global__ __launch_bounds__(kThreads) void k_ev_coset(typename fpk::VecOf<T, 1>::type const *__restrict__ states_T,
                                                       uint64_t dim, uint32_t d, uint32_t k, uint32_t n_cos,
                                                       fpt::TileGroup const *__restrict__ groups,
                                                       uint64_t const *__restrict__ btab,
                                                       EvChunk const *__restrict__ chunks,
                                                       EvStr const *__restrict__ strs, uint64_t n_states,
                                                       uint64_t n_strings, double *__restrict__ partial)
{
    typedef typename fpk::VecOf<T, 1>::type V;
    typedef typename fpk::RowAcc<T>::type A;
    extern __shared__ __align__(16) unsigned char smem[];
    V *col = reinterpret_cast<V *>(smem);
    __shared__ uint32_t next_chunk;
    uint64_t const t = blockIdx.x;
    uint32_t const g = blockIdx.z / n_cos, q = blockIdx.z % n_cos;
    fpt::TileGroup const grp = groups[g];
    uint64_t const cb = fpt::coset_base(q, k, grp);
    uint32_t const rows = 1u << d;
    uint32_t const lowmask = (1u << k) - 1u;
    V const *src = states_T + t * dim;
    uint64_t const *bt = btab + grp.btab_off;
    for (uint32_t i = threadIdx.x; i < rows; i += kThreads)
        col[i] = src[cb ^ bt[i >> k] ^ uint64_t(i & lowmask)];
    uint32_t const n_reg = grp.n_chunks - grp.n_wht;
    uint32_t const per = (n_reg + gridDim.y - 1) / gridDim.y;
    uint32_t const c_begin = grp.chunk_first + grp.n_wht + blockIdx.y * per;
    uint32_t const c_last = grp.chunk_first + grp.n_chunks;
    uint32_t const c_end = c_begin + per < c_last ? c_begin + per : c_last;
    // the group's first n_wht chunks are transform classes (wht_class), one per wave at a
    // time; a wave then moves on to the per-string chunks without waiting for the others
    __shared__ uint32_t next_wht;
    if (threadIdx.x == 0)
    {
        next_wht = blockIdx.y;
        next_chunk = c_begin;
    }
    __syncthreads();
    uint32_t const lane = threadIdx.x & 63u;
    uint32_t const wave = threadIdx.x >> 6;
    if (grp.n_wht && wave < kWhtWaves)
    {
        A *scr = reinterpret_cast<A *>(smem + size_t(rows) * sizeof(V) + wave * kWhtPartBytes);
        while (true)
        {
            uint32_t c = 0;
            if (lane == 0)
                c = atomicAdd(&next_wht, gridDim.y);
            c = __builtin_amdgcn_readfirstlane(__shfl(c, 0, 64));
            if (c >= grp.n_wht)
                break;
            EvChunk const ch = chunks[grp.chunk_first + c];
            wht_class<T>(col, d, lane, ch, strs, scr, [&](uint32_t s, A v) {
                EvStr const st = strs[ch.first + s];
                double const dv = double(v) * st.scale;
                partial[(uint64_t(q) * n_strings + st.out) * n_states + t] = (__popc(q & st.zc) & 1) ? -dv : dv;
            });
        }
    }
    uint32_t const n_blocks64 = rows >> 6;
    while (true)
    {
        uint32_t c = 0;
        if (lane == 0)
            c = atomicAdd(&next_chunk, 1u);
        c = __builtin_amdgcn_readfirstlane(__shfl(c, 0, 64));
        if (c >= c_end)
            break;
        EvChunk const ch = chunks[c];
        uint32_t zhi[kChunk], lsign[kChunk];
        A acc[kChunk];
        uint32_t const lr = lane_row(ch.mode, ch.h, lane);
#pragma unroll
        for (int s = 0; s < kChunk; ++s)
        {
            acc[s] = A(0);
            zhi[s] = 0;
            lsign[s] = 0;
            if (s < int(ch.count))
            {
                EvStr const st = strs[ch.first + s];
                zhi[s] = st.z >> 6;
                lsign[s] = uint32_t(__popc(lr & st.z & 127u) & 1) << 31;
            }
        }
        chunk_rows<T, 4>(col, lane, ch, n_blocks64, zhi, acc);
#pragma unroll
        for (int s = 0; s < kChunk; ++s)
        {
            if (s < int(ch.count))
            {
                A const sum = wave_sum(flip_sign(acc[s], lsign[s]));
                if (lane == 0)
                {
                    EvStr const st = strs[ch.first + s];
                    double const v = double(sum) * st.scale;
                    partial[(uint64_t(q) * n_strings + st.out) * n_states + t] = (__popc(q & st.zc) & 1) ? -v : v;
                }
            }
        }
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_sum_cosets(double const *__restrict__ partial, uint32_t n_cos, uint64_t total,
                                                    T *__restrict__ ev)
{
    uint64_t const i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= total)
        return;
    double s = partial[i];
    for (uint32_t q = 1; q < n_cos; ++q)
        s += partial[uint64_t(q) * total + i];
    ev[i] = T(s);
}

} // namespace fpl
