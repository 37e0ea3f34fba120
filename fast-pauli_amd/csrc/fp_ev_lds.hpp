// fp_ev_lds.hpp — per-string expectation values with the state held in LDS.
//
// For n <= 14 (complex64) or n <= 13 (complex128) one state column fits in a
// workgroup's LDS (160 KiB on gfx950).  A block loads column t once (from a
// (states x rows) transposed copy, so the load is contiguous) and evaluates
// every string against it; the state is read from HBM once per block instead
// of once per string.
//
// Strings are grouped on the host into chunks of up to kChunk strings sharing
// the same x mask, so the two LDS reads and the product w = conj(psi_j) psi_{j^x}
// of a row serve every string of the chunk.  Rows are enumerated as
// j = u * 64 + lane with u uniform across the wave, which splits the sign
// (-1)^{popcount(j & z)} into a per-lane bit (low 6 bits of z, fixed for the
// chunk) and a wave-uniform bit (popcount(u & z >> 6), scalar ALU).
//   mode kAllRows: every row j contributes eps * sign * G(w_j), G = Re w (nY even)
//                  or Im w (nY odd), eps = +1 for nY = 0, 1 mod 4, -1 otherwise;
//                  used when x = 0 or the highest set bit of x is below 6.
//   mode kPairs:   rows with bit h of j clear contribute 2 eps sign G(w_j) (the
//                  partner row's term is the same); used when h >= 6, where the
//                  leader rows are whole 64-row blocks u with bit h-6 clear.
// Per-lane row sums use RowAcc<T> (FP32 for complex64); the wave reduction and
// the final scale are done in the same precision and written to ev(s, t).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fp_ev_tiled.hpp"
#include "fp_kernels.hpp"

namespace fpl
{

constexpr int kChunk = 8;
constexpr int kThreads = 1024;

enum : uint32_t
{
    kAllRows = 0,
    kPairs = 1
};

struct EvChunk
{
    uint32_t x;     // x mask (n <= 14: fits 32 bits)
    uint32_t h;     // highest set bit of x (kPairs only)
    uint32_t mode;  // kAllRows / kPairs
    uint32_t count; // strings in this chunk (1..kChunk)
    uint32_t first; // index of the first EvStr
    uint32_t pad[3];
};

struct EvStr
{
    uint32_t z;
    uint32_t imag; // 1: G = Im w, 0: G = Re w
    uint32_t out;  // output row (original string index)
    uint32_t zc;   // coset kernel: z compressed onto the coset-index bits (0 otherwise)
    double scale;  // eps (kAllRows) or 2 eps (kPairs)
};

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

// Rows of one chunk with exactly CNT strings.  The wave-uniform row sign bit
// (popcount(u & zhi) parity) is applied as a scalar +-1 factor; the per-lane bit
// is applied by the caller once per chunk.  Re w / Im w per string is a select
// on a wave-uniform flag.
template <class T> __device__ __forceinline__ T xor_sign(T v, uint32_t bit31)
{
    if constexpr (sizeof(T) == 4)
        return __uint_as_float(__float_as_uint(v) ^ bit31);
    else
        return __longlong_as_double(__double_as_longlong(v) ^ (static_cast<long long>(bit31) << 32));
}

// Rows of one chunk with exactly CNT strings.  Iterations q = q0 | k (q0 a
// multiple of U, k < U) visit row blocks u(q) = insert(q) = insert(q0) | insert(k)
// (insert_zero_bit at hh is linear over GF(2)), so the wave-uniform sign bit
// parity(u & zhi) splits into a per-chunk constant per (string, k), held as a
// +-1 factor, and one flip bit per (string, q0).  The per-lane bit is applied by
// the caller once per chunk.  Re w / Im w per string is a select on a
// wave-uniform flag.
template <class T, int U, int CNT>
__device__ __forceinline__ void chunk_rows_n(typename fpk::VecOf<T, 1>::type const *col, uint32_t lane,
                                             EvChunk const &ch, uint32_t n_blocks64, uint32_t const (&zhi)[kChunk],
                                             bool const (&imag)[kChunk], typename fpk::RowAcc<T>::type (&acc)[kChunk])
{
    typedef typename fpk::RowAcc<T>::type A;
    uint32_t const x = ch.x;
    bool const pairs = ch.mode == kPairs;
    uint32_t const hh = pairs ? ch.h - 6 : 0;
    uint32_t const lo_mask = (1u << hh) - 1;
    uint32_t const n_iter = pairs ? (n_blocks64 >> 1) : n_blocks64;
    auto ins = [&](uint32_t q) { return pairs ? (((q >> hh) << (hh + 1)) | (q & lo_mask)) : q; };
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
        typename fpk::VecOf<T, 1>::type a[U], b[U];
#pragma unroll
        for (int k = 0; k < U; ++k)
        {
            uint32_t const j = ((uq0 | ins(uint32_t(k))) << 6) | lane;
            a[k] = col[j];
            b[k] = col[j ^ x];
        }
        uint32_t flip[CNT];
#pragma unroll
        for (int s = 0; s < CNT; ++s)
            flip[s] = uint32_t(__popc(uq0 & zhi[s]) & 1) << 31;
#pragma unroll
        for (int k = 0; k < U; ++k)
        {
            A const ar = A(a[k][0]), ai = A(a[k][1]), br = A(b[k][0]), bi = A(b[k][1]);
            A const gre = fma(ar, br, ai * bi);
            A const gim = fma(ar, bi, -(ai * br));
#pragma unroll
            for (int s = 0; s < CNT; ++s)
                acc[s] = fma(xor_sign(sgk[s][k], flip[s]), imag[s] ? gim : gre, acc[s]);
        }
    }
    for (uint32_t q = n_full; q < n_iter; ++q) // fewer than U row blocks left
    {
        uint32_t const u = ins(q);
        uint32_t const j = (u << 6) | lane;
        typename fpk::VecOf<T, 1>::type const av = col[j], bv = col[j ^ x];
        A const ar = A(av[0]), ai = A(av[1]), br = A(bv[0]), bi = A(bv[1]);
        A const gre = fma(ar, br, ai * bi);
        A const gim = fma(ar, bi, -(ai * br));
#pragma unroll
        for (int s = 0; s < CNT; ++s)
        {
            A const sg = (__popc(u & zhi[s]) & 1) ? A(-1) : A(1);
            acc[s] = fma(sg, imag[s] ? gim : gre, acc[s]);
        }
    }
}

template <class T, int U>
__device__ __forceinline__ void chunk_rows(typename fpk::VecOf<T, 1>::type const *col, uint32_t lane, EvChunk const &ch,
                                           uint32_t n_blocks64, uint32_t const (&zhi)[kChunk],
                                           bool const (&imag)[kChunk], typename fpk::RowAcc<T>::type (&acc)[kChunk])
{
    switch (ch.count)
    {
    case 1: chunk_rows_n<T, U, 1>(col, lane, ch, n_blocks64, zhi, imag, acc); break;
    case 2: chunk_rows_n<T, U, 2>(col, lane, ch, n_blocks64, zhi, imag, acc); break;
    case 3: chunk_rows_n<T, U, 3>(col, lane, ch, n_blocks64, zhi, imag, acc); break;
    case 4: chunk_rows_n<T, U, 4>(col, lane, ch, n_blocks64, zhi, imag, acc); break;
    case 5: chunk_rows_n<T, U, 5>(col, lane, ch, n_blocks64, zhi, imag, acc); break;
    case 6: chunk_rows_n<T, U, 6>(col, lane, ch, n_blocks64, zhi, imag, acc); break;
    case 7: chunk_rows_n<T, U, 7>(col, lane, ch, n_blocks64, zhi, imag, acc); break;
    default: chunk_rows_n<T, U, kChunk>(col, lane, ch, n_blocks64, zhi, imag, acc); break;
    }
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_ev_lds(typename fpk::VecOf<T, 1>::type const *__restrict__ states_T,
                                                     uint32_t dim, EvChunk const *__restrict__ chunks,
                                                     uint32_t n_chunks, EvStr const *__restrict__ strs,
                                                     uint64_t n_states, T *__restrict__ ev)
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
    // chunks [c_begin, c_end) of this block (gridDim.y splits the chunk list)
    uint32_t const per = (n_chunks + gridDim.y - 1) / gridDim.y;
    uint32_t const c_begin = blockIdx.y * per;
    uint32_t const c_end = c_begin + per < n_chunks ? c_begin + per : n_chunks;
    if (threadIdx.x == 0)
        next_chunk = c_begin;
    __syncthreads();
    uint32_t const lane = threadIdx.x & 63u;
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
        bool imag[kChunk];
        A acc[kChunk];
#pragma unroll
        for (int s = 0; s < kChunk; ++s)
        {
            acc[s] = A(0);
            zhi[s] = 0;
            lsign[s] = 0;
            imag[s] = false;
            if (s < int(ch.count))
            {
                EvStr const st = strs[ch.first + s];
                zhi[s] = st.z >> 6;
                lsign[s] = uint32_t(__popc(lane & st.z & 63u) & 1) << 31;
                imag[s] = st.imag != 0;
            }
        }
        chunk_rows<T, 4>(col, lane, ch, n_blocks64, zhi, imag, acc);
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
    uint32_t const per = (grp.n_chunks + gridDim.y - 1) / gridDim.y;
    uint32_t const c_begin = grp.chunk_first + blockIdx.y * per;
    uint32_t const c_last = grp.chunk_first + grp.n_chunks;
    uint32_t const c_end = c_begin + per < c_last ? c_begin + per : c_last;
    if (threadIdx.x == 0)
        next_chunk = c_begin;
    __syncthreads();
    uint32_t const lane = threadIdx.x & 63u;
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
        bool imag[kChunk];
        A acc[kChunk];
#pragma unroll
        for (int s = 0; s < kChunk; ++s)
        {
            acc[s] = A(0);
            zhi[s] = 0;
            lsign[s] = 0;
            imag[s] = false;
            if (s < int(ch.count))
            {
                EvStr const st = strs[ch.first + s];
                zhi[s] = st.z >> 6;
                lsign[s] = uint32_t(__popc(lane & st.z & 63u) & 1) << 31;
                imag[s] = st.imag != 0;
            }
        }
        chunk_rows<T, 4>(col, lane, ch, n_blocks64, zhi, imag, acc);
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
