// fp_algebra.hpp — SummedPauliOp::square on the device
// (fast_pauli/cpp/include/__summed_pauli_op.hpp:197-268).
//
//   A_k^2 = sum_{i,j} h(i,k) h(j,k) P_i P_j = sum_a h'(a,k) P_a,
//   h'(a,k) = sum over the pairs (i, j) with P_i P_j = phase_ij P_a of phase_ij h(i,k) h(j,k),
//
// over the strings of weight <= W = min(n, 2 * max weight) listed in the order of
// calculate_pauli_strings_max_weight (__pauli_helpers.hpp:99-153).  Each product
// term couples one pair to one output string, so the contraction has no dense
// shared index for the matrix cores; it is a gather and segmented sum:
//   1. k_square_pairs: pair p = i * N + j -> key a (the product's rank in that
//      order) and value (i << 33 | j << 2 | phase exponent), phase_ij = i^e
//      (N^2 < 2^31: i and j fit 31 bits);
//   2. a stable radix sort by key (pairs of one a stay in (i, j) order);
//   3. k_segment_starts: first sorted position of every a (binary search);
//   4. k_square_contract: one thread per (a, k) sums its segment in (i, j)
//      order, (phase * h(i,k)) * h(j,k) as std::complex
//      multiplies (no fused multiply-add), which is the reference's operation
//      order: results are IEEE-equal to its sequential loop.
//
// Bit layout: a string is (x, z) with character c on bit n-1-c (X: x, Z: z, Y: both).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fp_kernels.hpp"

namespace fpalg
{

// Exponent e of the phase i^e in P(x1, z1) P(x2, z2) = i^e P(x1 ^ x2, z1 ^ z2), letter
// by letter: X*Y = iZ, Y*Z = iX, Z*X = iY take +1, the reversed orders -1.
__host__ __device__ inline uint32_t product_phase(uint64_t x1, uint64_t z1, uint64_t x2, uint64_t z2)
{
    uint64_t const lx = x1 & ~z1, ly = x1 & z1, lz = ~x1 & z1; // letters of the left string
    uint64_t const rx = x2 & ~z2, ry = x2 & z2, rz = ~x2 & z2; // ... of the right one
    int const plus = __builtin_popcountll(lx & ry) + __builtin_popcountll(ly & rz) + __builtin_popcountll(lz & rx);
    int const minus = __builtin_popcountll(lx & rz) + __builtin_popcountll(ly & rx) + __builtin_popcountll(lz & ry);
    return uint32_t(plus - minus) & 3u;
}

// Rank tables of the enumeration of strings of weight <= W on n qubits:
// binom(m, r) for m < 64, r <= W, and the first index of each weight.
struct RankTables
{
    uint32_t W;
    uint64_t const *binom; // [64][W + 1]
    uint64_t const *first; // [W + 1]
    uint64_t const *pow3;  // [W + 1]
};

// Index of the string (x, z) in calculate_pauli_strings_max_weight(n, W) order:
//   first(w) + word * C(n, w) + (C(n, w) - 1 - colex(support)),
// word: the letters in character order as base-3 digits (X = 0, Y = 1, Z = 2, the
// first character most significant); colex(support) = sum_t C(b_t, t + 1) over the
// support's bit indices b_0 < b_1 < ... (supports run from the largest bit pattern
// down, as the reference's prev_permutation walk over its bitmask does).
__host__ __device__ inline uint64_t string_rank(uint64_t x, uint64_t z, uint32_t n, RankTables const &t)
{
    uint64_t m = x | z;
    uint32_t const w = uint32_t(__builtin_popcountll(m));
    uint64_t word = 0, colex = 0;
    for (uint32_t s = 0; m != 0; ++s)
    {
        uint32_t const b = uint32_t(__builtin_ctzll(m));
        m &= m - 1;
        uint32_t const digit = ((z >> b) & 1u) ? (((x >> b) & 1u) ? 1u : 2u) : 0u;
        word += digit * t.pow3[s];
        colex += t.binom[b * (t.W + 1) + s + 1];
    }
    uint64_t const c_nw = t.binom[n * (t.W + 1) + w];
    return t.first[w] + word * c_nw + (c_nw - 1 - colex);
}

__global__ __launch_bounds__(256) void k_square_pairs(uint64_t const *__restrict__ xs, uint64_t const *__restrict__ zs,
                                                      uint32_t n_strings, uint32_t n, RankTables t,
                                                      uint32_t *__restrict__ keys, uint64_t *__restrict__ vals)
{
    uint32_t const p = blockIdx.x * 256u + threadIdx.x;
    if (p >= n_strings * n_strings) // < 2^31
        return;
    uint32_t const i = p / n_strings, j = p - i * n_strings;
    uint64_t const x1 = xs[i], z1 = zs[i], x2 = xs[j], z2 = zs[j];
    keys[p] = uint32_t(string_rank(x1 ^ x2, z1 ^ z2, n, t));
    vals[p] = (uint64_t(i) << 33) | (uint64_t(j) << 2) | product_phase(x1, z1, x2, z2);
}

// starts[a] = first position of key a in the sorted keys (a = 0 .. n_out inclusive).
__global__ __launch_bounds__(256) void k_segment_starts(uint32_t const *__restrict__ keys, uint64_t n_keys,
                                                        uint64_t n_out, uint64_t *__restrict__ starts)
{
    uint64_t const a = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (a > n_out)
        return;
    uint64_t lo = 0, hi = n_keys;
    while (lo < hi)
    {
        uint64_t const mid = lo + (hi - lo) / 2;
        if (uint64_t(keys[mid]) < a)
            lo = mid + 1;
        else
            hi = mid;
    }
    starts[a] = lo;
}

// recs[a] = (start | count << 32, first value) of the segment of a, so that the
// contraction reaches its first pair with one load (n_keys < 2^31).
__global__ __launch_bounds__(256) void k_segment_records(uint64_t const *__restrict__ starts,
                                                         uint64_t const *__restrict__ vals, uint64_t n_out,
                                                         ulonglong2 *__restrict__ recs)
{
    uint64_t const a = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (a >= n_out)
        return;
    uint64_t const s = starts[a], e = starts[a + 1];
    recs[a] = make_ulonglong2(s | ((e - s) << 32), e > s ? vals[s] : 0ull);
}

// out(a, k) = sum over the segment of a of (i^e * h(i,k)) * h(j,k), in sorted order.
// Threads stride over the (a, k) outputs, k fastest (a resident grid: the outputs
// are many and short, so the waves stay resident instead of being launched one
// per 64 outputs); SMALL: every index < 2^32 (32-bit division).  Each output sums
// its own segment in sorted order.
// A thread's chain of dependent loads is its segment record, then the two h values of
// each pair (later pairs' records are independent of the first pair's h loads).
template <class T, bool SMALL>
__global__ __launch_bounds__(256) void k_square_contract(uint64_t const *__restrict__ vals,
                                                         ulonglong2 const *__restrict__ recs,
                                                         typename fpk::VecOf<T, 1>::type const *__restrict__ h,
                                                         uint64_t n_ops, uint64_t n_out,
                                                         typename fpk::VecOf<T, 1>::type *__restrict__ out)
{
    typedef typename fpk::VecOf<T, 1>::type V;
    uint64_t const n_work = n_out * n_ops;
    uint64_t const stride = uint64_t(gridDim.x) * 256;
    for (uint64_t idx = uint64_t(blockIdx.x) * 256 + threadIdx.x; idx < n_work; idx += stride)
    {
        uint64_t a, k;
        if constexpr (SMALL)
        {
            uint32_t const a32 = uint32_t(idx) / uint32_t(n_ops);
            a = a32;
            k = uint32_t(idx) - a32 * uint32_t(n_ops);
        }
        else
        {
            a = idx / n_ops;
            k = idx - a * n_ops;
        }
        T ar = T(0), ai = T(0);
        ulonglong2 const rec = recs[a];
        uint64_t const s0 = rec.x & 0xffffffffull, cnt = rec.x >> 32;
        for (uint64_t c = 0; c < cnt; ++c)
        {
            uint64_t const v = c == 0 ? rec.y : vals[s0 + c];
            uint32_t const ph = uint32_t(v) & 3u;
            uint64_t const i = v >> 33, j = (v >> 2) & 0x7fffffffu;
            // i^ph as a complex number with exact components
            T const pr = ph == 0 ? T(1) : (ph == 2 ? T(-1) : T(0));
            T const pi = ph == 1 ? T(1) : (ph == 3 ? T(-1) : T(0));
            V const hi = h[i * n_ops + k], hj = h[j * n_ops + k];
            T tr, ti, ur, ui;
            fpk::cmul_exact<T>(pr, pi, hi[0], hi[1], tr, ti);
            fpk::cmul_exact<T>(tr, ti, hj[0], hj[1], ur, ui);
            ar = fpk::add_exact<T>(ar, ur);
            ai = fpk::add_exact<T>(ai, ui);
        }
        V r;
        r[0] = ar;
        r[1] = ai;
        __builtin_nontemporal_store(r, &out[idx]);
    }
}

} // namespace fpalg
