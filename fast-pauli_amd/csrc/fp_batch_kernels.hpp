// fp_batch_kernels.hpp — one persistent launch over many operators of uneven
// size (qubit counts, string counts, state counts, apply or expectation).
//
// The host flattens every operator into work items (tiles) and orders them
// largest first; the persistent blocks pull items from a device counter until
// the list is exhausted, so every wave reaches the exit condition.
//   apply tile:        a run of output vectors; all strings accumulated in
//                      registers in order s = 0..N-1 (the same arithmetic as
//                      the single-operator kernel, so results are bit-exact
//                      with it).
//   expectation tile:  a range of row pairs for all strings; FP64 partials per
//                      state.  The block that completes an operator's last tile
//                      (device-scope counter) reduces its partials in tile
//                      order, so for a given schedule results are identical run
//                      to run; another tile size regroups the FP64 partial sums
//                      (last-bit differences).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fp_kernels.hpp"

namespace fpb
{

using fpk::kBlock;

enum : uint32_t
{
    kApply = 0,
    kExpect = 1
};

struct BatchOp
{
    void const *in;
    void *out;
    uint64_t const *xs;
    uint64_t const *zs;
    void const *cf;                  // folded coefficients (apply), complex of dtype
    fpk::StringMeta const *meta;     // expectation
    double const *scale;             // c_s * kappa_s (expectation)
    uint64_t n_vec;                  // vectors in the batch (apply)
    uint64_t vpr;                    // vectors per row
    uint64_t n_pairs;                // row pairs (expectation)
    uint64_t n_states;
    uint64_t partial_base;           // first partial element (expectation): tile i, state t at base + i*n_states + t
    uint32_t n_strings;
    uint32_t shift;                  // log2(vpr) when pow2
    uint32_t pow2;
    uint32_t nc;                     // complexes per vector (1 or 2)
    uint32_t dtype;                  // 0 complex128, 1 complex64
    uint32_t kind;                   // kApply / kExpect
    uint32_t accumulate;
    uint32_t n_tiles;                // tiles of this operator
    uint32_t ct, rt;                 // expectation thread layout
};

struct WorkItem
{
    uint32_t op;
    uint32_t tile;   // tile index within the operator
    uint64_t begin;  // first vector (apply) or first pair (expectation)
    uint64_t count;  // vectors or pairs in the tile
};

__device__ __forceinline__ double load_coherent(double const *p)
{
    unsigned long long const bits =
        __hip_atomic_load(reinterpret_cast<unsigned long long const *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double(static_cast<long long>(bits));
}

template <class T, int NC>
__device__ void apply_tile(BatchOp const &op, uint64_t begin, uint64_t count)
{
    typedef typename fpk::VecOf<T, NC>::type V;
    auto const *in = static_cast<V const *>(op.in);
    auto *out = static_cast<V *>(op.out);
    auto const *cf = static_cast<T const *>(op.cf);
    uint64_t const end = begin + count;
    constexpr int U = 2;
    for (uint64_t base = begin + threadIdx.x; base < end; base += uint64_t(kBlock) * U)
    {
        uint64_t row[U], col[U];
        bool live[U];
        V acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            uint64_t const v = base + uint64_t(u) * kBlock;
            live[u] = v < end;
            uint64_t const vv = live[u] ? v : begin;
            if (op.pow2)
            {
                row[u] = vv >> op.shift;
                col[u] = vv & (op.vpr - 1);
            }
            else
            {
                row[u] = vv / op.vpr;
                col[u] = vv - row[u] * op.vpr;
            }
            acc[u] = (op.accumulate && live[u]) ? out[v] : V(0);
        }
        for (uint32_t s = 0; s < op.n_strings; ++s)
        {
            uint64_t const x = op.xs[s];
            uint64_t const z = op.zs[s];
            T const cr = cf[2 * s], ci = cf[2 * s + 1];
            V src[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                src[u] = live[u] ? in[(row[u] ^ x) * op.vpr + col[u]] : V(0);
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                uint32_t const sg = fpk::parity64(row[u] & z);
                T const a = sg ? -cr : cr;
                T const b = sg ? -ci : ci;
#pragma unroll
                for (int c = 0; c < NC; ++c)
                {
                    T orr, oi;
                    fpk::cmul_exact<T>(a, b, src[u][2 * c], src[u][2 * c + 1], orr, oi);
                    acc[u][2 * c] = fpk::add_exact<T>(acc[u][2 * c], orr);
                    acc[u][2 * c + 1] = fpk::add_exact<T>(acc[u][2 * c + 1], oi);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (live[u])
                out[base + uint64_t(u) * kBlock] = acc[u];
    }
}

// Expectation tile: pairs [begin, begin + count) of all strings; writes one FP64
// complex partial per state; the last tile of the operator reduces them.
template <class T, int NC>
__device__ void expect_tile(BatchOp const &op, uint32_t tile, uint64_t begin, uint64_t count, fpk::Dbl2 *partials,
                            double *red)
{
    typedef typename fpk::RowAcc<T>::type A;
    auto const *states = static_cast<typename fpk::VecOf<T, NC>::type const *>(op.in);
    uint32_t const ct = op.ct, rt = op.rt;
    uint32_t const tx = threadIdx.x % ct, ty = threadIdx.x / ct;
    uint64_t const p1 = begin + count;
    for (uint64_t col0 = 0; col0 < op.vpr; col0 += ct)
    {
        uint64_t const col = col0 + tx;
        bool const col_live = ty < rt && col < op.vpr;
        double er[NC], ei[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c)
            er[c] = ei[c] = 0.0;
        for (uint32_t s = 0; s < op.n_strings; ++s)
        {
            fpk::StringMeta const sm = op.meta[s];
            A R[NC];
            fpk::ev_rows<T, NC, 4>(states, op.vpr, col, col_live, begin, p1, ty, rt, sm, R);
            double const sr = op.scale[2 * s], si = op.scale[2 * s + 1];
#pragma unroll
            for (int c = 0; c < NC; ++c)
            {
                er[c] += sr * double(R[c]);
                ei[c] += si * double(R[c]);
            }
        }
        double v[2 * NC];
#pragma unroll
        for (int c = 0; c < NC; ++c)
        {
            v[2 * c] = er[c];
            v[2 * c + 1] = ei[c];
        }
        fpk::reduce_over_rows<double, 2 * NC>(v, red, tx, ty, ct, rt);
        if (ty == 0 && col_live)
        {
#pragma unroll
            for (int c = 0; c < NC; ++c)
            {
                uint64_t const t = col * NC + c;
                if (t < op.n_states)
                    partials[op.partial_base + uint64_t(tile) * op.n_states + t] = fpk::Dbl2{v[2 * c], v[2 * c + 1]};
            }
        }
    }
}

// The block that completes an operator's last expectation tile sums its partials
// (tiles in order) into the output.
template <class T>
__device__ void expect_finish(BatchOp const &op, fpk::Dbl2 const *partials)
{
    auto *out = static_cast<T *>(op.out);
    for (uint64_t t = threadIdx.x; t < op.n_states; t += kBlock)
    {
        double ar = 0.0, ai = 0.0;
        for (uint32_t i = 0; i < op.n_tiles; ++i)
        {
            fpk::Dbl2 const *p = partials + op.partial_base + uint64_t(i) * op.n_states + t;
            ar += load_coherent(&p->re);
            ai += load_coherent(&p->im);
        }
        if (op.accumulate)
        {
            out[2 * t] = T(double(out[2 * t]) + ar);
            out[2 * t + 1] = T(double(out[2 * t + 1]) + ai);
        }
        else
        {
            out[2 * t] = T(ar);
            out[2 * t + 1] = T(ai);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_batch_reset(uint32_t *counter, uint32_t *done, uint32_t n_ops)
{
    uint32_t const i = blockIdx.x * kBlock + threadIdx.x;
    if (i == 0)
        *counter = 0;
    if (i < n_ops)
        done[i] = 0;
}

__global__ __launch_bounds__(kBlock) void k_batch_persistent(BatchOp const *__restrict__ ops,
                                                             WorkItem const *__restrict__ items, uint32_t n_items,
                                                             uint32_t *counter, uint32_t *done,
                                                             fpk::Dbl2 *partials)
{
    __shared__ uint32_t s_item;
    __shared__ uint32_t s_last;
    __shared__ double red[4 * kBlock];
    while (true)
    {
        if (threadIdx.x == 0)
            s_item = atomicAdd(counter, 1u);
        __syncthreads();
        uint32_t const w = s_item;
        __syncthreads();
        if (w >= n_items)
            break;
        WorkItem const wi = items[w];
        BatchOp const &op = ops[wi.op];
        if (op.kind == kApply)
        {
            if (op.dtype == 0)
                apply_tile<double, 1>(op, wi.begin, wi.count);
            else if (op.nc == 2)
                apply_tile<float, 2>(op, wi.begin, wi.count);
            else
                apply_tile<float, 1>(op, wi.begin, wi.count);
        }
        else
        {
            if (op.dtype == 0)
                expect_tile<double, 1>(op, wi.tile, wi.begin, wi.count, partials, red);
            else if (op.nc == 2)
                expect_tile<float, 2>(op, wi.tile, wi.begin, wi.count, partials, red);
            else
                expect_tile<float, 1>(op, wi.tile, wi.begin, wi.count, partials, red);
            // publish this tile's partials; the last finisher of the operator reduces
            __threadfence();
            __syncthreads();
            if (threadIdx.x == 0)
                s_last = (atomicAdd(&done[wi.op], 1u) == op.n_tiles - 1) ? 1u : 0u;
            __syncthreads();
            if (s_last)
            {
                __threadfence();
                if (op.dtype == 0)
                    expect_finish<double>(op, partials);
                else
                    expect_finish<float>(op, partials);
            }
        }
        __syncthreads();
    }
}

} // namespace fpb
