// fp_dense.hpp — dense and sparse materialisation of Pauli strings and sums on
// the device (debugging and interop with dense linear algebra).
//
//   sparse repr of string s:   cols[j] = j ^ x_s,  vals[j] = m_s(j)
//   dense string (assign):     out(j, j ^ x_s) = m_s(j)
//   dense sum (accumulate):    out(k, j, j ^ x_s) += h(s, k) * m_s(j),  s = 0, 1, ... in order
// with m_s(j) = (-i)^{nY_s} (-1)^{popcount(j & z_s)}.  m is formed as the
// reference forms it (a table entry, negated when the parity is odd, so -(0, -1)
// is (-0, 1)), the product h * m as std::complex does (cmul_exact), and one
// thread owns a row of the output and adds the strings in their order, so each
// element sees the reference's operations in the reference's order: results are
// IEEE-equal to fast_pauli/cpp/include/__pauli_string.hpp:549-560,
// __pauli_op.hpp:558-571 and __summed_pauli_op.hpp:644-665.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fp_kernels.hpp"

namespace fpd
{

// (-i)^{nY} * (-1)^{odd}: table entry, negated component-wise when odd
template <class T> __device__ __forceinline__ void pauli_value(uint32_t ny, bool odd, T &vr, T &vi)
{
    uint32_t const q = ny & 3u;
    vr = q == 0 ? T(1) : (q == 2 ? T(-1) : T(0));
    vi = q == 1 ? T(-1) : (q == 3 ? T(1) : T(0));
    if (odd)
    {
        vr = -vr;
        vi = -vi;
    }
}

// Coefficient c_s from the folded form cf_s = c_s (-i)^{nY_s} the operator set
// keeps (an exact swap and/or negation, undone exactly here).
template <class T> __device__ __forceinline__ void unfold(T fr, T fi, uint32_t ny, T &cr, T &ci)
{
    switch (ny & 3u)
    {
    case 0: cr = fr, ci = fi; break;
    case 1: cr = -fi, ci = fr; break; // cf = (ci, -cr)
    case 2: cr = -fr, ci = -fi; break;
    default: cr = fi, ci = -fr; break; // cf = (-ci, cr)
    }
}

template <class T>
__global__ __launch_bounds__(fpk::kBlock) void k_sparse_repr(uint64_t *__restrict__ cols,
                                                             typename fpk::VecOf<T, 1>::type *__restrict__ vals,
                                                             uint64_t dim, uint64_t x, uint64_t z, uint32_t ny)
{
    typedef typename fpk::VecOf<T, 1>::type V;
    for (uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; j < dim; j += uint64_t(gridDim.x) * blockDim.x)
    {
        T vr, vi;
        pauli_value<T>(ny, fpk::parity64(j & z) != 0, vr, vi);
        cols[j] = j ^ x;
        vals[j] = V{vr, vi};
    }
}

template <class T>
__global__ __launch_bounds__(fpk::kBlock) void k_fill_zero(typename fpk::VecOf<T, 1>::type *__restrict__ out, uint64_t n)
{
    typedef typename fpk::VecOf<T, 1>::type V;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        out[i] = V{T(0), T(0)};
}

// Dense matrices of K operators, (K x dim x dim) row-major.  Thread per (k, row j).
// ASSIGN: one string (index s0), out(j, j ^ x) = m(j), no coefficient.
// Otherwise strings [0, n_str) with coefficient h(s, k) = coef[s * cstride + k]
// (interleaved complex), folded by (-i)^{nY_s} when `folded`.
template <class T, bool ASSIGN>
__global__ __launch_bounds__(fpk::kBlock) void k_to_tensor(typename fpk::VecOf<T, 1>::type *__restrict__ out,
                                                           uint32_t n_qubits, uint64_t n_rows,
                                                           uint64_t const *__restrict__ xs,
                                                           uint64_t const *__restrict__ zs,
                                                           uint32_t const *__restrict__ nys, uint32_t s0,
                                                           uint32_t n_str, T const *__restrict__ coef,
                                                           uint32_t cstride, int folded)
{
    typedef typename fpk::VecOf<T, 1>::type V;
    uint64_t const dim = uint64_t(1) << n_qubits;
    for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < n_rows;
         r += uint64_t(gridDim.x) * blockDim.x)
    {
        uint64_t const k = r >> n_qubits;
        uint64_t const j = r & (dim - 1u);
        V *row = out + r * dim;
        if constexpr (ASSIGN)
        {
            T vr, vi;
            pauli_value<T>(nys[s0], fpk::parity64(j & zs[s0]) != 0, vr, vi);
            row[j ^ xs[s0]] = V{vr, vi};
        }
        else
        {
            for (uint32_t s = 0; s < n_str; ++s)
            {
                uint32_t const ny = nys[s];
                T vr, vi;
                pauli_value<T>(ny, fpk::parity64(j & zs[s]) != 0, vr, vi);
                uint64_t const ci_idx = 2u * (uint64_t(s) * cstride + k);
                T cr = coef[ci_idx], ci = coef[ci_idx + 1];
                if (folded)
                    unfold<T>(cr, ci, ny, cr, ci);
                T pr, pi;
                fpk::cmul_exact<T>(cr, ci, vr, vi, pr, pi);
                V &o = row[j ^ xs[s]];
                V const cur = o;
                o = V{fpk::add_exact<T>(cur[0], pr), fpk::add_exact<T>(cur[1], pi)};
            }
        }
    }
}

} // namespace fpd
