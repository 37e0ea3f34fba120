// fp_oracle.cpp — CPU oracle for the operator-application hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: the parity tests, the
// smoke check and the CPU-baseline leg of the benchmark load it.  No product
// code path may call it; the HIP library fails loudly instead of falling back.
//
// It restates, in plain C++20 + OpenMP, the loops of the reference library
// (qognitive/fast-pauli, header-only C++ under fast_pauli/cpp/include/):
//   * get_sparse_repr            __pauli_string.hpp:49-118  (the incremental
//     PauliComposer construction, *not* the closed form the GPU uses, so the
//     two are independent and are cross-checked by the tests)
//   * PauliString::apply          __pauli_string.hpp:296-341
//   * PauliString::apply_batch    __pauli_string.hpp:377-436
//   * PauliString::expectation_value  __pauli_string.hpp:470-538
//   * PauliOp::apply              __pauli_op.hpp:413-468
//   * PauliOp::expectation_value  __pauli_op.hpp:496-549
//   * SummedPauliOp::apply        __summed_pauli_op.hpp:287-349
//   * SummedPauliOp::apply_weighted __summed_pauli_op.hpp:378-503
//   * SummedPauliOp::expectation_value __summed_pauli_op.hpp:530-614
// with the same schedules, per-thread scratch and reduction order, so that it
// doubles as the reference's CPU baseline.  `par` selects the OpenMP branch as
// is_parallel_execution_policy_v does (__type_traits.hpp:51-53).
//
// Strings are passed as letter codes (0=I, 1=X, 2=Y, 3=Z; __pauli.hpp:41), one
// row of n_qubits codes per string, character 0 first.  Complex arrays are
// interleaved (re, im) in row-major (dim x n_states) order.
//
// Parity status: every entry point below is pinned by the tests in
// tests/test_oracle.py (dense Kronecker brute force, the reference's
// known-answer cases, and invariants).

#include <omp.h>

#include <complex>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

namespace
{

template <class T> using cplx = std::complex<T>;

// __pauli_string.hpp:49-118.  The tensor product is taken right to left, so
// the last character owns bit 0 of the row index.
template <class T>
void sparse_repr(size_t n, uint8_t const *codes, std::vector<size_t> &k, std::vector<cplx<T>> &m)
{
    if (n == 0)
    {
        k.clear();
        m.clear();
        return;
    }
    size_t const dim = size_t(1) << n;
    k.assign(dim, 0);
    m.assign(dim, cplx<T>(0));

    size_t n_y = 0;
    for (size_t i = 0; i < n; ++i)
        n_y += codes[i] == 2;

    // code of the letter that owns bit l
    auto letter = [&](size_t l) { return codes[n - 1 - l]; };
    auto off_diagonal = [](uint8_t c) -> size_t { return (c == 1 || c == 2) ? 1 : 0; };

    size_t k0 = 0;
    for (size_t l = 0; l < n; ++l)
        k0 += (size_t(1) << l) * off_diagonal(letter(l));
    k[0] = k0;

    cplx<T> const first_values[4] = {cplx<T>(1, 0), cplx<T>(0, -1), cplx<T>(-1, 0), cplx<T>(0, 1)};
    m[0] = first_values[n_y % 4];

    for (size_t l = 0; l < n; ++l)
    {
        uint8_t const c = letter(l);
        T const eps = (c == 2 || c == 3) ? T(-1) : T(1);
        bool const flips = off_diagonal(c);
        size_t const half = size_t(1) << l;
        for (size_t li = half; li < 2 * half; ++li)
        {
            k[li] = flips ? k[li - half] - half : k[li - half] + half;
            m[li] = m[li - half] * eps;
        }
    }
}

// __pauli_string.hpp:312-341
template <class T>
void string_apply(bool par, size_t n, uint8_t const *codes, cplx<T> c, cplx<T> const *in, cplx<T> *out)
{
    std::vector<size_t> k;
    std::vector<cplx<T>> m;
    sparse_repr<T>(n, codes, k, m);
    size_t const dim = k.size();
    if (par)
    {
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < dim; ++i)
            out[i] += c * m[i] * in[k[i]];
    }
    else
    {
        for (size_t i = 0; i < dim; ++i)
            out[i] += c * m[i] * in[k[i]];
    }
}

// __pauli_string.hpp:396-436
template <class T>
void string_apply_batch(bool par, size_t n, uint8_t const *codes, cplx<T> c, cplx<T> const *in, cplx<T> *out,
                        size_t n_states)
{
    std::vector<size_t> k;
    std::vector<cplx<T>> m;
    sparse_repr<T>(n, codes, k, m);
    size_t const dim = k.size();
    if (par)
    {
#pragma omp parallel for schedule(static) collapse(2)
        for (size_t i = 0; i < dim; ++i)
        {
            for (size_t t = 0; t < n_states; ++t)
            {
                cplx<T> const c_m_i = c * m[i];
                out[i * n_states + t] += c_m_i * in[k[i] * n_states + t];
            }
        }
    }
    else
    {
        for (size_t i = 0; i < dim; ++i)
        {
            cplx<T> const c_m_i = c * m[i];
            cplx<T> const *row = in + k[i] * n_states;
            for (size_t t = 0; t < n_states; ++t)
                out[i * n_states + t] += c_m_i * row[t];
        }
    }
}

// __pauli_string.hpp:488-538
template <class T>
void string_expectation(bool par, size_t n, uint8_t const *codes, cplx<T> c, cplx<T> const *states, cplx<T> *out,
                        size_t n_states)
{
    std::vector<size_t> k;
    std::vector<cplx<T>> m;
    sparse_repr<T>(n, codes, k, m);
    size_t const dim = k.size();
    if (par)
    {
        size_t const n_threads = omp_get_max_threads();
        std::vector<cplx<T>> partial(n_states * n_threads); // (t, thread)
#pragma omp parallel for schedule(static) collapse(2)
        for (size_t t = 0; t < n_states; ++t)
        {
            for (size_t i = 0; i < dim; ++i)
            {
                cplx<T> const c_m_i = c * m[i];
                size_t const tid = omp_get_thread_num();
                partial[t * n_threads + tid] +=
                    std::conj(states[i * n_states + t]) * c_m_i * states[k[i] * n_states + t];
            }
        }
        for (size_t t = 0; t < n_states; ++t)
            for (size_t tid = 0; tid < n_threads; ++tid)
                out[t] += partial[t * n_threads + tid];
    }
    else
    {
        for (size_t i = 0; i < dim; ++i)
        {
            cplx<T> const c_m_i = c * m[i];
            for (size_t t = 0; t < n_states; ++t)
                out[t] += std::conj(states[i * n_states + t]) * c_m_i * states[k[i] * n_states + t];
        }
    }
}

// __pauli_op.hpp:413-468.  The par branch gives every thread a full copy of
// the output and reduces over threads in thread order; each thread runs the
// seq apply_batch (no policy is passed at __pauli_op.hpp:442).
template <class T>
void op_apply(bool par, size_t n, size_t n_strings, uint8_t const *codes, cplx<T> const *coeffs, cplx<T> const *in,
              cplx<T> *out, size_t n_states)
{
    size_t const dim = n ? size_t(1) << n : 0;
    if (par)
    {
        size_t const n_threads = omp_get_max_threads();
        std::vector<cplx<T>> scratch(n_threads * dim * n_states);
#pragma omp parallel
        {
#pragma omp for schedule(static)
            for (size_t s = 0; s < n_strings; ++s)
            {
                size_t const tid = omp_get_thread_num();
                string_apply_batch<T>(false, n, codes + s * n, coeffs[s], in, scratch.data() + tid * dim * n_states,
                                      n_states);
            }
#pragma omp for schedule(static) collapse(2)
            for (size_t i = 0; i < dim; ++i)
                for (size_t t = 0; t < n_states; ++t)
                    for (size_t th = 0; th < n_threads; ++th)
                        out[i * n_states + t] += scratch[(th * dim + i) * n_states + t];
        }
    }
    else
    {
        for (size_t s = 0; s < n_strings; ++s)
            string_apply_batch<T>(false, n, codes + s * n, coeffs[s], in, out, n_states);
    }
}

// __pauli_op.hpp:496-549
template <class T>
void op_expectation(bool par, size_t n, size_t n_strings, uint8_t const *codes, cplx<T> const *coeffs,
                    cplx<T> const *states, cplx<T> *out, size_t n_states)
{
    if (par)
    {
        size_t const n_threads = omp_get_max_threads();
        std::vector<cplx<T>> partial(n_threads * n_states); // (thread, t)
#pragma omp parallel for schedule(static)
        for (size_t s = 0; s < n_strings; ++s)
        {
            size_t const tid = omp_get_thread_num();
            string_expectation<T>(false, n, codes + s * n, coeffs[s], states, partial.data() + tid * n_states,
                                  n_states);
        }
#pragma omp parallel for schedule(static)
        for (size_t t = 0; t < n_states; ++t)
            for (size_t th = 0; th < n_threads; ++th)
                out[t] += partial[th * n_states + t];
    }
    else
    {
        for (size_t s = 0; s < n_strings; ++s)
            string_expectation<T>(false, n, codes + s * n, coeffs[s], states, out, n_states);
    }
}

// __summed_pauli_op.hpp:530-614.  coeffs is (n_strings, n_ops) row-major,
// out is (n_ops, n_states).
template <class T>
void summed_expectation(bool par, size_t n, size_t n_strings, size_t n_ops, uint8_t const *codes,
                        cplx<T> const *coeffs, cplx<T> const *states, cplx<T> *out, size_t n_states)
{
    std::vector<cplx<T>> ev(n_strings * n_states);
    if (par)
    {
#pragma omp parallel
        {
#pragma omp for schedule(static)
            for (size_t j = 0; j < n_strings; ++j)
                string_expectation<T>(false, n, codes + j * n, cplx<T>(1), states, ev.data() + j * n_states,
                                      n_states);
#pragma omp for collapse(2) schedule(static)
            for (size_t k = 0; k < n_ops; ++k)
                for (size_t t = 0; t < n_states; ++t)
                    for (size_t j = 0; j < n_strings; ++j)
                        out[k * n_states + t] += coeffs[j * n_ops + k] * ev[j * n_states + t];
        }
    }
    else
    {
        for (size_t j = 0; j < n_strings; ++j)
            string_expectation<T>(false, n, codes + j * n, cplx<T>(1), states, ev.data() + j * n_states, n_states);
        for (size_t t = 0; t < n_states; ++t)
            for (size_t k = 0; k < n_ops; ++k)
                for (size_t j = 0; j < n_strings; ++j)
                    out[k * n_states + t] += coeffs[j * n_ops + k] * ev[j * n_states + t];
    }
}

// __summed_pauli_op.hpp:287-349 (per-string coefficient = row sum over ops)
template <class T>
void summed_apply(bool par, size_t n, size_t n_strings, size_t n_ops, uint8_t const *codes, cplx<T> const *coeffs,
                  cplx<T> const *in, cplx<T> *out, size_t n_states)
{
    size_t const dim = n ? size_t(1) << n : 0;
    if (par)
    {
        size_t const n_threads = omp_get_max_threads();
        std::vector<cplx<T>> scratch(n_threads * dim * n_states);
#pragma omp parallel
        {
#pragma omp for schedule(static)
            for (size_t j = 0; j < n_strings; ++j)
            {
                cplx<T> summed(0, 0);
                for (size_t k = 0; k < n_ops; ++k)
                    summed += coeffs[j * n_ops + k];
                size_t const tid = omp_get_thread_num();
                string_apply_batch<T>(false, n, codes + j * n, summed, in, scratch.data() + tid * dim * n_states,
                                      n_states);
            }
#pragma omp for schedule(static) collapse(2)
            for (size_t i = 0; i < dim; ++i)
                for (size_t t = 0; t < n_states; ++t)
                    for (size_t th = 0; th < n_threads; ++th)
                        out[i * n_states + t] += scratch[(th * dim + i) * n_states + t];
        }
    }
    else
    {
        for (size_t j = 0; j < n_strings; ++j)
        {
            cplx<T> summed(0, 0);
            for (size_t k = 0; k < n_ops; ++k)
                summed += coeffs[j * n_ops + k];
            string_apply_batch<T>(false, n, codes + j * n, summed, in, out, n_states);
        }
    }
}

// __summed_pauli_op.hpp:378-503.  data is real, (n_ops, n_states).
template <class T>
void summed_apply_weighted(bool par, size_t n, size_t n_strings, size_t n_ops, uint8_t const *codes,
                           cplx<T> const *coeffs, cplx<T> const *in, T const *data, cplx<T> *out, size_t n_states)
{
    size_t const dim = n ? size_t(1) << n : 0;
    std::vector<cplx<T>> weighted(n_strings * n_states); // (j, t)
    if (par)
    {
        size_t const n_threads = omp_get_max_threads();
        std::vector<cplx<T>> scratch(n_threads * dim * n_states);
#pragma omp parallel
        {
#pragma omp for schedule(dynamic) collapse(2)
            for (size_t j = 0; j < n_strings; ++j)
                for (size_t t = 0; t < n_states; ++t)
                    for (size_t k = 0; k < n_ops; ++k)
                        weighted[j * n_states + t] += coeffs[j * n_ops + k] * data[k * n_states + t];

            std::vector<cplx<T>> psi_j(dim * n_states);
            cplx<T> *local = scratch.data() + size_t(omp_get_thread_num()) * dim * n_states;
#pragma omp for schedule(dynamic)
            for (size_t j = 0; j < n_strings; ++j)
            {
                std::fill(psi_j.begin(), psi_j.end(), cplx<T>(0));
                string_apply_batch<T>(false, n, codes + j * n, cplx<T>(1), in, psi_j.data(), n_states);
                for (size_t l = 0; l < dim; ++l)
                    for (size_t t = 0; t < n_states; ++t)
                        local[l * n_states + t] += psi_j[l * n_states + t] * weighted[j * n_states + t];
            }
#pragma omp for schedule(dynamic) collapse(2)
            for (size_t l = 0; l < dim; ++l)
                for (size_t t = 0; t < n_states; ++t)
                    for (size_t th = 0; th < n_threads; ++th)
                        out[l * n_states + t] += scratch[(th * dim + l) * n_states + t];
        }
    }
    else
    {
        std::vector<cplx<T>> psi_j(dim * n_states);
        for (size_t j = 0; j < n_strings; ++j)
            for (size_t t = 0; t < n_states; ++t)
                for (size_t k = 0; k < n_ops; ++k)
                    weighted[j * n_states + t] += coeffs[j * n_ops + k] * data[k * n_states + t];
        for (size_t j = 0; j < n_strings; ++j)
        {
            std::fill(psi_j.begin(), psi_j.end(), cplx<T>(0));
            string_apply_batch<T>(false, n, codes + j * n, cplx<T>(1), in, psi_j.data(), n_states);
            for (size_t l = 0; l < dim; ++l)
                for (size_t t = 0; t < n_states; ++t)
                    out[l * n_states + t] += psi_j[l * n_states + t] * weighted[j * n_states + t];
        }
    }
}

template <class T> cplx<T> *C(void *p)
{
    return static_cast<cplx<T> *>(p);
}
template <class T> cplx<T> const *CC(void const *p)
{
    return static_cast<cplx<T> const *>(p);
}

} // namespace

// ---------------------------------------------------------------------------
// extern "C" surface used by oracle/oracle.py.  `dtype` 0 = complex128,
// 1 = complex64 (the reference's float instantiation).  Outputs are
// accumulated (+=), exactly as the reference's C++ API does.
// ---------------------------------------------------------------------------
extern "C"
{

int fpo_max_threads(void)
{
    return omp_get_max_threads();
}

void fpo_set_threads(int n)
{
    omp_set_num_threads(n);
}

// Writes dim column indices and dim complex128 values.
void fpo_sparse_repr(size_t n, uint8_t const *codes, uint64_t *k_out, double *m_out)
{
    std::vector<size_t> k;
    std::vector<cplx<double>> m;
    sparse_repr<double>(n, codes, k, m);
    for (size_t i = 0; i < k.size(); ++i)
    {
        k_out[i] = k[i];
        m_out[2 * i] = m[i].real();
        m_out[2 * i + 1] = m[i].imag();
    }
}

#define FPO_DISPATCH(dtype, call_d, call_f)                                                                            \
    do                                                                                                                 \
    {                                                                                                                  \
        if ((dtype) == 0)                                                                                              \
        {                                                                                                              \
            call_d;                                                                                                    \
        }                                                                                                              \
        else                                                                                                           \
        {                                                                                                              \
            call_f;                                                                                                    \
        }                                                                                                              \
    } while (0)

void fpo_string_apply(int dtype, int par, size_t n, uint8_t const *codes, double c_re, double c_im, void const *in,
                      void *out)
{
    FPO_DISPATCH(dtype,
                 string_apply<double>(par, n, codes, cplx<double>(c_re, c_im), CC<double>(in), C<double>(out)),
                 string_apply<float>(par, n, codes, cplx<float>(float(c_re), float(c_im)), CC<float>(in),
                                     C<float>(out)));
}

void fpo_string_apply_batch(int dtype, int par, size_t n, uint8_t const *codes, double c_re, double c_im,
                            void const *in, void *out, size_t n_states)
{
    FPO_DISPATCH(dtype,
                 string_apply_batch<double>(par, n, codes, cplx<double>(c_re, c_im), CC<double>(in), C<double>(out),
                                            n_states),
                 string_apply_batch<float>(par, n, codes, cplx<float>(float(c_re), float(c_im)), CC<float>(in),
                                           C<float>(out), n_states));
}

void fpo_string_expectation(int dtype, int par, size_t n, uint8_t const *codes, double c_re, double c_im,
                            void const *states, void *out, size_t n_states)
{
    FPO_DISPATCH(dtype,
                 string_expectation<double>(par, n, codes, cplx<double>(c_re, c_im), CC<double>(states),
                                            C<double>(out), n_states),
                 string_expectation<float>(par, n, codes, cplx<float>(float(c_re), float(c_im)), CC<float>(states),
                                           C<float>(out), n_states));
}

void fpo_op_apply(int dtype, int par, size_t n, size_t n_strings, uint8_t const *codes, void const *coeffs,
                  void const *in, void *out, size_t n_states)
{
    FPO_DISPATCH(dtype,
                 op_apply<double>(par, n, n_strings, codes, CC<double>(coeffs), CC<double>(in), C<double>(out),
                                  n_states),
                 op_apply<float>(par, n, n_strings, codes, CC<float>(coeffs), CC<float>(in), C<float>(out),
                                 n_states));
}

void fpo_op_expectation(int dtype, int par, size_t n, size_t n_strings, uint8_t const *codes, void const *coeffs,
                        void const *states, void *out, size_t n_states)
{
    FPO_DISPATCH(dtype,
                 op_expectation<double>(par, n, n_strings, codes, CC<double>(coeffs), CC<double>(states),
                                        C<double>(out), n_states),
                 op_expectation<float>(par, n, n_strings, codes, CC<float>(coeffs), CC<float>(states), C<float>(out),
                                       n_states));
}

void fpo_summed_expectation(int dtype, int par, size_t n, size_t n_strings, size_t n_ops, uint8_t const *codes,
                            void const *coeffs, void const *states, void *out, size_t n_states)
{
    FPO_DISPATCH(dtype,
                 summed_expectation<double>(par, n, n_strings, n_ops, codes, CC<double>(coeffs), CC<double>(states),
                                            C<double>(out), n_states),
                 summed_expectation<float>(par, n, n_strings, n_ops, codes, CC<float>(coeffs), CC<float>(states),
                                           C<float>(out), n_states));
}

void fpo_summed_apply(int dtype, int par, size_t n, size_t n_strings, size_t n_ops, uint8_t const *codes,
                      void const *coeffs, void const *in, void *out, size_t n_states)
{
    FPO_DISPATCH(dtype,
                 summed_apply<double>(par, n, n_strings, n_ops, codes, CC<double>(coeffs), CC<double>(in),
                                      C<double>(out), n_states),
                 summed_apply<float>(par, n, n_strings, n_ops, codes, CC<float>(coeffs), CC<float>(in), C<float>(out),
                                     n_states));
}

void fpo_summed_apply_weighted(int dtype, int par, size_t n, size_t n_strings, size_t n_ops, uint8_t const *codes,
                               void const *coeffs, void const *in, void const *data, void *out, size_t n_states)
{
    FPO_DISPATCH(dtype,
                 summed_apply_weighted<double>(par, n, n_strings, n_ops, codes, CC<double>(coeffs), CC<double>(in),
                                               static_cast<double const *>(data), C<double>(out), n_states),
                 summed_apply_weighted<float>(par, n, n_strings, n_ops, codes, CC<float>(coeffs), CC<float>(in),
                                              static_cast<float const *>(data), C<float>(out), n_states));
}

} // extern "C"
