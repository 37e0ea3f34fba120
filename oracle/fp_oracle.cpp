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
//   * Pauli / PauliString products  __pauli.hpp:110-178, __pauli_string.hpp:228-247
//   * calculate_pauli_strings_max_weight  __pauli_helpers.hpp:33-153
//   * SummedPauliOp::square         __summed_pauli_op.hpp:197-268
//   * PauliOp * PauliOp             __pauli_op.hpp:237-281 (one thread)
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

#include <algorithm>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <map>
#include <tuple>
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

// ---------------------------------------------------------------------------
// String algebra
// ---------------------------------------------------------------------------

// __pauli.hpp:110-178: the product of two letters (0 = I, 1 = X, 2 = Y, 3 = Z) and
// its phase, as the multiplication table of the Pauli matrices (row: left letter).
struct LetterProduct
{
    double re, im;
    uint8_t code;
};
constexpr LetterProduct kLetterTable[4][4] = {
    {{1, 0, 0}, {1, 0, 1}, {1, 0, 2}, {1, 0, 3}},
    {{1, 0, 1}, {1, 0, 0}, {0, 1, 3}, {0, -1, 2}},
    {{1, 0, 2}, {0, -1, 3}, {1, 0, 0}, {0, 1, 1}},
    {{1, 0, 3}, {0, 1, 2}, {0, -1, 1}, {1, 0, 0}},
};

// __pauli_string.hpp:228-247: letter by letter, the phase multiplied up in complex<double>.
cplx<double> string_mul(size_t n, uint8_t const *a, uint8_t const *b, uint8_t *out)
{
    cplx<double> phase = 1;
    for (size_t q = 0; q < n; ++q)
    {
        LetterProduct const &lp = kLetterTable[a[q]][b[q]];
        phase *= cplx<double>(lp.re, lp.im);
        out[q] = lp.code;
    }
    return phase;
}

// __pauli_helpers.hpp:33-58: the words of `weight` non-identity letters, each new
// letter appended (the first letter most significant), X < Y < Z.
std::vector<std::vector<uint8_t>> nontrivial_words(size_t weight)
{
    if (weight == 0)
        return {};
    std::vector<std::vector<uint8_t>> words{{1}, {2}, {3}};
    for (size_t i = 1; i < weight; ++i)
    {
        std::vector<std::vector<uint8_t>> next;
        for (auto const &w : words)
            for (uint8_t l = 1; l <= 3; ++l)
            {
                auto e = w;
                e.push_back(l);
                next.push_back(std::move(e));
            }
        words = std::move(next);
    }
    return words;
}

// __pauli_helpers.hpp:68-89: k of n positions, from the bitmask with k leading ones
// through every previous permutation of it.
std::vector<std::vector<size_t>> idx_combinations(size_t n, size_t k)
{
    std::vector<std::vector<size_t>> out;
    std::vector<int> mask(n, 0);
    for (size_t i = 0; i < k && i < n; ++i)
        mask[i] = 1;
    do
    {
        std::vector<size_t> c;
        for (size_t i = 0; i < n; ++i)
            if (mask[i])
                c.push_back(i);
        out.push_back(std::move(c));
    } while (std::prev_permutation(mask.begin(), mask.end()));
    return out;
}

// __pauli_helpers.hpp:99-153: every string of weight 0..max_weight; within a weight, words
// outer and positions inner.
std::vector<std::vector<uint8_t>> strings_max_weight(size_t n, size_t max_weight)
{
    std::vector<std::vector<uint8_t>> out;
    for (size_t w = 0; w <= max_weight; ++w)
    {
        if (w == 0)
        {
            out.emplace_back(n, uint8_t(0));
            continue;
        }
        auto const words = nontrivial_words(w);
        auto const combos = idx_combinations(n, w);
        for (auto const &word : words)
            for (auto const &c : combos)
            {
                std::vector<uint8_t> str(n, 0);
                for (size_t t = 0; t < c.size(); ++t)
                    str[c[t]] = word[t];
                out.push_back(std::move(str));
            }
    }
    return out;
}

// __summed_pauli_op.hpp:197-268.  coeffs is (N x K); returns the squared strings and
// their (A x K) coefficients.  Part 2 builds T_aij serially in (i, j) order; part 4
// sums each (a, k) over T_a in that order (OpenMP over (a, k), as the reference).
template <class T>
void summed_square(size_t n, size_t N, size_t K, uint8_t const *codes, cplx<T> const *coeffs,
                   std::vector<std::vector<uint8_t>> &sq, std::vector<cplx<T>> &coeffs_sq)
{
    size_t weight = 0;
    for (size_t i = 0; i < N; ++i)
    {
        size_t w = 0;
        for (size_t q = 0; q < n; ++q)
            w += codes[i * n + q] != 0;
        weight = std::max(weight, w);
    }
    sq = strings_max_weight(n, std::min(n, 2 * weight));
    std::map<std::vector<uint8_t>, size_t> index;
    for (size_t a = 0; a < sq.size(); ++a)
        index[sq[a]] = a;
    std::vector<std::vector<std::tuple<size_t, size_t, cplx<T>>>> t_aij(sq.size());
    std::vector<uint8_t> prod(n);
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j)
        {
            cplx<T> const phase(string_mul(n, codes + i * n, codes + j * n, prod.data()));
            t_aij[index.at(prod)].emplace_back(i, j, phase);
        }
    std::vector<cplx<T>> ct(K * N);
    for (size_t i = 0; i < N; ++i)
        for (size_t k = 0; k < K; ++k)
            ct[k * N + i] = coeffs[i * K + k];
    coeffs_sq.assign(sq.size() * K, cplx<T>(0));
#pragma omp parallel for schedule(dynamic) collapse(2)
    for (size_t a = 0; a < sq.size(); ++a)
        for (size_t k = 0; k < K; ++k)
            for (auto const &[i, j, phase] : t_aij[a])
                coeffs_sq[a * K + k] += phase * ct[k * N + i] * ct[k * N + j];
}

// __pauli_op.hpp:237-281 on one thread: products in (i, j) order, equal strings merged
// (phase * c1_i * c2_j added in that order).  The reference lists the merged strings in
// the iteration order of per-thread unordered_maps, which is unspecified; here they come
// in order of first occurrence.
void op_mul(size_t n, size_t N1, uint8_t const *codes1, cplx<double> const *c1, size_t N2, uint8_t const *codes2,
            cplx<double> const *c2, std::vector<std::vector<uint8_t>> &strs, std::vector<cplx<double>> &coeffs)
{
    std::map<std::vector<uint8_t>, size_t> index;
    std::vector<uint8_t> prod(n);
    for (size_t i = 0; i < N1; ++i)
        for (size_t j = 0; j < N2; ++j)
        {
            cplx<double> const phase = string_mul(n, codes1 + i * n, codes2 + j * n, prod.data());
            cplx<double> const c = phase * c1[i] * c2[j];
            auto const it = index.find(prod);
            if (it == index.end())
            {
                index.emplace(prod, strs.size());
                strs.push_back(prod);
                coeffs.push_back(c);
            }
            else
                coeffs[it->second] += c;
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

// Writes the strings of calculate_pauli_strings_max_weight(n, max_weight) (codes, n per
// string) when `codes_out` is not NULL; returns their count.
size_t fpo_strings_max_weight(size_t n, size_t max_weight, uint8_t *codes_out)
{
    auto const v = strings_max_weight(n, max_weight);
    if (codes_out)
        for (size_t a = 0; a < v.size(); ++a)
            std::memcpy(codes_out + a * n, v[a].data(), n);
    return v.size();
}

// Product of two strings: writes the product's codes and its phase (re, im).
void fpo_string_mul(size_t n, uint8_t const *a, uint8_t const *b, uint8_t *codes_out, double *phase_out)
{
    cplx<double> const p = string_mul(n, a, b, codes_out);
    phase_out[0] = p.real();
    phase_out[1] = p.imag();
}

// SummedPauliOp::square: returns the count A of squared strings; when the outputs are
// not NULL and A <= capacity, writes their codes (A x n) and coefficients (A x K
// complex of `dtype`).
size_t fpo_summed_square(int dtype, size_t n, size_t N, size_t K, uint8_t const *codes, void const *coeffs_nk,
                         size_t capacity, uint8_t *codes_out, void *coeffs_out)
{
    std::vector<std::vector<uint8_t>> sq;
    auto emit = [&](auto const &c) {
        if (sq.size() > capacity)
            return;
        if (codes_out)
            for (size_t a = 0; a < sq.size(); ++a)
                std::memcpy(codes_out + a * n, sq[a].data(), n);
        if (coeffs_out)
            std::memcpy(coeffs_out, c.data(), c.size() * sizeof(c[0]));
    };
    if (dtype == 0)
    {
        std::vector<cplx<double>> c;
        summed_square<double>(n, N, K, codes, CC<double>(coeffs_nk), sq, c);
        emit(c);
    }
    else
    {
        std::vector<cplx<float>> c;
        summed_square<float>(n, N, K, codes, CC<float>(coeffs_nk), sq, c);
        emit(c);
    }
    return sq.size();
}

// PauliOp * PauliOp (complex128): returns the count of merged strings; when the outputs
// are not NULL, writes their codes and coefficients (first-occurrence order).
size_t fpo_op_mul(size_t n, size_t N1, uint8_t const *codes1, double const *c1, size_t N2, uint8_t const *codes2,
                  double const *c2, uint8_t *codes_out, double *coeffs_out)
{
    std::vector<std::vector<uint8_t>> strs;
    std::vector<cplx<double>> coeffs;
    op_mul(n, N1, codes1, reinterpret_cast<cplx<double> const *>(c1), N2, codes2,
           reinterpret_cast<cplx<double> const *>(c2), strs, coeffs);
    if (codes_out)
        for (size_t a = 0; a < strs.size(); ++a)
            std::memcpy(codes_out + a * n, strs[a].data(), n);
    if (coeffs_out)
        std::memcpy(coeffs_out, coeffs.data(), coeffs.size() * sizeof(cplx<double>));
    return strs.size();
}

} // extern "C"
