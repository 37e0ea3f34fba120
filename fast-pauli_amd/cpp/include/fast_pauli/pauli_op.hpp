// pauli_op.hpp — PauliOp<T, H> = sum_s h_s P_s with its hot path on the GPU.
//
// Keeps the reference's type, members and method signatures
// (fast_pauli/cpp/include/__pauli_op.hpp:38-592).  apply and
// expectation_value accumulate into the caller's output like the reference.
// apply accumulates strings in order s = 0..N-1 for every element (the
// reference's sequential order), so its results are bit-exact with the
// sequential CPU loops for any policy.
#pragma once

#include <algorithm>
#include <complex>
#include <execution>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "device.hpp"
#include "mdspan.hpp"
#include "pauli_string.hpp"

namespace fast_pauli
{

namespace detail
{
template <class T>
void op_apply(fp_hip_ops *ops, std::complex<T> const *in, std::complex<T> *out, std::size_t dim, std::size_t n_states,
              bool accumulate)
{
    std::size_t const bytes = dim * n_states * sizeof(std::complex<T>);
    void const *d_in = hip::stage_in(0, in, bytes);
    hip::StagedOut o(1, out, bytes, accumulate);
    hip::check(fp_hip_op_apply(ops, hip::dtype_of<T>(), d_in, o.dev(), dim, n_states, accumulate ? 1 : 0, nullptr));
    o.finish();
}

template <class T>
void op_expectation(fp_hip_ops *ops, std::complex<T> const *states, std::complex<T> *out, std::size_t dim,
                    std::size_t n_states, bool accumulate)
{
    void const *d_in = hip::stage_in(0, states, dim * n_states * sizeof(std::complex<T>));
    hip::StagedOut o(1, out, n_states * sizeof(std::complex<T>), accumulate);
    hip::check(
        fp_hip_op_expectation(ops, hip::dtype_of<T>(), d_in, o.dev(), dim, n_states, accumulate ? 1 : 0, nullptr));
    o.finish();
}
} // namespace detail

template <std::floating_point T, typename H = std::complex<T>> struct PauliOp
{
    std::vector<H> coeffs;
    std::vector<PauliString> pauli_strings;

    PauliOp() = default;

    PauliOp(std::vector<std::string> const &strings)
    {
        for (auto const &s : strings)
            pauli_strings.push_back(PauliString(s));
        coeffs.resize(pauli_strings.size(), H(1.0));
        validate_pauli_strings(pauli_strings);
    }

    PauliOp(std::vector<PauliString> strings) : coeffs(strings.size(), H(1.0)), pauli_strings(std::move(strings))
    {
        validate_pauli_strings(pauli_strings);
    }

    PauliOp(std::vector<H> coefficients, std::vector<PauliString> strings)
        : coeffs(std::move(coefficients)), pauli_strings(std::move(strings))
    {
        if (coeffs.size() != pauli_strings.size())
            throw std::invalid_argument("coeffs and pauli_strings must have the same size");
        validate_pauli_strings(pauli_strings);
    }

    PauliOp(PauliOp const &o) : coeffs(o.coeffs), pauli_strings(o.pauli_strings)
    {
    }
    PauliOp &operator=(PauliOp const &o)
    {
        coeffs = o.coeffs;
        pauli_strings = o.pauli_strings;
        return *this;
    }
    PauliOp(PauliOp &&) noexcept = default;
    PauliOp &operator=(PauliOp &&) noexcept = default;

    std::size_t dim() const
    {
        return pauli_strings.empty() ? 0 : pauli_strings[0].dim();
    }
    std::size_t n_qubits() const
    {
        return pauli_strings.empty() ? 0 : pauli_strings[0].n_qubits();
    }
    std::size_t n_pauli_strings() const
    {
        return pauli_strings.size();
    }

    // ---- algebra (host side; __pauli_op.hpp:142-338) -------------------------

    void scale(std::complex<T> factor)
    {
        for (auto &c : coeffs)
            c *= factor;
    }

    template <class M>
        requires MdspanOfRank<M, 1>
    void scale(M factors)
    {
        if (factors.size() != n_pauli_strings())
            throw std::invalid_argument("factors must have the same length as the number of PauliStrings");
        for (std::size_t i = 0; i < coeffs.size(); ++i)
            coeffs[i] *= factors.data_handle()[i];
    }

    PauliOp operator-() const
    {
        std::vector<H> neg(coeffs.size());
        std::transform(coeffs.begin(), coeffs.end(), neg.begin(), [](H const &c) { return -c; });
        return PauliOp(std::move(neg), pauli_strings);
    }

    friend PauliOp operator*(PauliOp const &op, PauliString const &ps)
    {
        if (op.dim() != ps.dim())
            throw std::invalid_argument("PauliStrings must have same size as PauliOp");
        std::vector<H> c(op.coeffs);
        std::vector<PauliString> s;
        s.reserve(op.n_pauli_strings());
        for (std::size_t i = 0; i < op.n_pauli_strings(); ++i)
        {
            auto [phase, p] = op.pauli_strings[i] * ps;
            c[i] *= phase;
            s.push_back(std::move(p));
        }
        return PauliOp(std::move(c), std::move(s));
    }

    friend PauliOp operator*(PauliString const &ps, PauliOp const &op)
    {
        if (op.dim() != ps.dim())
            throw std::invalid_argument("PauliStrings must have same size as PauliOp");
        std::vector<H> c(op.coeffs);
        std::vector<PauliString> s;
        s.reserve(op.n_pauli_strings());
        for (std::size_t i = 0; i < op.n_pauli_strings(); ++i)
        {
            auto [phase, p] = ps * op.pauli_strings[i];
            c[i] *= phase;
            s.push_back(std::move(p));
        }
        return PauliOp(std::move(c), std::move(s));
    }

    // Product of two operators with duplicate strings merged
    // (__pauli_op.hpp:237-281).  Terms appear in order of first occurrence.
    friend PauliOp operator*(PauliOp const &lhs, PauliOp const &rhs)
    {
        if (lhs.dim() != rhs.dim())
            throw std::invalid_argument("Mismatched dimensions for provided PauliOp");
        std::vector<H> c;
        std::vector<PauliString> s;
        auto add = [&](auto &where, auto const &key, std::complex<double> phase, PauliString &&p, std::size_t i,
                       std::size_t j) {
            H const cij = H(phase) * lhs.coeffs[i] * rhs.coeffs[j];
            auto it = where.find(key);
            if (it == where.end())
            {
                where.emplace(key, s.size());
                s.push_back(std::move(p));
                c.push_back(cij);
            }
            else
            {
                c[it->second] += cij;
            }
        };
        std::size_t const n = lhs.n_qubits();
        if (n <= 32)
        {
            // products keyed by their (x, z) bit masks (letter order is fixed by n)
            std::unordered_map<std::uint64_t, std::size_t> where;
            for (std::size_t i = 0; i < lhs.n_pauli_strings(); ++i)
                for (std::size_t j = 0; j < rhs.n_pauli_strings(); ++j)
                {
                    auto [phase, p] = lhs.pauli_strings[i] * rhs.pauli_strings[j];
                    std::uint64_t x = 0, z = 0;
                    for (std::size_t q = 0; q < n; ++q)
                    {
                        x |= std::uint64_t(p.paulis[q].code == 1 || p.paulis[q].code == 2) << q;
                        z |= std::uint64_t(p.paulis[q].code == 2 || p.paulis[q].code == 3) << q;
                    }
                    add(where, (x << 32) | z, phase, std::move(p), i, j);
                }
        }
        else
        {
            std::map<std::string, std::size_t> where;
            for (std::size_t i = 0; i < lhs.n_pauli_strings(); ++i)
                for (std::size_t j = 0; j < rhs.n_pauli_strings(); ++j)
                {
                    auto [phase, p] = lhs.pauli_strings[i] * rhs.pauli_strings[j];
                    std::string const key = p.str();
                    add(where, key, phase, std::move(p), i, j);
                }
        }
        return PauliOp(std::move(c), std::move(s));
    }

    void extend(PauliString pauli_str, std::complex<T> coeff, bool dedupe = true)
    {
        if (pauli_str.dim() != dim())
            throw std::invalid_argument("PauliStrings must have same size as PauliOp");
        if (dedupe)
            for (std::size_t i = 0; i < pauli_strings.size(); ++i)
                if (pauli_strings[i] == pauli_str)
                {
                    coeffs[i] += coeff;
                    return;
                }
        coeffs.push_back(coeff);
        pauli_strings.push_back(std::move(pauli_str));
    }

    void extend(PauliOp const &other)
    {
        if (other.dim() != dim())
            throw std::invalid_argument("Mismatched dimensions for provided PauliOp");
        pauli_strings.insert(pauli_strings.end(), other.pauli_strings.begin(), other.pauli_strings.end());
        coeffs.insert(coeffs.end(), other.coeffs.begin(), other.coeffs.end());
    }

    // ---- hot path -------------------------------------------------------------

    template <class MOut, class MIn>
        requires MdspanLike<MOut> && MdspanLike<MIn> && (MOut::rank() == MIn::rank())
    void apply(MOut new_states, MIn const states) const
    {
        apply(std::execution::seq, new_states, states);
    }

    // 1-D: reshaped to (dim, 1) (__pauli_op.hpp:376-383); 2-D: __pauli_op.hpp:413-468
    template <execution_policy Policy, class MOut, class MIn>
        requires MdspanLike<MOut> && MdspanLike<MIn> && (MOut::rank() == MIn::rank()) && (MIn::rank() <= 2)
    void apply(Policy &&, MOut new_states, MIn const states) const
    {
        std::size_t const n_states = MIn::rank() == 2 ? std::size_t(states.extent(MIn::rank() - 1)) : 1;
        std::size_t const out_states = MOut::rank() == 2 ? std::size_t(new_states.extent(MOut::rank() - 1)) : 1;
        if (states.extent(0) != dim())
            throw std::invalid_argument("[PauliOp] state size must match the dimension of the operators");
        if (states.extent(0) != new_states.extent(0) || n_states != out_states)
            throw std::invalid_argument("[PauliOp] new_states must have the same dimensions as states");
        detail::op_apply<T>(device_ops(), states.data_handle(), new_states.data_handle(), dim(), n_states, true);
    }

    template <class MOut, class MIn>
        requires MdspanOfRank<MOut, 1> && MdspanOfRank<MIn, 2>
    void expectation_value(MOut expectation_vals_out, MIn states) const
    {
        expectation_value(std::execution::seq, expectation_vals_out, states);
    }

    // __pauli_op.hpp:496-549: out[t] += sum_s h_s <psi_t|P_s|psi_t>
    template <execution_policy Policy, class MOut, class MIn>
        requires MdspanOfRank<MOut, 1> && MdspanOfRank<MIn, 2>
    void expectation_value(Policy &&, MOut expectation_vals_out, MIn states) const
    {
        if (states.extent(0) != dim())
            throw std::invalid_argument("[PauliOp] state size must match the dimension of the operators");
        if (expectation_vals_out.extent(0) != states.extent(1))
            throw std::invalid_argument("[PauliOp] expectation_vals_out shape must match the number of states");
        detail::op_expectation<T>(device_ops(), states.data_handle(), expectation_vals_out.data_handle(), dim(),
                                  states.extent(1), true);
    }

    // Dense matrix (__pauli_op.hpp:558-571), accumulated: on the host, or on the
    // device when the view is device memory.
    template <class M>
        requires MdspanOfRank<M, 2>
    void to_tensor(M output) const
    {
        if (!pauli_strings.empty() && hip::is_device_view(output.data_handle()))
        {
            if (output.extent(0) != dim() || output.extent(1) != dim())
                throw std::invalid_argument("Output tensor must have the same dimensions as the PauliOp");
            hip::check(fp_hip_op_to_tensor(device_ops(), hip::dtype_of<T>(), output.data_handle(), dim(), 1, nullptr));
            hip::check(fp_hip_stream_synchronize(nullptr));
            return;
        }
        for (std::size_t i = 0; i < pauli_strings.size(); ++i)
        {
            auto [cols, vals] = get_sparse_repr<T>(pauli_strings[i].paulis);
            for (std::size_t j = 0; j < dim(); ++j)
                output(j, cols[j]) += std::complex<T>(coeffs[i]) * vals[j];
        }
    }

    static void validate_pauli_strings(std::vector<PauliString> const &ps)
    {
        if (ps.empty())
            return;
        std::size_t const n = ps[0].n_qubits();
        for (auto const &p : ps)
            if (p.n_qubits() != n)
                throw std::invalid_argument("All PauliStrings must have the same size");
    }

    // Device copy of the strings and coefficients (rebuilt when either changes).
    fp_hip_ops *device_ops() const
    {
        if (pauli_strings.empty())
            throw std::invalid_argument("PauliOp has no Pauli strings");
        std::size_t const n = n_qubits();
        std::vector<std::uint8_t> codes(pauli_strings.size() * n);
        for (std::size_t s = 0; s < pauli_strings.size(); ++s)
            for (std::size_t i = 0; i < n; ++i)
                codes[s * n + i] = pauli_strings[s].paulis[i].code;
        std::vector<double> c(2 * coeffs.size());
        for (std::size_t s = 0; s < coeffs.size(); ++s)
        {
            std::complex<double> const v(coeffs[s]);
            c[2 * s] = v.real();
            c[2 * s + 1] = v.imag();
        }
        return cache_->get(codes, c,
                           [&] { return std::make_unique<hip::Ops>(n, pauli_strings.size(), codes.data(), c.data()); });
    }

  private:
    std::shared_ptr<hip::OpsCache> cache_ = std::make_shared<hip::OpsCache>();
};

} // namespace fast_pauli
