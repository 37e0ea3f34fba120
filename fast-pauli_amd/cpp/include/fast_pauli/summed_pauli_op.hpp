// summed_pauli_op.hpp — SummedPauliOp<T>: K operators A_k = sum_j h(j,k) P_j
// sharing one string set, hot path on the GPU.
//
// Keeps the reference's members and signatures
// (fast_pauli/cpp/include/__summed_pauli_op.hpp:37-666): pauli_strings,
// coeffs_raw and the (n_pauli_strings x n_operators) view `coeffs`.
// expectation_value runs the per-string expectation values and then the
// (operators x strings) x (strings x states) contraction on MFMA; apply and
// apply_weighted are bit-exact with the reference's sequential loops.
#pragma once

#include <complex>
#include <cstring>
#include <execution>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "device.hpp"
#include "helpers.hpp"
#include "mdspan.hpp"
#include "pauli_op.hpp"
#include "pauli_string.hpp"

namespace fast_pauli
{

template <std::floating_point T> struct SummedPauliOp
{
    template <std::size_t N> using Tensor = mdspan<std::complex<T>, dextents<std::size_t, N>>;

    std::vector<PauliString> pauli_strings;
    std::vector<std::complex<T>> coeffs_raw;
    Tensor<2> coeffs; // (n_pauli_strings, n_operators)

    // Precision of the expectation-value contraction: FP_CONTRACT_AUTO (FP64 for
    // double, FP32 for float), FP_CONTRACT_FP32 or FP_CONTRACT_BF16X3.
    int contraction = FP_CONTRACT_AUTO;

    SummedPauliOp() noexcept = default;

    SummedPauliOp(std::vector<PauliString> const &strings, std::vector<std::complex<T>> const &raw)
        : pauli_strings(strings), coeffs_raw(raw)
    {
        if (pauli_strings.empty())
            throw std::invalid_argument("SummedPauliOp needs at least one PauliString");
        std::size_t const n_ops = coeffs_raw.size() / pauli_strings.size();
        coeffs = Tensor<2>(coeffs_raw.data(), pauli_strings.size(), n_ops);
        check_ctor_inputs();
    }

    template <class M>
        requires MdspanOfRank<M, 2>
    SummedPauliOp(std::vector<PauliString> const &strings, M const c) : pauli_strings(strings)
    {
        adopt(c);
    }

    template <class M>
        requires MdspanOfRank<M, 2>
    SummedPauliOp(std::vector<std::string> const &strings, M const c)
    {
        for (auto const &s : strings)
            pauli_strings.push_back(PauliString(s));
        adopt(c);
    }

    SummedPauliOp(SummedPauliOp const &o)
        : pauli_strings(o.pauli_strings), coeffs_raw(o.coeffs_raw), contraction(o.contraction)
    {
        coeffs = Tensor<2>(coeffs_raw.data(), o.coeffs.extent(0), o.coeffs.extent(1));
    }
    SummedPauliOp &operator=(SummedPauliOp const &o)
    {
        pauli_strings = o.pauli_strings;
        coeffs_raw = o.coeffs_raw;
        contraction = o.contraction;
        coeffs = Tensor<2>(coeffs_raw.data(), o.coeffs.extent(0), o.coeffs.extent(1));
        return *this;
    }

    std::size_t dim() const noexcept
    {
        return pauli_strings.empty() ? 0 : pauli_strings[0].dim();
    }
    std::size_t n_qubits() const noexcept
    {
        return pauli_strings.empty() ? 0 : pauli_strings[0].n_qubits();
    }
    std::size_t n_operators() const noexcept
    {
        return coeffs.extent(1);
    }
    std::size_t n_pauli_strings() const noexcept
    {
        return pauli_strings.size();
    }

    // ---- hot path ---------------------------------------------------------------

    template <class MOut, class MIn>
        requires MdspanOfRank<MOut, 2> && MdspanOfRank<MIn, 2>
    void expectation_value(MOut out, MIn states) const
    {
        expectation_value(std::execution::seq, out, states);
    }

    // __summed_pauli_op.hpp:530-614: out(k,t) += sum_j h(j,k) <psi_t|P_j|psi_t>
    template <execution_policy Policy, class MOut, class MIn>
        requires MdspanOfRank<MOut, 2> && MdspanOfRank<MIn, 2>
    void expectation_value(Policy &&, MOut out, MIn states) const
    {
        std::size_t const n_data = states.extent(1);
        if (out.extent(0) != n_operators())
            throw std::invalid_argument("expectation_vals_out must have the same number of operators (" +
                                        std::to_string(out.extent(0)) + ") as the SummedPauliOp (" +
                                        std::to_string(n_operators()) + ")");
        if (states.extent(0) != dim())
            throw std::invalid_argument("states must have the same dimension (" + std::to_string(states.extent(0)) +
                                        ") as the SummedPauliOp (" + std::to_string(dim()) + ")");
        if (out.extent(1) != n_data)
            throw std::invalid_argument("expectation_vals_out must have the same number of states (" +
                                        std::to_string(out.extent(1)) + ") as the input states (" +
                                        std::to_string(n_data) + ")");
        expectation_value_raw(states.data_handle(), out.data_handle(), n_data, true);
    }

    template <class MOut, class MIn>
        requires MdspanOfRank<MOut, 2> && MdspanOfRank<MIn, 2>
    void apply(MOut new_states, MIn states) const
    {
        apply(std::execution::seq, new_states, states);
    }

    // __summed_pauli_op.hpp:287-349: new_states += sum_j (sum_k h(j,k)) P_j states
    template <execution_policy Policy, class MOut, class MIn>
        requires MdspanOfRank<MOut, 2> && MdspanOfRank<MIn, 2>
    void apply(Policy &&, MOut new_states, MIn states) const
    {
        if (states.extent(0) != new_states.extent(0) || states.extent(1) != new_states.extent(1))
            throw std::invalid_argument("new_states must have the same dimensions as states");
        if (states.extent(0) != dim())
            throw std::invalid_argument("state size must match the dimension of the operators");
        apply_raw(states.data_handle(), new_states.data_handle(), states.extent(1), true);
    }

    template <class MOut, class MIn, class MData>
        requires MdspanOfRank<MOut, 2> && MdspanOfRank<MIn, 2> && MdspanOfRank<MData, 2>
    void apply_weighted(MOut new_states, MIn states, MData data) const
    {
        apply_weighted(std::execution::seq, new_states, states, data);
    }

    // __summed_pauli_op.hpp:378-503: new_states(:,t) += sum_j (sum_k h(j,k) x(k,t)) P_j states(:,t)
    template <execution_policy Policy, class MOut, class MIn, class MData>
        requires MdspanOfRank<MOut, 2> && MdspanOfRank<MIn, 2> && MdspanOfRank<MData, 2>
    void apply_weighted(Policy &&, MOut new_states, MIn states, MData data) const
    {
        if (states.extent(0) != new_states.extent(0) || states.extent(1) != new_states.extent(1))
            throw std::invalid_argument("new_states must have the same dimensions as states");
        if (data.extent(0) != n_operators() || data.extent(1) != states.extent(1))
            throw std::invalid_argument("data(k,t) must have the same number of operators as the SummedPauliOp and "
                                        "the same number of states as the input states");
        if (states.extent(0) != dim())
            throw std::invalid_argument("state size must match the dimension of the operators");
        using D = std::remove_const_t<std::remove_pointer_t<decltype(data.data_handle())>>;
        std::vector<T> x(data.size());
        for (std::size_t i = 0; i < x.size(); ++i)
            x[i] = T(static_cast<D>(data.data_handle()[i]));
        apply_weighted_raw(states.data_handle(), x.data(), new_states.data_handle(), states.extent(1), true);
    }

    // ---- raw entry points (used by the Python surface with fresh outputs) ----------

    void expectation_value_raw(std::complex<T> const *states, std::complex<T> *out, std::size_t n_data,
                               bool accumulate) const
    {
        void const *d_in = hip::stage_in(0, states, dim() * n_data * sizeof(std::complex<T>));
        hip::StagedOut o(1, out, n_operators() * n_data * sizeof(std::complex<T>), accumulate);
        hip::check(fp_hip_summed_expectation(device_ops(), hip::dtype_of<T>(), contraction, d_in, o.dev(), dim(),
                                             n_data, accumulate ? 1 : 0, nullptr));
        o.finish();
    }

    void apply_raw(std::complex<T> const *states, std::complex<T> *out, std::size_t n_data, bool accumulate) const
    {
        std::size_t const bytes = dim() * n_data * sizeof(std::complex<T>);
        void const *d_in = hip::stage_in(0, states, bytes);
        hip::StagedOut o(1, out, bytes, accumulate);
        hip::check(fp_hip_summed_apply(device_ops(), hip::dtype_of<T>(), d_in, o.dev(), dim(), n_data,
                                       accumulate ? 1 : 0, nullptr));
        o.finish();
    }

    void apply_weighted_raw(std::complex<T> const *states, T const *data, std::complex<T> *out, std::size_t n_data,
                            bool accumulate) const
    {
        std::size_t const bytes = dim() * n_data * sizeof(std::complex<T>);
        void const *d_in = hip::stage_in(0, states, bytes);
        void const *d_data = hip::stage_in(2, data, n_operators() * n_data * sizeof(T));
        hip::StagedOut o(1, out, bytes, accumulate);
        hip::check(fp_hip_summed_apply_weighted(device_ops(), hip::dtype_of<T>(), d_in, d_data, o.dev(), dim(),
                                                n_data, accumulate ? 1 : 0, nullptr));
        o.finish();
    }

    // ---- host-side helpers (__summed_pauli_op.hpp:197-268, 621-665) -------------------

    std::vector<PauliOp<T>> split() const
    {
        std::vector<PauliOp<T>> ops;
        ops.reserve(n_operators());
        for (std::size_t k = 0; k < n_operators(); ++k)
        {
            std::vector<std::complex<T>> c(n_pauli_strings());
            for (std::size_t i = 0; i < n_pauli_strings(); ++i)
                c[i] = coeffs(i, k);
            ops.emplace_back(std::move(c), pauli_strings);
        }
        return ops;
    }

    // A_k^2 = sum_{i,j} h_ik h_jk P_i P_j, collected over all strings of weight
    // <= min(n, 2 * max weight) in calculate_pauli_strings_max_weight order
    // (__summed_pauli_op.hpp:197-268).  The pair products and the contraction run on
    // the device (fp_hip_summed_square), IEEE-equal to the reference's sequential loop.
    SummedPauliOp square() const
    {
        std::size_t max_w = 0;
        for (auto const &p : pauli_strings)
            max_w = std::max<std::size_t>(max_w, p.weight);
        std::vector<PauliString> sq = calculate_pauli_strings_max_weight(n_qubits(), std::min(n_qubits(), 2 * max_w));
        std::size_t const K = n_operators();
        std::size_t count = 0;
        hip::check(fp_hip_summed_square(device_ops(), hip::dtype_of<T>(), nullptr, &count, nullptr));
        if (count != sq.size())
            throw std::runtime_error("square: the device and the host enumerate different string counts");
        std::vector<std::complex<T>> out(count * K);
        if (!out.empty())
        {
            hip::Buffer d(out.size() * sizeof(std::complex<T>));
            hip::check(fp_hip_summed_square(device_ops(), hip::dtype_of<T>(), d.get(), &count, nullptr));
            hip::d2h(out.data(), d.get(), out.size() * sizeof(std::complex<T>));
        }
        return SummedPauliOp(sq, Tensor<2>(out.data(), count, K));
    }

    // Dense matrices (__summed_pauli_op.hpp:644-665), accumulated: on the host, or
    // on the device when the view is device memory.
    template <class M>
        requires MdspanOfRank<M, 3>
    void to_tensor(M A_k_out) const
    {
        if (!pauli_strings.empty() && hip::is_device_view(A_k_out.data_handle()))
        {
            if (A_k_out.extent(0) != n_operators() || A_k_out.extent(1) != dim() || A_k_out.extent(2) != dim())
                throw std::invalid_argument("Output tensor must be (n_operators, dim, dim)");
            hip::check(
                fp_hip_summed_to_tensor(device_ops(), hip::dtype_of<T>(), A_k_out.data_handle(), dim(), 1, nullptr));
            hip::check(fp_hip_stream_synchronize(nullptr));
            return;
        }
        for (std::size_t k = 0; k < n_operators(); ++k)
            for (std::size_t i = 0; i < n_pauli_strings(); ++i)
            {
                auto [cols, vals] = get_sparse_repr<T>(pauli_strings[i].paulis);
                for (std::size_t j = 0; j < dim(); ++j)
                    A_k_out(k, j, cols[j]) += coeffs(i, k) * vals[j];
            }
    }

    fp_hip_ops *device_ops() const
    {
        if (pauli_strings.empty())
            throw std::invalid_argument("SummedPauliOp has no Pauli strings");
        std::size_t const n = n_qubits(), N = n_pauli_strings(), K = n_operators();
        std::vector<std::uint8_t> codes(N * n);
        for (std::size_t s = 0; s < N; ++s)
            for (std::size_t i = 0; i < n; ++i)
                codes[s * n + i] = pauli_strings[s].paulis[i].code;
        std::vector<double> m(2 * N * K);
        for (std::size_t i = 0; i < N * K; ++i)
        {
            m[2 * i] = double(coeffs_raw[i].real());
            m[2 * i + 1] = double(coeffs_raw[i].imag());
        }
        return cache_->get(codes, m, [&] {
            auto ops = std::make_unique<hip::Ops>(n, N, codes.data(), nullptr);
            ops->set_matrix(K, m.data());
            return ops;
        });
    }

  private:
    template <class M> void adopt(M const &c)
    {
        if (pauli_strings.empty())
            throw std::invalid_argument("SummedPauliOp needs at least one PauliString");
        coeffs_raw.assign(c.data_handle(), c.data_handle() + c.extent(0) * c.extent(1));
        coeffs = Tensor<2>(coeffs_raw.data(), c.extent(0), c.extent(1));
        check_ctor_inputs();
    }

    void check_ctor_inputs() const
    {
        std::size_t const n = pauli_strings.front().n_qubits();
        for (auto const &p : pauli_strings)
            if (p.n_qubits() != n)
                throw std::invalid_argument("All PauliStrings must have the same size");
        if (coeffs.extent(0) != pauli_strings.size())
            throw std::invalid_argument("The number of PauliStrings must match the number of rows in the coeffs matrix");
    }

    std::shared_ptr<hip::OpsCache> cache_ = std::make_shared<hip::OpsCache>();
};

} // namespace fast_pauli
