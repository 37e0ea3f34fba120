// pauli_string.hpp — PauliString with its hot-path methods on the GPU.
//
// Keeps the reference's type and method signatures
// (fast_pauli/cpp/include/__pauli_string.hpp:126-561): members `weight` and
// `paulis`; apply (1-D), apply_batch (2-D, states as columns of a row-major
// (dim x n_states) array) and expectation_value, each with and without an
// execution policy, accumulating (+=) into the caller's output.  Every policy
// runs on the GPU through the C ABI; results are independent of the policy.
#pragma once

#include <complex>
#include <cstdint>
#include <execution>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "device.hpp"
#include "mdspan.hpp"
#include "pauli.hpp"

namespace fast_pauli
{

template <class P>
concept execution_policy = std::is_execution_policy_v<std::remove_cvref_t<P>>;

namespace detail
{
template <class M> struct complex_of
{
    using element = std::remove_const_t<typename std::remove_pointer_t<decltype(std::declval<M>().data_handle())>>;
    using value = typename element::value_type;
};
template <class M> using real_of_t = typename complex_of<M>::value;

// Masks of a letter sequence: character i acts on bit (n-1-i) of the row index
// (__pauli_string.hpp:52-54).
inline void masks_of(std::vector<Pauli> const &paulis, std::uint64_t &x, std::uint64_t &z)
{
    std::size_t const n = paulis.size();
    x = z = 0;
    for (std::size_t i = 0; i < n; ++i)
    {
        std::uint64_t const bit = std::uint64_t(1) << (n - 1 - i);
        std::uint8_t const c = paulis[i].code;
        if (c == 1 || c == 2)
            x |= bit;
        if (c == 2 || c == 3)
            z |= bit;
    }
}

// --- raw device calls shared by PauliString / PauliOp / SummedPauliOp ---
template <class T>
void string_apply_batch(fp_hip_ops *ops, std::size_t idx, std::complex<T> const *in, std::complex<T> *out,
                        std::size_t dim, std::size_t n_states, std::complex<T> c, bool accumulate)
{
    std::size_t const bytes = dim * n_states * sizeof(std::complex<T>);
    void const *d_in = hip::stage_in(0, in, bytes);
    hip::StagedOut o(1, out, bytes, accumulate);
    hip::check(fp_hip_string_apply_batch(ops, idx, hip::dtype_of<T>(), d_in, o.dev(), dim, n_states, double(c.real()),
                                         double(c.imag()), accumulate ? 1 : 0, nullptr));
    o.finish();
}

template <class T>
void string_expectation(fp_hip_ops *ops, std::size_t idx, std::complex<T> const *states, std::complex<T> *out,
                        std::size_t dim, std::size_t n_states, std::complex<T> c, bool accumulate)
{
    void const *d_in = hip::stage_in(0, states, dim * n_states * sizeof(std::complex<T>));
    hip::StagedOut o(1, out, n_states * sizeof(std::complex<T>), accumulate);
    hip::check(fp_hip_string_expectation(ops, idx, hip::dtype_of<T>(), d_in, o.dev(), dim, n_states,
                                         double(c.real()), double(c.imag()), accumulate ? 1 : 0, nullptr));
    o.finish();
}
} // namespace detail

struct PauliString
{
    std::uint8_t weight = 0;
    std::vector<Pauli> paulis;

    PauliString() noexcept = default;

    PauliString(std::vector<Pauli> ps) : weight(0), paulis(std::move(ps))
    {
        for (auto const &p : paulis)
            weight += p.code > 0;
    }

    PauliString(std::string const &str) : weight(0)
    {
        paulis.reserve(str.size());
        for (char const ch : str)
        {
            switch (ch)
            {
            case 'I':
            case 'X':
            case 'Y':
            case 'Z':
                paulis.emplace_back(ch);
                weight += ch != 'I';
                break;
            default:
                throw std::invalid_argument(std::string("Invalid Pauli character ") + ch);
            }
        }
    }

    PauliString(char const *str) : PauliString(std::string(str))
    {
    }

    PauliString(PauliString const &o) : weight(o.weight), paulis(o.paulis)
    {
    }
    PauliString &operator=(PauliString const &o)
    {
        weight = o.weight;
        paulis = o.paulis;
        return *this;
    }
    PauliString(PauliString &&) noexcept = default;
    PauliString &operator=(PauliString &&) noexcept = default;

    friend bool operator==(PauliString const &a, PauliString const &b)
    {
        return a.weight == b.weight && a.paulis == b.paulis;
    }
    friend auto operator<=>(PauliString const &a, PauliString const &b)
    {
        if (auto c = a.weight <=> b.weight; c != 0)
            return c;
        return a.paulis <=> b.paulis;
    }

    // Product of two strings letter by letter, with the accumulated phase
    // (__pauli_string.hpp:228-247).
    friend std::pair<std::complex<double>, PauliString> operator*(PauliString const &lhs, PauliString const &rhs)
    {
        if (lhs.dim() != rhs.dim())
            throw std::invalid_argument("PauliStrings must have the same size");
        std::complex<double> phase = 1;
        std::vector<Pauli> out;
        out.reserve(lhs.n_qubits());
        for (std::size_t i = 0; i < lhs.n_qubits(); ++i)
        {
            auto [ph, p] = lhs.paulis[i] * rhs.paulis[i];
            phase *= ph;
            out.push_back(p);
        }
        return {phase, PauliString(std::move(out))};
    }

    std::size_t n_qubits() const noexcept
    {
        return paulis.size();
    }
    std::size_t dim() const noexcept
    {
        return paulis.empty() ? 0 : std::size_t(1) << paulis.size();
    }

    std::string str() const
    {
        std::string s;
        s.reserve(paulis.size());
        for (auto const &p : paulis)
            s.push_back(p.symbol());
        return s;
    }

    std::vector<std::uint8_t> codes() const
    {
        std::vector<std::uint8_t> c(paulis.size());
        for (std::size_t i = 0; i < paulis.size(); ++i)
            c[i] = paulis[i].code;
        return c;
    }

    // ---- hot path -----------------------------------------------------------

    template <class MOut, class MIn>
        requires MdspanOfRank<MOut, 1> && MdspanOfRank<MIn, 1>
    void apply(MOut new_states, MIn states, std::complex<detail::real_of_t<MIn>> c = 1.0) const
    {
        apply(std::execution::seq, new_states, states, c);
    }

    // __pauli_string.hpp:312-341: new_states[i] += c * P states
    template <execution_policy Policy, class MOut, class MIn>
        requires MdspanOfRank<MOut, 1> && MdspanOfRank<MIn, 1>
    void apply(Policy &&, MOut new_states, MIn states, std::complex<detail::real_of_t<MIn>> c = 1.0) const
    {
        if (states.extent(0) != dim())
            throw std::invalid_argument("Input vector size must match the number of qubits");
        if (states.extent(0) != new_states.extent(0))
            throw std::invalid_argument("new_states must have the same dimensions as states");
        using T = detail::real_of_t<MIn>;
        detail::string_apply_batch<T>(device_ops(), 0, states.data_handle(), new_states.data_handle(), dim(), 1, c,
                                      true);
    }

    template <class MOut, class MIn>
        requires MdspanOfRank<MOut, 2> && MdspanOfRank<MIn, 2>
    void apply_batch(MOut new_states_T, MIn const states_T, std::complex<detail::real_of_t<MIn>> const c) const
    {
        apply_batch(std::execution::seq, new_states_T, states_T, c);
    }

    // __pauli_string.hpp:396-436: new_states_T(i, t) += c * m_i * states_T(k_i, t)
    template <execution_policy Policy, class MOut, class MIn>
        requires MdspanOfRank<MOut, 2> && MdspanOfRank<MIn, 2>
    void apply_batch(Policy &&, MOut new_states_T, MIn const states_T,
                     std::complex<detail::real_of_t<MIn>> const c) const
    {
        check_apply_batch(new_states_T, states_T);
        using T = detail::real_of_t<MIn>;
        detail::string_apply_batch<T>(device_ops(), 0, states_T.data_handle(), new_states_T.data_handle(), dim(),
                                      states_T.extent(1), c, true);
    }

    template <class MOut, class MIn>
        requires MdspanOfRank<MOut, 1> && MdspanOfRank<MIn, 2>
    void expectation_value(MOut expectation_vals_out, MIn states,
                           std::complex<detail::real_of_t<MIn>> const c = 1.0) const
    {
        expectation_value(std::execution::seq, expectation_vals_out, states, c);
    }

    // __pauli_string.hpp:488-538: out[t] += c <psi_t|P|psi_t>
    template <execution_policy Policy, class MOut, class MIn>
        requires MdspanOfRank<MOut, 1> && MdspanOfRank<MIn, 2>
    void expectation_value(Policy &&, MOut expectation_vals_out, MIn states,
                           std::complex<detail::real_of_t<MIn>> const c = 1.0) const
    {
        if (states.extent(0) != dim())
            throw std::invalid_argument("[PauliString] states shape (" + std::to_string(states.extent(0)) +
                                        ") must match the dimension of the operators (" + std::to_string(dim()) +
                                        ")");
        if (expectation_vals_out.extent(0) != states.extent(1))
            throw std::invalid_argument("[PauliString] expectation_vals_out shape must match the number of states");
        using T = detail::real_of_t<MIn>;
        detail::string_expectation<T>(device_ops(), 0, states.data_handle(), expectation_vals_out.data_handle(), dim(),
                                      states.extent(1), c, true);
    }

    // Dense matrix (debugging helper, __pauli_string.hpp:549-560): on the host, or
    // on the device when the view is device memory.
    template <std::floating_point T, class M>
        requires MdspanOfRank<M, 2>
    void to_tensor(M output) const
    {
        if (output.extent(0) != dim() || output.extent(1) != dim())
            throw std::invalid_argument("Output tensor must have the same dimensions as the PauliString");
        if (hip::is_device_view(output.data_handle()))
        {
            hip::check(fp_hip_string_to_tensor(device_ops(), 0, hip::dtype_of<T>(), output.data_handle(), dim(), nullptr));
            hip::check(fp_hip_stream_synchronize(nullptr));
            return;
        }
        std::uint64_t x, z;
        detail::masks_of(paulis, x, z);
        std::size_t ny = 0;
        for (auto const &p : paulis)
            ny += p.code == 2;
        std::complex<T> const base[4] = {{1, 0}, {0, -1}, {-1, 0}, {0, 1}}; // (-i)^k
        for (std::size_t j = 0; j < dim(); ++j)
        {
            std::complex<T> m = base[ny & 3];
            if (__builtin_popcountll(std::uint64_t(j) & z) & 1)
                m = -m;
            output(j, std::size_t(std::uint64_t(j) ^ x)) = m;
        }
    }

    // The device copy of this single string (coefficient 1).
    fp_hip_ops *device_ops() const
    {
        if (paulis.empty())
            throw std::invalid_argument("PauliString is empty");
        auto const c = codes();
        static std::vector<double> const one{1.0, 0.0};
        return cache_->get(c, one, [&] { return std::make_unique<hip::Ops>(c.size(), 1, c.data(), one.data()); });
    }

  private:
    template <class MOut, class MIn> void check_apply_batch(MOut const &out, MIn const &in) const
    {
        if (in.extent(0) != dim())
            throw std::invalid_argument("[PauliString] states shape (" + std::to_string(in.extent(0)) +
                                        ") must match the dimension of the operators (" + std::to_string(dim()) +
                                        ")");
        if (in.extent(0) != out.extent(0) || in.extent(1) != out.extent(1))
            throw std::invalid_argument("[PauliString] new_states must have the same dimensions as states");
    }

    std::shared_ptr<hip::OpsCache> cache_ = std::make_shared<hip::OpsCache>();
};

inline std::string to_string(PauliString const &ps)
{
    return ps.str();
}

// get_sparse_repr (__pauli_string.hpp:49-118): k[j] = j ^ x,
// m[j] = (-i)^{nY} (-1)^{popcount(j & z)}; computed from the closed form.
template <std::floating_point T>
std::tuple<std::vector<std::size_t>, std::vector<std::complex<T>>> get_sparse_repr(std::vector<Pauli> const &paulis)
{
    std::size_t const n = paulis.size();
    if (n == 0)
        return {};
    std::size_t const d = std::size_t(1) << n;
    std::uint64_t x, z;
    detail::masks_of(paulis, x, z);
    std::size_t ny = 0;
    for (auto const &p : paulis)
        ny += p.code == 2;
    std::complex<T> const base[4] = {{1, 0}, {0, -1}, {-1, 0}, {0, 1}};
    std::vector<std::size_t> k(d);
    std::vector<std::complex<T>> m(d);
    for (std::size_t j = 0; j < d; ++j)
    {
        k[j] = std::size_t(std::uint64_t(j) ^ x);
        m[j] = (__builtin_popcountll(std::uint64_t(j) & z) & 1) ? -base[ny & 3] : base[ny & 3];
    }
    return {std::move(k), std::move(m)};
}

} // namespace fast_pauli

template <> struct std::hash<fast_pauli::PauliString>
{
    std::size_t operator()(fast_pauli::PauliString const &key) const
    {
        return std::hash<std::string>()(key.str());
    }
};
