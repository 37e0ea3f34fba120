// helpers.hpp — enumeration of Pauli strings by weight (host side).
//
// Same functions and output order as the reference
// (fast_pauli/cpp/include/__pauli_helpers.hpp:36-153): non-trivial letter
// words in lexicographic X < Y < Z order, support sets in lexicographic order,
// words outer and supports inner.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "pauli_string.hpp"

namespace fast_pauli
{

inline std::vector<std::string> get_nontrivial_paulis(std::size_t const weight)
{
    if (weight == 0)
        return {};
    std::vector<std::string> words{"X", "Y", "Z"};
    for (std::size_t i = 1; i < weight; ++i)
    {
        std::vector<std::string> next;
        next.reserve(words.size() * 3);
        for (auto const &w : words)
            for (char const c : {'X', 'Y', 'Z'})
                next.push_back(w + c);
        words = std::move(next);
    }
    return words;
}

// All k-subsets of {0..n-1}, each ascending, in lexicographic order.
inline std::vector<std::vector<std::size_t>> idx_combinations(std::size_t const n, std::size_t const k)
{
    std::vector<std::vector<std::size_t>> out;
    if (k > n)
        return out;
    std::vector<std::size_t> c(k);
    for (std::size_t i = 0; i < k; ++i)
        c[i] = i;
    while (true)
    {
        out.push_back(c);
        // advance the rightmost index that can still move
        std::size_t i = k;
        while (i > 0 && c[i - 1] == n - k + (i - 1))
            --i;
        if (i == 0)
            break;
        ++c[i - 1];
        for (std::size_t j = i; j < k; ++j)
            c[j] = c[j - 1] + 1;
    }
    return out;
}

inline std::vector<PauliString> calculate_pauli_strings(std::size_t const n_qubits, std::size_t const weight)
{
    if (weight == 0)
        return {PauliString(std::string(n_qubits, 'I'))};
    auto const words = get_nontrivial_paulis(weight);
    auto const supports = idx_combinations(n_qubits, weight);
    std::vector<PauliString> out;
    out.reserve(words.size() * supports.size());
    for (auto const &w : words)
        for (auto const &sup : supports)
        {
            std::string s(n_qubits, 'I');
            for (std::size_t k = 0; k < sup.size(); ++k)
                s[sup[k]] = w[k];
            out.emplace_back(s);
        }
    return out;
}

inline std::vector<PauliString> calculate_pauli_strings_max_weight(std::size_t n_qubits, std::size_t weight)
{
    std::vector<PauliString> out;
    for (std::size_t w = 0; w <= weight; ++w)
    {
        auto ps = calculate_pauli_strings(n_qubits, w);
        out.insert(out.end(), ps.begin(), ps.end());
    }
    return out;
}

} // namespace fast_pauli
