// pauli.hpp — one-byte Pauli letter with its product table.
//
// Same type and behaviour as the reference's fast_pauli::Pauli
// (fast_pauli/cpp/include/__pauli.hpp:39-212): code 0..3 = I, X, Y, Z;
// construction from a code or a symbol (std::invalid_argument otherwise);
// product with phase; 2x2 dense form.
#pragma once

#include <complex>
#include <concepts>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "mdspan.hpp"

namespace fast_pauli
{

struct Pauli
{
    std::uint8_t code; // 0: I, 1: X, 2: Y, 3: Z

    constexpr Pauli() : code(0)
    {
    }

    template <class T>
        requires(std::integral<T> && !std::same_as<T, char>)
    constexpr Pauli(T const c) : code(static_cast<std::uint8_t>(c))
    {
        if (c < 0 || c > 3)
            throw std::invalid_argument("Pauli code must be 0, 1, 2, or 3");
    }

    constexpr Pauli(char const symbol) : code(0)
    {
        switch (symbol)
        {
        case 'I':
            code = 0;
            break;
        case 'X':
            code = 1;
            break;
        case 'Y':
            code = 2;
            break;
        case 'Z':
            code = 3;
            break;
        default:
            throw std::invalid_argument("Invalid Pauli matrix symbol");
        }
    }

    Pauli(Pauli const &) = default;
    Pauli &operator=(Pauli const &) noexcept = default;
    friend auto operator<=>(Pauli const &, Pauli const &) = default;

    char symbol() const
    {
        return "IXYZ"[code & 3];
    }

    // sigma_a sigma_b = phase * sigma_c.  For distinct non-identity a, b the
    // product is +i sigma_c when (a, b, c) is a cyclic order of (X, Y, Z) and
    // -i sigma_c otherwise; equal letters give I, and I is neutral.
    friend std::pair<std::complex<double>, Pauli> operator*(Pauli const &lhs, Pauli const &rhs)
    {
        std::uint8_t const a = lhs.code, b = rhs.code;
        if (a == 0)
            return {1.0, Pauli(int(b))};
        if (b == 0)
            return {1.0, Pauli(int(a))};
        if (a == b)
            return {1.0, Pauli(0)};
        std::uint8_t const c = std::uint8_t(6 - a - b); // the remaining letter
        bool const cyclic = (b == (a % 3) + 1);           // X->Y, Y->Z, Z->X
        return {cyclic ? std::complex<double>(0, 1) : std::complex<double>(0, -1), Pauli(int(c))};
    }

    template <std::floating_point T, class M>
        requires MdspanOfRank<M, 2>
    void to_tensor(M output) const
    {
        if (output.extent(0) != 2 || output.extent(1) != 2)
            throw std::invalid_argument("Output tensor must have (2, 2) shape");
        using C = std::complex<T>;
        C const I(0, 1);
        C v[4];
        switch (code)
        {
        case 0:
            v[0] = 1, v[1] = 0, v[2] = 0, v[3] = 1;
            break;
        case 1:
            v[0] = 0, v[1] = 1, v[2] = 1, v[3] = 0;
            break;
        case 2:
            v[0] = 0, v[1] = -I, v[2] = I, v[3] = 0;
            break;
        default:
            v[0] = 1, v[1] = 0, v[2] = 0, v[3] = -1;
            break;
        }
        for (int i = 0; i < 4; ++i)
            output.data_handle()[i] = v[i];
    }
};

inline std::string to_string(Pauli const &p)
{
    return std::string(1, p.symbol());
}

} // namespace fast_pauli
