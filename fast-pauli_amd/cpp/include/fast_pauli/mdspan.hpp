// mdspan.hpp — a minimal non-owning, row-major (layout_right) multidimensional
// view with the member names of std::mdspan that the operator API uses:
// extent(r), rank(), size(), data_handle(), operator()(i...), operator[](i).
//
// The reference API is written against std::mdspan / std::dextents
// (fast_pauli/cpp/include/__pauli_string.hpp:396-399).  libstdc++ 11 ships no
// <mdspan>, so the host API accepts any view that models MdspanLike (this
// class, or std::mdspan where the standard library provides it).
#pragma once

#include <array>
#include <concepts>
#include <cstddef>
#include <type_traits>

namespace fast_pauli
{

template <class IndexType, std::size_t Rank> struct dextents
{
    static constexpr std::size_t rank()
    {
        return Rank;
    }
    std::array<IndexType, Rank> e{};
};

template <class T, class Extents> class mdspan;

template <class T, class IndexType, std::size_t Rank> class mdspan<T, dextents<IndexType, Rank>>
{
  public:
    using element_type = T;
    using index_type = IndexType;
    using extents_type = dextents<IndexType, Rank>;

    mdspan() = default;

    template <class... I>
        requires(sizeof...(I) == Rank && (std::convertible_to<I, IndexType> && ...))
    mdspan(T *p, I... ext) : ptr_(p), ext_{{static_cast<IndexType>(ext)...}}
    {
    }

    mdspan(T *p, std::array<IndexType, Rank> const &ext) : ptr_(p), ext_{ext}
    {
    }

    // const view from a mutable one
    template <class U>
        requires(std::is_same_v<std::remove_const_t<T>, U> && std::is_const_v<T>)
    mdspan(mdspan<U, dextents<IndexType, Rank>> const &o) : ptr_(o.data_handle())
    {
        for (std::size_t r = 0; r < Rank; ++r)
            ext_.e[r] = o.extent(r);
    }

    static constexpr std::size_t rank()
    {
        return Rank;
    }
    IndexType extent(std::size_t r) const
    {
        return ext_.e[r];
    }
    std::size_t size() const
    {
        std::size_t s = 1;
        for (std::size_t r = 0; r < Rank; ++r)
            s *= ext_.e[r];
        return s;
    }
    T *data_handle() const
    {
        return ptr_;
    }

    template <class... I>
        requires(sizeof...(I) == Rank)
    T &operator()(I... idx) const
    {
        std::array<IndexType, Rank> const ii{{static_cast<IndexType>(idx)...}};
        std::size_t off = 0;
        for (std::size_t r = 0; r < Rank; ++r)
            off = off * ext_.e[r] + ii[r];
        return ptr_[off];
    }

    T &operator[](IndexType i) const
        requires(Rank == 1)
    {
        return ptr_[i];
    }

  private:
    T *ptr_ = nullptr;
    extents_type ext_{};
};

template <class T, class... I> mdspan(T *, I...) -> mdspan<T, dextents<std::size_t, sizeof...(I)>>;

// Anything with the std::mdspan observers the API needs.
template <class M>
concept MdspanLike = requires(M const &m) {
    { m.data_handle() };
    { m.extent(0) } -> std::convertible_to<std::size_t>;
    { M::rank() } -> std::convertible_to<std::size_t>;
};

template <class M, std::size_t R>
concept MdspanOfRank = MdspanLike<M> && (M::rank() == R);

} // namespace fast_pauli
