// device.hpp — the C++ host side of the C ABI (include/fast_pauli_hip.h).
//
// Maps fp_status codes to the exceptions the reference throws
// (std::invalid_argument for shape/letter errors, fast_pauli/cpp/include/
// __pauli_op.hpp:340-351), owns device resources RAII-style, and stages host
// views through device buffers.  A view whose data already lives in device
// memory is used in place, so callers that keep states resident on the GPU
// pay no transfer.
#pragma once

#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "fast_pauli_hip.h"

namespace fast_pauli::hip
{

inline std::string last_error()
{
    char buf[1024];
    fp_hip_last_error(buf, sizeof(buf));
    return std::string(buf);
}

inline void check(fp_status st)
{
    switch (st)
    {
    case FP_OK:
        return;
    case FP_ERR_INVALID_ARGUMENT:
        throw std::invalid_argument(last_error());
    case FP_ERR_OUT_OF_MEMORY:
        throw std::runtime_error("out of device memory: " + last_error());
    default:
        throw std::runtime_error("fast_pauli HIP error " + std::to_string(st) + ": " + last_error());
    }
}

inline void ensure_init()
{
    static std::once_flag once;
    std::call_once(once, [] { check(fp_hip_init(0)); });
}

template <class T> constexpr int dtype_of()
{
    static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>, "T must be double or float");
    return std::is_same_v<T, double> ? FP_C128 : FP_C64;
}

inline bool is_device_pointer(void const *p)
{
    int d = 0;
    check(fp_hip_is_device_pointer(p, &d));
    return d != 0;
}

// Whether p is device memory, without initialising a device first (false where no
// device is present): host-only helpers such as to_tensor stay usable there.
inline bool is_device_view(void const *p)
{
    int d = 0;
    return fp_hip_is_device_pointer(p, &d) == FP_OK && d != 0;
}

// Owning device allocation.
class Buffer
{
  public:
    Buffer() = default;
    explicit Buffer(std::size_t bytes)
    {
        ensure_init();
        check(fp_hip_malloc(&ptr_, bytes));
        bytes_ = bytes;
    }
    Buffer(Buffer const &) = delete;
    Buffer &operator=(Buffer const &) = delete;
    Buffer(Buffer &&o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_)
    {
        o.ptr_ = nullptr;
        o.bytes_ = 0;
    }
    Buffer &operator=(Buffer &&o) noexcept
    {
        if (this != &o)
        {
            reset();
            ptr_ = o.ptr_;
            bytes_ = o.bytes_;
            o.ptr_ = nullptr;
            o.bytes_ = 0;
        }
        return *this;
    }
    ~Buffer()
    {
        reset();
    }
    void reset()
    {
        if (ptr_)
            (void)fp_hip_free(ptr_);
        ptr_ = nullptr;
        bytes_ = 0;
    }
    void *get() const
    {
        return ptr_;
    }
    std::size_t bytes() const
    {
        return bytes_;
    }

  private:
    void *ptr_ = nullptr;
    std::size_t bytes_ = 0;
};

// Per-thread staging buffers, grown on demand and reused across calls.
inline void *staging(int slot, std::size_t bytes)
{
    thread_local std::vector<Buffer> pool(4);
    Buffer &b = pool.at(std::size_t(slot));
    if (b.bytes() < bytes)
    {
        b.reset();
        b = Buffer(bytes);
    }
    return b.get();
}

inline void h2d(void *dst, void const *src, std::size_t bytes)
{
    check(fp_hip_memcpy(dst, src, bytes, FP_H2D, nullptr));
}
inline void d2h(void *dst, void const *src, std::size_t bytes)
{
    check(fp_hip_memcpy(dst, src, bytes, FP_D2H, nullptr));
}

// A read-only operand: the device address of a host or device array.
inline void const *stage_in(int slot, void const *p, std::size_t bytes)
{
    ensure_init();
    if (is_device_pointer(p))
        return p;
    void *d = staging(slot, bytes);
    h2d(d, p, bytes);
    return d;
}

// An output operand.  `load` copies the current contents in (for += semantics).
class StagedOut
{
  public:
    StagedOut(int slot, void *p, std::size_t bytes, bool load) : host_(p), bytes_(bytes)
    {
        ensure_init();
        if (is_device_pointer(p))
        {
            dev_ = p;
            host_ = nullptr;
        }
        else
        {
            dev_ = staging(slot, bytes);
            if (load)
                h2d(dev_, host_, bytes_);
        }
    }
    void *dev() const
    {
        return dev_;
    }
    void finish()
    {
        if (host_)
            d2h(host_, dev_, bytes_);
        else
            check(fp_hip_stream_synchronize(nullptr));
    }

  private:
    void *host_;
    void *dev_;
    std::size_t bytes_;
};

// Device-resident operator set.
class Ops
{
  public:
    Ops(std::size_t n_qubits, std::size_t n_strings, std::uint8_t const *codes, double const *coeffs)
    {
        ensure_init();
        check(fp_hip_ops_create(&h_, n_qubits, n_strings, codes, coeffs));
    }
    Ops(Ops const &) = delete;
    Ops &operator=(Ops const &) = delete;
    ~Ops()
    {
        if (h_)
            (void)fp_hip_ops_destroy(h_);
    }
    void set_matrix(std::size_t n_operators, double const *coeffs_nk)
    {
        check(fp_hip_ops_set_matrix(h_, n_operators, coeffs_nk));
    }
    fp_hip_ops *get() const
    {
        return h_;
    }

  private:
    fp_hip_ops *h_ = nullptr;
};

// Caches the device copy of an operator; rebuilt whenever the host-side
// letters or coefficients differ from the ones it was built from.
class OpsCache
{
  public:
    template <class Build>
    fp_hip_ops *get(std::vector<std::uint8_t> const &codes, std::vector<double> const &coeffs, Build &&build)
    {
        std::lock_guard<std::mutex> lock(mu_);
        if (!ops_ || codes != codes_ || coeffs != coeffs_)
        {
            ops_ = build();
            codes_ = codes;
            coeffs_ = coeffs;
        }
        return ops_->get();
    }

  private:
    std::mutex mu_;
    std::unique_ptr<Ops> ops_;
    std::vector<std::uint8_t> codes_;
    std::vector<double> coeffs_;
};

} // namespace fast_pauli::hip
