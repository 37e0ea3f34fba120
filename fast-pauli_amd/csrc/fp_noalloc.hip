// fp_noalloc.hip — PauliOp::expectation_value for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).
//
// fp_hip_op_expectation keeps its reduction partials in per-stream scratch and
// its string-group tables in a device copy per operator, both allocated on
// demand.  The variant here runs the same kernels with both of them in one
// static device array per GPU, `scratch`, and allocates no device memory.  In
// host code the array's name only stands in for it: kernels receive the
// device address that hipGetSymbolAddress returns for the current device,
// looked up once per device when the library is initialised on it (or on the
// first call there).
#include "fast_pauli_hip.h"

#include <hip/hip_runtime.h>

#include <mutex>

#include "fp_runtime.hpp"

using namespace fprt;

namespace
{

constexpr size_t kScratchBytes = FP_NOALLOC_SCRATCH_BYTES;

// String-group tables first, reduction partials after them.  Every GPU the
// library runs on holds its own copy until the process ends.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch_ptr[kMaxDevices] = {};

} // namespace

namespace fprt
{

fp_status noalloc_setup()
{
    int dev = 0;
    FP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(FP_ERR_UNSUPPORTED, "device %d is beyond the %d devices the static workspace tracks", dev,
                    kMaxDevices);
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch_ptr[dev])
        return FP_OK;
    void *ptr = nullptr;
    hipError_t const e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(FP_ERR_RUNTIME, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                    hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
    return FP_OK;
}

} // namespace fprt

extern "C" fp_status fp_hip_op_expectation_noalloc(fp_hip_ops const *ops, int dtype, void const *states, void *out,
                                                   size_t dim, size_t n_states, int accumulate, void *stream)
{
    fp_status st = check_shape(ops, dtype, states, out, dim, n_states);
    if (st != FP_OK)
        return st;
    if ((st = noalloc_setup()) != FP_OK)
        return st;
    int dev = 0;
    FP_HIP(hipGetDevice(&dev));
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        base = g_scratch_ptr[dev];
    }
    Workspace const ws{base, kScratchBytes};
    hipStream_t const s = static_cast<hipStream_t>(stream);
    bool tiled = false;
    if ((st = choose_op_expect_tiled(ops, dtype, n_states, &tiled)) != FP_OK)
        return st;
    if (tiled)
        return run_op_expect_tiled(ops, dtype, states, out, dim, n_states, accumulate, s, &ws);
    return run_op_expect(ops, dtype, states, out, dim, n_states, 0, ops->n_strings, ops->d_scale, 1.0, 0.0,
                         accumulate, s, &ws);
}
