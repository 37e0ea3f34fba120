// fp_algebra.hip — SummedPauliOp::square on the device (C ABI fp_hip_summed_square):
// pair products and their ranks, a stable radix sort by rank, segment starts,
// and the per-(string, operator) contraction (kernels in fp_algebra.hpp).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "fast_pauli_hip.h"
#include "fp_algebra.hpp"
#include "fp_runtime.hpp"

using namespace fprt;

namespace
{

// binom(m, r) for m < 64, r <= W; first[w] = number of strings of weight < w on n
// qubits; pow3[w] = 3^w.  One thread: the tables are a few thousand entries.
__global__ void k_rank_tables(uint32_t n, uint32_t W, uint64_t *binom, uint64_t *first, uint64_t *pow3)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    uint32_t const cols = W + 1;
    for (uint32_t m = 0; m < 64; ++m)
        for (uint32_t r = 0; r <= W; ++r)
            binom[m * cols + r] = r == 0 ? 1 : (m == 0 ? 0 : binom[(m - 1) * cols + r - 1] + binom[(m - 1) * cols + r]);
    uint64_t p = 1, f = 0;
    for (uint32_t w = 0; w <= W; ++w)
    {
        pow3[w] = p;
        first[w] = f;
        f += p * binom[n * cols + w];
        p *= 3;
    }
}

uint64_t binom_host(uint32_t m, uint32_t r)
{
    if (r > m)
        return 0;
    unsigned __int128 v = 1;
    for (uint32_t i = 1; i <= r; ++i)
        v = v * (m - r + i) / i;
    return v > ~uint64_t(0) ? ~uint64_t(0) : uint64_t(v);
}

size_t round256(size_t b)
{
    return (b + 255) / 256 * 256;
}

} // namespace

extern "C" fp_status fp_hip_summed_square(fp_hip_ops const *ops, int dtype, void *out, size_t *n_out_strings,
                                          void *stream)
{
    if (!ops || !n_out_strings)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL operator set or output count");
    if (dtype != FP_C128 && dtype != FP_C64)
        return fail(FP_ERR_INVALID_ARGUMENT, "unknown dtype %d", dtype);
    if (ops->n_operators == 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "SummedPauliOp has no coefficient matrix");
    uint32_t const n = uint32_t(ops->n_qubits);
    uint64_t const N = ops->n_strings, K = ops->n_operators;
    uint32_t w_max = 0;
    for (size_t s = 0; s < N; ++s)
        w_max = std::max<uint32_t>(w_max, uint32_t(__builtin_popcountll(ops->x[s] | ops->z[s])));
    uint32_t const W = std::min<uint32_t>(n, 2 * w_max);
    // A = sum_{w <= W} 3^w C(n, w): the squared operator's string count
    unsigned __int128 total = 0, p3 = 1;
    for (uint32_t w = 0; w <= W; ++w, p3 *= 3)
        total += p3 * binom_host(n, w);
    if (total > 0xffffffffu)
        return fail(FP_ERR_UNSUPPORTED, "the square has more than 2^32 - 1 strings (n = %u, weight <= %u)", n, W);
    uint64_t const A = uint64_t(total);
    *n_out_strings = size_t(A);
    if (!out)
        return FP_OK;
    if (N * N > uint64_t(0x7fffffff))
        return fail(FP_ERR_UNSUPPORTED, "%llu x %llu string pairs exceed one sort (2^31 - 1 items)",
                    static_cast<unsigned long long>(N), static_cast<unsigned long long>(N));
    if ((reinterpret_cast<uintptr_t>(out) & (dtype == FP_C128 ? 15u : 7u)) != 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "out must be aligned to the complex element size");
    uint64_t const n_pairs = N * N;
    hipStream_t const st = static_cast<hipStream_t>(stream);
    StreamGuard guard(st);
    int const end_bit = std::max(1, 64 - __builtin_clzll(A)); // keys < A
    size_t cub_bytes = 0;
    FP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, cub_bytes, static_cast<uint32_t const *>(nullptr),
                                              static_cast<uint32_t *>(nullptr), static_cast<uint64_t const *>(nullptr),
                                              static_cast<uint64_t *>(nullptr), int(n_pairs), 0, end_bit, st));
    size_t const cols = W + 1;
    size_t const o_binom = 0, o_first = round256(64 * cols * 8), o_pow3 = o_first + round256(cols * 8);
    size_t const o_kin = o_pow3 + round256(cols * 8), o_kout = o_kin + round256(n_pairs * 4);
    size_t const o_vin = o_kout + round256(n_pairs * 4), o_vout = o_vin + round256(n_pairs * 8);
    size_t const o_starts = o_vout + round256(n_pairs * 8), o_recs = o_starts + round256((A + 1) * 8);
    size_t const o_cub = o_recs + round256(A * 16);
    void *scratch = nullptr;
    fp_status s = scratch_get(st, o_cub + cub_bytes, &scratch);
    if (s != FP_OK)
        return s;
    auto *base = static_cast<unsigned char *>(scratch);
    auto *binom = reinterpret_cast<uint64_t *>(base + o_binom);
    auto *first = reinterpret_cast<uint64_t *>(base + o_first);
    auto *pow3 = reinterpret_cast<uint64_t *>(base + o_pow3);
    auto *k_in = reinterpret_cast<uint32_t *>(base + o_kin);
    auto *k_out = reinterpret_cast<uint32_t *>(base + o_kout);
    auto *v_in = reinterpret_cast<uint64_t *>(base + o_vin);
    auto *v_out = reinterpret_cast<uint64_t *>(base + o_vout);
    auto *starts = reinterpret_cast<uint64_t *>(base + o_starts);
    auto *recs = reinterpret_cast<ulonglong2 *>(base + o_recs);
    hipLaunchKernelGGL(k_rank_tables, dim3(1), dim3(64), 0, st, n, W, binom, first, pow3);
    FP_LAUNCHED();
    fpalg::RankTables const tab{W, binom, first, pow3};
    hipLaunchKernelGGL(fpalg::k_square_pairs, dim3(uint32_t((n_pairs + 255) / 256)), dim3(256), 0, st, ops->d_x,
                       ops->d_z, uint32_t(N), n, tab, k_in, v_in);
    FP_LAUNCHED();
    FP_HIP(hipcub::DeviceRadixSort::SortPairs(base + o_cub, cub_bytes, k_in, k_out, v_in, v_out, int(n_pairs), 0,
                                              end_bit, st));
    hipLaunchKernelGGL(fpalg::k_segment_starts, dim3(uint32_t((A + 1 + 255) / 256)), dim3(256), 0, st, k_out,
                       n_pairs, A, starts);
    FP_LAUNCHED();
    hipLaunchKernelGGL(fpalg::k_segment_records, dim3(uint32_t((A + 255) / 256)), dim3(256), 0, st, starts, v_out, A,
                       recs);
    FP_LAUNCHED();
    uint64_t const n_work = A * K;
    if (n_work / 256 >= uint64_t(0x7fffffff))
        return fail(FP_ERR_UNSUPPORTED, "the square's coefficient matrix is too large for one launch");
    // a grid of FAST_PAULI_SQUARE_BLOCKS_PER_CU blocks per CU (default 64) whose threads
    // stride over the outputs: the contraction is bound by the latency of its dependent
    // loads, so more waves help (n = 14, 2000 strings, 200 operators: 8 blocks per CU
    // 6.92 ms, 16 6.45, 32 6.16, 64 5.95, 128 5.93)
    static int const cus = [] {
        int dev = 0, n = 256;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            n = prop.multiProcessorCount;
        return n;
    }();
    char const *bpc_env = std::getenv("FAST_PAULI_SQUARE_BLOCKS_PER_CU");
    uint64_t const bpc = bpc_env && std::atoi(bpc_env) > 0 ? uint64_t(std::atoi(bpc_env)) : 64;
    dim3 const grid(uint32_t(std::min<uint64_t>((n_work + 255) / 256, uint64_t(cus) * bpc)));
    bool const small = n_work + uint64_t(grid.x) * 256 <= uint64_t(0xffffffffu);
    typedef fpk::VecOf<double, 1>::type V2;
    typedef fpk::VecOf<float, 1>::type F2;
    auto const *h128 = reinterpret_cast<V2 const *>(ops->d_h128);
    auto const *h64 = reinterpret_cast<F2 const *>(ops->d_h64);
    if (dtype == FP_C128)
    {
        auto const kfn = small ? &fpalg::k_square_contract<double, true> : &fpalg::k_square_contract<double, false>;
        hipLaunchKernelGGL(kfn, grid, dim3(256), 0, st, v_out, recs, h128, K, A, static_cast<V2 *>(out));
    }
    else
    {
        auto const kfn = small ? &fpalg::k_square_contract<float, true> : &fpalg::k_square_contract<float, false>;
        hipLaunchKernelGGL(kfn, grid, dim3(256), 0, st, v_out, recs, h64, K, A, static_cast<F2 *>(out));
    }
    FP_LAUNCHED();
    return FP_OK;
}
