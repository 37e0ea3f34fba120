// fp_mfma.hpp — real GEMM on the gfx950 matrix cores (MFMA), used for the
// SummedPauliOp contraction out(k,t) = sum_j h(j,k) ev(j,t)
// (fast_pauli/cpp/include/__summed_pauli_op.hpp:579-591), i.e. C = A * B with
// A = h^T (n_operators x n_strings), B = ev (n_strings x n_states).
//
// Three precisions:
//   FP64   v_mfma_f64_16x16x4_f64          (plain / oracle mode)
//   FP32   v_mfma_f32_16x16x4_f32
//   BF16x3 v_mfma_f32_16x16x16_bf16: each FP32 operand is split into
//          hi = bf16(x), lo = bf16(x - hi); C = hi*hi + hi*lo + lo*hi with FP32
//          accumulation (the dropped lo*lo term is ~2^-16 relative).
//
// Tiling: a 256-thread block computes a 32 x 32 tile of C as 2 x 2 waves of
// 16 x 16 MFMA tiles; BK-deep slabs of A and B are staged through LDS with
// coalesced loads (zero-filled at the edges).  Operand fragment layout for
// the 16x16xK instructions with one block: lane l holds A[l % 16][k0 + l/16]
// (and B[k0 + l/16][l % 16]); the f32 accumulator holds C[4*(l/16) + r][l % 16]
// in register r, the f64 one C[l/16 + 4*r][l % 16].  For the bf16 instruction
// each lane holds 4 consecutive k.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fast_pauli_hip.h"

namespace fpm
{

inline thread_local char const *g_err = "";
inline char const *last_error()
{
    return g_err;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 32; // block tile (M and N)
constexpr int kPad = 1;

__device__ __forceinline__ uint16_t bf16_rne(float x)
{
    uint32_t const u = __float_as_uint(x);
    uint32_t const r = 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t((u + r) >> 16);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h)
{
    return __uint_as_float(uint32_t(h) << 16);
}

template <class T>
__device__ __forceinline__ T load_or_zero(T const *__restrict__ p, uint64_t r, uint64_t c, uint64_t rows,
                                          uint64_t cols)
{
    return (r < rows && c < cols) ? p[r * cols + c] : T(0);
}

// FP64 / FP32: K-step 4 per instruction.
template <class T, int BK>
__global__ __launch_bounds__(256) void k_gemm_mfma(T const *__restrict__ A, T const *__restrict__ B, T *__restrict__ C,
                                                   uint64_t M, uint64_t N, uint64_t K)
{
    __shared__ T As[kTile][BK + kPad];
    __shared__ T Bs[BK][kTile + kPad];
    uint32_t const tid = threadIdx.x;
    uint32_t const lane = tid & 63u;
    uint32_t const wave = tid >> 6;
    uint32_t const wm = (wave >> 1) * 16, wn = (wave & 1) * 16;
    uint64_t const m0 = uint64_t(blockIdx.y) * kTile, n0 = uint64_t(blockIdx.x) * kTile;
    typedef typename std::conditional<sizeof(T) == 8, f64x4, f32x4>::type Acc;
    Acc acc = {0, 0, 0, 0};
    for (uint64_t k0 = 0; k0 < K; k0 += BK)
    {
        for (uint32_t i = tid; i < kTile * BK; i += 256)
        {
            uint32_t const r = i / BK, c = i % BK;
            As[r][c] = load_or_zero(A, m0 + r, k0 + c, M, K);
        }
        for (uint32_t i = tid; i < BK * kTile; i += 256)
        {
            uint32_t const r = i / kTile, c = i % kTile;
            Bs[r][c] = load_or_zero(B, k0 + r, n0 + c, K, N);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4)
        {
            T const a = As[wm + (lane & 15u)][kk + (lane >> 4)];
            T const b = Bs[kk + (lane >> 4)][wn + (lane & 15u)];
            if constexpr (sizeof(T) == 8)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
            else
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
    {
        // f32: register r holds row 4*(lane/16) + r; f64: row lane/16 + 4*r
        uint64_t const row = m0 + wm + (sizeof(T) == 8 ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r);
        uint64_t const col = n0 + wn + (lane & 15u);
        if (row < M && col < N)
            C[row * N + col] = acc[r];
    }
}

// BF16 split precision, K-step 16 per instruction.
template <int BK>
__global__ __launch_bounds__(256) void k_gemm_bf16x3(float const *__restrict__ A, float const *__restrict__ B,
                                                     float *__restrict__ C, uint64_t M, uint64_t N, uint64_t K)
{
    // hi / lo planes; K contiguous per row for A, per column for B (transposed
    // in LDS so each lane reads 4 consecutive k as one 8-byte load)
    __shared__ uint16_t Ah[kTile][BK + 4], Al[kTile][BK + 4];
    __shared__ uint16_t Bh[kTile][BK + 4], Bl[kTile][BK + 4];
    uint32_t const tid = threadIdx.x;
    uint32_t const lane = tid & 63u;
    uint32_t const wave = tid >> 6;
    uint32_t const wm = (wave >> 1) * 16, wn = (wave & 1) * 16;
    uint64_t const m0 = uint64_t(blockIdx.y) * kTile, n0 = uint64_t(blockIdx.x) * kTile;
    f32x4 acc = {0, 0, 0, 0};
    for (uint64_t k0 = 0; k0 < K; k0 += BK)
    {
        for (uint32_t i = tid; i < kTile * BK; i += 256)
        {
            uint32_t const r = i / BK, c = i % BK;
            float const v = load_or_zero(A, m0 + r, k0 + c, M, K);
            uint16_t const h = bf16_rne(v);
            Ah[r][c] = h;
            Al[r][c] = bf16_rne(v - bf16_to_f32(h));
        }
        for (uint32_t i = tid; i < BK * kTile; i += 256)
        {
            uint32_t const r = i / kTile, c = i % kTile; // r = k, c = n
            float const v = load_or_zero(B, k0 + r, n0 + c, K, N);
            uint16_t const h = bf16_rne(v);
            Bh[c][r] = h;
            Bl[c][r] = bf16_rne(v - bf16_to_f32(h));
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16)
        {
            uint32_t const ar = wm + (lane & 15u), bc = wn + (lane & 15u);
            uint32_t const kb = kk + 4 * (lane >> 4);
            s16x4 ah, al, bh, bl;
#pragma unroll
            for (int e = 0; e < 4; ++e)
            {
                ah[e] = short(Ah[ar][kb + e]);
                al[e] = short(Al[ar][kb + e]);
                bh[e] = short(Bh[bc][kb + e]);
                bl[e] = short(Bl[bc][kb + e]);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
    {
        uint64_t const row = m0 + wm + 4 * (lane >> 4) + r;
        uint64_t const col = n0 + wn + (lane & 15u);
        if (row < M && col < N)
            C[row * N + col] = acc[r];
    }
}

inline fp_status gemm(int mode, size_t m, size_t n, size_t k, void const *a, void const *b, void *c,
                      hipStream_t stream)
{
    dim3 const grid(uint32_t((n + kTile - 1) / kTile), uint32_t((m + kTile - 1) / kTile));
    switch (mode)
    {
    case FP_CONTRACT_FP64:
        hipLaunchKernelGGL((k_gemm_mfma<double, 16>), grid, dim3(256), 0, stream, static_cast<double const *>(a),
                           static_cast<double const *>(b), static_cast<double *>(c), uint64_t(m), uint64_t(n),
                           uint64_t(k));
        break;
    case FP_CONTRACT_FP32:
        hipLaunchKernelGGL((k_gemm_mfma<float, 32>), grid, dim3(256), 0, stream, static_cast<float const *>(a),
                           static_cast<float const *>(b), static_cast<float *>(c), uint64_t(m), uint64_t(n),
                           uint64_t(k));
        break;
    case FP_CONTRACT_BF16X3:
        hipLaunchKernelGGL((k_gemm_bf16x3<32>), grid, dim3(256), 0, stream, static_cast<float const *>(a),
                           static_cast<float const *>(b), static_cast<float *>(c), uint64_t(m), uint64_t(n),
                           uint64_t(k));
        break;
    default:
        g_err = "unknown contraction mode";
        return FP_ERR_INVALID_ARGUMENT;
    }
    hipError_t const e = hipGetLastError();
    if (e != hipSuccess)
    {
        g_err = hipGetErrorString(e);
        return FP_ERR_RUNTIME;
    }
    return FP_OK;
}

} // namespace fpm
