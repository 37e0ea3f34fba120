// fp_mfma.hpp — real GEMM on the gfx950 matrix cores (MFMA), used for the
// SummedPauliOp contraction out(k,t) = sum_j h(j,k) ev(j,t)
// (fast_pauli/cpp/include/__summed_pauli_op.hpp:579-591), i.e. C = A * B with
// A = h^T (n_operators x n_strings), B = ev (n_strings x n_states).
//
// Three precisions:
//   FP64   v_mfma_f64_16x16x4_f64          (plain / oracle mode)
//   FP32   v_mfma_f32_16x16x4_f32
//   BF16x3 v_mfma_f32_16x16x32_bf16 (gfx950): each FP32 operand is split into
//          hi = bf16(x), lo = bf16(x - hi); C = lo*hi + hi*lo + hi*hi with FP32
//          accumulation (the dropped lo*lo term is ~2^-16 relative).
//
// Tiling: a 256-thread block computes a 64 x 64 tile of C; its 4 waves (2 x 2)
// each own 32 x 32 as 2 x 2 MFMA tiles of 16 x 16.  BK-deep slabs of A and B
// go through double-buffered LDS: the next slab is loaded into registers while
// the current one feeds the MFMAs.  The K range is split across gridDim.z when
// the tile grid alone would not fill the chip; split partials are summed in
// fixed order by a second kernel, so results are deterministic.
//
// Operand layouts of the 16x16xK single-block instructions: lane l holds
// A[l % 16][k0 + l/16] and B[k0 + l/16][l % 16] (the bf16 form: 8 consecutive
// k per lane, k0 + 8*(l/16) + e); the f32 accumulator holds C[4*(l/16) + r][l % 16]
// in register r, the f64 one C[l/16 + 4*r][l % 16].  A and B take their k runs
// from the same LDS offsets, so a product term pairs A and B at one k whatever
// order the instruction sums a lane's eight k in.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdlib>

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
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kBM = 64, kBN = 64;

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

enum Mode
{
    kF64 = 0,
    kF32 = 1,
    kBF16x3 = 2
};

template <int MODE> struct Traits;
template <> struct Traits<kF64>
{
    typedef double T;
    typedef f64x4 Acc;
    static constexpr int BK = 16;
};
template <> struct Traits<kF32>
{
    typedef float T;
    typedef f32x4 Acc;
    static constexpr int BK = 16;
};
template <> struct Traits<kBF16x3>
{
    typedef float T;
    typedef f32x4 Acc;
    static constexpr int BK = 32;
};

// LDS slab: A as [BM][BK] (K contiguous), B as [BN][BK] (K contiguous, i.e.
// transposed) so every lane reads its k-run contiguously.  BF16x3 stores hi and
// lo planes of both, rows padded to 80 bytes: 16-byte aligned runs, and the 16
// rows one read instruction touches start 20 dwords apart, on distinct banks.
template <int MODE> struct Slab
{
    typedef typename Traits<MODE>::T T;
    static constexpr int BK = Traits<MODE>::BK;
    T a[kBM][BK + 1];
    T b[kBN][BK + 1];
};
template <> struct __align__(16) Slab<kBF16x3>
{
    static constexpr int BK = Traits<kBF16x3>::BK;
    uint16_t ah[kBM][BK + 8], al[kBM][BK + 8];
    uint16_t bh[kBN][BK + 8], bl[kBN][BK + 8];
};

template <int MODE>
__global__ __launch_bounds__(256) void k_gemm(typename Traits<MODE>::T const *__restrict__ A,
                                              typename Traits<MODE>::T const *__restrict__ B,
                                              typename Traits<MODE>::T *__restrict__ C, uint64_t M, uint64_t N,
                                              uint64_t K, uint64_t k_per_split)
{
    typedef typename Traits<MODE>::T T;
    typedef typename Traits<MODE>::Acc Acc;
    constexpr int BK = Traits<MODE>::BK;
    constexpr int A_PER = kBM * BK / 256; // elements of A (and of B) per thread per slab
    __shared__ Slab<MODE> slab[2];

    uint32_t const tid = threadIdx.x;
    uint32_t const lane = tid & 63u;
    uint32_t const wave = tid >> 6;
    uint32_t const wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    uint64_t const m0 = uint64_t(blockIdx.y) * kBM, n0 = uint64_t(blockIdx.x) * kBN;
    uint64_t const kb = uint64_t(blockIdx.z) * k_per_split;
    uint64_t const ke = kb + k_per_split < K ? kb + k_per_split : K;

    Acc acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
            acc[i][j] = Acc{0, 0, 0, 0};

    // per-thread global -> register prefetch of one slab
    T ra[A_PER], rb[A_PER];
    auto fetch = [&](uint64_t k0) {
#pragma unroll
        for (int e = 0; e < A_PER; ++e)
        {
            uint32_t const i = tid + 256u * e;
            uint32_t const r = i / BK, c = i % BK; // A: row r (M), column c (K)
            uint64_t const gm = m0 + r, gk = k0 + c;
            ra[e] = (gm < M && gk < ke) ? A[gm * K + gk] : T(0);
            uint32_t const br = i / kBN, bc = i % kBN; // B: row br (K), column bc (N)
            uint64_t const gk2 = k0 + br, gn = n0 + bc;
            rb[e] = (gk2 < ke && gn < N) ? B[gk2 * N + gn] : T(0);
        }
    };
    auto stage = [&](Slab<MODE> &s) {
#pragma unroll
        for (int e = 0; e < A_PER; ++e)
        {
            uint32_t const i = tid + 256u * e;
            uint32_t const r = i / BK, c = i % BK;
            uint32_t const br = i / kBN, bc = i % kBN;
            if constexpr (MODE == kBF16x3)
            {
                uint16_t const h = bf16_rne(ra[e]);
                s.ah[r][c] = h;
                s.al[r][c] = bf16_rne(ra[e] - bf16_to_f32(h));
                uint16_t const g = bf16_rne(rb[e]);
                s.bh[bc][br] = g;
                s.bl[bc][br] = bf16_rne(rb[e] - bf16_to_f32(g));
            }
            else
            {
                s.a[r][c] = ra[e];
                s.b[bc][br] = rb[e];
            }
        }
    };

    int buf = 0;
    if (kb < ke)
    {
        fetch(kb);
        stage(slab[0]);
    }
    __syncthreads();
    for (uint64_t k0 = kb; k0 < ke; k0 += BK)
    {
        bool const more = k0 + BK < ke;
        if (more)
            fetch(k0 + BK);
        Slab<MODE> const &s = slab[buf];
        if constexpr (MODE == kBF16x3)
        {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 32)
            {
                uint32_t const kq = kk + 8 * (lane >> 4);
                bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
                for (int i = 0; i < 2; ++i)
                {
                    uint32_t const ar = wm + 16 * i + (lane & 15u);
                    uint32_t const bc = wn + 16 * i + (lane & 15u);
                    // rows are 80 bytes and kq is a multiple of 8: 16-byte aligned
                    ah[i] = *reinterpret_cast<bf16x8 const *>(&s.ah[ar][kq]);
                    al[i] = *reinterpret_cast<bf16x8 const *>(&s.al[ar][kq]);
                    bh[i] = *reinterpret_cast<bf16x8 const *>(&s.bh[bc][kq]);
                    bl[i] = *reinterpret_cast<bf16x8 const *>(&s.bl[bc][kq]);
                }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                    {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
                    }
            }
        }
        else
        {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 4)
            {
                T a[2], b[2];
#pragma unroll
                for (int i = 0; i < 2; ++i)
                {
                    a[i] = s.a[wm + 16 * i + (lane & 15u)][kk + (lane >> 4)];
                    b[i] = s.b[wn + 16 * i + (lane & 15u)][kk + (lane >> 4)];
                }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                    {
                        if constexpr (MODE == kF64)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
                        else
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
                    }
            }
        }
        if (more)
            stage(slab[buf ^ 1]);
        __syncthreads();
        buf ^= 1;
    }
    // C (or this split's partial plane)
    T *out = C + uint64_t(blockIdx.z) * M * N;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
            {
                uint32_t const rr = MODE == kF64 ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
                uint64_t const row = m0 + wm + 16 * i + rr;
                uint64_t const col = n0 + wn + 16 * j + (lane & 15u);
                if (row < M && col < N)
                    out[row * N + col] = acc[i][j][r];
            }
}

template <class T>
__global__ __launch_bounds__(256) void k_sum_splits(T const *__restrict__ parts, T *__restrict__ C, uint64_t mn,
                                                    uint32_t splits)
{
    uint64_t const i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= mn)
        return;
    T s = parts[i];
    for (uint32_t p = 1; p < splits; ++p)
        s += parts[uint64_t(p) * mn + i];
    C[i] = s;
}

// Splits of K that bring the block count near `target_blocks`
// (FAST_PAULI_GEMM_BLOCKS, default 512: two per CU), keeping at least 4 slabs per split.
inline uint32_t choose_splits(size_t m, size_t n, size_t k, int bk)
{
    size_t target = 512;
    if (char const *v = std::getenv("FAST_PAULI_GEMM_BLOCKS"))
        if (long const t = std::strtol(v, nullptr, 10); t > 0)
            target = size_t(t);
    size_t const tiles = ((m + kBM - 1) / kBM) * ((n + kBN - 1) / kBN);
    size_t splits = 1;
    while (tiles * splits * 2 <= target && k / (splits * 2) >= size_t(4 * bk))
        splits *= 2;
    return uint32_t(splits);
}

// Scratch bytes needed by gemm() for split partials (0 when it writes C directly).
inline size_t gemm_scratch_bytes(int mode, size_t m, size_t n, size_t k)
{
    int const bk = mode == FP_CONTRACT_BF16X3 ? Traits<kBF16x3>::BK : 16;
    uint32_t const splits = choose_splits(m, n, k, bk);
    if (splits == 1)
        return 0;
    return size_t(splits) * m * n * (mode == FP_CONTRACT_FP64 ? sizeof(double) : sizeof(float));
}

template <int MODE>
fp_status launch_gemm(size_t m, size_t n, size_t k, void const *a, void const *b, void *c, void *scratch,
                      hipStream_t stream)
{
    typedef typename Traits<MODE>::T T;
    constexpr int BK = Traits<MODE>::BK;
    uint32_t const splits = scratch ? choose_splits(m, n, k, BK) : 1;
    uint64_t k_per = (k + splits - 1) / splits;
    k_per = (k_per + BK - 1) / BK * BK;
    dim3 const grid(uint32_t((n + kBN - 1) / kBN), uint32_t((m + kBM - 1) / kBM), splits);
    T *dst = splits > 1 ? static_cast<T *>(scratch) : static_cast<T *>(c);
    hipLaunchKernelGGL((k_gemm<MODE>), grid, dim3(256), 0, stream, static_cast<T const *>(a),
                       static_cast<T const *>(b), dst, uint64_t(m), uint64_t(n), uint64_t(k), k_per);
    if (splits > 1)
    {
        uint64_t const mn = uint64_t(m) * n;
        hipLaunchKernelGGL((k_sum_splits<T>), dim3(uint32_t((mn + 255) / 256)), dim3(256), 0, stream,
                           static_cast<T const *>(scratch), static_cast<T *>(c), mn, splits);
    }
    hipError_t const e = hipGetLastError();
    if (e != hipSuccess)
    {
        g_err = hipGetErrorString(e);
        return FP_ERR_RUNTIME;
    }
    return FP_OK;
}

// C(m x n) = A(m x k) * B(k x n), row-major.  `scratch` (gemm_scratch_bytes) enables
// split-K; NULL computes without splitting.
inline fp_status gemm(int mode, size_t m, size_t n, size_t k, void const *a, void const *b, void *c, void *scratch,
                      hipStream_t stream)
{
    switch (mode)
    {
    case FP_CONTRACT_FP64:
        return launch_gemm<kF64>(m, n, k, a, b, c, scratch, stream);
    case FP_CONTRACT_FP32:
        return launch_gemm<kF32>(m, n, k, a, b, c, scratch, stream);
    case FP_CONTRACT_BF16X3:
        return launch_gemm<kBF16x3>(m, n, k, a, b, c, scratch, stream);
    default:
        g_err = "unknown contraction mode";
        return FP_ERR_INVALID_ARGUMENT;
    }
}

} // namespace fpm
