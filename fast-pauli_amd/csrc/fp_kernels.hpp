// fp_kernels.hpp — gfx950 kernels of the operator-application hot path.
//
// Layout: a batch of states is row-major (dim x n_states).  A kernel moves it
// as 16-byte (or 8-byte) "vectors": one complex128, two complex64 of the same
// row (NC = 2) or one complex64 (NC = 1, odd n_states).  vpr = vectors per row.
//
// A Pauli string is carried as bit masks: x has bit b set where the letter on
// bit b is X or Y, z where it is Y or Z, nY = popcount(x & z).  Row j of P
// has its one non-zero in column j ^ x with value
//     m_j = (-i)^{nY} * (-1)^{popcount(j & z)}
// which is the closed form of the PauliComposer construction the reference
// materialises per call (fast_pauli/cpp/include/__pauli_string.hpp:49-118).
// The host folds (-i)^{nY} into the coefficient (cf = c * (-i)^{nY}, an
// exact swap/negation), so a kernel only needs the parity of popcount(j & z).
//
// Apply kernels avoid fused multiply-add so that c*m_j*psi rounds exactly as
// the reference's std::complex arithmetic does: the products are bit-exact.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fpk
{

constexpr int kBlock = 256;

template <class T, int NC> struct VecOf;
template <> struct VecOf<double, 1>
{
    typedef double type __attribute__((ext_vector_type(2)));
};
template <> struct VecOf<float, 2>
{
    typedef float type __attribute__((ext_vector_type(4)));
};
template <> struct VecOf<float, 1>
{
    typedef float type __attribute__((ext_vector_type(2)));
};

// Accumulator used for per-thread row sums of expectation values: FP64 for
// complex128 states, FP32 for complex64 states.
template <class T> struct RowAcc;
template <> struct RowAcc<double>
{
    typedef double type;
};
template <> struct RowAcc<float>
{
    typedef float type;
};

__device__ __forceinline__ uint64_t insert_zero_bit(uint64_t p, uint32_t h)
{
    uint64_t const low = p & ((uint64_t(1) << h) - 1);
    return ((p >> h) << (h + 1)) | low;
}

__device__ __forceinline__ uint32_t parity64(uint64_t v)
{
    return __popcll(v) & 1u;
}

// Row index of a vector for a non power-of-two row length: a double-precision
// reciprocal estimate corrected by one compare each way (v < 2^53).
__device__ __forceinline__ uint64_t row_of(uint64_t v, uint64_t vpr, double inv_vpr)
{
    uint64_t r = uint64_t(double(v) * inv_vpr);
    uint64_t rv = r * vpr;
    if (rv > v)
    {
        --r;
    }
    else if (rv + vpr <= v)
    {
        ++r;
    }
    return r;
}

// out = (a + bi) * (pr + pi i), rounded like the reference: two products per
// component and one add, no contraction.
template <class T> __device__ __forceinline__ void cmul_exact(T a, T b, T pr, T pi, T &orr, T &oi)
{
#pragma clang fp contract(off)
    T const t0 = a * pr;
    T const t1 = b * pi;
    T const t2 = a * pi;
    T const t3 = b * pr;
    orr = t0 - t1;
    oi = t2 + t3;
}

template <class T> __device__ __forceinline__ T add_exact(T a, T b)
{
#pragma clang fp contract(off)
    return a + b;
}

// ---------------------------------------------------------------------------
// K1: single-string apply.  out(j,:) (+)= cf * (-1)^{popcount(j&z)} * in(j^x,:)
// Each thread moves U vectors spaced one block apart (the launch uses U = 1);
// reads of the partner row j^x stay coalesced because the row is contiguous
// along the states axis.
// ---------------------------------------------------------------------------
template <class T, int NC, int U, bool POW2, bool ACC>
__global__ __launch_bounds__(kBlock) void k_string_apply(typename VecOf<T, NC>::type const *__restrict__ in,
                                                         typename VecOf<T, NC>::type *__restrict__ out,
                                                         uint64_t v_begin, uint64_t n_vec, uint32_t vpr_shift,
                                                         uint64_t vpr, double inv_vpr, uint64_t x, uint64_t z, T cr,
                                                         T ci)
{
    typedef typename VecOf<T, NC>::type V;
    uint64_t const base = v_begin + uint64_t(blockIdx.x) * (kBlock * U) + threadIdx.x;
    V src[U];
    uint32_t sgn[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
        uint64_t const v = base + uint64_t(u) * kBlock;
        if (v < n_vec)
        {
            uint64_t row, sv;
            if constexpr (POW2)
            {
                row = v >> vpr_shift;
                sv = v ^ (x << vpr_shift);
            }
            else
            {
                row = row_of(v, vpr, inv_vpr);
                sv = (row ^ x) * vpr + (v - row * vpr);
            }
            src[u] = __builtin_nontemporal_load(in + sv);
            sgn[u] = parity64(row & z);
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
        uint64_t const v = base + uint64_t(u) * kBlock;
        if (v < n_vec)
        {
            T const a = sgn[u] ? -cr : cr;
            T const b = sgn[u] ? -ci : ci;
            V r;
#pragma unroll
            for (int c = 0; c < NC; ++c)
            {
                T orr, oi;
                cmul_exact<T>(a, b, src[u][2 * c], src[u][2 * c + 1], orr, oi);
                r[2 * c] = orr;
                r[2 * c + 1] = oi;
            }
            if constexpr (ACC)
            {
                V const o = out[v];
#pragma unroll
                for (int c = 0; c < 2 * NC; ++c)
                    r[c] = add_exact<T>(o[c], r[c]);
            }
            __builtin_nontemporal_store(r, out + v);
        }
    }
}

// ---------------------------------------------------------------------------
// K3: multi-string apply (PauliOp::apply).  Each output vector accumulates
// sum_s cf_s (-1)^{popcount(j&z_s)} in(j^x_s,:) in registers, strings in order
// s = 0..N-1 (the reference's sequential order), then is written once.
// String masks are wave-uniform (scalar loads); no scratch copies of the
// output (the reference's OpenMP form keeps n_threads of them,
// fast_pauli/cpp/include/__pauli_op.hpp:427-429).
// ---------------------------------------------------------------------------
template <class T, int NC, int U, bool POW2, bool ACC>
__global__ __launch_bounds__(kBlock) void k_op_apply(typename VecOf<T, NC>::type const *__restrict__ in,
                                                     typename VecOf<T, NC>::type *__restrict__ out, uint64_t v_begin,
                                                     uint64_t n_vec, uint32_t vpr_shift, uint64_t vpr, double inv_vpr,
                                                     uint64_t const *__restrict__ xs,
                                                     uint64_t const *__restrict__ zs, T const *__restrict__ cf,
                                                     uint32_t n_strings)
{
    typedef typename VecOf<T, NC>::type V;
    uint64_t const base = v_begin + uint64_t(blockIdx.x) * (kBlock * U) + threadIdx.x;
    uint64_t row[U], col[U];
    V acc[U];
    bool live[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
        uint64_t const v = base + uint64_t(u) * kBlock;
        live[u] = v < n_vec;
        uint64_t const vv = live[u] ? v : 0;
        if constexpr (POW2)
        {
            row[u] = vv >> vpr_shift;
            col[u] = vv & ((uint64_t(1) << vpr_shift) - 1);
        }
        else
        {
            row[u] = row_of(vv, vpr, inv_vpr);
            col[u] = vv - row[u] * vpr;
        }
        if constexpr (ACC)
        {
            if (live[u])
                acc[u] = out[v];
            else
                acc[u] = V(0);
        }
        else
        {
            acc[u] = V(0);
        }
    }
    for (uint32_t s = 0; s < n_strings; ++s)
    {
        uint64_t const x = xs[s];
        uint64_t const z = zs[s];
        T const cr = cf[2 * s];
        T const ci = cf[2 * s + 1];
        V src[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            uint64_t const sv = POW2 ? (((row[u] ^ x) << vpr_shift) | col[u]) : ((row[u] ^ x) * vpr + col[u]);
            src[u] = live[u] ? in[sv] : V(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            uint32_t const sg = parity64(row[u] & z);
            T const a = sg ? -cr : cr;
            T const b = sg ? -ci : ci;
#pragma unroll
            for (int c = 0; c < NC; ++c)
            {
                T orr, oi;
                cmul_exact<T>(a, b, src[u][2 * c], src[u][2 * c + 1], orr, oi);
                acc[u][2 * c] = add_exact<T>(acc[u][2 * c], orr);
                acc[u][2 * c + 1] = add_exact<T>(acc[u][2 * c + 1], oi);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
        uint64_t const v = base + uint64_t(u) * kBlock;
        if (live[u])
            __builtin_nontemporal_store(acc[u], out + v);
    }
}

// ---------------------------------------------------------------------------
// K2/K4: expectation values.
//
// Rows are visited in pairs (j, j^x) so each string reads the state once.
// With h the highest set bit of x, leader rows have bit h clear; pair p maps
// to leader j = insert_zero_bit(p, h).  Because m_{j^x} = (-1)^{nY} m_j, the
// pair's contribution is
//     m_j w + m_{j^x} conj(w),  w = conj(psi_j) psi_{j^x}
//   = (-1)^{popcount(j&z)} * kappa * (Re w  if nY even,  Im w  if nY odd)
// with kappa = 2(-1)^{nY/2} (nY even), +2 (nY = 1 mod 4), -2 (nY = 3 mod 4).
// A string with x = 0 (mode SELF) pairs rows (2p, 2p+1) and adds
// sigma_j |psi_j|^2 for both, kappa = 1.  The expectation value of a Pauli
// string is therefore a real sum; kappa and the coefficient are applied once
// per string on the block's row total.
//
// Thread layout: tid = ty * ct + tx; tx walks the vectors of a row (column
// tile blockIdx.y), ty walks pairs.  Strings [blockIdx.z*spb, ...) are looped
// inside the block.  Row sums use RowAcc<T> (FP32 for complex64); the
// per-string scaled totals and the cross-block reduction are FP64.
// ---------------------------------------------------------------------------
enum : uint32_t
{
    kModeSelf = 0,
    kModeRe = 1,
    kModeIm = 2
};

struct StringMeta
{
    uint64_t x;
    uint64_t z;
    uint32_t hbit; // highest set bit of x (0 for x == 0)
    uint32_t mode; // kModeSelf / kModeRe / kModeIm
};

template <class T, int NC, int U>
__device__ __forceinline__ void ev_rows(typename VecOf<T, NC>::type const *__restrict__ states, uint64_t vpr,
                                        uint64_t col, bool col_live, uint64_t p0, uint64_t p1, uint32_t ty,
                                        uint32_t rt, StringMeta const &sm, typename RowAcc<T>::type (&R)[NC])
{
    typedef typename VecOf<T, NC>::type V;
    typedef typename RowAcc<T>::type A;
#pragma unroll
    for (int c = 0; c < NC; ++c)
        R[c] = A(0);
    if (!col_live)
        return;
    uint64_t const xp = sm.x ? sm.x : 1;
    A const self_sign2 = (sm.z & 1) ? A(-1) : A(1);
    for (uint64_t pb = p0 + ty; pb < p1; pb += uint64_t(rt) * U)
    {
        V a[U], b[U];
        uint64_t jj[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            uint64_t const p = pb + uint64_t(u) * rt;
            uint64_t const j = insert_zero_bit(p < p1 ? p : p0, sm.hbit);
            jj[u] = j;
            if (p < p1)
            {
                a[u] = states[j * vpr + col];
                b[u] = states[(j ^ xp) * vpr + col];
            }
            else
            {
                a[u] = V(0);
                b[u] = V(0);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            bool const neg = parity64(jj[u] & sm.z);
#pragma unroll
            for (int c = 0; c < NC; ++c)
            {
                A const ar = A(a[u][2 * c]), ai = A(a[u][2 * c + 1]);
                A const br = A(b[u][2 * c]), bi = A(b[u][2 * c + 1]);
                A g;
                if (sm.mode == kModeRe)
                    g = ar * br + ai * bi;
                else if (sm.mode == kModeIm)
                    g = ar * bi - ai * br;
                else
                    g = (ar * ar + ai * ai) + self_sign2 * (br * br + bi * bi);
                R[c] += neg ? -g : g;
            }
        }
    }
}

// Block-wide sum over ty for every (tx, c); result valid in threads with
// ty == 0.  `red` holds B * NC values of type S (B: threads per block).
template <class S, int NC, int B = kBlock>
__device__ __forceinline__ void reduce_over_rows(S (&v)[NC], S *red, uint32_t tx, uint32_t ty, uint32_t ct,
                                                 uint32_t rt)
{
    uint32_t const tid = threadIdx.x;
#pragma unroll
    for (int c = 0; c < NC; ++c)
        red[c * B + tid] = v[c];
    __syncthreads();
    if (ty == 0 && tx < ct)
    {
#pragma unroll
        for (int c = 0; c < NC; ++c)
        {
            S s = red[c * B + tx];
            for (uint32_t r = 1; r < rt; ++r)
                s += red[c * B + r * ct + tx];
            v[c] = s;
        }
    }
    __syncthreads();
}

struct Dbl2
{
    double re, im;
};

// Sums c_s * kappa_s * (row sums) over the block's strings and rows and writes
// one complex FP64 partial per state: partial[(bz*gx + bx)*n_states + t].
template <class T, int NC, int U>
__global__ __launch_bounds__(kBlock) void k_op_expect(typename VecOf<T, NC>::type const *__restrict__ states,
                                                      uint64_t vpr, uint32_t ct, uint32_t rt, uint64_t n_pairs,
                                                      uint64_t ppb, StringMeta const *__restrict__ meta,
                                                      double const *__restrict__ scale, uint32_t n_strings,
                                                      uint32_t spb, uint64_t n_states, Dbl2 *__restrict__ partial)
{
    typedef typename RowAcc<T>::type A;
    __shared__ double red[2 * NC * kBlock];
    uint32_t const tx = threadIdx.x % ct;
    uint32_t const ty = threadIdx.x / ct;
    bool const thread_live = ty < rt;
    uint64_t const col = uint64_t(blockIdx.y) * ct + tx;
    bool const col_live = thread_live && col < vpr;
    uint64_t const p0 = uint64_t(blockIdx.x) * ppb;
    uint64_t const p1 = (p0 + ppb < n_pairs) ? p0 + ppb : n_pairs;
    uint32_t const s0 = blockIdx.z * spb;
    uint32_t const s1 = (s0 + spb < n_strings) ? s0 + spb : n_strings;

    double er[NC], ei[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
        er[c] = ei[c] = 0.0;
    for (uint32_t s = s0; s < s1; ++s)
    {
        StringMeta const sm = meta[s];
        A R[NC];
        ev_rows<T, NC, U>(states, vpr, col, col_live, p0, p1, ty, rt, sm, R);
        double const sr = scale[2 * s], si = scale[2 * s + 1];
#pragma unroll
        for (int c = 0; c < NC; ++c)
        {
            er[c] += sr * double(R[c]);
            ei[c] += si * double(R[c]);
        }
    }
    double v[2 * NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
    {
        v[2 * c] = er[c];
        v[2 * c + 1] = ei[c];
    }
    reduce_over_rows<double, 2 * NC>(v, red, tx, ty, ct, rt);
    if (ty == 0 && col_live)
    {
        uint64_t const part = uint64_t(blockIdx.z) * gridDim.x + blockIdx.x;
#pragma unroll
        for (int c = 0; c < NC; ++c)
        {
            uint64_t const t = col * NC + c;
            if (t < n_states)
                partial[part * n_states + t] = Dbl2{v[2 * c], v[2 * c + 1]};
        }
    }
}

// Per-string variant: partial_real[(s*gx + bx)*n_states + t] = block row total
// of string s (kappa applied in the final reduction).
template <class T, int NC, int U>
__global__ __launch_bounds__(kBlock) void k_strings_expect(typename VecOf<T, NC>::type const *__restrict__ states,
                                                           uint64_t vpr, uint32_t ct, uint32_t rt, uint64_t n_pairs,
                                                           uint64_t ppb, StringMeta const *__restrict__ meta,
                                                           uint32_t n_strings, uint32_t spb, uint64_t n_states,
                                                           double *__restrict__ partial)
{
    typedef typename RowAcc<T>::type A;
    __shared__ double red[NC * kBlock];
    uint32_t const tx = threadIdx.x % ct;
    uint32_t const ty = threadIdx.x / ct;
    bool const thread_live = ty < rt;
    uint64_t const col = uint64_t(blockIdx.y) * ct + tx;
    bool const col_live = thread_live && col < vpr;
    uint64_t const p0 = uint64_t(blockIdx.x) * ppb;
    uint64_t const p1 = (p0 + ppb < n_pairs) ? p0 + ppb : n_pairs;
    uint32_t const s0 = blockIdx.z * spb;
    uint32_t const s1 = (s0 + spb < n_strings) ? s0 + spb : n_strings;
    for (uint32_t s = s0; s < s1; ++s)
    {
        StringMeta const sm = meta[s];
        A R[NC];
        ev_rows<T, NC, U>(states, vpr, col, col_live, p0, p1, ty, rt, sm, R);
        double v[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c)
            v[c] = double(R[c]);
        reduce_over_rows<double, NC>(v, red, tx, ty, ct, rt);
        if (ty == 0 && col_live)
        {
            uint64_t const part = uint64_t(s) * gridDim.x + blockIdx.x;
#pragma unroll
            for (int c = 0; c < NC; ++c)
            {
                uint64_t const t = col * NC + c;
                if (t < n_states)
                    partial[part * n_states + t] = v[c];
            }
        }
    }
}

// out[t] (+)= sum_i partial[i*n_states + t], FP64, fixed order (per-thread
// strided sums, then an LDS tree).  One block per state.
template <class OutT>
__global__ __launch_bounds__(kBlock) void k_reduce_partials(Dbl2 const *__restrict__ partial, uint64_t n_parts,
                                                            uint64_t n_states, OutT *__restrict__ out, int accumulate,
                                                            double post_re, double post_im)
{
    __shared__ double sr[kBlock], si[kBlock];
    uint64_t const t = blockIdx.x;
    double ar = 0.0, ai = 0.0;
    for (uint64_t i = threadIdx.x; i < n_parts; i += kBlock)
    {
        Dbl2 const p = partial[i * n_states + t];
        ar += p.re;
        ai += p.im;
    }
    sr[threadIdx.x] = ar;
    si[threadIdx.x] = ai;
    __syncthreads();
    for (uint32_t w = kBlock / 2; w > 0; w >>= 1)
    {
        if (threadIdx.x < w)
        {
            sr[threadIdx.x] += sr[threadIdx.x + w];
            si[threadIdx.x] += si[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        // optional post-multiplication by a complex scalar (the string coefficient)
        double const vr = sr[0] * post_re - si[0] * post_im;
        double const vi = sr[0] * post_im + si[0] * post_re;
        if (accumulate)
        {
            out[2 * t] = OutT(double(out[2 * t]) + vr);
            out[2 * t + 1] = OutT(double(out[2 * t + 1]) + vi);
        }
        else
        {
            out[2 * t] = OutT(vr);
            out[2 * t + 1] = OutT(vi);
        }
    }
}

// ev[s*n_states + t] = kappa_s * sum_b partial[(s*gx + b)*n_states + t]
template <class EvT>
__global__ __launch_bounds__(kBlock) void k_reduce_string_partials(double const *__restrict__ partial, uint32_t gx,
                                                                   uint64_t n_states, uint64_t n_strings,
                                                                   double const *__restrict__ kappa,
                                                                   EvT *__restrict__ ev)
{
    uint64_t const i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n_strings * n_states)
        return;
    uint64_t const s = i / n_states;
    uint64_t const t = i - s * n_states;
    double acc = 0.0;
    for (uint32_t b = 0; b < gx; ++b)
        acc += partial[(s * gx + b) * n_states + t];
    ev[i] = EvT(kappa[s] * acc);
}

// ---------------------------------------------------------------------------
// Seeded synthetic states: counter-based splitmix64 + Box-Muller.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t splitmix64_at(uint64_t seed, uint64_t i)
{
    uint64_t zz = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ull;
    zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBull;
    return zz ^ (zz >> 31);
}

__device__ __forceinline__ double u01_at(uint64_t seed, uint64_t i)
{
    return (double(splitmix64_at(seed, i) >> 11) + 0.5) * 0x1.0p-53;
}

// normal number k: uniforms 2*(k/2), 2*(k/2)+1; even k -> cos branch, odd -> sin
__device__ __forceinline__ double normal_at(uint64_t seed, uint64_t k)
{
    uint64_t const pair = k >> 1;
    double const u1 = u01_at(seed, 2 * pair);
    double const u2 = u01_at(seed, 2 * pair + 1);
    double const r = sqrt(-2.0 * log(u1));
    double const ang = 6.283185307179586476925286766559 * u2;
    return (k & 1) ? r * sin(ang) : r * cos(ang);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_fill_normal(T *__restrict__ data, uint64_t n_real, uint64_t seed)
{
    uint64_t const i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n_real)
        data[i] = T(normal_at(seed, i));
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_colnorm_partial(T const *__restrict__ data, uint64_t dim,
                                                            uint64_t n_states, uint64_t rows_per_block,
                                                            double *__restrict__ partial)
{
    // thread -> (state t = tid % n_states-tile); simple and only used at set-up
    uint64_t const t = uint64_t(blockIdx.y) * kBlock + threadIdx.x;
    if (t >= n_states)
        return;
    uint64_t const r0 = uint64_t(blockIdx.x) * rows_per_block;
    uint64_t const r1 = (r0 + rows_per_block < dim) ? r0 + rows_per_block : dim;
    double acc = 0.0;
    for (uint64_t r = r0; r < r1; ++r)
    {
        double const re = double(data[2 * (r * n_states + t)]);
        double const im = double(data[2 * (r * n_states + t) + 1]);
        acc += re * re + im * im;
    }
    partial[uint64_t(blockIdx.x) * n_states + t] = acc;
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_colnorm_final(R const *__restrict__ partial, uint64_t n_parts,
                                                          uint64_t n_states, R *__restrict__ inv_norm)
{
    uint64_t const t = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= n_states)
        return;
    R acc = 0;
    for (uint64_t i = 0; i < n_parts; ++i)
        acc += partial[i * n_states + t];
    inv_norm[t] = acc > 0 ? R(1) / sqrt(acc) : R(1);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_scale_columns(T *__restrict__ data, uint64_t n_elem, uint64_t n_states,
                                                          double const *__restrict__ inv_norm)
{
    uint64_t const i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n_elem)
        return;
    uint64_t const t = i % n_states;
    double const s = inv_norm[t];
    data[2 * i] = T(double(data[2 * i]) * s);
    data[2 * i + 1] = T(double(data[2 * i + 1]) * s);
}

template <class In, class Out>
__global__ __launch_bounds__(kBlock) void k_convert(In const *__restrict__ in, Out *__restrict__ out, uint64_t n)
{
    uint64_t const i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n)
        out[i] = Out(in[i]);
}

// out (complex, interleaved) (+)= (re plane, im plane) of the contraction;
// im plane absent (zero) when the coefficient matrix is real.
template <class In, class Out>
__global__ __launch_bounds__(kBlock) void k_pack_complex(In const *__restrict__ planes, Out *__restrict__ out,
                                                         uint64_t n, int imag_zero, int accumulate)
{
    uint64_t const i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    double const re = double(planes[i]);
    double const im = imag_zero ? 0.0 : double(planes[n + i]);
    if (accumulate)
    {
        out[2 * i] = Out(double(out[2 * i]) + re);
        out[2 * i + 1] = Out(double(out[2 * i + 1]) + im);
    }
    else
    {
        out[2 * i] = Out(re);
        out[2 * i + 1] = Out(im);
    }
}

// ---------------------------------------------------------------------------
// SummedPauliOp::apply_weighted (fast_pauli/cpp/include/__summed_pauli_op.hpp:378-503)
//
// Step 1: w(j,t) = sum_k h(j,k) x(k,t), k in order, product then add (the
//         reference's sequential loop, no contraction).
// Step 2: out(l,t) (+)= sum_j (m_j(l) psi(l^x_j, t)) * w(j,t), strings in
//         order j = 0..N-1; m_j psi is an exact sign flip / re-im swap.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlock) void k_weighted_coeffs(T const *__restrict__ h, T const *__restrict__ x,
                                                            T *__restrict__ w, uint64_t n_strings, uint64_t n_ops,
                                                            uint64_t n_states)
{
#pragma clang fp contract(off)
    uint64_t const i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n_strings * n_states)
        return;
    uint64_t const j = i / n_states;
    uint64_t const t = i - j * n_states;
    T wr = T(0), wi = T(0);
    for (uint64_t k = 0; k < n_ops; ++k)
    {
        T const d = x[k * n_states + t];
        T const pr = h[2 * (j * n_ops + k)] * d;
        T const pi = h[2 * (j * n_ops + k) + 1] * d;
        wr = wr + pr;
        wi = wi + pi;
    }
    w[2 * i] = wr;
    w[2 * i + 1] = wi;
}

// R rows per thread (rows row0 + u * dim / R, one column): the per-(string, state)
// coefficient w(s, t) is read once for the R output elements that use it, and the
// R partner reads per string are independent.  Each element sums its strings in
// order, as the reference's sequential loop does.
template <class T, int NC, bool POW2, bool ACC, int R>
__global__ __launch_bounds__(kBlock) void k_op_apply_weighted(typename VecOf<T, NC>::type const *__restrict__ in,
                                                              typename VecOf<T, NC>::type *__restrict__ out,
                                                              uint64_t v_begin, uint64_t n_vec, uint32_t vpr_shift,
                                                              uint64_t vpr, double inv_vpr,
                                                              uint64_t const *__restrict__ xs,
                                                              uint64_t const *__restrict__ zs,
                                                              uint32_t const *__restrict__ nys,
                                                              T const *__restrict__ w, uint32_t n_strings,
                                                              uint64_t n_states)
{
    typedef typename VecOf<T, NC>::type V;
    // n_vec / R thread slots; slot v covers vectors v + u * (n_vec / R)
    uint64_t const n_slot = n_vec / R;
    uint64_t const v = v_begin + uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (v >= n_slot)
        return;
    uint64_t row0, col;
    if constexpr (POW2)
    {
        row0 = v >> vpr_shift;
        col = v & ((uint64_t(1) << vpr_shift) - 1);
    }
    else
    {
        row0 = row_of(v, vpr, inv_vpr);
        col = v - row0 * vpr;
    }
    uint64_t const row_step = n_slot / vpr; // dim / R
    V acc[R];
#pragma unroll
    for (int u = 0; u < R; ++u)
        acc[u] = ACC ? out[v + uint64_t(u) * n_slot] : V(0);
    for (uint32_t s = 0; s < n_strings; ++s)
    {
        uint64_t const x = xs[s];
        uint64_t const z = zs[s];
        uint32_t const ny = nys[s] & 3u;
        T wr[NC], wi[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c)
        {
            uint64_t const t = col * NC + c;
            wr[c] = w[2 * (uint64_t(s) * n_states + t)];
            wi[c] = w[2 * (uint64_t(s) * n_states + t) + 1];
        }
        V p[R];
#pragma unroll
        for (int u = 0; u < R; ++u)
        {
            uint64_t const row = row0 + uint64_t(u) * row_step;
            uint64_t const sv = POW2 ? (((row ^ x) << vpr_shift) | col) : ((row ^ x) * vpr + col);
            p[u] = in[sv];
        }
#pragma unroll
        for (int u = 0; u < R; ++u)
        {
            uint64_t const row = row0 + uint64_t(u) * row_step;
            uint32_t const sg = parity64(row & z);
            // m = (-i)^{ny} * (-1)^{sg}: (re, im) of m * (pr + i pi) without rounding
#pragma unroll
            for (int c = 0; c < NC; ++c)
            {
                T pr = p[u][2 * c], pi = p[u][2 * c + 1];
                T ar, ai;
                switch (ny)
                {
                case 0:
                    ar = pr, ai = pi;
                    break;
                case 1:
                    ar = pi, ai = -pr;
                    break;
                case 2:
                    ar = -pr, ai = -pi;
                    break;
                default:
                    ar = -pi, ai = pr;
                    break;
                }
                if (sg)
                    ar = -ar, ai = -ai;
                T orr, oi;
                cmul_exact<T>(ar, ai, wr[c], wi[c], orr, oi);
                acc[u][2 * c] = add_exact<T>(acc[u][2 * c], orr);
                acc[u][2 * c + 1] = add_exact<T>(acc[u][2 * c + 1], oi);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < R; ++u)
        out[v + uint64_t(u) * n_slot] = acc[u];
}

// Plain streaming copy (16-byte vectors, non-temporal): the reference point for
// achievable HBM bandwidth and for scaling hardware byte counters.
template <int U>
__global__ __launch_bounds__(kBlock) void k_copy(VecOf<double, 1>::type const *__restrict__ in,
                                                 VecOf<double, 1>::type *__restrict__ out, uint64_t v_begin,
                                                 uint64_t n_vec)
{
    typedef VecOf<double, 1>::type V;
    uint64_t const base = v_begin + uint64_t(blockIdx.x) * (kBlock * U) + threadIdx.x;
    V r[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
        uint64_t const v = base + uint64_t(u) * kBlock;
        if (v < n_vec)
            r[u] = __builtin_nontemporal_load(in + v);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
        uint64_t const v = base + uint64_t(u) * kBlock;
        if (v < n_vec)
            __builtin_nontemporal_store(r[u], out + v);
    }
}

} // namespace fpk
