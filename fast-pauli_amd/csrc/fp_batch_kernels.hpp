// fp_batch_kernels.hpp — one persistent launch over many operators of uneven
// size (qubit counts, string counts, state counts, apply or expectation).
//
// The host flattens every operator into work items and orders them; the
// persistent blocks pull items from a device counter until the list is
// exhausted, so every wave reaches the exit condition.
//   apply tile:        a run of output vectors; all strings accumulated in
//                      registers in order s = 0..N-1 (the same arithmetic as
//                      the single-operator kernel, so results are bit-exact
//                      with it).
//   coset apply item:  cosets [q0, q0 + count) of one string group of a large
//                      operator: the group's state tiles go through LDS and the
//                      group's strings are added, in order, to the output
//                      elements (the coset-tiled apply kernel's arithmetic, so
//                      also bit-exact).  Groups of one operator run in order:
//                      an item of group g > 0 waits until every item of group
//                      g - 1 has published its outputs (device-scope counter).
//                      The outputs are read and written with agent-scope
//                      atomic accesses, which bypass the XCDs' non-coherent L2
//                      copies; an item's threads wait for their stores to
//                      complete before the counter is incremented, and a waiting
//                      item issues its output loads only after it has observed
//                      the final count, so no L2 write-back or invalidation is
//                      needed per item (those cost about 20 us each).
//   expectation tile:  a range of row pairs for all strings; FP64 partials per
//                      state.  The block that completes an operator's last tile
//                      (device-scope counter) reduces its partials in tile
//                      order, so for a given schedule results are identical run
//                      to run; another tile size regroups the FP64 partial sums
//                      (last-bit differences).
// Waits cannot deadlock: the host lists every item of group g - 1 before any
// item of group g, items are claimed in list order, and a claimed item is held
// by a running block, so the earliest unfinished item never waits.  The wait is
// still bounded (kSpinLimit sleeps); an item that gives up sets the batch's
// error word and runs anyway, so the kernel always exits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fp_apply_tiled.hpp"
#include "fp_ev_tiled.hpp"
#include "fp_kernels.hpp"

namespace fpb
{

using fpk::kBlock;

// threads per block of the persistent kernel: kBlock without coset items, 512
// (the coset-tiled apply kernel's block, one 64 KiB tile) with them
constexpr int kTiledBlock = fpa::kThreads;
constexpr uint32_t kSpinLimit = 1u << 22; // x s_sleep 8 (~0.25 us): about one second

enum : uint32_t
{
    kApply = 0,
    kExpect = 1
};

struct BatchOp
{
    void const *in;
    void *out;
    uint64_t const *xs;
    uint64_t const *zs;
    void const *cf;                  // folded coefficients (apply), complex of dtype
    fpk::StringMeta const *meta;     // expectation
    double const *scale;             // c_s * kappa_s (expectation)
    uint64_t n_vec;                  // vectors in the batch (apply)
    uint64_t vpr;                    // vectors per row
    uint64_t n_pairs;                // row pairs (expectation)
    uint64_t n_states;
    uint64_t partial_base;           // first partial element (expectation): tile i, state t at base + i*n_states + t
    uint32_t n_strings;
    uint32_t shift;                  // log2(vpr) when pow2
    uint32_t pow2;
    uint32_t nc;                     // complexes per vector (1 or 2)
    uint32_t dtype;                  // 0 complex128, 1 complex64
    uint32_t kind;                   // kApply / kExpect
    uint32_t accumulate;
    uint32_t n_tiles;                // tiles of this operator
    uint32_t ct, rt;                 // expectation thread layout
    // coset-tiled apply (coset != 0): the operator set's string-group tables
    fpt::TileGroup const *groups;
    uint64_t const *btab;
    fpa::ApplyStr const *astrs;
    uint64_t n_cosets;               // cosets of 2^d rows
    uint32_t coset;                  // 1: this operator runs as coset items (nc: states per slot)
    uint32_t lsb, k, d;              // tile: 2^d rows (the low k row bits among them) x 2^lsb states
    uint32_t sshift;                 // log2(n_states) when a power of two, else 64
    uint32_t dep_base;               // group g's completion counter: deps[dep_base + g]
    uint32_t items_per_group;        // coset items of every group
    uint32_t pad;
};

struct WorkItem
{
    uint32_t op;
    uint32_t tile;   // tile index within the operator; coset items: the string group
    uint64_t begin;  // first vector (apply), first pair (expectation) or first coset
    uint64_t count;  // vectors, pairs or cosets in the item
    uint32_t chunk;  // coset items: states [chunk << lsb, (chunk + 1) << lsb)
    uint32_t pad;
};

__device__ __forceinline__ double load_coherent(double const *p)
{
    unsigned long long const bits =
        __hip_atomic_load(reinterpret_cast<unsigned long long const *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double(static_cast<long long>(bits));
}

// Device-coherent accesses to a coset item's outputs (64-bit relaxed atomics at
// agent scope, which bypass the XCDs' non-coherent L2 copies): one group's stores
// reach the coherence point, where the next group's loads read them, without an
// L2 write-back or invalidation per item.
template <class V> __device__ __forceinline__ V load_coherent_vec(V const *p)
{
    static_assert(sizeof(V) % 8 == 0, "64-bit pieces");
    constexpr int W = int(sizeof(V) / 8);
    unsigned long long w[W];
#pragma unroll
    for (int i = 0; i < W; ++i)
        w[i] = __hip_atomic_load(reinterpret_cast<unsigned long long const *>(p) + i, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
    V v;
    __builtin_memcpy(&v, w, sizeof(V));
    return v;
}

template <class V> __device__ __forceinline__ void store_coherent_vec(V *p, V const &v)
{
    static_assert(sizeof(V) % 8 == 0, "64-bit pieces");
    constexpr int W = int(sizeof(V) / 8);
    unsigned long long w[W];
    __builtin_memcpy(w, &v, sizeof(V));
#pragma unroll
    for (int i = 0; i < W; ++i)
        __hip_atomic_store(reinterpret_cast<unsigned long long *>(p) + i, w[i], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

template <class T, int NC, int B>
__device__ void apply_tile(BatchOp const &op, uint64_t begin, uint64_t count)
{
    typedef typename fpk::VecOf<T, NC>::type V;
    auto const *in = static_cast<V const *>(op.in);
    auto *out = static_cast<V *>(op.out);
    auto const *cf = static_cast<T const *>(op.cf);
    uint64_t const end = begin + count;
    constexpr int U = 2;
    for (uint64_t base = begin + threadIdx.x; base < end; base += uint64_t(B) * U)
    {
        uint64_t row[U], col[U];
        bool live[U];
        V acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            uint64_t const v = base + uint64_t(u) * B;
            live[u] = v < end;
            uint64_t const vv = live[u] ? v : begin;
            if (op.pow2)
            {
                row[u] = vv >> op.shift;
                col[u] = vv & (op.vpr - 1);
            }
            else
            {
                row[u] = vv / op.vpr;
                col[u] = vv - row[u] * op.vpr;
            }
            acc[u] = (op.accumulate && live[u]) ? out[v] : V(0);
        }
        for (uint32_t s = 0; s < op.n_strings; ++s)
        {
            uint64_t const x = op.xs[s];
            uint64_t const z = op.zs[s];
            T const cr = cf[2 * s], ci = cf[2 * s + 1];
            V src[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                src[u] = live[u] ? in[(row[u] ^ x) * op.vpr + col[u]] : V(0);
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
                uint32_t const sg = fpk::parity64(row[u] & z);
                T const a = sg ? -cr : cr;
                T const b = sg ? -ci : ci;
#pragma unroll
                for (int c = 0; c < NC; ++c)
                {
                    T orr, oi;
                    fpk::cmul_exact<T>(a, b, src[u][2 * c], src[u][2 * c + 1], orr, oi);
                    acc[u][2 * c] = fpk::add_exact<T>(acc[u][2 * c], orr);
                    acc[u][2 * c + 1] = fpk::add_exact<T>(acc[u][2 * c + 1], oi);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (live[u])
                out[base + uint64_t(u) * B] = acc[u];
    }
}

// Coset apply item: cosets [q0, q0 + count) of string group g, states
// [chunk << lsb, (chunk + 1) << lsb).  A slot holds P states of one row (P = 2:
// complex64 with an even state count, two neighbouring states moved by 16-byte
// loads, stores and LDS reads); a row of the chunk has 2^lsbx slots, lsbx =
// lsb - log2 P.  Slot m of thread tid is tile slot tid + m * B: local row
// a_m = a_lane + (m << rm) (a_lane = tid >> lsbx < 2^rm, disjoint bits) and
// states P tx .. P tx + P - 1, tx = tid & (2^lsbx - 1).  The local-to-global row map
// (low k bits kept, the rest through the group's table bt of span combinations)
// is linear over XOR, and rm >= k, so the row of slot m in the coset with base
// cb is cb ^ roff(a_lane) ^ u_m with u_m = bt[(m << rm) >> k] the same in every
// lane: one offset register per thread instead of one per slot.
// For a string with local masks (A, Z): partner row a_m ^ A and sign
// parity(a_m & Z) ^ parity(q & zc), q the coset index; the coefficient is the
// folded c_s (-i)^{nY_s}.  Each output element sees out + term_0 + term_1 + ...
// in the operator's string order (cmul_exact, add_exact), as in the
// single-operator kernels.  Requires 2^n * S < 2^32 (32-bit element offsets).
template <class T, int B, int P>
__device__ void apply_coset_item(BatchOp const &op, uint32_t g, uint32_t chunk, uint64_t q0, uint64_t count,
                                 unsigned char *smem)
{
    static_assert(P == 1 || (P == 2 && sizeof(T) == 4), "two states per slot: complex64 only");
    typedef typename fpk::VecOf<T, P>::type V; // one slot
    constexpr uint32_t kSlots = fpa::kTileBytes / sizeof(V);
    constexpr uint32_t EPT = kSlots / B;
    constexpr uint32_t kLog2B = B == 512 ? 9u : (B == 256 ? 8u : 10u);
    static_assert((1u << kLog2B) == uint32_t(B), "power-of-two block");
    V *tile = reinterpret_cast<V *>(smem);
    V const *src = static_cast<V const *>(op.in);
    V *out = static_cast<V *>(op.out);

    fpt::TileGroup const grp = op.groups[g];
    uint64_t const *bt = op.btab + grp.btab_off;
    uint32_t const n_str = grp.n_chunks; // apply tables: the group's strings, in order
    fpa::ApplyStr const *gs = op.astrs + grp.chunk_first;
    uint32_t const lsb = op.lsb, k = op.k, d = op.d, sshift = op.sshift;
    uint32_t const n_states = uint32_t(op.n_states);
    uint32_t const tid = threadIdx.x;
    uint32_t const lsbx = P == 2 ? lsb - 1u : lsb;
    uint32_t const tx = tid & ((1u << lsbx) - 1u);
    uint32_t const st0 = (chunk << lsb) + uint32_t(P) * tx; // this thread's first state
    bool const t_live = st0 < n_states;                    // P = 2: n_states is even
    uint32_t const tile_slots = (1u << d) << lsbx;
    bool const full = tile_slots == kSlots && (chunk << lsb) + (1u << lsb) <= n_states;
    bool const s_pow2 = sshift < 64u;
    bool const init_zero = g == 0 && !op.accumulate;
    uint32_t const rm = kLog2B - lsbx;
    uint32_t const a_lane = tid >> lsbx;
    uint32_t const lane_mask = (1u << rm) - 1u;
    uint32_t const roff = a_lane < (1u << d) ? uint32_t(bt[a_lane >> k] ^ (a_lane & ((1u << k) - 1u))) : 0u;
    uint32_t u[EPT];
#pragma unroll
    for (uint32_t m = 0; m < EPT; ++m)
        u[m] = (m << rm) < (1u << d) ? uint32_t(bt[(m << rm) >> k]) : 0u;
    auto slot_live = [&](uint32_t m) { return full || (tid + m * B < tile_slots && t_live); };
    // slot index (in units of P states) of slot m in the coset whose base, xor this
    // lane's offset, is cbl
    auto index_of = [&](uint32_t cbl, uint32_t m) -> uint32_t {
        uint32_t const row = cbl ^ u[m];
        if constexpr (P == 2)
            return s_pow2 ? ((row << (sshift - 1u)) | (st0 >> 1)) : row * (n_states >> 1) + (st0 >> 1);
        else
            return s_pow2 ? ((row << sshift) | st0) : row * n_states + st0;
    };

    V reg[EPT];
    auto fetch = [&](uint32_t cbl) {
#pragma unroll
        for (uint32_t m = 0; m < EPT; ++m)
        {
            reg[m] = V(0);
            if (slot_live(m))
                reg[m] = __builtin_nontemporal_load(&src[index_of(cbl, m)]);
        }
    };
    // base(q + 1) = ((base(q) | ~cmask) + base(1)) & cmask: the pivot bits, set to
    // 1, pass the carry through
    uint64_t const cmask = fpt::coset_base(op.n_cosets - 1u, k, grp);
    uint64_t const cstep = fpt::coset_base(1, k, grp);
    uint64_t cb = fpt::coset_base(q0, k, grp);
    fetch(uint32_t(cb) ^ roff);
    // the next coset's state values are requested before this coset's strings are
    // evaluated, its output values just before the tile store (registers)
    for (uint64_t q = q0; q < q0 + count; ++q)
    {
        uint32_t const cbl = uint32_t(cb) ^ roff;
        V acc[EPT];
#pragma unroll
        for (uint32_t m = 0; m < EPT; ++m)
        {
            acc[m] = V(0);
            if (!init_zero && slot_live(m))
                acc[m] = load_coherent_vec(&out[index_of(cbl, m)]);
        }
        __syncthreads(); // the previous tile is no longer read
#pragma unroll
        for (uint32_t m = 0; m < EPT; ++m)
        {
            uint32_t const e = tid + m * B;
            if (full || e < tile_slots)
                tile[e] = reg[m];
        }
        __syncthreads();
        uint64_t const cbn = ((cb | ~cmask) + cstep) & cmask;
        if (q + 1 < q0 + count)
            fetch(uint32_t(cbn) ^ roff);
        for (uint32_t s = 0; s < n_str; ++s)
        {
            fpa::ApplyStr const st = gs[s];
            uint32_t const cs = uint32_t(__popcll(q & st.zc) & 1u);
            T cr, ci;
            if constexpr (sizeof(T) == 8)
            {
                cr = st.cr;
                ci = st.ci;
            }
            else
            {
                cr = st.fr;
                ci = st.fi;
            }
            uint32_t const A_lane = st.A & lane_mask, A_m = st.A & ~lane_mask;
            V const *pl = tile + (((a_lane ^ A_lane) << lsbx) | tx);
            bool const lneg = ((uint32_t(__popc(a_lane & st.Z)) ^ cs) & 1u) != 0;
            T const lcr = lneg ? -cr : cr;
            T const lci = lneg ? -ci : ci;
#pragma unroll
            for (uint32_t m = 0; m < EPT; ++m)
            {
                uint32_t const mrow = m << rm;
                V const v = pl[(mrow ^ A_m) << lsbx];
                uint32_t const mbit = uint32_t(__popc(mrow & st.Z) & 1u) << 31;
                T const a = fpt::xor_sign_bit(lcr, mbit);
                T const b = fpt::xor_sign_bit(lci, mbit);
#pragma unroll
                for (int j = 0; j < P; ++j)
                {
                    T orr, oi;
                    fpk::cmul_exact<T>(a, b, v[2 * j], v[2 * j + 1], orr, oi);
                    acc[m][2 * j] = fpk::add_exact<T>(acc[m][2 * j], orr);
                    acc[m][2 * j + 1] = fpk::add_exact<T>(acc[m][2 * j + 1], oi);
                }
            }
        }
#pragma unroll
        for (uint32_t m = 0; m < EPT; ++m)
            if (slot_live(m))
                store_coherent_vec(&out[index_of(cbl, m)], acc[m]);
        cb = cbn;
    }
}

// Expectation tile: pairs [begin, begin + count) of all strings; writes one FP64
// complex partial per state; the last tile of the operator reduces them.
// `red` holds 4 * B doubles.
template <class T, int NC, int B>
__device__ void expect_tile(BatchOp const &op, uint32_t tile, uint64_t begin, uint64_t count, fpk::Dbl2 *partials,
                            double *red)
{
    typedef typename fpk::RowAcc<T>::type A;
    auto const *states = static_cast<typename fpk::VecOf<T, NC>::type const *>(op.in);
    uint32_t const ct = op.ct, rt = op.rt;
    uint32_t const tx = threadIdx.x % ct, ty = threadIdx.x / ct;
    uint64_t const p1 = begin + count;
    for (uint64_t col0 = 0; col0 < op.vpr; col0 += ct)
    {
        uint64_t const col = col0 + tx;
        bool const col_live = ty < rt && col < op.vpr;
        double er[NC], ei[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c)
            er[c] = ei[c] = 0.0;
        for (uint32_t s = 0; s < op.n_strings; ++s)
        {
            fpk::StringMeta const sm = op.meta[s];
            A R[NC];
            fpk::ev_rows<T, NC, 4>(states, op.vpr, col, col_live, begin, p1, ty, rt, sm, R);
            double const sr = op.scale[2 * s], si = op.scale[2 * s + 1];
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
        fpk::reduce_over_rows<double, 2 * NC, B>(v, red, tx, ty, ct, rt);
        if (ty == 0 && col_live)
        {
#pragma unroll
            for (int c = 0; c < NC; ++c)
            {
                uint64_t const t = col * NC + c;
                if (t < op.n_states)
                    partials[op.partial_base + uint64_t(tile) * op.n_states + t] = fpk::Dbl2{v[2 * c], v[2 * c + 1]};
            }
        }
    }
}

// The block that completes an operator's last expectation tile sums its partials
// (tiles in order) into the output.
template <class T, int B>
__device__ void expect_finish(BatchOp const &op, fpk::Dbl2 const *partials)
{
    auto *out = static_cast<T *>(op.out);
    for (uint64_t t = threadIdx.x; t < op.n_states; t += B)
    {
        double ar = 0.0, ai = 0.0;
        for (uint32_t i = 0; i < op.n_tiles; ++i)
        {
            fpk::Dbl2 const *p = partials + op.partial_base + uint64_t(i) * op.n_states + t;
            ar += load_coherent(&p->re);
            ai += load_coherent(&p->im);
        }
        if (op.accumulate)
        {
            out[2 * t] = T(double(out[2 * t]) + ar);
            out[2 * t + 1] = T(double(out[2 * t + 1]) + ai);
        }
        else
        {
            out[2 * t] = T(ar);
            out[2 * t + 1] = T(ai);
        }
    }
}

// Zeroes the work counter and the n completion counters that follow it.
__global__ __launch_bounds__(kBlock) void k_batch_reset(uint32_t *counter, uint32_t *done, uint32_t n)
{
    uint32_t const i = blockIdx.x * kBlock + threadIdx.x;
    if (i == 0)
        *counter = 0;
    if (i < n)
        done[i] = 0;
}

// B threads per block; dynamic LDS: 4 * B doubles for expectation items and, with
// coset items (B = kTiledBlock), one 64 KiB tile (the two share the buffer).
// `done`: per-operator expectation counters, then the coset groups' counters;
// `error`: set (never cleared by the kernel) when a coset item gave up waiting.
template <int B>
__global__ __launch_bounds__(B, B == kTiledBlock ? 4 : 2) void k_batch_persistent(BatchOp const *__restrict__ ops,
                                                        WorkItem const *__restrict__ items, uint32_t n_items,
                                                        uint32_t *counter, uint32_t *done, fpk::Dbl2 *partials,
                                                        uint32_t *error)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t s_item;
    __shared__ uint32_t s_last;
    double *red = reinterpret_cast<double *>(smem);
    while (true)
    {
        if (threadIdx.x == 0)
            s_item = atomicAdd(counter, 1u);
        __syncthreads();
        // the same in every lane: uniform, so that the item's and operator's fields
        // are scalar loads held in scalar registers
        uint32_t const w = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (w >= n_items)
            break;
        WorkItem const wi = items[w];
        BatchOp const &op = ops[wi.op];
        if (op.kind == kApply && op.coset)
        {
            if constexpr (B == kTiledBlock)
            {
                if (wi.tile > 0)
                {
                    // wait for group tile - 1 of this operator (bounded)
                    if (threadIdx.x == 0)
                    {
                        uint32_t const *c = done + op.dep_base + wi.tile - 1;
                        uint32_t spins = 0;
                        while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < op.items_per_group)
                        {
                            if (++spins > kSpinLimit)
                            {
                                __hip_atomic_store(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                break;
                            }
                            __builtin_amdgcn_s_sleep(8);
                        }
                    }
                    __syncthreads();
                }
                if (op.dtype == 0)
                    apply_coset_item<double, B, 1>(op, wi.tile, wi.chunk, wi.begin, wi.count, smem);
                else if (op.nc == 2)
                    apply_coset_item<float, B, 2>(op, wi.tile, wi.chunk, wi.begin, wi.count, smem);
                else
                    apply_coset_item<float, B, 1>(op, wi.tile, wi.chunk, wi.begin, wi.count, smem);
                // count this item once every thread's output stores have completed (they
                // are device-coherent, so no cache write-back is needed to publish them)
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
                if (threadIdx.x == 0)
                    __hip_atomic_fetch_add(done + op.dep_base + wi.tile, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        else if (op.kind == kApply)
        {
            if (op.dtype == 0)
                apply_tile<double, 1, B>(op, wi.begin, wi.count);
            else if (op.nc == 2)
                apply_tile<float, 2, B>(op, wi.begin, wi.count);
            else
                apply_tile<float, 1, B>(op, wi.begin, wi.count);
        }
        else
        {
            if (op.dtype == 0)
                expect_tile<double, 1, B>(op, wi.tile, wi.begin, wi.count, partials, red);
            else if (op.nc == 2)
                expect_tile<float, 2, B>(op, wi.tile, wi.begin, wi.count, partials, red);
            else
                expect_tile<float, 1, B>(op, wi.tile, wi.begin, wi.count, partials, red);
            // publish this tile's partials; the last finisher of the operator reduces
            __threadfence();
            __syncthreads();
            if (threadIdx.x == 0)
                s_last = (atomicAdd(&done[wi.op], 1u) == op.n_tiles - 1) ? 1u : 0u;
            __syncthreads();
            if (s_last)
            {
                __threadfence();
                if (op.dtype == 0)
                    expect_finish<double, B>(op, partials);
                else
                    expect_finish<float, B>(op, partials);
            }
        }
        __syncthreads();
    }
}

} // namespace fpb
