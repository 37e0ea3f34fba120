// fp_runtime.hpp — pieces of the runtime shared by the library's translation
// units: error reporting, per-stream scratch, the operator-set object and the
// vector view of a row-major batch.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <list>
#include <mutex>
#include <vector>

#include "fast_pauli_hip.h"
#include "fp_ev_lds.hpp"
#include "fp_apply_tiled.hpp"
#include "fp_ev_tiled.hpp"
#include "fp_kernels.hpp"
#include "fp_tiles_host.hpp"

namespace fprt
{

// Per-stream device scratch of at least `bytes`; the pointer stays valid until
// the next, larger request on the same stream.  Callers hold a StreamGuard for
// the stream until the kernels that use the buffer are enqueued.
fp_status scratch_get(hipStream_t stream, size_t bytes, void **out);

// Serialises the library's entry points on one (device, stream): an entry point
// that uses the stream's scratch holds the guard until its kernels are enqueued,
// so that calls from several host threads on one stream neither interleave
// their kernels nor free a buffer another call is about to use.
class StreamGuard
{
  public:
    explicit StreamGuard(hipStream_t stream);
    ~StreamGuard();
    StreamGuard(StreamGuard const &) = delete;
    StreamGuard &operator=(StreamGuard const &) = delete;

  private:
    std::recursive_mutex *mu_;
};

// Device workspace supplied by the caller instead of the stream's scratch (the
// allocation-free entry point passes its static buffer).
struct Workspace
{
    void *base;
    size_t bytes;
};

// LDS bytes one workgroup may use on the current device (160 KiB on gfx950), from
// the device's attributes, cached per device.
size_t lds_bytes();
// LDS bytes of one state column in the LDS-resident expectation kernels: the
// workgroup's LDS less room for their static counters.
inline size_t column_lds_bytes()
{
    return lds_bytes() - 256;
}

// Smallest class of strings (equal x, equal half of w) that the per-string EV
// matrix kernels evaluate through a Walsh-Hadamard transform, or
// FAST_PAULI_WHT_MIN (0 disables the transform).  The transform trades vector
// and scalar instructions for LDS traffic: the complex128 coset kernel, whose
// loop is LDS-bound, gains from the larger classes only (32 strings and up; at n = 14,
// 2000 weight-3 strings, 1000 states: 2.01 ms against 2.21 ms with 24); the complex64
// LDS kernel gains from 12 (1.03 ms; 1.05 ms with 32, 1.31 ms with none).
inline uint32_t wht_min_strings(uint32_t dflt = fpl::kWhtMin)
{
    char const *v = std::getenv("FAST_PAULI_WHT_MIN");
    if (!v)
        return dflt;
    long const n = std::strtol(v, nullptr, 10);
    return n <= 0 ? 0xffffffffu : uint32_t(n);
}

// Blocks per (state column, coset) that share out the per-string chunks of the
// LDS-resident EV matrix kernels (each block loads its column or coset and takes every
// gy-th chunk): floor(256 / (n_states * n_cosets)), at least 1, at most `max_chunks`;
// FAST_PAULI_EV_SPLIT overrides.  Measured at n = 14, 2000 weight-3 strings (best split
// and its time, against the former ceil(1024 / n_states) rule): complex128 coset kernel
// (2 cosets) 16 states 8 / 0.065 ms against 0.220, 100 states 1 / 0.204 against 0.465,
// 1000 states 1 / 1.80 against 2.02; complex64 whole-column kernel 16 states 8-16 /
// 0.074, 100 states 2-4 / 0.150 against 0.248 with 1, 1000 states 1 / 1.005.
inline uint32_t ev_matrix_split(size_t n_states, size_t n_cosets, size_t max_chunks)
{
    size_t const per = std::max<size_t>(n_states, 1) * std::max<size_t>(n_cosets, 1);
    size_t gy = std::max<size_t>(1, std::min<size_t>(max_chunks, 256 / per));
    if (char const *v = std::getenv("FAST_PAULI_EV_SPLIT"))
    {
        long const e = std::strtol(v, nullptr, 10);
        if (e > 0)
            gy = size_t(std::min<long>(e, 64));
    }
    return uint32_t(gy);
}

// A launch stays below 2^32 work-items: at most 2^23 blocks of 256 threads.
constexpr uint64_t kMaxLaunchBlocks = uint64_t(1) << 23;

// Launch geometry of the vector view of a (dim x n_states) batch.
struct VecGeom
{
    int nc;         // complexes per vector
    uint64_t vpr;   // vectors per row
    uint64_t n_vec; // total vectors
    bool pow2;
    uint32_t shift;
    double inv_vpr;
};

VecGeom vec_geom(int dtype, size_t dim, size_t n_states, void const *a, void const *b);

// String-group tables of one kind and geometry (k, dh), cached per operator
// set: a host image (groups | B tables | chunks | strings) and its device copy.
// Re-targeting the operator set invalidates them but keeps both buffers for
// reuse, so re-targeted sets (the serving loop) neither allocate nor free.
struct TileTables
{
    int kind = 0;
    uint32_t k = 0, dh = 0, n_groups = 0, max_chunks = 0;
    uint32_t wht_min = 0; // per-string EV matrix tables: smallest transform class
    bool valid = false;     // built for the current strings and coefficients
    bool on_device = false; // the device copy matches the host image
    std::vector<unsigned char> h;
    size_t o_btab = 0, o_chunks = 0, o_strs = 0, o_sh = 0;
    void *d_mem = nullptr;
    size_t cap = 0;
    hipEvent_t uploaded = nullptr; // completes when the last upload has read `h`
    fpt::TileGroup *groups = nullptr;
    uint64_t *btab = nullptr;
    void *chunks = nullptr;
    void *strs = nullptr;
    fpt::ShStr *sh = nullptr; // shared-leader string records (expectation tables)
};

// `wht_min`: for kTilesEvMatrix, the smallest class evaluated through a transform (part of the key)
fp_status tile_tables(fp_hip_ops const *ops, int kind, uint32_t k, uint32_t dh, TileTables **out,
                      uint32_t wht_min = 0);
fp_status tile_device(fp_hip_ops const *ops, TileTables *t, hipStream_t stream);
// The tables were cut for `geo`, and one tile of `geo` fits `tile_bytes` of LDS.
fp_status check_tile_fit(TileTables const *tt, TileGeom const &geo, int dtype, size_t tile_bytes);
void invalidate_tiles(fp_hip_ops *ops);
// Per-string EV matrix over cosets: for columns that do not fit in LDS whole.
bool ev_coset_eligible(fp_hip_ops const *ops, int dtype, size_t dim, size_t n_states);
size_t ev_coset_scratch_bytes(fp_hip_ops const *ops, int dtype, size_t dim, size_t n_states);
fp_status run_ev_coset(fp_hip_ops const *ops, int dtype, void const *states, void *ev, size_t dim, size_t n_states,
                       void *scratch, hipStream_t st);
void release_tiles(fp_hip_ops *ops);
fp_status choose_op_expect_tiled(fp_hip_ops const *ops, int dtype, size_t n_states, bool *use_tiled);
// Coset-tiled apply: `tt` (and its geometry) when the state has at least `min_bytes`,
// the set at least 8 strings, and its string groups' traffic (about three state passes
// each) is at most 0.8 of the streaming form's (one per string); else `tt` = null.
// `kind`: kTilesOpApply (PauliOp::apply), kTilesSummedApply (SummedPauliOp::apply, the
// row-summed coefficients) or kTilesWeightedApply (SummedPauliOp::apply_weighted).
fp_status apply_tiled_plan(fp_hip_ops const *ops, int dtype, size_t n_states, size_t min_bytes, int kind,
                           TileGeom *geo, TileTables **tt);
// The single-operator entry points take the tiled form from 64 MiB of states: below
// that the caches serve the streaming kernel's repeated row reads.
constexpr size_t kApplyTiledMinBytes = size_t(64) << 20;
fp_status choose_op_apply_tiled(fp_hip_ops const *ops, int dtype, size_t n_states, bool *use_tiled,
                                int kind = kTilesOpApply);
// `weights` (kTilesWeightedApply only): the (strings x states) complex weights w(s, t)
fp_status run_op_apply_tiled(fp_hip_ops const *ops, int dtype, void const *states, void *out, size_t dim,
                             size_t n_states, int accumulate, hipStream_t st, int kind = kTilesOpApply,
                             void const *weights = nullptr);
// The launches of run_op_apply_tiled (one per string group) on tables already on the
// device; enqueues nothing else, so a graph capture may record it.
fp_status launch_op_apply_tiled(TileTables const *tt, TileGeom const &geo, int dtype, void const *states, void *out,
                                size_t dim, size_t n_states, int accumulate, hipStream_t st,
                                void const *weights = nullptr);
// Sets the tiled apply kernels' LDS limit (once per process).
fp_status apply_tiled_attributes();
// With `ws`, the string-group tables and the partials go into the workspace (no
// allocation; fails with FP_ERR_INVALID_ARGUMENT when they cannot fit), else the
// tables' device copy and the stream's scratch.
fp_status run_op_expect_tiled(fp_hip_ops const *ops, int dtype, void const *states, void *out, size_t dim,
                              size_t n_states, int accumulate, hipStream_t st, Workspace const *ws = nullptr);
// The streaming PauliOp expectation (strings [s_begin, s_begin + count), per-string
// complex scales, out[t] (+)= post * sum), partials in `ws` or the stream's scratch.
fp_status run_op_expect(fp_hip_ops const *ops, int dtype, void const *states, void *out, size_t dim, size_t n_states,
                        size_t s_begin, size_t count, double const *scale_table, double post_re, double post_im,
                        int accumulate, hipStream_t st, Workspace const *ws = nullptr);
// Shape checks shared by the entry points.
fp_status check_shape(fp_hip_ops const *ops, int dtype, void const *a, void const *b, size_t dim, size_t n_states);
// Sets up the allocation-free entry point's static buffer on the current device
// (called by fp_hip_init; idempotent).
fp_status noalloc_setup();

} // namespace fprt

#define FP_HIP(call)                                                                                                   \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t const e_ = (call);                                                                                  \
        if (e_ != hipSuccess)                                                                                          \
            return fprt::fail(e_ == hipErrorOutOfMemory ? FP_ERR_OUT_OF_MEMORY : FP_ERR_RUNTIME, "%s failed: %s (%s:%d)",    \
                        #call, hipGetErrorString(e_), __FILE__, __LINE__);                                             \
    } while (0)

#define FP_LAUNCHED()                                                                                                  \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t const e_ = hipGetLastError();                                                                       \
        if (e_ != hipSuccess)                                                                                          \
            return fprt::fail(FP_ERR_RUNTIME, "kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), __FILE__,          \
                        __LINE__);                                                                                     \
    } while (0)

// ---------------------------------------------------------------------------
// Operator sets
// ---------------------------------------------------------------------------
struct fp_hip_ops
{
    size_t n_qubits = 0;
    size_t n_strings = 0;
    size_t n_operators = 0;
    int device = 0;
    std::vector<uint64_t> x, z;
    std::vector<uint32_t> ny;
    std::vector<double> coeffs;  // complex128, n_strings
    std::vector<double> cf128;   // folded c_s (-i)^{nY_s}, complex128 (host copy of d_cf128)
    std::vector<float> cf64;     // the same rounded to complex64 (host copy of d_cf64)
    std::vector<double> matrix;  // complex128, n_strings x n_operators
    // Per-string tables live in one device arena (re-targeted in place by
    // fp_hip_ops_assign, reallocated only when it must grow) mirrored on the host.
    void *d_arena = nullptr;
    size_t arena_bytes = 0;
    std::vector<unsigned char> h_arena;
    hipEvent_t upload_done = nullptr; // completes when the last upload of h_arena has been read
    // device copies (pointers into the arena)
    uint64_t *d_x = nullptr, *d_z = nullptr;
    fpk::StringMeta *d_meta = nullptr;
    double *d_cf128 = nullptr;   // c_s * (-i)^{nY_s}, complex128
    float *d_cf64 = nullptr;     // same, complex64
    double *d_scale = nullptr;   // c_s * kappa_s (complex, FP64): PauliOp EV
    double *d_kappa_c = nullptr; // (kappa_s, 0): PauliString EV (coefficient applied after)
    double *d_kappa = nullptr;   // kappa_s (real): per-string EV matrix
    fpl::EvChunk *d_ev_chunks = nullptr; // strings grouped by x for the LDS expectation kernel
    fpl::EvStr *d_ev_strs = nullptr;
    uint32_t n_ev_chunks = 0;
    uint32_t n_ev_wht = 0; // the first n_ev_wht chunks are transform classes
    double *d_rowsum128 = nullptr; // SummedPauliOp::apply folded coefficients (complex128)
    float *d_rowsum64 = nullptr;
    std::vector<double> rowsum128; // host copies of the same (the tiled path's tables)
    std::vector<float> rowsum64;
    // contraction operands: W^T as (n_operators x n_strings) real and imaginary parts
    double *d_wr64 = nullptr, *d_wi64 = nullptr;
    float *d_wr32 = nullptr, *d_wi32 = nullptr;
    bool matrix_imag_zero = true;
    // h(j, k) interleaved complex, (n_strings x n_operators), for apply_weighted
    double *d_h128 = nullptr;
    float *d_h64 = nullptr;
    uint32_t *d_ny = nullptr;
    // string groups of the tiled expectation kernel, built on first use per geometry
    mutable std::mutex tiles_mu;
    mutable std::list<fprt::TileTables> tiles;
    // bumped whenever strings, coefficients or the matrix change, so that
    // batches built over this set can tell that their tables are stale
    uint64_t generation = 0;
};
