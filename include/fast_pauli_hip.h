/*
 * fast_pauli_hip.h — C ABI of the MI355X (gfx950) operator-application library
 * (libfastpauli_hip.so).
 *
 * This is the drop-in boundary.  Everything above it (the C++ host API in
 * fast-pauli_amd/cpp/include/fast_pauli/ and the Python package built on it)
 * reaches the HIP kernels only through these functions.  Signatures use plain
 * pointers and sizes; no HIP, torch or C++ types cross the boundary and no C++
 * exception escapes it.  Every function returns an fp_status; the message of
 * the last failure on the calling thread is available from
 * fp_hip_last_error().
 *
 * Conventions
 *  - A batch of states is a row-major (dim x n_states) complex array, element
 *    (j, t) at j * n_states + t, exactly the C-contiguous layout the reference
 *    enforces (fast_pauli/cpp/include/__nb_helpers.hpp:55-74).  dim = 2^n.
 *  - Complex values are interleaved (re, im): FP_C128 = two doubles,
 *    FP_C64 = two floats.
 *  - Pauli strings are given as letter codes (0 = I, 1 = X, 2 = Y, 3 = Z,
 *    fast_pauli/cpp/include/__pauli.hpp:41), n_qubits codes per string,
 *    character 0 first; character i acts on bit (n-1-i) of the row index
 *    (fast_pauli/cpp/include/__pauli_string.hpp:52-54).
 *  - Compute entry points take DEVICE pointers and are asynchronous on the
 *    given HIP stream (NULL = the null stream).  `accumulate` = 1 adds into
 *    the output (out += result, the reference C++ API's semantics); 0
 *    overwrites it (what the reference's Python layer gets from its fresh,
 *    zeroed outputs).
 *  - Library-owned scratch (reduction partials) is kept per stream.  Calls
 *    on one stream are ordered; calls on different streams may overlap.
 */
#ifndef FAST_PAULI_HIP_H
#define FAST_PAULI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int fp_status;
enum
{
    FP_OK = 0,
    FP_ERR_INVALID_ARGUMENT = 1, /* maps to std::invalid_argument / ValueError */
    FP_ERR_RUNTIME = 2,          /* HIP runtime failure */
    FP_ERR_OUT_OF_MEMORY = 3,
    FP_ERR_NO_DEVICE = 4,
    FP_ERR_UNSUPPORTED = 5
};

/* element types */
enum
{
    FP_C128 = 0, /* complex128, FP64 arithmetic (the reference's precision) */
    FP_C64 = 1   /* complex64, FP32 arithmetic with FP32 accumulation and an FP64 final cross-row reduction */
};

/* precision of SummedPauliOp's (operators x strings) x (strings x states) contraction */
enum
{
    FP_CONTRACT_FP64 = 0,      /* FP64 MFMA (plain / oracle mode) */
    FP_CONTRACT_FP32 = 1,      /* FP32 MFMA */
    FP_CONTRACT_BF16X3 = 2,    /* BF16 split-precision MFMA: hi*hi + hi*lo + lo*hi, FP32 accumulation */
    FP_CONTRACT_AUTO = 3       /* FP64 for FP_C128 states, FP32 for FP_C64 */
};

/* memcpy kinds */
enum
{
    FP_H2D = 0,
    FP_D2H = 1,
    FP_D2D = 2
};

typedef struct fp_hip_ops fp_hip_ops; /* device-resident set of Pauli strings and coefficients */

/* ---- library and device --------------------------------------------------- */
const char *fp_hip_version(void);
fp_status fp_hip_init(int device);                 /* selects and initialises `device` */
fp_status fp_hip_device_count(int *count);
size_t fp_hip_last_error(char *buf, size_t len);   /* copies the last message; returns its full length */

/* ---- memory and streams (so callers need no HIP headers) ------------------ */
fp_status fp_hip_malloc(void **ptr, size_t bytes);
fp_status fp_hip_free(void *ptr);
fp_status fp_hip_memcpy(void *dst, void const *src, size_t bytes, int kind, void *stream);
fp_status fp_hip_memset(void *ptr, int value, size_t bytes, void *stream);
fp_status fp_hip_stream_create(void **stream);
fp_status fp_hip_stream_destroy(void *stream);
fp_status fp_hip_stream_synchronize(void *stream);
fp_status fp_hip_device_synchronize(void);
fp_status fp_hip_is_device_pointer(void const *ptr, int *is_device);
fp_status fp_hip_mem_info(size_t *free_bytes, size_t *total_bytes);
/* HIP events, for timing work on the stream it is launched on */
fp_status fp_hip_event_create(void **event);
fp_status fp_hip_event_destroy(void *event);
fp_status fp_hip_event_record(void *event, void *stream);
fp_status fp_hip_event_synchronize(void *event);
fp_status fp_hip_event_query(void *event, int *done); /* done = 1 once the work before the record has finished */
fp_status fp_hip_event_elapsed_ms(void *start, void *stop, float *ms);

/* ---- seeded synthetic inputs ------------------------------------------------
 * Fills a (dim x n_states) batch with re/im ~ N(0,1) from counter-based
 * splitmix64 + Box-Muller (element e, component c uses stream index 2e+c) and,
 * if normalise != 0, scales every column to unit 2-norm (FP64 norms). */
fp_status fp_hip_fill_states(int dtype, void *states, size_t dim, size_t n_states, uint64_t seed, int normalise,
                             void *stream);

/* Plain device-to-device copy with a streaming kernel (16-byte aligned, multiple of 16 bytes):
 * the achievable-bandwidth reference for the roofline and for scaling byte counters. */
fp_status fp_hip_copy(void *dst, void const *src, size_t bytes, void *stream);

/* Element-wise conversion between complex128 and complex64 batches on the device
 * (round to nearest when narrowing).  n_elements complex values. */
fp_status fp_hip_convert(int src_dtype, void const *src, int dst_dtype, void *dst, size_t n_elements, void *stream);

/* ---- operator sets ------------------------------------------------------------
 * A set of n_strings Pauli strings on n_qubits (1..62) qubits with one complex128
 * coefficient per string (NULL = all ones).  Replaces the per-call
 * get_sparse_repr materialisation (fast_pauli/cpp/include/__pauli_string.hpp:49-118):
 * the kernels evaluate column index j^x and phase (-i)^{nY} (-1)^{popcount(j&z)}
 * from per-string bit masks instead. */
fp_status fp_hip_ops_create(fp_hip_ops **ops, size_t n_qubits, size_t n_strings, uint8_t const *codes,
                            double const *coeffs);
fp_status fp_hip_ops_destroy(fp_hip_ops *ops);
/* Re-target an existing set to new strings and coefficients, uploading on `stream`
 * (asynchronously when the host tables were not reused by a call still in flight);
 * allocates only when the set grows.  No reference counterpart: lets a serving loop
 * reuse operator sets without allocating.  Drops an attached coefficient matrix.
 * Work that uses the previous strings must have completed, or be ordered before the
 * next use of the set (the same stream): device tables are rewritten in place. */
fp_status fp_hip_ops_assign(fp_hip_ops *ops, size_t n_qubits, size_t n_strings, uint8_t const *codes,
                            double const *coeffs, void *stream);
/* Replace the per-string coefficients (PauliOp::scale, extend, ...). */
fp_status fp_hip_ops_set_coeffs(fp_hip_ops *ops, double const *coeffs);
/* Attach SummedPauliOp's coefficient matrix h(j, k): complex128, (n_strings x n_operators) row-major,
 * the layout of fast_pauli/cpp/include/__summed_pauli_op.hpp:45. */
fp_status fp_hip_ops_set_matrix(fp_hip_ops *ops, size_t n_operators, double const *coeffs_nk);
fp_status fp_hip_ops_info(fp_hip_ops const *ops, size_t *n_qubits, size_t *n_strings, size_t *n_operators);

/* ---- hot path ---------------------------------------------------------------- */

/* PauliString::apply_batch (fast_pauli/cpp/include/__pauli_string.hpp:396-436) and, with
 * n_states = 1, PauliString::apply (:312-341): out(j,t) (+)= c * P_s psi, P_s = string `string_index`.
 * Bit-exact with the reference's FP arithmetic (no fused multiply-add). */
fp_status fp_hip_string_apply_batch(fp_hip_ops const *ops, size_t string_index, int dtype, void const *states,
                                    void *out, size_t dim, size_t n_states, double c_re, double c_im,
                                    int accumulate, void *stream);

/* PauliString::expectation_value (fast_pauli/cpp/include/__pauli_string.hpp:488-538):
 * out[t] (+)= c * <psi_t| P_s |psi_t>, out has n_states complex elements of `dtype`. */
fp_status fp_hip_string_expectation(fp_hip_ops const *ops, size_t string_index, int dtype, void const *states,
                                    void *out, size_t dim, size_t n_states, double c_re, double c_im,
                                    int accumulate, void *stream);

/* PauliOp::apply (fast_pauli/cpp/include/__pauli_op.hpp:376-468): out (+)= sum_s h_s P_s psi.
 * Strings are accumulated in order s = 0..N-1 per element, which is the reference's
 * sequential order, so results are bit-exact with it. */
fp_status fp_hip_op_apply(fp_hip_ops const *ops, int dtype, void const *states, void *out, size_t dim,
                          size_t n_states, int accumulate, void *stream);

/* PauliOp::expectation_value (fast_pauli/cpp/include/__pauli_op.hpp:496-549):
 * out[t] (+)= sum_s h_s <psi_t| P_s |psi_t>. */
fp_status fp_hip_op_expectation(fp_hip_ops const *ops, int dtype, void const *states, void *out, size_t dim,
                                size_t n_states, int accumulate, void *stream);

/* Size of the static device workspace of fp_hip_op_expectation_noalloc (per GPU). */
#define FP_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)

/* fp_hip_op_expectation without any device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same kernels, same results as fp_hip_op_expectation (identical bits
 * whenever the partials fit beside the tables without fewer blocks; otherwise a regrouped FP64
 * sum); the reduction partials and the string-group tables live in one static
 * FP_NOALLOC_SCRATCH_BYTES device array per GPU instead of library-allocated memory.  Shapes
 * whose tables and partials (tiled: groups x n_states x 16 B at least; streaming:
 * n_states x 16 B at least) exceed it fail with FP_ERR_INVALID_ARGUMENT before any kernel runs.
 * The first call for an operator and a (dtype, n_states) geometry builds the host-side string
 * groups, as fp_hip_op_expectation does, and caches them in the operator; the tables are copied
 * into the workspace on `stream` by every call.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (on different streams, or from several host threads) race. */
fp_status fp_hip_op_expectation_noalloc(fp_hip_ops const *ops, int dtype, void const *states, void *out,
                                        size_t dim, size_t n_states, int accumulate, void *stream);

/* How fp_hip_op_expectation evaluates these shapes.  No counterpart in the reference:
 * diagnostics and roofline accounting for callers that measure the kernel.
 *   kernel 0 (FP_EV_STREAM): every string reads the whole state once.
 *   kernel 1 (FP_EV_TILED):  strings are grouped so that the x masks of a group span a subspace
 *     of dimension tile_rows_log2 (with its low bits); each coset of it is a tile of
 *     2^tile_rows_log2 rows x states_per_block states, loaded once into LDS per group.
 * hbm_bytes: state bytes the chosen kernel reads per call, (strings or groups) x 2^n x S x B.
 * shared_groups: tiled groups whose strings all pair rows with one common local bit clear, so
 *   that one LDS read of a leader row serves every string (the fastest instance).
 * hoisted_groups: other tiled groups whose waves each own at most one single-string unit, set
 *   up once per block; the remaining groups set units up per tile. */
enum
{
    FP_EV_STREAM = 0,
    FP_EV_TILED = 1
};
typedef struct fp_hip_ev_plan
{
    int kernel;
    uint32_t n_groups;
    uint32_t tile_rows_log2;
    uint32_t states_per_block;
    uint64_t hbm_bytes;
    uint32_t hoisted_groups;
    uint32_t shared_groups;
} fp_hip_ev_plan;
fp_status fp_hip_op_expectation_plan(fp_hip_ops const *ops, int dtype, size_t dim, size_t n_states,
                                     fp_hip_ev_plan *plan);

/* Which kernel fp_hip_op_apply runs for these shapes, in the same record (no reference
 * counterpart: for measurement and triage).  kernel 0 (FP_EV_STREAM): every string reads the
 * state once and the output is written once.  kernel 1 (FP_EV_TILED): groups of consecutive
 * strings, one launch each, every group reading the state and reading and writing the output
 * (the first group does not read it unless accumulating).  hbm_bytes counts those reads and
 * writes; hoisted_groups and shared_groups are 0. */
fp_status fp_hip_op_apply_plan(fp_hip_ops const *ops, int dtype, size_t dim, size_t n_states, int accumulate,
                               fp_hip_ev_plan *plan);
/* The same for fp_hip_summed_apply (the row-summed coefficients) and for
 * fp_hip_summed_apply_weighted (per-state weights; the weight computation's bytes are not
 * counted); no reference counterparts. */
fp_status fp_hip_summed_apply_plan(fp_hip_ops const *ops, int dtype, size_t dim, size_t n_states, int accumulate,
                                   fp_hip_ev_plan *plan);
fp_status fp_hip_summed_apply_weighted_plan(fp_hip_ops const *ops, int dtype, size_t dim, size_t n_states,
                                            int accumulate, fp_hip_ev_plan *plan);

/* Per-string expectation values without coefficients, as a REAL (n_strings x n_states) matrix
 * ev(j,t) = <psi_t|P_j|psi_t> (FP64 for FP_C128, FP32 for FP_C64).  Pauli strings are Hermitian,
 * so the imaginary part is exactly zero.  Stage 1 of SummedPauliOp::expectation_value
 * (fast_pauli/cpp/include/__summed_pauli_op.hpp:572-577). */
fp_status fp_hip_strings_expectation_matrix(fp_hip_ops const *ops, int dtype, void const *states, void *ev,
                                            size_t dim, size_t n_states, void *stream);

/* SummedPauliOp::expectation_value (fast_pauli/cpp/include/__summed_pauli_op.hpp:530-614):
 * out(k,t) (+)= sum_j h(j,k) <psi_t|P_j|psi_t>; out is (n_operators x n_states) of `dtype`.
 * `contraction` selects the MFMA precision of stage 2 (FP_CONTRACT_*). */
fp_status fp_hip_summed_expectation(fp_hip_ops const *ops, int dtype, int contraction, void const *states,
                                    void *out, size_t dim, size_t n_states, int accumulate, void *stream);

/* SummedPauliOp::apply (fast_pauli/cpp/include/__summed_pauli_op.hpp:287-349):
 * out (+)= sum_j (sum_k h(j,k)) P_j psi. */
fp_status fp_hip_summed_apply(fp_hip_ops const *ops, int dtype, void const *states, void *out, size_t dim,
                              size_t n_states, int accumulate, void *stream);

/* SummedPauliOp::apply_weighted (fast_pauli/cpp/include/__summed_pauli_op.hpp:378-503):
 * out(:,t) (+)= sum_j (sum_k h(j,k) x(k,t)) P_j psi(:,t); `data` x is a real (n_operators x n_states)
 * row-major device array (double for FP_C128, float for FP_C64). */
fp_status fp_hip_summed_apply_weighted(fp_hip_ops const *ops, int dtype, void const *states, void const *data,
                                       void *out, size_t dim, size_t n_states, int accumulate, void *stream);

/* SummedPauliOp::square (fast_pauli/cpp/include/__summed_pauli_op.hpp:197-268): the operators'
 * squares A_k^2 = sum_a h'(a,k) P_a over the strings of weight <= W = min(n, 2 * the set's largest
 * weight), in the order of calculate_pauli_strings_max_weight(n, W)
 * (fast_pauli/cpp/include/__pauli_helpers.hpp:147-153).  *n_out_strings receives their count;
 * with out = NULL nothing else happens (a size query).  out: (n_out_strings x n_operators) complex of
 * `dtype` on the device, overwritten.  Each h'(a,k) sums phase_ij h(i,k) h(j,k) over the pairs with
 * P_i P_j = phase_ij P_a in (i, j) order, as (phase * h(i,k)) * h(j,k), the reference's operations in
 * its order, so results are IEEE-equal to its sequential loop.  Limits: n_strings^2 < 2^31 and
 * n_out_strings < 2^32 (FP_ERR_UNSUPPORTED beyond). */
fp_status fp_hip_summed_square(fp_hip_ops const *ops, int dtype, void *out, size_t *n_out_strings, void *stream);

/* ---- dense and sparse materialisation on the device (debugging, interop) --------
 * Values are IEEE-equal to the reference's: m(j) = (-i)^{nY} (-1)^{popcount(j & z)} formed
 * as the reference forms it, c * m as std::complex multiplies, strings added in order. */

/* get_sparse_repr (fast_pauli/cpp/include/__pauli_string.hpp:49-118) of string `string_index`:
 * cols[j] = j ^ x (dim uint64 entries), vals[j] = m(j) (dim complex values of `dtype`). */
fp_status fp_hip_string_sparse_repr(fp_hip_ops const *ops, size_t string_index, int dtype, uint64_t *cols,
                                    void *vals, size_t dim, void *stream);

/* PauliString::to_tensor (fast_pauli/cpp/include/__pauli_string.hpp:549-560): out(j, j ^ x) = m(j)
 * for every row j of the (dim x dim) row-major matrix `out`; other elements are left unchanged. */
fp_status fp_hip_string_to_tensor(fp_hip_ops const *ops, size_t string_index, int dtype, void *out, size_t dim,
                                  void *stream);

/* PauliOp::to_tensor (fast_pauli/cpp/include/__pauli_op.hpp:558-571): out (dim x dim) (+)=
 * sum_s c_s P_s, strings in order (accumulate = 0 clears out first). */
fp_status fp_hip_op_to_tensor(fp_hip_ops const *ops, int dtype, void *out, size_t dim, int accumulate,
                              void *stream);

/* SummedPauliOp::to_tensor (fast_pauli/cpp/include/__summed_pauli_op.hpp:644-665): out
 * (n_operators x dim x dim) (+)= sum_s h(s,k) P_s for each operator k, strings in order. */
fp_status fp_hip_summed_to_tensor(fp_hip_ops const *ops, int dtype, void *out, size_t dim, int accumulate,
                                  void *stream);

/* ---- uneven batches: many operators in one persistent, graph-captured launch ----
 * Entries may differ in qubit count, string count, state count, dtype and kind.
 * Outputs of one entry must not overlap any other entry's buffers.  The first
 * launch (or the first after a change) cuts the entries into tiles, orders them,
 * and captures [counter reset, persistent kernel] into a hipGraph; later launches
 * replay the graph.  No reference counterpart (the reference runs one operator
 * per call). */
typedef struct fp_hip_batch fp_hip_batch;
enum
{
    FP_BATCH_APPLY = 0,  /* PauliOp::apply:  out (+)= sum_s h_s P_s psi */
    FP_BATCH_EXPECT = 1  /* PauliOp::expectation_value: out[t] (+)= sum_s h_s <psi_t|P_s|psi_t> */
};
enum
{
    FP_BATCH_ORDER_LARGEST_FIRST = 0,
    FP_BATCH_ORDER_AS_ADDED = 1
};
typedef struct fp_hip_batch_schedule
{
    uint32_t tile_vectors;  /* 16-byte vectors per apply tile */
    uint64_t ev_tile_work;  /* strings x pairs x vectors-per-row per expectation tile */
    uint32_t blocks_per_cu; /* persistent blocks per compute unit */
    int order;              /* FP_BATCH_ORDER_* */
    uint32_t coset_tiles;   /* coset apply items: 64 KiB tiles (cosets) per item, >= 1 */
    uint32_t coset_min_kib; /* apply entries with states of at least this many KiB run as
                               coset items (string groups sharing LDS tiles) when that moves
                               fewer bytes; 0: never */
    uint32_t tiled_min_kib; /* apply entries with states of at least this many KiB whose string
                               groups pay (as in fp_hip_op_apply) run as coset-tiled launches of
                               their own, one per group, after the persistent kernel in the same
                               graph; they take precedence over coset items; 0: never */
} fp_hip_batch_schedule;

fp_status fp_hip_batch_create(fp_hip_batch **batch);
fp_status fp_hip_batch_destroy(fp_hip_batch *batch);
fp_status fp_hip_batch_clear(fp_hip_batch *batch);
fp_status fp_hip_batch_add(fp_hip_batch *batch, int kind, fp_hip_ops const *ops, int dtype, void const *states,
                           void *out, size_t dim, size_t n_states, int accumulate);
fp_status fp_hip_batch_set_schedule(fp_hip_batch *batch, fp_hip_batch_schedule const *schedule);
fp_status fp_hip_batch_get_schedule(fp_hip_batch const *batch, fp_hip_batch_schedule *schedule);
fp_status fp_hip_batch_info(fp_hip_batch const *batch, size_t *n_entries, size_t *n_tiles, size_t *n_blocks);
/* Of the last build: entries run as coset items, their string groups and items, and the
 * plan's state/output bytes per launch (coset entries: 3 per string group, less the first
 * group's output read unless accumulating; streaming apply: strings + 1 (+ 1 accumulating);
 * expectation: one state pass per string). */
fp_status fp_hip_batch_coset_info(fp_hip_batch const *batch, size_t *n_entries, size_t *n_groups, size_t *n_items,
                                  double *plan_bytes);
/* Of the last build: entries run as coset-tiled launches of their own, and their string
 * groups (one launch each). */
fp_status fp_hip_batch_tiled_info(fp_hip_batch const *batch, size_t *n_entries, size_t *n_groups);
/* Waits for the batch's last launch; FP_ERR_RUNTIME if a coset item of any launch since the
 * last build gave up waiting for the previous string group (an internal error: its outputs
 * are then unspecified). */
fp_status fp_hip_batch_check(fp_hip_batch *batch);
fp_status fp_hip_batch_launch(fp_hip_batch *batch, void *stream);
/* enabled = 1 (default): launches replay a captured hipGraph; 0: kernels are enqueued
 * directly (cheaper when every launch has new entries, as in a serving loop). */
fp_status fp_hip_batch_set_graph(fp_hip_batch *batch, int enabled);
/* Times schedule candidates (coordinate descent, `reps` launches each) and keeps the fastest.
 * Trial launches write the outputs. */
fp_status fp_hip_batch_autotune(fp_hip_batch *batch, void *stream, int reps, fp_hip_batch_schedule *chosen,
                                float *best_ms);

/* ---- serving: continuous batching of apply / expectation requests ----------------
 * A scheduler thread admits queued requests (first fit, arrival order) whose device
 * memory fits under `hbm_budget_bytes`, and runs each admitted group as one uneven
 * batch on one of `n_slots` streams; requests arriving meanwhile go into the next free
 * slot, so batches overlap on the device.  States are device-resident: either the
 * caller's device pointer (`states`) or created on the device from `state_seed`
 * (normalised N(0,1) columns).  Latency is measured from submit to completion.
 * No reference counterpart. */
typedef struct fp_hip_server fp_hip_server;
typedef struct fp_serve_config
{
    size_t hbm_budget_bytes; /* admission limit for request buffers */
    int n_slots;             /* concurrent batches (streams), 1..16 */
    int max_batch;           /* requests per batch */
    int poll_us;             /* completion polling interval while batches run */
} fp_serve_config;
typedef struct fp_serve_request
{
    int kind;              /* FP_BATCH_APPLY or FP_BATCH_EXPECT */
    int dtype;             /* FP_C64 or FP_C128 */
    size_t n_qubits;
    size_t n_strings;
    uint8_t const *codes;  /* n_strings x n_qubits letter codes (copied at submit) */
    double const *coeffs;  /* complex128 per string, or NULL for ones (copied) */
    size_t n_states;
    uint64_t state_seed;   /* states created on the device when `states` is NULL */
    void const *states;    /* optional caller-owned device states (dim x n_states) */
    int keep_output;       /* apply: keep the output on the device for fp_hip_server_fetch */
} fp_serve_request;
typedef struct fp_serve_result
{
    int done;
    fp_status status;
    double latency_s; /* submit -> completion */
    double queue_s;   /* submit -> admission */
    uint64_t batch_index;
} fp_serve_result;
typedef struct fp_serve_stats
{
    uint64_t submitted, completed, failed, batches, batched_requests, queued;
    size_t peak_bytes_in_use, bytes_allocated;
    int max_concurrent_batches;
} fp_serve_stats;

fp_status fp_hip_server_create(fp_hip_server **server, fp_serve_config const *config);
fp_status fp_hip_server_destroy(fp_hip_server *server); /* drains the queue first */
fp_status fp_hip_server_submit(fp_hip_server *server, fp_serve_request const *request, uint64_t *id);
/* Waits up to timeout_s (negative: forever); result->done = 0 on timeout. */
fp_status fp_hip_server_wait(fp_hip_server *server, uint64_t id, double timeout_s, fp_serve_result *result);
/* Copies an expectation result (n_states complex) or a kept apply output (dim x n_states). */
fp_status fp_hip_server_fetch(fp_hip_server *server, uint64_t id, void *host_out, size_t bytes);
fp_status fp_hip_server_release(fp_hip_server *server, uint64_t id);
fp_status fp_hip_server_stats(fp_hip_server *server, fp_serve_stats *stats);

/* ---- real GEMM on MFMA (stage 2 of SummedPauliOp, exposed for testing) ---------
 * C(MxN) = A(MxK) * B(KxN), all row-major, FP64 for FP_CONTRACT_FP64 else FP32 operands/results. */
fp_status fp_hip_gemm(int contraction, size_t m, size_t n, size_t k, void const *a, void const *b, void *c,
                      void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FAST_PAULI_HIP_H */
