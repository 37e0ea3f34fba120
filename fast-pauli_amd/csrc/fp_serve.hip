// fp_serve.hip — a continuous-batching request loop for apply / expectation
// requests of different sizes, with HBM-budget admission.
//
// A scheduler thread owns `n_slots` streams.  Whenever a slot is free it admits
// queued requests (first fit, in arrival order) whose device memory fits under
// the budget, re-targets pooled operator sets to their strings, creates their
// device-resident states, and runs them together as one uneven batch
// (fp_batch.hip) on the slot's stream.  Requests that arrive while other
// batches are on the device are admitted into the next free slot, so batches
// overlap on the GPU.  Device buffers come from a size-class pool and operator
// sets are reused, so steady-state serving performs no device allocation.
// No reference counterpart: the reference library has no request loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fast_pauli_hip.h"
#include "fp_runtime.hpp"

using fprt::fail;

namespace
{

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Largest state or output buffer of one request (2^50 bytes): far beyond any
// device, and small enough that size classes and budget sums cannot overflow.
constexpr size_t kMaxRequestBytes = size_t(1) << 50;

struct Block
{
    void *ptr = nullptr;
    size_t cls = 0; // size class in bytes (power of two)
};

// Size-class cache of device buffers under a byte budget.
class BufferPool
{
  public:
    explicit BufferPool(size_t budget) : budget_(budget)
    {
    }

    // The power-of-two class holding `bytes` (at least 4 KiB); requests are
    // limited to kMaxRequestBytes at submission, so the class never overflows.
    static size_t size_class(size_t bytes)
    {
        size_t c = 4096;
        while (c < bytes && c <= (SIZE_MAX >> 1))
            c <<= 1;
        return c;
    }

    // Returns false when the budget cannot hold the block now.
    bool get(size_t bytes, Block &out, fp_status &err)
    {
        err = FP_OK;
        size_t const c = size_class(bytes);
        auto it = free_.find(c);
        if (it != free_.end() && !it->second.empty())
        {
            out = Block{it->second.back(), c};
            it->second.pop_back();
            in_use_ += c;
            return true;
        }
        if (allocated_ + c > budget_)
            trim(allocated_ + c - budget_);
        if (allocated_ + c > budget_)
            return false;
        void *p = nullptr;
        hipError_t const e = hipMalloc(&p, c);
        if (e != hipSuccess)
        {
            (void)hipGetLastError();
            trim(allocated_); // give cached memory back and retry once
            if (hipMalloc(&p, c) != hipSuccess)
            {
                (void)hipGetLastError();
                err = FP_ERR_OUT_OF_MEMORY;
                return false;
            }
        }
        allocated_ += c;
        in_use_ += c;
        out = Block{p, c};
        return true;
    }

    void put(Block const &b)
    {
        if (!b.ptr)
            return;
        free_[b.cls].push_back(b.ptr);
        in_use_ -= b.cls;
    }

    size_t in_use() const
    {
        return in_use_;
    }
    size_t allocated() const
    {
        return allocated_;
    }
    size_t budget() const
    {
        return budget_;
    }

    ~BufferPool()
    {
        trim(allocated_);
    }

  private:
    // Frees cached (unused) blocks until `want` bytes are released or none remain.
    void trim(size_t want)
    {
        size_t freed = 0;
        for (auto &kv : free_)
        {
            while (!kv.second.empty() && freed < want)
            {
                (void)hipFree(kv.second.back());
                kv.second.pop_back();
                allocated_ -= kv.first;
                freed += kv.first;
            }
        }
    }

    size_t budget_;
    size_t allocated_ = 0;
    size_t in_use_ = 0;
    std::map<size_t, std::vector<void *>> free_;
};

enum Phase
{
    kQueued,
    kRunning,
    kDone,
    kFailed
};

struct Request
{
    uint64_t id = 0;
    fp_serve_request spec{};
    std::vector<uint8_t> codes;
    std::vector<double> coeffs;
    Phase phase = kQueued;
    double t_submit = 0, t_start = 0, t_done = 0;
    Block states, out;
    size_t reserve = 0;
    fp_hip_ops *ops = nullptr;
    std::vector<unsigned char> host_result; // expectation values (small), copied at completion
    bool keep_output = false;
    fp_status status = FP_OK;
    std::string error;
    uint64_t batch_index = 0;
};

struct Slot
{
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    fp_hip_batch *batch = nullptr;
    std::vector<Request *> reqs;
    bool busy = false;
};

size_t elem_bytes(int dtype)
{
    return dtype == FP_C128 ? 16 : 8;
}

} // namespace

struct fp_hip_server
{
    fp_serve_config cfg{};
    int device = 0;
    std::unique_ptr<BufferPool> pool;
    std::vector<Slot> slots;
    std::vector<fp_hip_ops *> ops_free;
    std::vector<fp_hip_ops *> ops_all;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    // statistics
    uint64_t n_submitted = 0, n_done = 0, n_failed = 0, n_batches = 0, n_batched_requests = 0;
    size_t peak_in_use = 0;
    int max_concurrent_batches = 0;
    std::string fatal;
};

namespace
{

fp_hip_ops *take_ops(fp_hip_server *s, Request const &r, hipStream_t stream, fp_status &st)
{
    fp_hip_ops *o = nullptr;
    if (!s->ops_free.empty())
    {
        o = s->ops_free.back();
        s->ops_free.pop_back();
        st = fp_hip_ops_assign(o, r.spec.n_qubits, r.spec.n_strings, r.codes.data(),
                               r.coeffs.empty() ? nullptr : r.coeffs.data(), stream);
        if (st != FP_OK)
        {
            s->ops_free.push_back(o);
            return nullptr;
        }
        return o;
    }
    st = fp_hip_ops_create(&o, r.spec.n_qubits, r.spec.n_strings, r.codes.data(),
                           r.coeffs.empty() ? nullptr : r.coeffs.data());
    if (st != FP_OK)
        return nullptr;
    s->ops_all.push_back(o);
    return o;
}

void finish_request(fp_hip_server *s, Request *r, fp_status st, char const *msg)
{
    r->t_done = now_s();
    r->status = st;
    if (st != FP_OK)
    {
        r->phase = kFailed;
        r->error = msg ? msg : "request failed";
        ++s->n_failed;
    }
    else
    {
        r->phase = kDone;
        ++s->n_done;
    }
}

void release_request_memory(fp_hip_server *s, Request *r)
{
    s->pool->put(r->states);
    r->states = Block{};
    if (!r->keep_output)
    {
        s->pool->put(r->out);
        r->out = Block{};
    }
    if (r->ops)
    {
        s->ops_free.push_back(r->ops);
        r->ops = nullptr;
    }
}

// Completed slots: copy small results, recycle memory.  Called with the lock held.
void reap(fp_hip_server *s)
{
    for (Slot &sl : s->slots)
    {
        if (!sl.busy)
            continue;
        hipError_t const q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady)
        {
            (void)hipGetLastError();
            continue;
        }
        fp_status st = FP_OK;
        char const *msg = nullptr;
        if (q != hipSuccess)
        {
            st = FP_ERR_RUNTIME;
            msg = hipGetErrorString(q);
        }
        for (Request *r : sl.reqs)
        {
            if (st == FP_OK && r->spec.kind == FP_BATCH_EXPECT)
            {
                size_t const bytes = r->spec.n_states * elem_bytes(r->spec.dtype);
                r->host_result.resize(bytes);
                if (hipMemcpy(r->host_result.data(), r->out.ptr, bytes, hipMemcpyDeviceToHost) != hipSuccess)
                {
                    st = FP_ERR_RUNTIME;
                    msg = "result copy failed";
                }
            }
            finish_request(s, r, st, msg);
            release_request_memory(s, r);
        }
        sl.reqs.clear();
        sl.busy = false;
    }
}

// Admits queued requests into one free slot and launches them as a batch.
void admit(fp_hip_server *s, Slot &sl)
{
    std::vector<Request *> take;
    for (auto it = s->queue.begin(); it != s->queue.end() && int(take.size()) < s->cfg.max_batch;)
    {
        Request *r = *it;
        size_t const esz = elem_bytes(r->spec.dtype);
        size_t const in_b = (size_t(1) << r->spec.n_qubits) * r->spec.n_states * esz;
        size_t const out_b = (r->spec.kind == FP_BATCH_APPLY ? (size_t(1) << r->spec.n_qubits) : 1) *
                             r->spec.n_states * esz;
        size_t const need = (r->spec.states ? 0 : BufferPool::size_class(in_b)) + BufferPool::size_class(out_b);
        if (need > s->pool->budget())
        {
            it = s->queue.erase(it);
            finish_request(s, r, FP_ERR_OUT_OF_MEMORY, "request exceeds the HBM budget");
            continue;
        }
        fp_status err = FP_OK;
        Block bin, bout;
        bool ok = true;
        if (!r->spec.states)
            ok = s->pool->get(in_b, bin, err);
        if (ok && !s->pool->get(out_b, bout, err))
        {
            s->pool->put(bin);
            bin = Block{};
            ok = false;
        }
        if (!ok)
        {
            if (err != FP_OK)
            {
                it = s->queue.erase(it);
                finish_request(s, r, err, "device allocation failed");
                continue;
            }
            ++it; // does not fit now; later requests may
            continue;
        }
        r->states = bin;
        r->out = bout;
        r->reserve = need;
        take.push_back(r);
        it = s->queue.erase(it);
    }
    if (take.empty())
        return;
    s->peak_in_use = std::max(s->peak_in_use, s->pool->in_use());
    fp_status st = fp_hip_batch_clear(sl.batch);
    double const t0 = now_s();
    // Requests whose strings group into shared coset tiles run through the tiled
    // kernels (expectation: one state read per group; apply on large states: one
    // state read and an output read + write per group) after the batch, on the
    // same stream; everything else goes into the persistent batch launch.
    std::vector<Request *> tiled;
    size_t n_batched = 0;
    for (Request *r : take)
    {
        r->phase = kRunning;
        r->t_start = t0;
        r->batch_index = s->n_batches;
        if (st == FP_OK)
        {
            r->ops = take_ops(s, *r, sl.stream, st);
        }
        void const *in = r->spec.states ? r->spec.states : r->states.ptr;
        size_t const dim = size_t(1) << r->spec.n_qubits;
        if (st == FP_OK && !r->spec.states)
            st = fp_hip_fill_states(r->spec.dtype, r->states.ptr, dim, r->spec.n_states, r->spec.state_seed, 1,
                                    sl.stream);
        bool use_tiled = false;
        if (st == FP_OK && r->spec.kind == FP_BATCH_EXPECT)
            st = fprt::choose_op_expect_tiled(r->ops, r->spec.dtype, r->spec.n_states, &use_tiled);
        else if (st == FP_OK)
            st = fprt::choose_op_apply_tiled(r->ops, r->spec.dtype, r->spec.n_states, &use_tiled);
        if (st == FP_OK && use_tiled)
            tiled.push_back(r);
        else if (st == FP_OK)
        {
            st = fp_hip_batch_add(sl.batch, r->spec.kind, r->ops, r->spec.dtype, in, r->out.ptr, dim,
                                  r->spec.n_states, 0);
            ++n_batched;
        }
        sl.reqs.push_back(r);
    }
    if (st == FP_OK && n_batched > 0)
        st = fp_hip_batch_launch(sl.batch, sl.stream);
    for (Request *r : tiled)
    {
        if (st != FP_OK)
            break;
        void const *in = r->spec.states ? r->spec.states : r->states.ptr;
        size_t const dim = size_t(1) << r->spec.n_qubits;
        st = r->spec.kind == FP_BATCH_EXPECT
                 ? fp_hip_op_expectation(r->ops, r->spec.dtype, in, r->out.ptr, dim, r->spec.n_states, 0, sl.stream)
                 : fp_hip_op_apply(r->ops, r->spec.dtype, in, r->out.ptr, dim, r->spec.n_states, 0, sl.stream);
    }
    if (st == FP_OK && hipEventRecord(sl.done, sl.stream) != hipSuccess)
        st = FP_ERR_RUNTIME;
    if (st != FP_OK)
    {
        std::string const msg = "batch launch failed";
        (void)hipStreamSynchronize(sl.stream);
        for (Request *r : sl.reqs)
        {
            finish_request(s, r, st, msg.c_str());
            release_request_memory(s, r);
        }
        sl.reqs.clear();
        return;
    }
    sl.busy = true;
    ++s->n_batches;
    s->n_batched_requests += take.size();
    int busy = 0;
    for (Slot const &x : s->slots)
        busy += x.busy ? 1 : 0;
    s->max_concurrent_batches = std::max(s->max_concurrent_batches, busy);
}

void worker_loop(fp_hip_server *s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lock(s->mu);
    while (true)
    {
        reap(s);
        for (Slot &sl : s->slots)
            if (!sl.busy && !s->queue.empty())
                admit(s, sl);
        s->cv_done.notify_all();
        bool any_busy = false;
        for (Slot const &sl : s->slots)
            any_busy = any_busy || sl.busy;
        if (s->stop && s->queue.empty() && !any_busy)
            break;
        // poll the device while batches run; otherwise sleep until a submission
        if (any_busy)
            s->cv_work.wait_for(lock, std::chrono::microseconds(s->cfg.poll_us));
        else
            s->cv_work.wait_for(lock, std::chrono::milliseconds(50));
    }
    s->cv_done.notify_all();
}

} // namespace

extern "C"
{

fp_status fp_hip_server_create(fp_hip_server **out, fp_serve_config const *cfg)
{
    if (!out || !cfg)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL server or config");
    if (cfg->hbm_budget_bytes == 0 || cfg->n_slots <= 0 || cfg->n_slots > 16 || cfg->max_batch <= 0)
        return fail(FP_ERR_INVALID_ARGUMENT, "invalid server config");
    auto s = std::make_unique<fp_hip_server>();
    s->cfg = *cfg;
    if (s->cfg.poll_us <= 0)
        s->cfg.poll_us = 20;
    FP_HIP(hipGetDevice(&s->device));
    s->pool = std::make_unique<BufferPool>(cfg->hbm_budget_bytes);
    s->slots.resize(size_t(cfg->n_slots));
    for (Slot &sl : s->slots)
    {
        FP_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        FP_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        fp_status st = fp_hip_batch_create(&sl.batch);
        if (st == FP_OK)
            st = fp_hip_batch_set_graph(sl.batch, 0);
        if (st != FP_OK)
            return st;
    }
    fp_hip_server *raw = s.release();
    raw->worker = std::thread(worker_loop, raw);
    *out = raw;
    return FP_OK;
}

fp_status fp_hip_server_submit(fp_hip_server *s, fp_serve_request const *req, uint64_t *id)
{
    if (!s || !req || !id)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL server, request or id");
    if (req->kind != FP_BATCH_APPLY && req->kind != FP_BATCH_EXPECT)
        return fail(FP_ERR_INVALID_ARGUMENT, "unknown request kind %d", req->kind);
    if (req->dtype != FP_C128 && req->dtype != FP_C64)
        return fail(FP_ERR_INVALID_ARGUMENT, "unknown dtype %d", req->dtype);
    if (req->n_qubits == 0 || req->n_qubits > 40 || req->n_strings == 0 || req->n_states == 0 || !req->codes)
        return fail(FP_ERR_INVALID_ARGUMENT, "empty or oversized request");
    if (req->n_strings > 0xffffffffu)
        return fail(FP_ERR_INVALID_ARGUMENT, "too many strings (%zu)", req->n_strings);
    {
        // state and output bytes must be representable, and below the largest pool class
        size_t const esz = elem_bytes(req->dtype);
        size_t const rows_out = req->kind == FP_BATCH_APPLY ? (size_t(1) << req->n_qubits) : 1;
        size_t in_b = 0, out_b = 0;
        if (__builtin_mul_overflow(size_t(1) << req->n_qubits, req->n_states, &in_b) ||
            __builtin_mul_overflow(in_b, esz, &in_b) || __builtin_mul_overflow(rows_out, req->n_states, &out_b) ||
            __builtin_mul_overflow(out_b, esz, &out_b) || in_b > kMaxRequestBytes || out_b > kMaxRequestBytes)
            return fail(FP_ERR_INVALID_ARGUMENT,
                        "request of %zu qubits x %zu states is too large to address (state or output above %zu bytes)",
                        req->n_qubits, req->n_states, kMaxRequestBytes);
    }
    for (size_t i = 0; i < req->n_qubits * req->n_strings; ++i)
        if (req->codes[i] > 3)
            return fail(FP_ERR_INVALID_ARGUMENT, "Pauli code must be 0, 1, 2, or 3");
    auto r = std::make_unique<Request>();
    r->spec = *req;
    r->codes.assign(req->codes, req->codes + req->n_qubits * req->n_strings);
    if (req->coeffs)
        r->coeffs.assign(req->coeffs, req->coeffs + 2 * req->n_strings);
    r->spec.codes = nullptr;
    r->spec.coeffs = nullptr;
    r->keep_output = req->keep_output != 0;
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> lock(s->mu);
        if (s->stop)
            return fail(FP_ERR_INVALID_ARGUMENT, "server is shutting down");
        r->id = s->next_id++;
        *id = r->id;
        s->queue.push_back(r.get());
        s->requests.emplace(r->id, std::move(r));
        ++s->n_submitted;
    }
    s->cv_work.notify_one();
    return FP_OK;
}

fp_status fp_hip_server_wait(fp_hip_server *s, uint64_t id, double timeout_s, fp_serve_result *res)
{
    if (!s || !res)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL server or result");
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end())
        return fail(FP_ERR_INVALID_ARGUMENT, "unknown request id %llu", static_cast<unsigned long long>(id));
    Request *r = it->second.get();
    auto const deadline = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
    while (r->phase == kQueued || r->phase == kRunning)
    {
        if (timeout_s >= 0 && s->cv_done.wait_until(lock, deadline) == std::cv_status::timeout &&
            (r->phase == kQueued || r->phase == kRunning))
        {
            res->done = 0;
            return FP_OK;
        }
        if (timeout_s < 0)
            s->cv_done.wait(lock);
    }
    res->done = 1;
    res->status = r->status;
    res->latency_s = r->t_done - r->t_submit;
    res->queue_s = r->t_start > 0 ? r->t_start - r->t_submit : 0.0;
    res->batch_index = r->batch_index;
    return FP_OK;
}

fp_status fp_hip_server_fetch(fp_hip_server *s, uint64_t id, void *host_out, size_t bytes)
{
    if (!s || !host_out)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL server or output");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end())
        return fail(FP_ERR_INVALID_ARGUMENT, "unknown request id");
    Request *r = it->second.get();
    if (r->phase != kDone)
        return fail(FP_ERR_INVALID_ARGUMENT, "request has not completed successfully");
    if (r->spec.kind == FP_BATCH_EXPECT)
    {
        if (bytes < r->host_result.size())
            return fail(FP_ERR_INVALID_ARGUMENT, "output buffer too small");
        std::memcpy(host_out, r->host_result.data(), r->host_result.size());
        return FP_OK;
    }
    if (!r->out.ptr)
        return fail(FP_ERR_INVALID_ARGUMENT, "apply output was not retained (set keep_output)");
    size_t const need = (size_t(1) << r->spec.n_qubits) * r->spec.n_states * elem_bytes(r->spec.dtype);
    if (bytes < need)
        return fail(FP_ERR_INVALID_ARGUMENT, "output buffer too small");
    FP_HIP(hipMemcpy(host_out, r->out.ptr, need, hipMemcpyDeviceToHost));
    return FP_OK;
}

fp_status fp_hip_server_release(fp_hip_server *s, uint64_t id)
{
    if (!s)
        return fail(FP_ERR_INVALID_ARGUMENT, "server is NULL");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end())
        return FP_OK;
    Request *r = it->second.get();
    if (r->phase == kQueued || r->phase == kRunning)
        return fail(FP_ERR_INVALID_ARGUMENT, "request still in flight");
    s->pool->put(r->out);
    r->out = Block{};
    s->requests.erase(it);
    s->cv_work.notify_one();
    return FP_OK;
}

fp_status fp_hip_server_stats(fp_hip_server *s, fp_serve_stats *out)
{
    if (!s || !out)
        return fail(FP_ERR_INVALID_ARGUMENT, "NULL server or stats");
    std::lock_guard<std::mutex> lock(s->mu);
    out->submitted = s->n_submitted;
    out->completed = s->n_done;
    out->failed = s->n_failed;
    out->batches = s->n_batches;
    out->batched_requests = s->n_batched_requests;
    out->queued = s->queue.size();
    out->peak_bytes_in_use = s->peak_in_use;
    out->bytes_allocated = s->pool->allocated();
    out->max_concurrent_batches = s->max_concurrent_batches;
    return FP_OK;
}

fp_status fp_hip_server_destroy(fp_hip_server *s)
{
    if (!s)
        return FP_OK;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable())
        s->worker.join();
    for (Slot &sl : s->slots)
    {
        if (sl.stream)
            (void)hipStreamSynchronize(sl.stream);
        if (sl.batch)
            (void)fp_hip_batch_destroy(sl.batch);
        if (sl.done)
            (void)hipEventDestroy(sl.done);
        if (sl.stream)
            (void)hipStreamDestroy(sl.stream);
    }
    for (auto &kv : s->requests)
        s->pool->put(kv.second->out);
    for (fp_hip_ops *o : s->ops_all)
        (void)fp_hip_ops_destroy(o);
    delete s;
    return FP_OK;
}

} // extern "C"
