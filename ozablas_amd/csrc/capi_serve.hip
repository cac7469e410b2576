// capi_serve.hip — the request server (include/ozablas_hip.h, oza_server_*): Scheme II GEMM
// requests are queued by any thread and executed in continuous batches.
//
// Two resident workspace arenas alternate: while the batch of one arena runs on the GPU, the
// worker thread admits the next requests (FIFO, as many as fit the arena's bytes and the batch
// limit), plans them with the grouped planner and enqueues them behind it on the server's
// stream, so the GPU does not idle between batches.  A completion thread waits for each batch's
// event in order, stamps its requests done and frees the arena.  The CRT tables stay resident
// in the context.  Results are those of the grouped kernels, i.e. bit-identical to
// oza_scheme2_gemm for the same request alone.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "capi_internal.hpp"
#include "group_internal.hpp"

using namespace oza;
using namespace oza::capi;

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    uint64_t id = 0;
    size_t M = 0, N = 0, K = 0;
    int S = 0, tile = 0;
    size_t bytes = 0;
    const double* A = nullptr;
    const double* B = nullptr;
    double* C = nullptr;
    double t_submit = 0.0, t_done = 0.0;
    bool done = false;
    int status = OZA_SUCCESS;
    std::string error;
};

struct Arena {
    char* ws = nullptr;  // workspace of one batch
    size_t cap = 0;
    char* meta_dev = nullptr;  // [problems | pointers | items] of one batch
    char* meta_host = nullptr;  // pinned staging of the same
    size_t meta_cap = 0;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<Request*> reqs;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

struct oza_server {
    oza_ctx ctx = nullptr;
    hipStream_t stream = nullptr;
    Arena arena[2];
    int last_arena = 1;
    int max_batch = 64;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request*> pending;
    std::deque<int> inflight;  // arenas in launch order
    std::unordered_map<uint64_t, Request*> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    bool worker_done = false;
    std::thread worker, completer;
    // statistics
    uint64_t batches = 0, served = 0, max_seen = 0;

    int free_arena() const {
        const int pref = 1 - last_arena;
        if (!arena[pref].busy) return pref;
        if (!arena[1 - pref].busy) return 1 - pref;
        return -1;
    }
};

namespace {

// Plans `batch` into arena `ar` and enqueues it on the server stream (caller holds no lock; the
// arena is reserved).  Returns an oza_status.
int launch_batch(oza_server* s, Arena& ar, const std::vector<Request*>& batch) {
    const int n = (int)batch.size();
    std::vector<size_t> M(n), N(n), K(n);
    std::vector<int> S(n), T(n);
    for (int i = 0; i < n; ++i) {
        M[i] = batch[i]->M;
        N[i] = batch[i]->N;
        K[i] = batch[i]->K;
        S[i] = batch[i]->S;
        T[i] = batch[i]->tile;
    }
    HostPlan plan;
    if (int r = build_plan(s->ctx->group_cost, n, M.data(), N.data(), K.data(), S.data(), 0, plan, T.data()))
        return r;
    if (plan.ws_bytes > ar.cap) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: batch exceeds the server arena.");
    relocate(plan, ar.ws);
    const size_t off_ptrs = align256(sizeof(GroupProblem) * n);
    const size_t off_items = off_ptrs + align256(sizeof(GroupPtrs) * n);
    const size_t bytes = off_items + sizeof(GItem) * std::max<size_t>(plan.items.size(), 1);
    if (bytes > ar.meta_cap) {  // grow (the arena is idle: its previous batch has completed)
        const size_t cap = std::max(bytes, 2 * ar.meta_cap);
        if (ar.meta_dev) OZA_HIP(hipFree(ar.meta_dev), "HIP: server free");
        if (ar.meta_host) OZA_HIP(hipHostFree(ar.meta_host), "HIP: server free");
        ar.meta_dev = ar.meta_host = nullptr;
        ar.meta_cap = 0;
        OZA_HIP(hipMalloc(&ar.meta_dev, cap), "HIP: server allocation");
        OZA_HIP(hipHostMalloc(&ar.meta_host, cap, hipHostMallocDefault), "HIP: server pinned allocation");
        ar.meta_cap = cap;
    }
    std::memcpy(ar.meta_host, plan.probs.data(), sizeof(GroupProblem) * n);
    GroupPtrs* ptrs = reinterpret_cast<GroupPtrs*>(ar.meta_host + off_ptrs);
    for (int i = 0; i < n; ++i) ptrs[i] = GroupPtrs{batch[i]->A, batch[i]->B, batch[i]->C};
    if (!plan.items.empty())
        std::memcpy(ar.meta_host + off_items, plan.items.data(), sizeof(GItem) * plan.items.size());
    OZA_HIP(hipMemcpyAsync(ar.meta_dev, ar.meta_host, bytes, hipMemcpyHostToDevice, s->stream),
            "HIP: server upload");
    DevPlan d;
    d.probs = reinterpret_cast<const GroupProblem*>(ar.meta_dev);
    d.ptrs = reinterpret_cast<const GroupPtrs*>(ar.meta_dev + off_ptrs);
    d.items = reinterpret_cast<const GItem*>(ar.meta_dev + off_items);
    d.keys = reinterpret_cast<uint32_t*>(ar.ws);
    d.nkeys = plan.nkeys;
    for (int st = 0; st < kGStages; ++st) {
        d.stage_off[st] = plan.stage_off[st];
        d.stage_n[st] = plan.stage_n[st];
    }
    OZA_HIP(enqueue_plan(d, s->ctx->crt_dev, s->stream, nullptr), "server batch launch");
    OZA_HIP(hipEventRecord(ar.done, s->stream), "event");
    return OZA_SUCCESS;
}

void finish(Request* r, double t, int status, const std::string& err) {
    r->t_done = t;
    r->status = status;
    r->error = err;
    r->done = true;
}

void worker_loop(oza_server* s) {
    (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv.wait(lk, [&] { return (s->stop && s->pending.empty()) || (!s->pending.empty() && s->free_arena() >= 0); });
        if (s->pending.empty()) break;  // stop requested and nothing left to dispatch
        const int a = s->free_arena();
        Arena& ar = s->arena[a];
        std::vector<Request*> batch;
        size_t bytes = 0;
        while (!s->pending.empty() && (int)batch.size() < s->max_batch) {
            Request* r = s->pending.front();
            if (!batch.empty() && bytes + r->bytes > ar.cap) break;  // FIFO admission by workspace bytes
            batch.push_back(r);
            bytes += r->bytes;
            s->pending.pop_front();
        }
        ar.busy = true;
        s->last_arena = a;
        lk.unlock();
        const int rc = launch_batch(s, ar, batch);
        const std::string err = rc == OZA_SUCCESS ? std::string() : std::string(oza_last_error());
        lk.lock();
        if (rc != OZA_SUCCESS) {
            const double t = now_ms();
            for (Request* r : batch) finish(r, t, rc, err);
            ar.busy = false;
            s->cv.notify_all();
            continue;
        }
        ar.reqs = std::move(batch);
        s->inflight.push_back(a);
        s->batches += 1;
        s->max_seen = std::max<uint64_t>(s->max_seen, ar.reqs.size());
        s->cv.notify_all();
    }
    s->worker_done = true;
    s->cv.notify_all();
}

void completer_loop(oza_server* s) {
    (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv.wait(lk, [&] { return !s->inflight.empty() || s->worker_done; });
        if (s->inflight.empty()) break;  // worker finished and every batch is accounted for
        Arena& ar = s->arena[s->inflight.front()];
        hipEvent_t ev = ar.done;
        lk.unlock();
        const hipError_t e = hipEventSynchronize(ev);
        const double t = now_ms();
        lk.lock();
        const int status = e == hipSuccess ? OZA_SUCCESS : OZA_HIP_ERROR;
        const std::string err = e == hipSuccess ? std::string() : std::string("HIP: server batch failed: ") +
                                                                      hipGetErrorString(e);
        for (Request* r : ar.reqs) finish(r, t, status, err);
        s->served += ar.reqs.size();
        ar.reqs.clear();
        ar.busy = false;
        s->inflight.pop_front();
        s->cv.notify_all();
    }
}

void release(oza_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    if (s->completer.joinable()) s->completer.join();
    DeviceGuard g(s->ctx->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto& ar : s->arena) {
        if (ar.ws) (void)hipFree(ar.ws);
        if (ar.meta_dev) (void)hipFree(ar.meta_dev);
        if (ar.meta_host) (void)hipHostFree(ar.meta_host);
        if (ar.done) (void)hipEventDestroy(ar.done);
    }
    if (s->stream) (void)hipStreamDestroy(s->stream);
    for (auto& kv : s->reqs) delete kv.second;
    delete s;
}

}  // namespace

extern "C" {

int oza_server_create(oza_ctx ctx, size_t workspace_bytes, int max_batch, oza_server_t* out) {
    if (int r = check_ctx(ctx)) return r;
    if (!out) return fail(OZA_INVALID_ARGUMENT, "null output");
    *out = nullptr;
    if (max_batch < 1) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: max_batch must be >= 1.");
    if (workspace_bytes < (size_t)2 << 20) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: server workspace too small.");
    oza_server* s = new (std::nothrow) oza_server;
    if (!s) return fail(OZA_ALLOC_FAILED, "server");
    s->ctx = ctx;
    s->max_batch = max_batch;
    DeviceGuard g(ctx->device);
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    for (auto& ar : s->arena) {
        ar.cap = workspace_bytes / 2 / 256 * 256;
        if (e == hipSuccess) e = hipMalloc(&ar.ws, ar.cap);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ar.done, hipEventDisableTiming);
        if (e == hipSuccess) {
            ar.meta_cap = (size_t)4 << 20;
            e = hipMalloc(&ar.meta_dev, ar.meta_cap);
            if (e == hipSuccess) e = hipHostMalloc(&ar.meta_host, ar.meta_cap, hipHostMallocDefault);
        }
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        const int st = e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? OZA_ALLOC_FAILED : OZA_HIP_ERROR;
        release(s);
        return st == OZA_ALLOC_FAILED ? fail(st, "HIP: server workspace allocation failed.")
                                      : hip_fail(e, "server creation");
    }
    s->worker = std::thread(worker_loop, s);
    s->completer = std::thread(completer_loop, s);
    *out = s;
    return OZA_SUCCESS;
}

int oza_server_destroy(oza_server_t s) {
    release(s);
    return OZA_SUCCESS;
}

int oza_server_submit(oza_server_t s, size_t M, size_t N, size_t K, int S, const double* A, const double* B,
                      double* C, uint64_t* ticket) {
    if (!s) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null server.");
    if (!ticket) return fail(OZA_INVALID_ARGUMENT, "null output");
    if (int r = check_problem(2, M, N, K, S)) return r;
    if (M > 0 && N > 0 && (!A || !B || !C)) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null matrix pointer.");
    Request* r = new (std::nothrow) Request;
    if (!r) return fail(OZA_ALLOC_FAILED, "request");
    r->M = M;
    r->N = N;
    r->K = K;
    r->S = S;
    r->A = A;
    r->B = B;
    r->C = C;
    std::unique_lock<std::mutex> lk(s->mu);
    r->tile = choose_tile(s->ctx->group_cost, (int64_t)M, (int64_t)N, (int64_t)round_up(K, kPadK), S);
    r->bytes = problem_bytes(M, N, K, S, r->tile);
    if (r->bytes + 256 > s->arena[0].cap) {
        const size_t need = r->bytes;
        delete r;
        return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: request needs " + std::to_string(need) +
                                              " workspace bytes; a server arena holds " +
                                              std::to_string(s->arena[0].cap) + ".");
    }
    if (s->stop) {
        delete r;
        return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: server is shutting down.");
    }
    r->id = s->next_id++;
    r->t_submit = now_ms();
    if (M == 0 || N == 0) finish(r, r->t_submit, OZA_SUCCESS, "");  // nothing to compute
    else s->pending.push_back(r);
    s->reqs[r->id] = r;
    *ticket = r->id;
    lk.unlock();
    s->cv.notify_all();
    return OZA_SUCCESS;
}

int oza_server_wait(oza_server_t s, uint64_t ticket, double* latency_ms) {
    if (!s) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null server.");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: unknown or already collected ticket.");
    Request* r = it->second;
    s->cv.wait(lk, [&] { return r->done; });
    const int status = r->status;
    const std::string err = r->error;
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    s->reqs.erase(it);
    delete r;
    if (status != OZA_SUCCESS) return fail(status, err);
    return OZA_SUCCESS;
}

int oza_server_drain(oza_server_t s) {
    if (!s) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null server.");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv.wait(lk, [&] { return s->pending.empty() && s->inflight.empty() && !s->arena[0].busy && !s->arena[1].busy; });
    return OZA_SUCCESS;
}

int oza_server_stats(oza_server_t s, double stats[4]) {
    if (!s || !stats) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null server or output.");
    std::lock_guard<std::mutex> lk(s->mu);
    stats[0] = (double)s->batches;
    stats[1] = (double)s->served;
    stats[2] = s->batches ? (double)s->served / (double)s->batches : 0.0;
    stats[3] = (double)s->max_seen;
    return OZA_SUCCESS;
}

}  // extern "C"
