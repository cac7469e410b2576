// capi_serve.hip — the request server (include/ozablas_hip.h, oza_server_*): Scheme II GEMM
// requests are queued by any thread and executed in continuous batches.
//
// Two resident workspace arenas, each with its own HIP stream, so two batches run on the device at
// once.  A worker thread admits queued requests into a free arena (as many as fit its bytes and
// the batch limit), plans them with the grouped planner and enqueues them on that arena's stream.
// Requests whose INT8 work (2 M N K S) exceeds a threshold are "large": a batch holds either large
// or small requests, at most one large batch runs at a time, and while it runs the other arena
// takes small requests only (in their order of arrival, passing the waiting large ones), so a small
// request never completes with a large batch.  A large batch's persistent GEMM can be held to a
// share of the CUs (oza_server_config), which leaves CUs to the small batches beside it; the
// default is all of them.  One completion thread per arena waits for
// its batch's event, stamps its requests done and frees the arena.  The CRT tables stay resident
// in the context.  Results are those of the grouped kernels, i.e. bit-identical to
// oza_scheme2_gemm for the same request alone, whatever batch or stream it ran in.
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
    int S = 0, tile = 0, split = 1;
    size_t bytes = 0;
    bool large = false;
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
    hipStream_t stream = nullptr;  // this arena's batches
    hipEvent_t done = nullptr;
    bool busy = false;     // reserved by the worker, or running
    bool running = false;  // enqueued: the completion thread waits for `done`
    bool large = false;    // the batch holds large requests
    std::vector<Request*> reqs;
    std::thread completer;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

struct oza_server {
    oza_ctx ctx = nullptr;
    Arena arena[2];
    int max_batch = 64;
    double large_ops = 1e12;    // INT8 ops above which a request is large
    double large_share = 1.0;   // CU share of a large batch's GEMM
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, Request*> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    bool worker_done = false;
    std::thread worker;
    // statistics
    uint64_t batches = 0, served = 0, max_seen = 0;

    bool large_in_flight() const { return (arena[0].busy && arena[0].large) || (arena[1].busy && arena[1].large); }
    int free_arena() const {
        if (!arena[0].busy) return 0;
        if (!arena[1].busy) return 1;
        return -1;
    }
    // A request the worker may admit now: any, or small ones only while a large batch runs.
    bool eligible() const {
        if (!large_in_flight()) return !pending.empty();
        for (const Request* r : pending)
            if (!r->large) return true;
        return false;
    }
};

namespace {

// Plans `batch` into arena `ar` and enqueues it on the arena's stream (caller holds no lock; the
// arena is reserved).  Returns an oza_status.
int launch_batch(oza_server* s, Arena& ar, const std::vector<Request*>& batch, double gemm_share) {
    const int n = (int)batch.size();
    std::vector<size_t> M(n), N(n), K(n);
    std::vector<int> S(n), T(n), P(n);
    for (int i = 0; i < n; ++i) {
        M[i] = batch[i]->M;
        N[i] = batch[i]->N;
        K[i] = batch[i]->K;
        S[i] = batch[i]->S;
        T[i] = batch[i]->tile;
        P[i] = batch[i]->split;
    }
    HostPlan plan;
    if (int r = build_plan(group_model(s->ctx), n, M.data(), N.data(), K.data(), S.data(), 0, plan, T.data(),
                           P.data()))
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
    OZA_HIP(hipMemcpyAsync(ar.meta_dev, ar.meta_host, bytes, hipMemcpyHostToDevice, ar.stream),
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
    OZA_HIP(enqueue_plan(d, s->ctx->crt_dev, ar.stream, nullptr, gemm_share), "server batch launch");
    OZA_HIP(hipEventRecord(ar.done, ar.stream), "event");
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
        s->cv.wait(lk, [&] { return (s->stop && s->pending.empty()) || (s->free_arena() >= 0 && s->eligible()); });
        if (s->pending.empty()) break;  // stop requested and nothing left to dispatch
        const int a = s->free_arena();
        Arena& ar = s->arena[a];
        // The batch's kind: that of the first eligible request (small only while a large batch runs).
        const bool small_only = s->large_in_flight();
        bool large = false;
        for (const Request* r : s->pending)
            if (!small_only || !r->large) {
                large = r->large;
                break;
            }
        std::vector<Request*> batch;
        size_t bytes = 0;
        for (auto it = s->pending.begin(); it != s->pending.end() && (int)batch.size() < s->max_batch;) {
            Request* r = *it;
            if (r->large != large) {  // the other kind keeps its place
                ++it;
                continue;
            }
            if (!batch.empty() && bytes + r->bytes > ar.cap) break;  // admission by workspace bytes
            batch.push_back(r);
            bytes += r->bytes;
            it = s->pending.erase(it);
        }
        ar.busy = true;
        ar.large = large;
        const double share = large ? s->large_share : 1.0;
        lk.unlock();
        const int rc = launch_batch(s, ar, batch, share);
        const std::string err = rc == OZA_SUCCESS ? std::string() : std::string(oza_last_error());
        lk.lock();
        if (rc != OZA_SUCCESS) {
            const double t = now_ms();
            for (Request* r : batch) finish(r, t, rc, err);
            ar.busy = false;
            ar.large = false;
            s->cv.notify_all();
            continue;
        }
        ar.reqs = std::move(batch);
        ar.running = true;
        s->batches += 1;
        s->max_seen = std::max<uint64_t>(s->max_seen, ar.reqs.size());
        s->cv.notify_all();
    }
    s->worker_done = true;
    s->cv.notify_all();
}

void completer_loop(oza_server* s, int a) {
    (void)hipSetDevice(s->ctx->device);
    Arena& ar = s->arena[a];
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv.wait(lk, [&] { return ar.running || (s->worker_done && !ar.busy); });
        if (!ar.running) break;  // the worker finished and this arena's batches are accounted for
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
        ar.running = false;
        ar.busy = false;
        ar.large = false;
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
    {  // a server whose worker never started has no batch in flight
        std::lock_guard<std::mutex> lk(s->mu);
        s->worker_done = true;
    }
    s->cv.notify_all();
    for (auto& ar : s->arena)
        if (ar.completer.joinable()) ar.completer.join();
    DeviceGuard g(s->ctx->device);
    for (auto& ar : s->arena) {
        if (ar.stream) (void)hipStreamSynchronize(ar.stream);
        if (ar.ws) (void)hipFree(ar.ws);
        if (ar.meta_dev) (void)hipFree(ar.meta_dev);
        if (ar.meta_host) (void)hipHostFree(ar.meta_host);
        if (ar.done) (void)hipEventDestroy(ar.done);
        if (ar.stream) (void)hipStreamDestroy(ar.stream);
    }
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
    hipError_t e = hipSuccess;
    for (auto& ar : s->arena) {
        ar.cap = workspace_bytes / 2 / 256 * 256;
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&ar.stream, hipStreamNonBlocking);
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
    for (int a = 0; a < 2; ++a) s->arena[a].completer = std::thread(completer_loop, s, a);
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
    // K split as if the request ran alone: a request that cannot fill the device by itself (small
    // M x N, long K) gets its latency cut whatever it is batched with.
    r->split = M && N ? alone_split(group_model(s->ctx), (int64_t)M, (int64_t)N, (int64_t)K, S, r->tile) : 1;
    r->bytes = problem_bytes(M, N, K, S, r->tile, r->split);
    r->large = 2.0 * (double)M * (double)N * (double)K * (double)S > s->large_ops;
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
    s->cv.wait(lk, [&] { return s->pending.empty() && !s->arena[0].busy && !s->arena[1].busy; });
    return OZA_SUCCESS;
}

int oza_server_config(oza_server_t s, double large_ops, double large_gemm_share) {
    if (!s) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null server.");
    if (!(large_ops >= 0.0) || !(large_gemm_share > 0.0 && large_gemm_share <= 1.0))
        return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: large_ops must be >= 0 and the GEMM share in (0, 1].");
    std::lock_guard<std::mutex> lk(s->mu);
    s->large_ops = large_ops;  // applies to requests submitted from now on
    s->large_share = large_gemm_share;
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
