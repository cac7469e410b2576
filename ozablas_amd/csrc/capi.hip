// capi.hip — the C ABI (include/ozablas_hip.h): contexts, workspace sizing, pipelines, stages.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <utility>

#include "capi_internal.hpp"

using namespace oza;
using namespace oza::capi;

namespace {

// Workspace of oza_gemm_noalloc calls, in a file-scope device array so that those calls allocate
// nothing.  Each device the library runs on holds its own copy until the process ends.  In host
// code `scratch` names a host-side stand-in, never the device memory: the device address is taken
// with hipGetSymbolAddress in oza_ctx_create, with the context's device current, and only that
// address (oza_context::scratch) is handed to kernels.
static __device__ __align__(256) uint8_t scratch[OZA_NOALLOC_SCRATCH_BYTES];

void ws_bytes(int scheme, size_t M, size_t N, size_t K, int S, size_t b[5]) {
    const Dims d = make_dims(M, N, K, S);
    b[0] = sizeof(int32_t) * (M ? M : 1);
    b[1] = sizeof(int32_t) * (N ? N : 1);
    b[2] = (size_t)S * d.Mp * d.Kp;
    b[3] = (size_t)S * d.Np * d.Kp;
    // Scheme II: uint8 digit planes, times the K split of the call (the partial digits of parts
    // 1.. follow the S planes); Scheme I: INT32 group-product planes (both blocked, Mp x Np), and
    // at least the reference's planar [S][M][N] INT32 size for the stage entry points.
    b[4] = scheme == 2 ? (size_t)S * d.Mp * d.Np * (size_t)scheme2_single_split(M, N, K, S)
                       : sizeof(int32_t) * (size_t)S * std::max((size_t)d.Mp * d.Np, (M ? M : 1) * (N ? N : 1));
}

}  // namespace

extern "C" {

const char* oza_version(void) { return "ozablas-hip 0.1 (gfx950)"; }

const char* oza_status_string(int s) {
    switch (s) {
        case OZA_SUCCESS: return "success";
        case OZA_INVALID_ARGUMENT: return "invalid argument";
        case OZA_ALLOC_FAILED: return "allocation failed";
        case OZA_HIP_ERROR: return "HIP runtime error";
        case OZA_NOT_SUPPORTED: return "not supported";
        default: return "unknown status";
    }
}

const char* oza_last_error(void) { return g_last_error.c_str(); }

int oza_get_device_count(int* count) {
    if (!count) return fail(OZA_INVALID_ARGUMENT, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return OZA_SUCCESS;
}

int oza_ctx_create(int device_id, oza_ctx* out) {
    if (!out) return fail(OZA_INVALID_ARGUMENT, "null output");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(OZA_HIP_ERROR, "HIP: no device available.");
    if (device_id < 0 || device_id >= n) return fail(OZA_HIP_ERROR, "HIP: Failed to set device.");
    oza_context* c = new (std::nothrow) oza_context();
    if (!c) return fail(OZA_ALLOC_FAILED, "context allocation");
    c->device = device_id;
    DeviceGuard g(device_id);
    hipError_t e = hipStreamCreate(&c->stream);
    if (e == hipSuccess) e = hipMalloc(&c->crt_dev, sizeof(CrtTables));
    if (e == hipSuccess) e = hipMemcpy(c->crt_dev, &host_crt_tables(), sizeof(CrtTables), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&c->stats_dev, 4 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c->queue_dev, 8 * oza_context::kMaxBands * sizeof(unsigned));
    for (int i = 0; i < 5 && e == hipSuccess; ++i) e = hipEventCreate(&c->ev[i]);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    for (int i = 0; i < oza_context::kMaxBands && e == hipSuccess; ++i)
        e = hipEventCreateWithFlags(&c->band_ev[i], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device_id);
    if (e == hipSuccess) e = group_gemm_slots(c->group_slots);
    if (e == hipSuccess) e = hipGetSymbolAddress(&c->scratch, HIP_SYMBOL(scratch));  // this device's copy
    if (e != hipSuccess) {
        oza_ctx_destroy(c);
        return hip_fail(e, "context creation");
    }
    *out = c;
    return OZA_SUCCESS;
}

int oza_ctx_destroy(oza_ctx c) {
    if (!c) return OZA_SUCCESS;
    {
        DeviceGuard g(c->device);
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        if (c->blas) rocblas_destroy_handle(c->blas);
        if (c->ring) {
            for (int i = 0; i < oza_context::kRing * 5; ++i)
                if (c->ring[i]) (void)hipEventDestroy(c->ring[i]);
            delete[] c->ring;
        }
        if (c->side) (void)hipStreamSynchronize(c->side);
        for (auto& e : c->ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : c->band_ev)
            if (e) (void)hipEventDestroy(e);
        if (c->join_ev) (void)hipEventDestroy(c->join_ev);
        if (c->side) (void)hipStreamDestroy(c->side);
        if (c->crt_dev) (void)hipFree(c->crt_dev);
        if (c->ws) (void)hipFree(c->ws);
        if (c->stats_dev) (void)hipFree(c->stats_dev);
        if (c->queue_dev) (void)hipFree(c->queue_dev);
        if (c->stream) (void)hipStreamDestroy(c->stream);
    }
    delete c;
    return OZA_SUCCESS;
}

int oza_ctx_device(oza_ctx ctx, int* dev) {
    if (int r = check_ctx(ctx)) return r;
    if (!dev) return fail(OZA_INVALID_ARGUMENT, "null output");
    *dev = ctx->device;
    return OZA_SUCCESS;
}

int oza_ctx_stream(oza_ctx ctx, void** st) {
    if (int r = check_ctx(ctx)) return r;
    if (!st) return fail(OZA_INVALID_ARGUMENT, "null output");
    *st = (void*)ctx->stream;
    return OZA_SUCCESS;
}

int oza_malloc(oza_ctx ctx, void** ptr, size_t bytes) {
    if (int r = check_ctx(ctx)) return r;
    if (!ptr) return fail(OZA_INVALID_ARGUMENT, "null output");
    DeviceGuard g(ctx->device);
    *ptr = nullptr;
    if (hipMalloc(ptr, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return fail(OZA_ALLOC_FAILED, "HIP: device allocation of " + std::to_string(bytes) + " bytes failed.");
    }
    return OZA_SUCCESS;
}

int oza_free(oza_ctx ctx, void* ptr) {
    if (int r = check_ctx(ctx)) return r;
    DeviceGuard g(ctx->device);
    if (ptr) OZA_HIP(hipFree(ptr), "HIP: free failed");
    return OZA_SUCCESS;
}

static int memcpy_kind(oza_ctx ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind,
                       const char* what) {
    if (int r = check_ctx(ctx)) return r;
    if (bytes == 0) return OZA_SUCCESS;
    if (!dst || !src) return fail(OZA_INVALID_ARGUMENT, "null pointer");
    DeviceGuard g(ctx->device);
    OZA_HIP(hipMemcpyAsync(dst, src, bytes, kind, ctx->stream), what);
    OZA_HIP(hipStreamSynchronize(ctx->stream), what);
    return OZA_SUCCESS;
}

int oza_memcpy_h2d(oza_ctx ctx, void* dst, const void* src, size_t bytes) {
    return memcpy_kind(ctx, dst, src, bytes, hipMemcpyHostToDevice, "HIP: Memcpy HostToDevice failed.");
}
int oza_memcpy_d2h(oza_ctx ctx, void* dst, const void* src, size_t bytes) {
    return memcpy_kind(ctx, dst, src, bytes, hipMemcpyDeviceToHost, "HIP: Memcpy DeviceToHost failed.");
}
int oza_memcpy_d2d(oza_ctx ctx, void* dst, const void* src, size_t bytes) {
    return memcpy_kind(ctx, dst, src, bytes, hipMemcpyDeviceToDevice, "HIP: Memcpy DeviceToDevice failed.");
}

int oza_synchronize(oza_ctx ctx) {
    if (int r = check_ctx(ctx)) return r;
    DeviceGuard g(ctx->device);
    OZA_HIP(hipDeviceSynchronize(), "HIP: Device synchronize failed.");
    OZA_HIP(hipGetLastError(), "HIP: kernel error");
    return OZA_SUCCESS;
}

int oza_stream_synchronize(oza_ctx ctx) {
    if (int r = check_ctx(ctx)) return r;
    DeviceGuard g(ctx->device);
    OZA_HIP(hipStreamSynchronize(ctx->stream), "HIP: stream synchronize failed.");
    OZA_HIP(hipGetLastError(), "HIP: kernel error");
    return OZA_SUCCESS;
}

int oza_blas_handle(oza_ctx ctx, void** handle) {
    if (int r = check_ctx(ctx)) return r;
    if (!handle) return fail(OZA_INVALID_ARGUMENT, "null output");
    std::lock_guard<std::mutex> lk(ctx->blas_mu);
    if (!ctx->blas) {
        DeviceGuard g(ctx->device);
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success)
            return fail(OZA_HIP_ERROR, "HIP: Failed to create rocBLAS handle.");
        rocblas_set_stream(h, ctx->stream);
        ctx->blas = h;
    }
    *handle = (void*)ctx->blas;
    return OZA_SUCCESS;
}

int oza_workspace_bytes(int scheme, size_t M, size_t N, size_t K, int S, size_t bytes[5]) {
    if (!bytes) return fail(OZA_INVALID_ARGUMENT, "null output");
    if (int r = check_problem(scheme, M, N, K, S)) return r;
    ws_bytes(scheme, M, N, K, S, bytes);
    return OZA_SUCCESS;
}

int oza_scheme1_beta(size_t K, int* beta) {
    if (!beta || K < 1) return fail(OZA_INVALID_ARGUMENT, "K must be >= 1");
    *beta = scheme1_beta(K);
    return OZA_SUCCESS;
}

int oza_scheme2_kparam(int S, size_t K, int* kp) {
    if (!kp || K < 1 || S < 1 || S > OZA_MAX_SLICES) return fail(OZA_INVALID_ARGUMENT, "bad slices or K");
    *kp = scheme2_kparam(S, K);
    return OZA_SUCCESS;
}

int oza_scheme2_slices_for_digits(size_t K, double digits, int* slices) {
    if (!slices || K < 1 || !(digits > 0.0)) return fail(OZA_INVALID_ARGUMENT, "bad K or digits");
    const int S = scheme2_slices_for_digits(K, digits);
    if (S == 0) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: no moduli count up to 20 reaches that accuracy.");
    *slices = S;
    return OZA_SUCCESS;
}

int oza_scheme2_slices_for_accuracy(size_t K, double digits, double spread, int* slices) {
    if (!slices || K < 1 || !(digits > 0.0) || !(spread >= 0.0))
        return fail(OZA_INVALID_ARGUMENT, "bad K, digits or spread");
    const int S = scheme2_slices_for_accuracy(K, digits, spread);
    if (S == 0) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: no moduli count up to 20 reaches that accuracy.");
    *slices = S;
    return OZA_SUCCESS;
}

// ------------------------------------------------------------------------ pipelines
// Row bands for overlapping step 4 with step 3 (both schemes; bands of whole 256-row tiles).
// Scheme II, automatic: the most bands (up to 4) for which every band's products still take at
// least 8 rounds of one tile per CU and the bands' partial last rounds cost at most one round in
// total.  Scheme I (items of g - 1 pair products, largest first): the most bands (up to 4) with at
// least 16 pair products per CU in every band.  1 (no overlap) when none qualifies.
static int pipeline_bands(const oza_context* ctx, int scheme, const Dims& d) {
    const int64_t tm = (d.M + kS2TileRows - 1) / kS2TileRows, tn = (d.N + kS2TileRows - 1) / kS2TileRows;
    const int64_t cus = std::max(ctx->cus, 1);
    if (ctx->bands > 0) return (int)std::min<int64_t>(ctx->bands, tm);
    if (scheme == 1) {
        const int64_t units = tn * (int64_t)d.S * (d.S + 1) / 2;  // pair products per tile row
        for (int nb = 4; nb >= 2; --nb)
            if (tm >= nb && (tm / nb) * units >= 16 * cus) return nb;
        return 1;
    }
    auto rounds = [&](int64_t items) { return (items + cus - 1) / cus; };
    const int64_t whole = rounds(tm * tn * d.S);
    for (int nb = 4; nb >= 2; --nb) {
        if (tm < nb) continue;
        int64_t sum = 0, least = INT64_MAX;
        for (int b = 0; b < nb; ++b) {
            const int64_t rows = tm * (b + 1) / nb - tm * b / nb;
            const int64_t r = rounds(rows * tn * d.S);
            sum += r;
            least = std::min(least, r);
        }
        if (least >= 8 && sum <= whole + 1) return nb;
    }
    return 1;
}

// Steps 3-4 in nb row bands: band b's products on the context stream, then its reconstruction on
// the side stream (one workgroup per CU, beside the next band's products); the last band's
// reconstruction runs on the context stream, which then joins the side stream.  `e3` (may be
// null) is recorded after the last products.
static int enqueue_banded(oza_ctx ctx, int scheme, const Dims& d, int nb, const int8_t* Asl, const int8_t* Bsl,
                          void* Ctc, const int32_t* shA, const int32_t* shB, const OutC& C, hipEvent_t e3) {
    hipStream_t st = ctx->stream;
    const CrtSet* cs = crt_for(ctx, d.S);
    const int beta = scheme1_beta((size_t)d.K);
    const int tm = (int)((d.M + kS2TileRows - 1) / kS2TileRows);
    for (int b = 0; b < nb; ++b) {
        const int t0 = (int)((int64_t)tm * b / nb), t1 = (int)((int64_t)tm * (b + 1) / nb);
        if (scheme == 2)
            OZA_HIP(launch_igemm_s2(d, Asl, Bsl, S2Out::Digits, Ctc, cs, st, t0, t1 - t0, ctx->queue_dev + 8 * b),
                    "step 3 launch");
        else
            OZA_HIP(launch_igemm_s1_blocked(d, Asl, Bsl, static_cast<int32_t*>(Ctc), st, ctx->queue_dev + 8 * b, t0,
                                            t1 - t0),
                    "step 3 launch");
        const int64_t row0 = (int64_t)t0 * kS2TileRows, rows = (int64_t)(t1 - t0) * kS2TileRows;
        const bool last = b + 1 == nb;
        hipStream_t rs = last ? st : ctx->side;
        if (!last) {
            OZA_HIP(hipEventRecord(ctx->band_ev[b], st), "event");
            OZA_HIP(hipStreamWaitEvent(ctx->side, ctx->band_ev[b], 0), "stream wait");
        } else if (e3) {
            OZA_HIP(hipEventRecord(e3, st), "event");
        }
        if (scheme == 2)
            OZA_HIP(launch_recon_s2_digits(d, static_cast<const uint8_t*>(Ctc), shA, shB, cs, C, rs, row0, rows, !last),
                    "step 4 launch");
        else
            OZA_HIP(launch_recon_s1_blocked(d, beta, static_cast<const int32_t*>(Ctc), shA, shB, C, rs, row0, rows,
                                            !last),
                    "step 4 launch");
    }
    OZA_HIP(hipEventRecord(ctx->join_ev, ctx->side), "event");
    OZA_HIP(hipStreamWaitEvent(st, ctx->join_ev, 0), "stream wait");
    return OZA_SUCCESS;
}

// The four steps for operands A (M logical rows), B (N logical rows) and output C.  The caller
// has validated the arguments and made the context's device current.
static int enqueue_pipeline(int scheme, oza_ctx ctx, size_t M, size_t N, size_t K, int S, const Operand& A,
                            const Operand& B, const OutC& C, int32_t* shA, int32_t* shB, int8_t* Asl, int8_t* Bsl,
                            void* Ctc, oza_timings* t) {
    const Dims d = make_dims(M, N, K, S);
    hipStream_t st = ctx->stream;
    hipEvent_t* ev = nullptr;  // events to record this call, if any
    if (t) ev = ctx->ev;
    else if (ctx->ring_on && ctx->ring_n < oza_context::kRing) ev = ctx->ring + 5 * ctx->ring_n++;
    if (ev) OZA_HIP(hipEventRecord(ev[0], st), "event");
    if (scheme == 2) {
        const int kp = scheme2_kparam(S, K);
        OZA_HIP(launch_stats(d, A, B, kp, shA, shB, st, ctx->queue_dev, 8 * oza_context::kMaxBands),
                "step 1 launch");
        if (ev) OZA_HIP(hipEventRecord(ev[1], st), "event");
        OZA_HIP(launch_slice_s2(d, A, B, shA, shB, Asl, Bsl, st), "step 2 launch");
        if (ev) OZA_HIP(hipEventRecord(ev[2], st), "event");
        const int nb = pipeline_bands(ctx, 2, d);
        if (nb > 1) {
            if (int r = enqueue_banded(ctx, 2, d, nb, Asl, Bsl, Ctc, shA, shB, C, ev ? ev[3] : nullptr)) return r;
        } else {
            // Few tiles and a long K: the K steps run in parts (partial digits summed mod m_s)
            const int split = scheme2_single_split(M, N, K, S);
            OZA_HIP(launch_igemm_s2(d, Asl, Bsl, S2Out::Digits, Ctc, crt_for(ctx, S), st, 0, -1, ctx->queue_dev, split),
                    "step 3 launch");
            OZA_HIP(launch_digits_reduce(d, static_cast<uint8_t*>(Ctc), split, crt_for(ctx, S), st), "step 3 launch");
            if (ev) OZA_HIP(hipEventRecord(ev[3], st), "event");
            OZA_HIP(launch_recon_s2_digits(d, static_cast<const uint8_t*>(Ctc), shA, shB, crt_for(ctx, S), C, st),
                    "step 4 launch");
        }
    } else {
        const int sbeta = scheme1_beta(K);
        OZA_HIP(launch_stats(d, A, B, 0, shA, shB, st, ctx->queue_dev, 8 * oza_context::kMaxBands),
                "step 1 launch");
        if (ev) OZA_HIP(hipEventRecord(ev[1], st), "event");
        OZA_HIP(launch_slice_s1(d, sbeta, A, B, shA, shB, Asl, Bsl, st), "step 2 launch");
        if (ev) OZA_HIP(hipEventRecord(ev[2], st), "event");
        const int nb = pipeline_bands(ctx, 1, d);
        if (nb > 1) {
            if (int r = enqueue_banded(ctx, 1, d, nb, Asl, Bsl, Ctc, shA, shB, C, ev ? ev[3] : nullptr)) return r;
        } else {
            OZA_HIP(launch_igemm_s1_blocked(d, Asl, Bsl, static_cast<int32_t*>(Ctc), st, ctx->queue_dev),
                    "step 3 launch");
            if (ev) OZA_HIP(hipEventRecord(ev[3], st), "event");
            OZA_HIP(launch_recon_s1_blocked(d, sbeta, static_cast<const int32_t*>(Ctc), shA, shB, C, st),
                    "step 4 launch");
        }
    }
    if (ev && !t) OZA_HIP(hipEventRecord(ev[4], st), "event");
    if (t) {
        OZA_HIP(hipEventRecord(ctx->ev[4], st), "event");
        OZA_HIP(hipEventSynchronize(ctx->ev[4]), "HIP: pipeline failed");
        OZA_HIP(hipGetLastError(), "HIP: kernel error");
        (void)hipEventElapsedTime(&t->step1_ms, ctx->ev[0], ctx->ev[1]);
        (void)hipEventElapsedTime(&t->step2_ms, ctx->ev[1], ctx->ev[2]);
        (void)hipEventElapsedTime(&t->step3_ms, ctx->ev[2], ctx->ev[3]);
        (void)hipEventElapsedTime(&t->step4_ms, ctx->ev[3], ctx->ev[4]);
        (void)hipEventElapsedTime(&t->total_ms, ctx->ev[0], ctx->ev[4]);
    }
    return OZA_SUCCESS;
}

static int run_pipeline(int scheme, oza_ctx ctx, size_t M, size_t N, size_t K, int S, const double* A,
                        const double* B, double* C, int32_t* shA, int32_t* shB, int8_t* Asl, int8_t* Bsl,
                        void* Ctc, oza_timings* t) {
    if (int r = check_ctx(ctx)) return r;
    if (int r = check_problem(scheme, M, N, K, S)) return r;
    if (M == 0 || N == 0) return OZA_SUCCESS;
    if (!A || !B || !C || !shA || !shB || !Asl || !Bsl || !Ctc)
        return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null matrix or workspace pointer.");
    DeviceGuard g(ctx->device);
    return enqueue_pipeline(scheme, ctx, M, N, K, S, Operand{A, (int64_t)K, true}, Operand{B, (int64_t)N, false},
                            OutC{C, (int64_t)N, 1.0, 0.0}, shA, shB, Asl, Bsl, Ctc, t);
}

int oza_scheme1_gemm(oza_ctx ctx, size_t M, size_t N, size_t K, int S, const double* A, const double* B, double* C,
                     int32_t* shA, int32_t* shB, int8_t* Asl, int8_t* Bsl, void* Ctc, oza_timings* t) {
    return run_pipeline(1, ctx, M, N, K, S, A, B, C, shA, shB, Asl, Bsl, Ctc, t);
}

int oza_scheme2_gemm(oza_ctx ctx, size_t M, size_t N, size_t K, int S, const double* A, const double* B, double* C,
                     int32_t* shA, int32_t* shB, int8_t* Asl, int8_t* Bsl, void* Ctc, oza_timings* t) {
    return run_pipeline(2, ctx, M, N, K, S, A, B, C, shA, shB, Asl, Bsl, Ctc, t);
}

int oza_ctx_set_bands(oza_ctx ctx, int bands) {
    if (int r = check_ctx(ctx)) return r;
    if (bands < 0 || bands > oza_context::kMaxBands)
        return fail(OZA_INVALID_ARGUMENT, "bands must be 0 (automatic) to " + std::to_string(oza_context::kMaxBands));
    ctx->bands = bands;
    return OZA_SUCCESS;
}

int oza_ctx_bands_for(oza_ctx ctx, int scheme, size_t M, size_t N, size_t K, int S, int* bands) {
    if (int r = check_ctx(ctx)) return r;
    if (!bands) return fail(OZA_INVALID_ARGUMENT, "null output");
    if (int r = check_problem(scheme, M, N, K, S)) return r;
    *bands = pipeline_bands(ctx, scheme, make_dims(M, N, K, S));
    return OZA_SUCCESS;
}

int oza_ctx_event_timing(oza_ctx ctx, int enable) {
    if (int r = check_ctx(ctx)) return r;
    DeviceGuard g(ctx->device);
    if (enable && !ctx->ring) {
        ctx->ring = new (std::nothrow) hipEvent_t[oza_context::kRing * 5]();
        if (!ctx->ring) return fail(OZA_ALLOC_FAILED, "event ring");
        for (int i = 0; i < oza_context::kRing * 5; ++i) OZA_HIP(hipEventCreate(&ctx->ring[i]), "event create");
    }
    ctx->ring_on = enable != 0;
    ctx->ring_n = 0;
    return OZA_SUCCESS;
}

int oza_ctx_event_times(oza_ctx ctx, double sums[5], int* count) {
    if (int r = check_ctx(ctx)) return r;
    if (!sums || !count) return fail(OZA_INVALID_ARGUMENT, "null output");
    DeviceGuard g(ctx->device);
    for (int j = 0; j < 5; ++j) sums[j] = 0.0;
    *count = ctx->ring_n;
    if (ctx->ring_n == 0) return OZA_SUCCESS;
    OZA_HIP(hipEventSynchronize(ctx->ring[5 * (ctx->ring_n - 1) + 4]), "event sync");
    for (int i = 0; i < ctx->ring_n; ++i) {
        hipEvent_t* e = ctx->ring + 5 * i;
        float ms;
        for (int j = 0; j < 4; ++j) {
            OZA_HIP(hipEventElapsedTime(&ms, e[j], e[j + 1]), "elapsed");
            sums[j] += ms;
        }
        OZA_HIP(hipEventElapsedTime(&ms, e[0], e[4]), "elapsed");
        sums[4] += ms;
    }
    ctx->ring_n = 0;
    return OZA_SUCCESS;
}

// ------------------------------------------------------------------------ BLAS-style front end
int oza_gemm_workspace_bytes(int scheme, size_t M, size_t N, size_t K, int S, size_t* bytes) {
    if (!bytes) return fail(OZA_INVALID_ARGUMENT, "null output");
    if (int r = check_problem(scheme, M, N, K ? K : 1, S)) return r;
    size_t b[5];
    ws_bytes(scheme, M, N, K ? K : 1, S, b);
    size_t tot = 0;
    for (size_t x : b) tot += (x + 255) / 256 * 256;
    *bytes = tot;
    return OZA_SUCCESS;
}

int oza_gemm(oza_ctx ctx, int scheme, int S, int layout, int transA, int transB, size_t M, size_t N, size_t K,
             double alpha, const double* A, size_t lda, const double* B, size_t ldb, double beta, double* C,
             size_t ldc, void* workspace, size_t workspace_bytes) {
    if (int r = check_ctx(ctx)) return r;
    if (layout != OZA_ROW_MAJOR && layout != OZA_COL_MAJOR) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: bad layout.");
    if ((transA != OZA_OP_N && transA != OZA_OP_T) || (transB != OZA_OP_N && transB != OZA_OP_T))
        return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: bad transpose flag.");
    // Column-major C = op(A) op(B) is row-major C^T = op(B)^T op(A)^T: swap the operands.
    const char* name_a = "lda";
    const char* name_b = "ldb";
    if (layout == OZA_COL_MAJOR) {
        std::swap(M, N);
        std::swap(A, B);
        std::swap(lda, ldb);
        std::swap(transA, transB);
        std::swap(name_a, name_b);
    }
    // Leading dimensions (BLAS argument checks, in the row-major view).
    const size_t a_cols = transA == OZA_OP_N ? K : M, b_cols = transB == OZA_OP_N ? N : K;
    if (lda < std::max<size_t>(1, a_cols)) return fail(OZA_INVALID_ARGUMENT, std::string("OzaBLAS: ") + name_a + " too small.");
    if (ldb < std::max<size_t>(1, b_cols)) return fail(OZA_INVALID_ARGUMENT, std::string("OzaBLAS: ") + name_b + " too small.");
    if (ldc < std::max<size_t>(1, N)) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: ldc too small.");
    if (M == 0 || N == 0) return OZA_SUCCESS;
    if (!C) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null C.");
    DeviceGuard g(ctx->device);
    if (K == 0 || alpha == 0.0) {  // C = beta * C (BLAS quick path: op(A) op(B) is not formed)
        if (beta == 1.0) return OZA_SUCCESS;
        OZA_HIP(launch_scale_c(OutC{C, (int64_t)ldc, 0.0, beta}, (int64_t)M, (int64_t)N, ctx->stream), "scale C");
        return OZA_SUCCESS;
    }
    if (int r = check_problem(scheme, M, N, K, S)) return r;
    if (!A || !B) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null matrix pointer.");
    size_t b[5];
    ws_bytes(scheme, M, N, K, S, b);
    size_t need = 0;
    for (size_t x : b) need += (x + 255) / 256 * 256;
    char* ws = static_cast<char*>(workspace);
    if (ws) {
        if (workspace_bytes < need)
            return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: workspace smaller than oza_gemm_workspace_bytes.");
    } else {  // the context's own workspace, grown on demand and kept for later calls
        std::lock_guard<std::mutex> lk(ctx->ws_mu);
        if (ctx->ws_bytes < need) {
            if (ctx->ws) {
                OZA_HIP(hipStreamSynchronize(ctx->stream), "HIP: workspace resize");
                OZA_HIP(hipFree(ctx->ws), "HIP: workspace resize");
                ctx->ws = nullptr;
                ctx->ws_bytes = 0;
            }
            if (hipMalloc(&ctx->ws, need) != hipSuccess) {
                (void)hipGetLastError();
                ctx->ws = nullptr;
                return fail(OZA_ALLOC_FAILED, "HIP: workspace allocation of " + std::to_string(need) + " bytes failed.");
            }
            ctx->ws_bytes = need;
        }
        ws = static_cast<char*>(ctx->ws);
    }
    char* parts[5];
    for (int i = 0; i < 5; ++i) {
        parts[i] = ws;
        ws += (b[i] + 255) / 256 * 256;
    }
    const Operand a{A, (int64_t)lda, transA == OZA_OP_N};   // op(A) rows: A rows, or A columns
    const Operand bb{B, (int64_t)ldb, transB == OZA_OP_T};  // op(B) columns: B columns, or B^T rows
    return enqueue_pipeline(scheme, ctx, M, N, K, S, a, bb, OutC{C, (int64_t)ldc, alpha, beta},
                            reinterpret_cast<int32_t*>(parts[0]), reinterpret_cast<int32_t*>(parts[1]),
                            reinterpret_cast<int8_t*>(parts[2]), reinterpret_cast<int8_t*>(parts[3]), parts[4],
                            nullptr);
}

int oza_gemm_noalloc(oza_ctx ctx, int scheme, int S, int layout, int transA, int transB, size_t M, size_t N, size_t K,
                     double alpha, const double* A, size_t lda, const double* B, size_t ldb, double beta, double* C,
                     size_t ldc) {
    if (int r = check_ctx(ctx)) return r;
    if (!ctx->scratch) return fail(OZA_HIP_ERROR, "OzaBLAS: no scratch array on this context's device.");
    // Problems that form a product must fit the scratch array; the C = beta * C quick path of
    // oza_gemm needs no workspace.  (oza_gemm repeats the argument checks.)
    if (K != 0 && alpha != 0.0 && M != 0 && N != 0 && check_problem(scheme, M, N, K, S) == OZA_SUCCESS) {
        size_t b[5];
        ws_bytes(scheme, M, N, K, S, b);
        size_t need = 0;
        for (size_t x : b) need += (x + 255) / 256 * 256;
        if (need > OZA_NOALLOC_SCRATCH_BYTES)
            return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: the problem needs " + std::to_string(need) +
                                                  " bytes of workspace; oza_gemm_noalloc holds " +
                                                  std::to_string((size_t)OZA_NOALLOC_SCRATCH_BYTES) +
                                                  " (call oza_gemm with a workspace instead).");
    }
    // A non-null workspace: oza_gemm then takes no allocating path.
    return oza_gemm(ctx, scheme, S, layout, transA, transB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                    ctx->scratch, OZA_NOALLOC_SCRATCH_BYTES);
}

// ------------------------------------------------------------------------ stages
#define OZA_STAGE_PROLOGUE(scheme)                                   \
    if (int r = check_ctx(ctx)) return r;                            \
    if (int r = check_problem(scheme, M, N, K, S)) return r;         \
    DeviceGuard g(ctx->device);                                      \
    const Dims d = make_dims(M, N, K, S);                            \
    hipStream_t st = ctx->stream;                                    \
    (void)d;                                                         \
    (void)st;

int oza_stage1_shifts(oza_ctx ctx, int scheme, size_t M, size_t N, size_t K, int S, const double* A,
                      const double* B, int32_t* shA, int32_t* shB) {
    OZA_STAGE_PROLOGUE(scheme)
    const int offset = scheme == 2 ? scheme2_kparam(S, K) : 0;
    OZA_HIP(launch_stats(d, Operand{A, (int64_t)K, true}, Operand{B, (int64_t)N, false}, offset, shA, shB, st),
            "step 1 launch");
    return OZA_SUCCESS;
}

int oza_stage2_slices(oza_ctx ctx, int scheme, size_t M, size_t N, size_t K, int S, const double* A,
                      const double* B, const int32_t* shA, const int32_t* shB, int8_t* Asl, int8_t* Bsl) {
    OZA_STAGE_PROLOGUE(scheme)
    const Operand a{A, (int64_t)K, true}, b{B, (int64_t)N, false};
    if (scheme == 2)
        OZA_HIP(launch_slice_s2(d, a, b, shA, shB, Asl, Bsl, st), "step 2 launch");
    else
        OZA_HIP(launch_slice_s1(d, scheme1_beta(K), a, b, shA, shB, Asl, Bsl, st), "step 2 launch");
    return OZA_SUCCESS;
}

int oza_unpack_slices(oza_ctx ctx, size_t M, size_t N, size_t K, int S, const int8_t* Asl, const int8_t* Bsl,
                      int8_t* Ap, int8_t* Bp) {
    OZA_STAGE_PROLOGUE(2)
    if (Ap) OZA_HIP(launch_unpack_A(d, Asl, Ap, st), "unpack");
    if (Bp) OZA_HIP(launch_unpack_B(d, Bsl, Bp, st), "unpack");
    return OZA_SUCCESS;
}

int oza_stage3_s2_products(oza_ctx ctx, size_t M, size_t N, size_t K, int S, const int8_t* Asl,
                           const int8_t* Bsl, int32_t* Cp) {
    OZA_STAGE_PROLOGUE(2)
    OZA_HIP(launch_igemm_s2(d, Asl, Bsl, S2Out::Int32, Cp, crt_for(ctx, S), st), "step 3 launch");
    return OZA_SUCCESS;
}

int oza_stage3_s2_digits(oza_ctx ctx, size_t M, size_t N, size_t K, int S, const int8_t* Asl, const int8_t* Bsl,
                         void* Ctc) {
    OZA_STAGE_PROLOGUE(2)
    OZA_HIP(launch_igemm_s2(d, Asl, Bsl, S2Out::Digits, Ctc, crt_for(ctx, S), st), "step 3 launch");
    return OZA_SUCCESS;
}

int oza_unpack_digits(oza_ctx ctx, size_t M, size_t N, int S, const void* Ctc, uint8_t* planar) {
    const size_t K = 1;
    OZA_STAGE_PROLOGUE(2)
    if (!Ctc || !planar) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null digits pointer.");
    OZA_HIP(launch_digits_unpack(d, static_cast<const uint8_t*>(Ctc), planar, st), "unpack");
    return OZA_SUCCESS;
}

int oza_pack_digits(oza_ctx ctx, size_t M, size_t N, int S, const uint8_t* planar, void* Ctc) {
    const size_t K = 1;
    OZA_STAGE_PROLOGUE(2)
    if (!Ctc || !planar) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null digits pointer.");
    OZA_HIP(launch_digits_pack(d, planar, crt_for(ctx, S), static_cast<uint8_t*>(Ctc), st), "pack");
    return OZA_SUCCESS;
}

int oza_stage3_s1_groups(oza_ctx ctx, size_t M, size_t N, size_t K, int S, const int8_t* Asl, const int8_t* Bsl,
                         int32_t* T) {
    OZA_STAGE_PROLOGUE(1)
    OZA_HIP(launch_igemm_s1_groups(d, Asl, Bsl, T, st), "step 3 launch");
    return OZA_SUCCESS;
}

int oza_stage4_s2_products(oza_ctx ctx, size_t M, size_t N, size_t K, int S, const int32_t* Cp,
                           const int32_t* shA, const int32_t* shB, double* C) {
    OZA_STAGE_PROLOGUE(2)
    OZA_HIP(launch_recon_s2_int32(d, Cp, shA, shB, crt_for(ctx, S), C, st), "step 4 launch");
    return OZA_SUCCESS;
}

int oza_stage4_s2_digits(oza_ctx ctx, size_t M, size_t N, size_t K, int S, const void* Ctc, const int32_t* shA,
                         const int32_t* shB, double* C) {
    OZA_STAGE_PROLOGUE(2)
    OZA_HIP(launch_recon_s2_digits(d, static_cast<const uint8_t*>(Ctc), shA, shB, crt_for(ctx, S),
                                   OutC{C, (int64_t)N, 1.0, 0.0}, st),
            "step 4 launch");
    return OZA_SUCCESS;
}

int oza_stage4_s1_groups(oza_ctx ctx, size_t M, size_t N, size_t K, int S, const int32_t* T, const int32_t* shA,
                         const int32_t* shB, double* C) {
    OZA_STAGE_PROLOGUE(1)
    OZA_HIP(launch_recon_s1_groups(d, scheme1_beta(K), T, shA, shB, C, st), "step 4 launch");
    return OZA_SUCCESS;
}

// ------------------------------------------------------------------------ accuracy
int oza_dd_gemm(oza_ctx ctx, size_t M, size_t N, size_t K, const double* A, const double* B, double* Cref,
                double* Cabs) {
    if (int r = check_ctx(ctx)) return r;
    if (!A || !B || !Cref) return fail(OZA_INVALID_ARGUMENT, "null pointer");
    DeviceGuard g(ctx->device);
    OZA_HIP(launch_dd_gemm((int64_t)M, (int64_t)N, (int64_t)K, A, B, Cref, Cabs, ctx->stream), "dd gemm");
    return OZA_SUCCESS;
}

int oza_dgemm(oza_ctx ctx, size_t M, size_t N, size_t K, const double* A, const double* B, double* C) {
    void* h = nullptr;
    if (int r = oza_blas_handle(ctx, &h)) return r;
    DeviceGuard g(ctx->device);
    const double alpha = 1.0, beta = 0.0;
    // Row-major C = A B as column-major C^T = B^T A^T.
    rocblas_status s = rocblas_dgemm((rocblas_handle)h, rocblas_operation_none, rocblas_operation_none, (int)N,
                                     (int)M, (int)K, &alpha, B, (int)N, A, (int)K, &beta, C, (int)N);
    if (s != rocblas_status_success) return fail(OZA_HIP_ERROR, "rocblas_dgemm failed");
    return OZA_SUCCESS;
}

int oza_mfma_ceiling(oza_ctx ctx, int zero_operands, double out[2]) {
    if (int r = check_ctx(ctx)) return r;
    if (!out) return fail(OZA_INVALID_ARGUMENT, "null output");
    DeviceGuard g(ctx->device);
    constexpr int kIters = 30000;  // ~40 ms at the power-limited clock: long enough to settle it
    hipEvent_t e0 = ctx->ev[0], e1 = ctx->ev[4];
    OZA_HIP(launch_mfma_ceiling(kIters / 10, zero_operands != 0, (unsigned)ctx->cus, ctx->stats_dev, ctx->stream),
            "mfma ceiling");  // warm-up
    OZA_HIP(hipEventRecord(e0, ctx->stream), "event");
    OZA_HIP(launch_mfma_ceiling(kIters, zero_operands != 0, (unsigned)ctx->cus, ctx->stats_dev, ctx->stream),
            "mfma ceiling");
    OZA_HIP(hipEventRecord(e1, ctx->stream), "event");
    unsigned long long tmp[4];
    OZA_HIP(hipMemcpyAsync(tmp, ctx->stats_dev, sizeof(tmp), hipMemcpyDeviceToHost, ctx->stream), "copy");
    OZA_HIP(hipStreamSynchronize(ctx->stream), "sync");
    float ms = 0.f;
    OZA_HIP(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    const double ops = (double)ctx->cus * 8 /* waves */ * 32 /* MFMAs per iteration */ * 65536.0 * kIters;
    out[0] = ms > 0.f ? ops / (ms * 1e-3) / 1e12 : 0.0;
    out[1] = tmp[1] ? (double)tmp[0] / (double)tmp[1] * 0.1 : 0.0;  // cycles per 10 ns tick -> GHz
    return OZA_SUCCESS;
}

int oza_error_stats(oza_ctx ctx, size_t n, const double* exact, const double* actual, const double* absab,
                    double out[4]) {
    if (int r = check_ctx(ctx)) return r;
    if (!out || (n && (!exact || !actual))) return fail(OZA_INVALID_ARGUMENT, "null pointer");
    DeviceGuard g(ctx->device);
    OZA_HIP(launch_error_stats((int64_t)n, exact, actual, absab, ctx->stats_dev, ctx->stream), "error stats");
    double tmp[4];
    OZA_HIP(hipMemcpyAsync(tmp, ctx->stats_dev, sizeof(tmp), hipMemcpyDeviceToHost, ctx->stream), "copy");
    OZA_HIP(hipStreamSynchronize(ctx->stream), "sync");
    out[0] = tmp[0];
    out[1] = tmp[1];
    out[2] = tmp[2];
    unsigned long long cnt;
    std::memcpy(&cnt, &tmp[3], sizeof(cnt));
    out[3] = (double)cnt;
    return OZA_SUCCESS;
}

}  // extern "C"
