// capi_internal.hpp — state and helpers shared by the C-ABI translation units.
#pragma once

#include <rocblas/rocblas.h>

#include <cstring>
#include <mutex>
#include <string>

#include "group_internal.hpp"
#include "ozablas_hip.h"
#include "oza_internal.hpp"

struct oza_context {
    int device = 0;
    hipStream_t stream = nullptr;
    oza::CrtTables* crt_dev = nullptr;  // resident per device
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // Scheme II row bands: band b's reconstruction runs on `side` while band b+1's products run on
    // `stream`.  bands = 0 picks the count per call, 1 keeps the steps in sequence.
    static constexpr int kMaxBands = 8;
    int bands = 0;
    int cus = 0;
    hipStream_t side = nullptr;
    hipEvent_t band_ev[kMaxBands] = {};
    hipEvent_t join_ev = nullptr;
    double* stats_dev = nullptr;   // 4 doubles for oza_error_stats
    // Work counters of the step-3 GEMMs: 8 per row band (kMaxBands sets), cleared by each call's
    // step-1 key-clearing launch.
    unsigned* queue_dev = nullptr;
    rocblas_handle blas = nullptr;
    std::mutex blas_mu;
    // Optional per-call stage events recorded without synchronizing (for profiling a loop).
    static constexpr int kRing = 512;
    bool ring_on = false;
    int ring_n = 0;
    hipEvent_t* ring = nullptr;  // kRing * 5 events, created on first enable
    // Grouped GEMM cost model, CU-microseconds per tile: cost[2c] + cost[2c+1] * (K steps of
    // 128) for the tile edges 256, 128, 64 (c = 0, 1, 2).  oza_group_autotune measures it; the
    // defaults are one such measurement on an MI355X.
    double group_cost[6] = {oza::kDefaultGroupCost[0], oza::kDefaultGroupCost[1], oza::kDefaultGroupCost[2],
                            oza::kDefaultGroupCost[3], oza::kDefaultGroupCost[4], oza::kDefaultGroupCost[5]};
    bool group_tuned = false;
    int group_slots[3] = {0, 0, 0};  // resident workgroups per grouped GEMM tile shape
    // Workspace of oza_gemm calls that pass none (grown on demand, freed with the context).
    std::mutex ws_mu;
    void* ws = nullptr;
    size_t ws_bytes = 0;
    // Device address of the scratch array of oza_gemm_noalloc on this context's device (from
    // hipGetSymbolAddress at context creation; one array per device, shared by its contexts).
    void* scratch = nullptr;
};

namespace oza {
namespace capi {

inline thread_local std::string g_last_error;

// The grouped planner's view of a context's device.
inline GroupModel group_model(const oza_context* c) {
    GroupModel m;
    std::memcpy(m.cost, c->group_cost, sizeof(m.cost));
    m.cus = c->cus;
    for (int i = 0; i < 3; ++i) m.slots[i] = c->group_slots[i];
    return m;
}

inline int fail(int status, const std::string& msg) {
    g_last_error = msg;
    return status;
}

inline int hip_fail(hipError_t e, const char* what) {
    return fail(OZA_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

#define OZA_HIP(call, what)                              \
    do {                                                 \
        hipError_t _e = (call);                          \
        if (_e != hipSuccess) return hip_fail(_e, what); \
    } while (0)

struct DeviceGuard {  // make the context's device current for the duration of a call
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// Largest K for which the INT32 products are exact: Scheme II |C_s| <= K * 125^2, Scheme I
// |T_g| <= min(g-1, S) * K * 64^2 (reference analysis of symmetric_mod_8 and the slice bound).
inline bool k_exact(int scheme, size_t K, int S) {
    if (scheme == 2) return (unsigned long long)K * 125ull * 125ull < (1ull << 31);
    return (unsigned long long)K * (unsigned long long)S * 4096ull < (1ull << 31);
}

inline int check_problem(int scheme, size_t M, size_t N, size_t K, int S) {
    if (scheme != 1 && scheme != 2) return fail(OZA_INVALID_ARGUMENT, "scheme must be 1 or 2");
    if (scheme == 2 && (S < 1 || S > OZA_MAX_SLICES))
        return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: Scheme II supports between 1 and 20 slices.");
    if (scheme == 1 && (S < 1 || S > OZA_MAX_SLICES))
        return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: Scheme I supports between 1 and 20 slices.");
    if (K < 1) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: K must be >= 1.");
    if (!k_exact(scheme, K, S))
        return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: K too large for exact INT32 accumulation.");
    if (M > (1ull << 31) || N > (1ull << 31) || K > (1ull << 31))
        return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: dimension too large.");
    return OZA_SUCCESS;
}

inline int check_ctx(oza_ctx ctx) {
    if (!ctx) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null context.");
    return OZA_SUCCESS;
}

inline const CrtSet* crt_for(oza_ctx ctx, int S) { return &ctx->crt_dev->set[S - 1]; }

}  // namespace capi
}  // namespace oza
