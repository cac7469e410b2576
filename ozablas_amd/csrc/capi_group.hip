// capi_group.hip — grouped Scheme II (include/ozablas_hip.h, oza_group_*): planning of a batch of
// GEMMs into per-stage work lists, one owned workspace, launches (eager or HIP-graph replay) and
// the tile cost model.
#include <algorithm>
#include <climits>
#include <cstring>
#include <new>
#include <vector>

#include "capi_internal.hpp"
#include "group_internal.hpp"
#include "stats_core.hpp"

using namespace oza;
using namespace oza::capi;

struct oza_group {
    oza_ctx ctx = nullptr;
    int count = 0;
    HostPlan plan;  // descriptors hold device addresses after planning
    GroupProblem* probs_dev = nullptr;
    GroupPtrs* ptrs_dev = nullptr;
    GroupPtrs* staging = nullptr;  // pinned host copy of the pointer table
    std::vector<GroupPtrs> bound;  // pointers of the last upload
    bool bound_valid = false;
    bool upload_pending = false;
    GItem* items_dev = nullptr;
    char* ws = nullptr;
    hipEvent_t upload_done = nullptr;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    DevPlan dev() const {
        DevPlan d;
        d.probs = probs_dev;
        d.ptrs = ptrs_dev;
        d.items = items_dev;
        d.keys = reinterpret_cast<uint32_t*>(ws);
        d.nkeys = plan.nkeys;
        for (int s = 0; s < kGStages; ++s) {
            d.stage_off[s] = plan.stage_off[s];
            d.stage_n[s] = plan.stage_n[s];
        }
        return d;
    }
};

namespace {

void release(oza_group* g) {
    if (!g) return;
    DeviceGuard dg(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->upload_done) (void)hipEventDestroy(g->upload_done);
    for (auto& e : g->ev)
        if (e) (void)hipEventDestroy(e);
    if (g->probs_dev) (void)hipFree(g->probs_dev);
    if (g->ptrs_dev) (void)hipFree(g->ptrs_dev);
    if (g->items_dev) (void)hipFree(g->items_dev);
    if (g->ws) (void)hipFree(g->ws);
    if (g->staging) (void)hipHostFree(g->staging);
    delete g;
}

int plan(oza_group* g, const size_t* M, const size_t* N, const size_t* K, const int* S, int tile, const int* tiles,
         const int* splits) {
    oza_ctx ctx = g->ctx;
    const int n = g->count;
    if (int r = build_plan(group_model(ctx), n, M, N, K, S, tile, g->plan, tiles, splits)) return r;
    DeviceGuard dg(ctx->device);
    if (hipMalloc(&g->ws, g->plan.ws_bytes) != hipSuccess) {
        (void)hipGetLastError();
        g->ws = nullptr;
        return fail(OZA_ALLOC_FAILED, "HIP: group workspace allocation of " + std::to_string(g->plan.ws_bytes) +
                                          " bytes failed.");
    }
    relocate(g->plan, g->ws);
    const int64_t total = (int64_t)g->plan.items.size();
    OZA_HIP(hipMalloc(&g->probs_dev, sizeof(GroupProblem) * n), "HIP: group allocation");
    OZA_HIP(hipMalloc(&g->ptrs_dev, sizeof(GroupPtrs) * n), "HIP: group allocation");
    OZA_HIP(hipMalloc(&g->items_dev, sizeof(GItem) * std::max<int64_t>(total, 1)), "HIP: group allocation");
    OZA_HIP(hipHostMalloc(&g->staging, sizeof(GroupPtrs) * n, hipHostMallocDefault), "HIP: pinned allocation");
    OZA_HIP(hipEventCreate(&g->upload_done), "event create");
    for (auto& e : g->ev) OZA_HIP(hipEventCreate(&e), "event create");
    OZA_HIP(hipMemcpy(g->probs_dev, g->plan.probs.data(), sizeof(GroupProblem) * n, hipMemcpyHostToDevice),
            "HIP: group upload");
    if (total)
        OZA_HIP(hipMemcpy(g->items_dev, g->plan.items.data(), sizeof(GItem) * total, hipMemcpyHostToDevice),
                "HIP: group upload");
    g->bound.assign(n, GroupPtrs{nullptr, nullptr, nullptr});
    return OZA_SUCCESS;
}

int enqueue(oza_group* g, hipEvent_t* ev) {
    OZA_HIP(enqueue_plan(g->dev(), g->ctx->crt_dev, g->ctx->stream, ev), "group pipeline launch");
    return OZA_SUCCESS;
}

int bind(oza_group* g, const double* const* A, const double* const* B, double* const* C) {
    bool same = g->bound_valid;
    for (int i = 0; i < g->count; ++i) {
        const GroupProblem& P = g->plan.probs[i];
        if (P.M > 0 && P.N > 0 && (!A[i] || !B[i] || !C[i]))
            return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null matrix pointer in group problem " + std::to_string(i) + ".");
        same = same && g->bound[i].A == A[i] && g->bound[i].B == B[i] && g->bound[i].C == C[i];
    }
    if (same) return OZA_SUCCESS;
    hipStream_t st = g->ctx->stream;
    if (g->upload_pending) OZA_HIP(hipEventSynchronize(g->upload_done), "HIP: group upload");
    for (int i = 0; i < g->count; ++i) g->bound[i] = GroupPtrs{A[i], B[i], C[i]};
    std::memcpy(g->staging, g->bound.data(), sizeof(GroupPtrs) * g->count);
    OZA_HIP(hipMemcpyAsync(g->ptrs_dev, g->staging, sizeof(GroupPtrs) * g->count, hipMemcpyHostToDevice, st),
            "HIP: group pointer upload");
    OZA_HIP(hipEventRecord(g->upload_done, st), "event");
    g->upload_pending = true;
    g->bound_valid = true;
    return OZA_SUCCESS;
}

int capture(oza_group* g) {
    hipStream_t st = g->ctx->stream;
    OZA_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "HIP: graph capture");
    const int r = enqueue(g, nullptr);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    if (r != OZA_SUCCESS) {
        if (graph) (void)hipGraphDestroy(graph);
        return r;
    }
    if (e != hipSuccess) return hip_fail(e, "HIP: graph capture");
    hipGraphExec_t exec = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e2 != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return hip_fail(e2, "HIP: graph instantiate");
    }
    g->graph = graph;
    g->exec = exec;
    return OZA_SUCCESS;
}

int run(oza_group* g, const double* const* A, const double* const* B, double* const* C, int use_graph,
        oza_timings* t) {
    if (!g) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null group.");
    if (!A || !B || !C) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null pointer array.");
    DeviceGuard dg(g->ctx->device);
    if (int r = bind(g, A, B, C)) return r;
    hipStream_t st = g->ctx->stream;
    if (t) {
        if (int r = enqueue(g, g->ev)) return r;
        OZA_HIP(hipEventSynchronize(g->ev[4]), "HIP: group pipeline failed");
        OZA_HIP(hipGetLastError(), "HIP: kernel error");
        (void)hipEventElapsedTime(&t->step1_ms, g->ev[0], g->ev[1]);
        (void)hipEventElapsedTime(&t->step2_ms, g->ev[1], g->ev[2]);
        (void)hipEventElapsedTime(&t->step3_ms, g->ev[2], g->ev[3]);
        (void)hipEventElapsedTime(&t->step4_ms, g->ev[3], g->ev[4]);
        (void)hipEventElapsedTime(&t->total_ms, g->ev[0], g->ev[4]);
        return OZA_SUCCESS;
    }
    if (use_graph) {
        if (!g->exec)
            if (int r = capture(g)) return r;
        OZA_HIP(hipGraphLaunch(g->exec, st), "HIP: graph launch");
        return OZA_SUCCESS;
    }
    return enqueue(g, nullptr);
}

bool valid_tile(int t) { return t == 0 || t == 256 || t == 128 || t == 64; }

int create(oza_ctx ctx, int count, const size_t* M, const size_t* N, const size_t* K, const int* S, int tile,
           oza_group_t* out, const int* tiles = nullptr, const int* splits = nullptr) {
    if (int r = check_ctx(ctx)) return r;
    if (!out) return fail(OZA_INVALID_ARGUMENT, "null output");
    *out = nullptr;
    if (count < 1 || !M || !N || !K || !S) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: empty group.");
    if (!valid_tile(tile)) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: tile must be 0, 256, 128 or 64.");
    for (int i = 0; i < count; ++i) {
        if (int r = check_problem(2, M[i], N[i], K[i], S[i])) return r;
        if (tiles && !valid_tile(tiles[i]))
            return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: tile must be 0, 256, 128 or 64.");
        if (splits && (splits[i] < 0 || splits[i] > kSplitMax))
            return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: K split must be between 0 and " + std::to_string(kSplitMax) + ".");
    }
    oza_group* g = new (std::nothrow) oza_group;
    if (!g) return fail(OZA_ALLOC_FAILED, "group");
    g->ctx = ctx;
    g->count = count;
    if (int r = plan(g, M, N, K, S, tile, tiles, splits)) {
        release(g);
        return r;
    }
    *out = g;
    return OZA_SUCCESS;
}

}  // namespace


// ------------------------------------------------------------------------------------------
// Planner and launcher shared by groups and the request server.
// ------------------------------------------------------------------------------------------
namespace oza {

namespace {
constexpr int kStageThreads[kGStages] = {256, 256, core::kFuseThreads, 256, 512, 256, 256, 256, 256};
size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
}  // namespace

int choose_tile(const double* cost, int64_t M, int64_t N, int64_t Kp, int S) {
    int best = 0;
    double best_t = 0.0;
    for (int c = 0; c < 3; ++c) {
        const int64_t T = kGroupTiles[c];
        const double tiles = (double)((M + T - 1) / T) * (double)((N + T - 1) / T) * S;
        const double t = tiles * (cost[2 * c] + cost[2 * c + 1] * (double)(Kp / 128));
        if (c == 0 || t < best_t) {
            best = c;
            best_t = t;
        }
    }
    return kGroupTiles[best];
}

size_t problem_bytes(size_t M, size_t N, size_t K, int S, int tile, int split) {
    const size_t Mp = round_up(M, (size_t)tile), Np = round_up(N, (size_t)tile), Kp = round_up(K, kPadK);
    const size_t digits = (size_t)S * Mp * Np;
    return 4 * (round_up(M, 64) + round_up(N, 64)) + align256((size_t)S * Mp * Kp) + align256((size_t)S * Np * Kp) +
           align256(digits) + (split > 1 ? align256((size_t)(split - 1) * digits) : 0) + 256;
}

namespace {

// 128-byte K blocks per split and the number of non-empty splits for a requested split p.
void split_shape(int64_t kb, int p, int32_t& split, int32_t& kbs) {
    p = (int)std::max<int64_t>(1, std::min<int64_t>(p, kb));
    kbs = (int32_t)((kb + p - 1) / p);
    split = (int32_t)((kb + kbs - 1) / kbs);
}

// Split planner, per GEMM tile shape (one launch each).  The launch is modelled as static
// striding over its items sorted longest first: with G workgroups resident, workgroup b runs items
// b, b + G, ..., so the launch lasts about the sum over rounds of G items of each round's first
// (longest) item.  An item lasts its tile cost (cost model, CU-microseconds) times the number of
// workgroups sharing its CU: min(occupancy, items / CUs).  A K split of P shortens a problem's
// items and multiplies their count by P; the partial digits then cost a reduction pass over
// (P + 1) digit planes.  Candidates: no split, and every duration limit T = item duration of one
// of the 16 longest problems at a power-of-two split, under which each problem takes the smallest
// power-of-two split (at most kAutoSplitMax, at least two K blocks per split) whose items last at
// most T.  The candidate with the least modelled time wins, no split on ties.  Problems with a
// caller-given split (fixed) keep it.
void split_planner(const GroupModel& model, std::vector<GroupProblem>& probs, const std::vector<char>& fixed) {
    const int64_t cus = std::max(model.cus, 1);
    constexpr double kReduceBytesPerUs = 4.0e6;  // reduction stream rate (bytes per microsecond)
    constexpr double kReduceLaunchUs = 4.0;
    for (int c = 0; c < 3; ++c) {
        const double c0 = model.cost[2 * c], c1 = model.cost[2 * c + 1];
        const int64_t occ = std::max<int64_t>(1, model.slots[c] / cus);
        std::vector<size_t> mine;
        for (size_t i = 0; i < probs.size(); ++i)
            if (probs[i].tile == kGroupTiles[c] && probs[i].M > 0 && probs[i].N > 0) mine.push_back(i);
        if (mine.empty()) continue;
        auto tile_us = [&](int64_t kbs) { return c0 + c1 * (double)kbs; };
        auto cap_of = [&](int64_t kb) { return std::max<int64_t>(1, std::min<int64_t>(kAutoSplitMax, kb / 2)); };
        // the split of problem i under limit T (T < 0: none)
        auto pick = [&](size_t i, double T) {
            const GroupProblem& P = probs[i];
            if (fixed[i]) return (int)P.split;
            const int64_t kb = P.Kp / 128;
            if (T < 0.0) return 1;
            int p = 1;
            while (p * 2 <= cap_of(kb) && tile_us((kb + p - 1) / p) > T) p *= 2;
            return p;
        };
        auto model_us = [&](double T) {
            std::vector<std::pair<double, int64_t>> runs;  // (tile cost, items)
            int64_t items = 0;
            double red = 0.0;
            for (size_t i : mine) {
                const GroupProblem& P = probs[i];
                int32_t split, kbs;
                split_shape(P.Kp / 128, pick(i, T), split, kbs);
                const int64_t n = (int64_t)P.S * P.tiles_m * P.tiles_n * split;
                runs.push_back({tile_us(kbs), n});
                items += n;
                if (split > 1) red += (double)(split + 1) * P.S * P.Mp * P.Np / kReduceBytesPerUs;
            }
            const int64_t per_cu = std::max<int64_t>(1, std::min<int64_t>(occ, (items + cus - 1) / cus));
            const int64_t G = per_cu * cus;
            std::sort(runs.begin(), runs.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
            double t = 0.0;
            int64_t pos = 0, next = 0;  // next: index of the next round's first item
            for (const auto& r : runs) {
                while (next < pos + r.second) {
                    t += r.first * (double)per_cu;
                    next += G;
                }
                pos += r.second;
            }
            return t + (red > 0.0 ? red + kReduceLaunchUs : 0.0);
        };
        // candidate limits from the 16 longest problems
        std::vector<size_t> longest(mine);
        std::sort(longest.begin(), longest.end(), [&](size_t a, size_t b) {
            return tile_us(probs[a].Kp / 128) > tile_us(probs[b].Kp / 128);
        });
        if (longest.size() > 16) longest.resize(16);
        double best_T = -1.0, best = model_us(-1.0);
        for (size_t i : longest) {
            if (fixed[i]) continue;
            const int64_t kb = probs[i].Kp / 128;
            for (int p = 2; p <= cap_of(kb); p *= 2) {
                const double T = tile_us((kb + p - 1) / p);
                const double t = model_us(T);
                if (t < best * (1.0 - 1e-9)) {
                    best = t;
                    best_T = T;
                }
            }
        }
        for (size_t i : mine)
            if (!fixed[i]) split_shape(probs[i].Kp / 128, pick(i, best_T), probs[i].split, probs[i].kbs);
    }
}

}  // namespace

int alone_split(const GroupModel& model, int64_t M, int64_t N, int64_t K, int S, int tile) {
    std::vector<GroupProblem> one(1, GroupProblem{});
    GroupProblem& P = one[0];
    P.M = M;
    P.N = N;
    P.K = K;
    P.S = S;
    P.tile = tile;
    P.Kp = (int64_t)round_up((size_t)K, kPadK);
    P.Mp = (int64_t)round_up((size_t)M, (size_t)tile);
    P.Np = (int64_t)round_up((size_t)N, (size_t)tile);
    P.tiles_m = (int32_t)(P.Mp / tile);
    P.tiles_n = (int32_t)(P.Np / tile);
    P.split = 1;
    P.kbs = (int32_t)(P.Kp / 128);
    std::vector<char> fixed(1, 0);
    split_planner(model, one, fixed);
    return P.split;
}

int scheme2_single_split(size_t M, size_t N, size_t K, int S) {
    if (M == 0 || N == 0 || K == 0) return 1;
    GroupModel model;
    std::memcpy(model.cost, kDefaultGroupCost, sizeof(model.cost));
    model.cus = 256;  // MI355X
    model.slots[0] = 256;  // one 256-wide GEMM workgroup per CU; the other shapes are not used here
    model.slots[1] = model.slots[2] = 1024;
    return alone_split(model, (int64_t)M, (int64_t)N, (int64_t)K, S, 256);
}

int build_plan(const GroupModel& model, int n, const size_t* M, const size_t* N, const size_t* K, const int* S,
               int tile, HostPlan& out, const int* tiles, const int* splits) {
    using capi::fail;
    out.probs.assign(n, GroupProblem{});
    out.items.clear();
    // ---- shapes, tiles and splits
    std::vector<char> fixed(n, 0);
    for (int i = 0; i < n; ++i) {
        GroupProblem& P = out.probs[i];
        P.M = (int64_t)M[i];
        P.N = (int64_t)N[i];
        P.K = (int64_t)K[i];
        P.S = S[i];
        P.Kp = (int64_t)round_up((size_t)P.K, kPadK);
        P.tile = tiles && tiles[i] ? tiles[i] : tile ? tile : choose_tile(model.cost, P.M, P.N, P.Kp, P.S);
        P.Mp = (int64_t)round_up((size_t)P.M, (size_t)P.tile);
        P.Np = (int64_t)round_up((size_t)P.N, (size_t)P.tile);
        P.tiles_m = (int32_t)(P.Mp / P.tile);
        P.tiles_n = (int32_t)(P.Np / P.tile);
        fixed[i] = splits && splits[i] ? 1 : 0;
        split_shape(P.Kp / 128, fixed[i] ? splits[i] : 1, P.split, P.kbs);
    }
    split_planner(model, out.probs, fixed);
    // ---- workspace offsets (addresses are offsets from the base until relocate)
    size_t off = 0;
    for (int i = 0; i < n; ++i) {  // keys first, contiguous, so that one fill clears them
        GroupProblem& P = out.probs[i];
        P.shA = reinterpret_cast<int32_t*>(off);
        off += 4 * round_up(M[i], 64);
        P.shB = reinterpret_cast<int32_t*>(off);
        off += 4 * round_up(N[i], 64);
    }
    out.nkeys = (int64_t)(off / 4);
    off = align256(off);
    for (int i = 0; i < n; ++i) {
        GroupProblem& P = out.probs[i];
        P.kparam = scheme2_kparam(P.S, (size_t)P.K);
        P.wide = P.kparam > 72 ? 1 : 0;  // residue split limit, as in the single-problem slicing
        const core::StatShape h = core::stat_shape(P.M, P.N, P.K);
        P.chunksA = h.chunksA;
        P.nblkA = h.nblkA;
        P.colblkB = h.colblkB;
        P.nblkB = h.nblkB;
        P.slA = ((P.M + 31) / 32) * (P.Kp / 128);  // blocks holding a logical row (see make_slice_op)
        P.slB = ((P.N + 63) / 64) * (P.Kp / 128);
        P.nfA = 0;
        if (P.Kp <= core::kFuseMaxKp) {  // short rows of A: steps 1 and 2 in one pass (k_group_fuse_a)
            P.nfA = ((P.M + 31) / 32 * 32 + core::fuse_rows(P.Kp) - 1) / core::fuse_rows(P.Kp);
            P.nblkA = 0;
            P.slA = 0;
        }
        P.Asl = reinterpret_cast<int8_t*>(off);
        off = align256(off + (size_t)P.S * P.Mp * P.Kp);
        P.Bsl = reinterpret_cast<int8_t*>(off);
        off = align256(off + (size_t)P.S * P.Np * P.Kp);
        P.dg = reinterpret_cast<uint8_t*>(off);
        off = align256(off + (size_t)P.S * P.Mp * P.Np);
        P.dgp = nullptr;
        if (P.split > 1) {
            P.dgp = reinterpret_cast<uint8_t*>(off);
            off = align256(off + (size_t)(P.split - 1) * P.S * P.Mp * P.Np);
        }
    }
    out.ws_bytes = std::max<size_t>(off, 256);

    // ---- work lists
    std::vector<GItem> items[kGStages];
    auto add = [&](int stage, int prob, int64_t count) -> int {
        if (count > INT32_MAX) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: problem too large for a grouped launch.");
        for (int64_t l = 0; l < count; ++l) items[stage].push_back({prob, (int32_t)l});
        return OZA_SUCCESS;
    };
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    for (int i = 0; i < n; ++i) {
        const GroupProblem& P = out.probs[i];
        if (P.M == 0 || P.N == 0) continue;
        if (int r = add(kGStats, i, P.nblkA + P.nblkB)) return r;
        if (int r = add(kGFinalize, i, (P.M + P.N + 255) / 256)) return r;
        if (int r = add(kGFuseA, i, P.nfA)) return r;
        if (int r = add(kGSlice, i, P.slA + P.slB)) return r;
    }
    // GEMM: longest items (most K blocks per split) first, per tile shape.  A problem's items run
    // split-major, then modulus, then tile, so neighbouring items share operand panels.
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return out.probs[a].kbs > out.probs[b].kbs; });
    for (int i : order) {
        const GroupProblem& P = out.probs[i];
        if (P.M == 0 || P.N == 0) continue;
        const int stage = P.tile == 256 ? kGGemm256 : P.tile == 128 ? kGGemm128 : kGGemm64;
        if (int r = add(stage, i, (int64_t)P.split * P.S * P.tiles_m * P.tiles_n)) return r;
    }
    // partial-digit reduction of split problems: 4 KiB of each digit plane per item
    for (int i = 0; i < n; ++i) {
        const GroupProblem& P = out.probs[i];
        if (P.M == 0 || P.N == 0 || P.split < 2) continue;
        if (int r = add(kGReduce, i, (int64_t)P.S * P.Mp * P.Np / 4096)) return r;
    }
    // reconstruction: grouped by moduli count (one code path per run of blocks)
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return out.probs[a].S > out.probs[b].S; });
    for (int i : order) {
        const GroupProblem& P = out.probs[i];
        if (P.M == 0 || P.N == 0) continue;
        if (int r = add(kGRecon, i, ((P.M + 31) / 32) * ((P.N + 31) / 32))) return r;
    }
    int64_t total = 0;
    for (int s = 0; s < kGStages; ++s) {
        out.stage_off[s] = total;
        out.stage_n[s] = (int64_t)items[s].size();
        if (out.stage_n[s] * kStageThreads[s] > (int64_t)UINT32_MAX)
            return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: too much work for one launch per stage; split the group.");
        total += out.stage_n[s];
    }
    out.items.reserve(total);
    for (int s = 0; s < kGStages; ++s) out.items.insert(out.items.end(), items[s].begin(), items[s].end());
    return OZA_SUCCESS;
}

void relocate(HostPlan& plan, char* base) {
    for (auto& P : plan.probs) {
        P.shA = reinterpret_cast<int32_t*>(base + reinterpret_cast<size_t>(P.shA));
        P.shB = reinterpret_cast<int32_t*>(base + reinterpret_cast<size_t>(P.shB));
        P.Asl = reinterpret_cast<int8_t*>(base + reinterpret_cast<size_t>(P.Asl));
        P.Bsl = reinterpret_cast<int8_t*>(base + reinterpret_cast<size_t>(P.Bsl));
        P.dg = reinterpret_cast<uint8_t*>(base + reinterpret_cast<size_t>(P.dg));
        if (P.dgp) P.dgp = reinterpret_cast<uint8_t*>(base + reinterpret_cast<size_t>(P.dgp));
    }
}

hipError_t enqueue_plan(const DevPlan& d, const CrtTables* crt, hipStream_t st, hipEvent_t* ev, double gemm_share) {
    GroupLaunch L;
    L.probs = d.probs;
    L.ptrs = d.ptrs;
    L.crt = crt;
    L.rc = core::make_residue_consts(kMaxSlices);
    L.gemm_share = gemm_share;
    hipError_t e = hipSuccess;
    auto stage = [&](int s) {
        if (e != hipSuccess) return;
        L.items = d.items + d.stage_off[s];
        L.n = d.stage_n[s];
        e = launch_group_stage(s, L, st);
    };
    auto mark = [&](int i) {
        if (ev && e == hipSuccess) e = hipEventRecord(ev[i], st);
    };
    mark(0);
    if (e == hipSuccess) e = launch_group_zero(d.keys, d.nkeys, st);
    stage(kGStats);
    stage(kGFinalize);
    mark(1);
    stage(kGFuseA);  // steps 1 and 2 of short-K A operands, counted with step 2
    stage(kGSlice);
    mark(2);
    stage(kGGemm256);
    stage(kGGemm128);
    stage(kGGemm64);
    stage(kGReduce);
    mark(3);
    stage(kGRecon);
    mark(4);
    return e;
}

}  // namespace oza

extern "C" {

int oza_group_create(oza_ctx ctx, int count, const size_t* M, const size_t* N, const size_t* K, const int* slices,
                     int tile, oza_group_t* out) {
    return create(ctx, count, M, N, K, slices, tile, out);
}

int oza_group_create_split(oza_ctx ctx, int count, const size_t* M, const size_t* N, const size_t* K,
                           const int* slices, const int* tiles, const int* splits, oza_group_t* out) {
    return create(ctx, count, M, N, K, slices, 0, out, tiles, splits);
}

int oza_group_plan(const double cost[6], int cus, const int slots[3], int count, const size_t* M, const size_t* N,
                   const size_t* K, const int* slices, int* tiles, int* splits, size_t* workspace_bytes) {
    if (!cost || !slots || cus < 1 || count < 1 || !M || !N || !K || !slices || !tiles || !splits)
        return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: group plan needs costs, CUs, slots, shapes, tiles and splits.");
    GroupModel model;
    std::memcpy(model.cost, cost, sizeof(model.cost));
    model.cus = cus;
    for (int c = 0; c < 3; ++c) {
        if (slots[c] < 1) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: slots must be >= 1.");
        model.slots[c] = slots[c];
    }
    for (int i = 0; i < count; ++i) {
        if (int r = check_problem(2, M[i], N[i], K[i], slices[i])) return r;
        if (!valid_tile(tiles[i])) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: tile must be 0, 256, 128 or 64.");
        if (splits[i] < 0 || splits[i] > kSplitMax)
            return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: K split must be between 0 and " + std::to_string(kSplitMax) + ".");
    }
    HostPlan plan;
    if (int r = build_plan(model, count, M, N, K, slices, 0, plan, tiles, splits)) return r;
    for (int i = 0; i < count; ++i) {
        tiles[i] = plan.probs[i].tile;
        splits[i] = plan.probs[i].split;
    }
    if (workspace_bytes) *workspace_bytes = plan.ws_bytes;
    return OZA_SUCCESS;
}

int oza_group_splits(oza_group_t g, int* splits) {
    if (!g) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null group.");
    if (!splits) return fail(OZA_INVALID_ARGUMENT, "null output");
    for (int i = 0; i < g->count; ++i) splits[i] = g->plan.probs[i].split;
    return OZA_SUCCESS;
}

int oza_group_destroy(oza_group_t g) {
    release(g);
    return OZA_SUCCESS;
}

int oza_group_info(oza_group_t g, size_t* workspace_bytes, int* tiles) {
    if (!g) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: null group.");
    if (workspace_bytes) *workspace_bytes = g->plan.ws_bytes;
    if (tiles)
        for (int i = 0; i < g->count; ++i) tiles[i] = g->plan.probs[i].tile;
    return OZA_SUCCESS;
}

int oza_group_scheme2_gemm(oza_group_t g, const double* const* A, const double* const* B, double* const* C,
                           int use_graph, oza_timings* t) {
    return run(g, A, B, C, use_graph, t);
}

int oza_group_costs(oza_ctx ctx, double cost[6], int set) {
    if (int r = check_ctx(ctx)) return r;
    if (!cost) return fail(OZA_INVALID_ARGUMENT, "null cost array");
    if (set) {
        for (int i = 0; i < 6; ++i)
            if (!(cost[i] >= 0.0)) return fail(OZA_INVALID_ARGUMENT, "OzaBLAS: costs must be >= 0.");
        std::memcpy(ctx->group_cost, cost, sizeof(ctx->group_cost));
    } else {
        std::memcpy(cost, ctx->group_cost, sizeof(ctx->group_cost));
    }
    return OZA_SUCCESS;
}

int oza_group_autotune(oza_ctx ctx, double cost[6]) {
    if (int r = check_ctx(ctx)) return r;
    DeviceGuard dg(ctx->device);
    hipDeviceProp_t prop;
    OZA_HIP(hipGetDeviceProperties(&prop, ctx->device), "HIP: device properties");
    const double cus = (double)prop.multiProcessorCount;
    const size_t n = 4096, kmax = 1024;  // several tiles per CU even at edge 256: launch cost amortized
    double *A = nullptr, *B = nullptr, *C = nullptr;
    int rc = OZA_SUCCESS;
    if (hipMalloc(&A, sizeof(double) * n * kmax) != hipSuccess || hipMalloc(&B, sizeof(double) * kmax * n) != hipSuccess ||
        hipMalloc(&C, sizeof(double) * n * n) != hipSuccess) {
        (void)hipGetLastError();
        rc = fail(OZA_ALLOC_FAILED, "HIP: autotune allocation failed.");
    }
    if (rc == OZA_SUCCESS && (launch_fill_pattern(A, (int64_t)(n * kmax), 1u, ctx->stream) != hipSuccess ||
                              launch_fill_pattern(B, (int64_t)(kmax * n), 2u, ctx->stream) != hipSuccess))
        rc = hip_fail(hipGetLastError(), "autotune fill");
    double measured[6] = {0, 0, 0, 0, 0, 0};
    const size_t Ks[2] = {128, kmax};
    for (int c = 0; c < 3 && rc == OZA_SUCCESS; ++c) {
        double per_tile[2];
        for (int j = 0; j < 2 && rc == OZA_SUCCESS; ++j) {
            const size_t M = n, N = n, K = Ks[j];
            const int S = 8;
            oza_group_t g = nullptr;
            const int unsplit = 1;
            rc = create(ctx, 1, &M, &N, &K, &S, kGroupTiles[c], &g, nullptr, &unsplit);
            if (rc != OZA_SUCCESS) break;
            const double* a = A;
            const double* b = B;
            double* cc = C;
            oza_timings t;
            float best = 1e30f;
            for (int rep = 0; rep < 4 && rc == OZA_SUCCESS; ++rep) {
                rc = run(g, &a, &b, &cc, 0, &t);
                if (rep > 0) best = std::min(best, t.step3_ms);
            }
            const int T = kGroupTiles[c];
            const double tiles = (double)((n + T - 1) / T) * (double)((n + T - 1) / T) * S;
            per_tile[j] = best * 1e3 * cus / tiles;
            release(g);
        }
        if (rc != OZA_SUCCESS) break;
        const double steps_hi = (double)(kmax / 128);
        const double slope = std::max((per_tile[1] - per_tile[0]) / (steps_hi - 1.0), 1e-6);
        measured[2 * c + 1] = slope;
        measured[2 * c] = std::max(per_tile[0] - slope, 0.0);
    }
    (void)hipStreamSynchronize(ctx->stream);
    if (A) (void)hipFree(A);
    if (B) (void)hipFree(B);
    if (C) (void)hipFree(C);
    if (rc != OZA_SUCCESS) return rc;
    std::memcpy(ctx->group_cost, measured, sizeof(measured));
    ctx->group_tuned = true;
    if (cost) std::memcpy(cost, measured, sizeof(measured));
    return OZA_SUCCESS;
}

}  // extern "C"
