// k_group.hip — grouped Scheme II: every stage of the pipeline as one launch over the work items
// of all problems of a group.  The per-block bodies are the single-problem ones (stats_core.hpp,
// slice_core.hpp, igemm_core.hpp, recon_core.hpp), so a problem's bits do not depend on the
// company it keeps: each problem's C equals oza_scheme2_gemm's for the same M, N, K, S.
//
// Workgroup -> item: blockIdx.x indexes the stage's item list, built on the host (GEMM items in
// longest-first order and dealt to the XCDs in runs, see xcd_runs).
#include <algorithm>

#include "group_internal.hpp"
#include "igemm_core.hpp"
#include "recon_core.hpp"
#include "stats_core.hpp"

namespace oza {
namespace {

using namespace core;

using GCfg256 = IgemmCfgS2;
// 128 x 128 with 64-byte K stages: 32 KiB of LDS per workgroup, so four of them share a CU and hide
// each other's short-K fill and epilogue (grouped GEMM step -4.5 % against 128-byte stages).
using GCfg128 = IgemmCfg<128, 128, 64, 2, 2, 2, 4>;
using GCfg64 = IgemmCfg<64, 64, 128, 2, 2, 2, 4>;

__global__ void __launch_bounds__(256) k_group_zero(uint32_t* __restrict__ keys, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) keys[i] = 0;
}

__global__ void __launch_bounds__(kStatThreads) k_group_stats(GroupLaunch g) {
    const GItem it = g.items[blockIdx.x];
    const GroupProblem& P = g.probs[it.prob];
    const GroupPtrs q = g.ptrs[it.prob];
    if (it.local < P.nblkA) {
        const bool vec = (P.K % 2 == 0) && ((reinterpret_cast<uintptr_t>(q.A) & 15) == 0);
        uint32_t* key = reinterpret_cast<uint32_t*>(P.shA);
        if (vec) stats_A_block<true>(q.A, P.M, P.K, P.K, P.chunksA, key, it.local);
        else stats_A_block<false>(q.A, P.M, P.K, P.K, P.chunksA, key, it.local);
    } else {
        const bool vec = (P.N % 2 == 0) && ((reinterpret_cast<uintptr_t>(q.B) & 15) == 0);
        uint32_t* key = reinterpret_cast<uint32_t*>(P.shB);
        if (vec) stats_B_block<true>(q.B, P.K, P.N, P.N, P.colblkB, key, it.local - P.nblkA);
        else stats_B_block<false>(q.B, P.K, P.N, P.N, P.colblkB, key, it.local - P.nblkA);
    }
}

__global__ void __launch_bounds__(256) k_group_finalize(GroupLaunch g) {
    const GItem it = g.items[blockIdx.x];
    const GroupProblem& P = g.probs[it.prob];
    const int64_t i = (int64_t)it.local * 256 + threadIdx.x;
    if (i < P.M) {
        if (!P.nfA) key_to_shift(reinterpret_cast<uint32_t*>(P.shA) + i, P.kparam);  // fused: written as shifts
    } else if (i - P.M < P.N) {
        key_to_shift(reinterpret_cast<uint32_t*>(P.shB) + (i - P.M), P.kparam);
    }
}

// Steps 1 and 2 of A in one pass for problems with short K (fused_s2_A_block).
__global__ void __launch_bounds__(kFuseThreads) k_group_fuse_a(GroupLaunch g) {
    __shared__ uint32_t s_key[kFuseThreads];
    __shared__ int32_t s_shift[128];
    const GItem it = g.items[blockIdx.x];
    const GroupProblem& P = g.probs[it.prob];
    const GroupPtrs q = g.ptrs[it.prob];
    if (P.wide)
        fused_s2_A_block<true>(q.A, P.K, P.M, P.K, P.Mp, P.Kp, P.kparam, P.S, g.rc, P.shA, P.Asl, it.local, s_key,
                               s_shift);
    else
        fused_s2_A_block<false>(q.A, P.K, P.M, P.K, P.Mp, P.Kp, P.kparam, P.S, g.rc, P.shA, P.Asl, it.local, s_key,
                                s_shift);
}

__global__ void __launch_bounds__(kSliceThreads) k_group_slice(GroupLaunch g) {
    const GItem it = g.items[blockIdx.x];
    const GroupProblem& P = g.probs[it.prob];
    const GroupPtrs q = g.ptrs[it.prob];
    const int64_t kb = P.Kp / 128;
    if (it.local < P.slA) {
        const int64_t bx = it.local / kb, by = it.local - bx * kb;
        if (P.wide) slice_s2_A_block<true, true>(q.A, P.K, P.shA, P.M, P.K, P.Mp, P.Kp, P.S, g.rc, P.Asl, bx, by);
        else slice_s2_A_block<false, true>(q.A, P.K, P.shA, P.M, P.K, P.Mp, P.Kp, P.S, g.rc, P.Asl, bx, by);
    } else {
        const int64_t l = it.local - P.slA;
        const int64_t bx = l / kb, by = l - bx * kb;
        if (P.wide) slice_s2_B_block<true, true>(q.B, P.N, P.shB, P.K, P.N, P.Np, P.Kp, P.S, g.rc, P.Bsl, bx, by);
        else slice_s2_B_block<false, true>(q.B, P.N, P.shB, P.K, P.N, P.Np, P.Kp, P.S, g.rc, P.Bsl, bx, by);
    }
}

// Physical block -> item.  Workgroups are dealt round-robin to the 8 XCDs; this deals runs of
// kRun consecutive items instead, so each XCD works through its runs in list order (the list is
// longest-first, and every XCD gets an even share of long and short items) while the tiles of a
// run, which share operand panels, meet in one L2.  A bijection on [0, n): blocks of the last
// partial set of 8 runs map to themselves.
constexpr int64_t kRun = 16;
__device__ __forceinline__ int64_t xcd_runs(int64_t b, int64_t n) {
    const int64_t full = n / (8 * kRun) * (8 * kRun);
    if (b >= full) return b;
    const int64_t i = b >> 3, x = b & 7;
    return ((i / kRun) * 8 + x) * kRun + i % kRun;
}

// Persistent: workgroup b walks items b, b + gridDim.x, ... (the resident workgroups fill the
// chip once).  Its items form one chain through the LDS ring (mainloop_chain): the next item's
// first two stages are issued during the current item's last two iterations.
// An item's descriptor (operand and digit addresses, K steps, modulus) takes two dependent reads
// (item list, problem); the next item's is read before the current mainloop, so those reads
// overlap the MFMAs instead of delaying the next item's first loads.
template <class C>
__global__ void __launch_bounds__(C::NT) k_group_gemm(GroupLaunch g) {
    __shared__ __attribute__((aligned(1024))) int8_t smem[C::LDS_BYTES];
    int64_t w = blockIdx.x;
    if (w >= g.n) return;  // whole workgroup leaves together
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / C::NWN, wn = wave % C::NWN;
    struct Item {
        const int8_t* a;
        const int8_t* b;
        uint8_t* plane;  // digit plane of modulus s (a partial plane for splits j > 0)
        int64_t Kp, Np, M, N;
        int tm, tn, k0, KT;  // K steps [k0, k0 + KT) of BK bytes
        uint32_t m;
        double rcp;
    };
    constexpr int kSteps128 = 128 / C::BK;  // K steps per 128-byte K block
    auto decode = [&](int64_t wi) {
        Item t;
        const auto& it = as_constant(g.items)[xcd_runs(wi, g.n)];
        const uint32_t local = (uint32_t)it.local;
        const auto& P = as_constant(g.probs)[it.prob];
        const uint32_t tiles = (uint32_t)P.tiles_m * (uint32_t)P.tiles_n;  // < 2^31: items are int32
        const uint32_t per_split = (uint32_t)P.S * tiles;
        const uint32_t j = P.split > 1 ? local / per_split : 0u;
        const uint32_t rest = local - j * per_split;
        const int s = (int)(rest / tiles);
        grouped_tile<C>((int64_t)(rest - (uint32_t)s * tiles), P.tiles_m, P.tiles_n, t.tm, t.tn);
        t.Kp = P.Kp;
        t.Np = P.Np;
        t.M = P.M;
        t.N = P.N;
        t.k0 = (int)j * P.kbs * kSteps128;
        t.KT = min(P.kbs * kSteps128, (int)(P.Kp / C::BK) - t.k0);
        t.a = P.Asl + (int64_t)s * P.Mp * P.Kp + (int64_t)t.tm * C::BM * P.Kp;
        t.b = P.Bsl + (int64_t)s * P.Np * P.Kp + (int64_t)t.tn * C::BN * P.Kp;
        const int64_t plane = P.Mp * P.Np;
        t.plane = (j == 0 ? P.dg : P.dgp + (int64_t)(j - 1) * P.S * plane) + (int64_t)s * plane;
        const auto& cs = as_constant(g.crt)->set[P.S - 1];
        t.m = cs.m[s];
        t.rcp = cs.rcp[s];
        return t;
    };
    Item cur = decode(w);
    auto own_of = [&](const Item& t) {
        return [&t](int i) { return StageSrc{t.a, t.b, t.Kp, (t.k0 + i) * C::BK}; };
    };
    chain_start<C>(smem, 0, cur.KT, own_of(cur));  // later items' stages arrive through the chain
    int base = 0;  // ring slot of the current item's stage 0
    for (;;) {
        const int64_t w2 = w + gridDim.x;
        const bool more = w2 < g.n;
        Item nxt = cur;
        if (more) nxt = decode(w2);
        // Source of the next item's stage j; past its stages (or with no next item) a refetch of
        // the current item's last stage.
        auto tail = [&](int j) {
            const bool real = more && j < nxt.KT;
            return StageSrc{real ? nxt.a : cur.a, real ? nxt.b : cur.b, real ? nxt.Kp : cur.Kp,
                            (real ? nxt.k0 + j : cur.k0 + cur.KT - 1) * C::BK};
        };
        v16i acc[C::FM][C::FN];
        zero_acc<C>(acc);
        mainloop_chain<C>(acc, smem, base, cur.KT, own_of(cur), tail);
        store_digits<C, true, true>(acc, cur.plane, cur.Np, (int64_t)cur.tm * C::BM + wm * C::WTM,
                                    (int64_t)cur.tn * C::BN + wn * C::WTN, cur.M, cur.N, cur.m, cur.rcp, lane);
        if (!more) break;
        // the epilogue's stores are drained by the next item's first stage wait
        base = (base + cur.KT) & 1;
        w = w2;
        cur = nxt;
    }
    wait_vmcnt<0>();  // no direct load in flight into LDS when the workgroup leaves
}

// Workgroups of a GEMM shape that fit on the device at once (LDS and registers permitting),
// times `share` (at least one).
template <class C>
hipError_t gemm_grid(int64_t n, double share, unsigned& grid) {
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_group_gemm<C>, C::NT, 0);
    const int64_t resident = (int64_t)std::max(cus, 1) * std::max(per_cu, 1);
    const int64_t allowed = share >= 1.0 ? resident : std::max<int64_t>(1, (int64_t)(resident * share));
    grid = (unsigned)std::min<int64_t>(n, allowed);
    return e;
}

// Split-K reduction: adds the partial digits of splits 1..split-1 into the digit planes, mod m_s.
// Item = 4 KiB of one problem's S digit planes (Mp * Np is a multiple of 4096, so an item lies in
// one plane); thread t takes bytes 16t..16t+15.
__global__ void __launch_bounds__(256) k_group_reduce(GroupLaunch g) {
    const GItem it = g.items[blockIdx.x];
    const GroupProblem& P = g.probs[it.prob];
    const int64_t plane = P.Mp * P.Np;
    const int64_t off = (int64_t)it.local * 4096 + threadIdx.x * 16;
    const int s = (int)(off / plane);
    reduce_digits16(P.dg, P.dgp, (int64_t)P.S * plane, P.split - 1, g.crt->set[P.S - 1].m[s], off);
}

template <int S>
__device__ __forceinline__ void group_recon_S(const GroupProblem& P, double* C, const CrtSet* cs, int64_t R,
                                              int64_t Q) {
    recon_digits_block<S>(P.dg, P.Mp * P.Np, P.Np, P.shA, P.shB, P.M, P.N, cs, C, P.N, 1.0, 0.0, R, Q);
}

__global__ void __launch_bounds__(256) k_group_recon(GroupLaunch g) {
    const GItem it = g.items[blockIdx.x];
    const GroupProblem& P = g.probs[it.prob];
    double* C = g.ptrs[it.prob].C;
    const int64_t qn = (P.N + 31) >> 5;
    const int64_t R = it.local / qn, Q = it.local - R * qn;
    const CrtSet* cs = &g.crt->set[P.S - 1];
    switch (P.S) {  // uniform per workgroup; items are sorted by S on the host
        case 1: group_recon_S<1>(P, C, cs, R, Q); break;
        case 2: group_recon_S<2>(P, C, cs, R, Q); break;
        case 3: group_recon_S<3>(P, C, cs, R, Q); break;
        case 4: group_recon_S<4>(P, C, cs, R, Q); break;
        case 5: group_recon_S<5>(P, C, cs, R, Q); break;
        case 6: group_recon_S<6>(P, C, cs, R, Q); break;
        case 7: group_recon_S<7>(P, C, cs, R, Q); break;
        case 8: group_recon_S<8>(P, C, cs, R, Q); break;
        case 9: group_recon_S<9>(P, C, cs, R, Q); break;
        case 10: group_recon_S<10>(P, C, cs, R, Q); break;
        case 11: group_recon_S<11>(P, C, cs, R, Q); break;
        case 12: group_recon_S<12>(P, C, cs, R, Q); break;
        case 13: group_recon_S<13>(P, C, cs, R, Q); break;
        case 14: group_recon_S<14>(P, C, cs, R, Q); break;
        case 15: group_recon_S<15>(P, C, cs, R, Q); break;
        case 16: group_recon_S<16>(P, C, cs, R, Q); break;
        case 17: group_recon_S<17>(P, C, cs, R, Q); break;
        case 18: group_recon_S<18>(P, C, cs, R, Q); break;
        case 19: group_recon_S<19>(P, C, cs, R, Q); break;
        case 20: group_recon_S<20>(P, C, cs, R, Q); break;
        default: break;
    }
}

__global__ void __launch_bounds__(256) k_fill_pattern(double* __restrict__ p, int64_t n, uint32_t seed) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t x = (uint32_t)i * 2654435761u ^ seed;
    x ^= x >> 15;
    x *= 0x2c1b3c6du;
    x ^= x >> 12;
    x *= 0x297a2d39u;
    x ^= x >> 15;
    p[i] = (double)(int32_t)x * 0x1p-31;
}

}  // namespace

hipError_t launch_group_zero(uint32_t* keys, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_group_zero<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(keys, n);
    return hipGetLastError();
}

hipError_t launch_group_stage(int stage, const GroupLaunch& g, hipStream_t st) {
    if (g.n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)g.n;
    switch (stage) {
        case kGStats: k_group_stats<<<grid, kStatThreads, 0, st>>>(g); break;
        case kGFinalize: k_group_finalize<<<grid, 256, 0, st>>>(g); break;
        case kGFuseA: k_group_fuse_a<<<grid, kFuseThreads, 0, st>>>(g); break;
        case kGSlice: k_group_slice<<<grid, kSliceThreads, 0, st>>>(g); break;
        case kGGemm256: {
            unsigned gg = 0;
            if (hipError_t e = gemm_grid<GCfg256>(g.n, g.gemm_share, gg)) return e;
            k_group_gemm<GCfg256><<<gg, GCfg256::NT, 0, st>>>(g);
            break;
        }
        case kGGemm128: {
            unsigned gg = 0;
            if (hipError_t e = gemm_grid<GCfg128>(g.n, g.gemm_share, gg)) return e;
            k_group_gemm<GCfg128><<<gg, GCfg128::NT, 0, st>>>(g);
            break;
        }
        case kGGemm64: {
            unsigned gg = 0;
            if (hipError_t e = gemm_grid<GCfg64>(g.n, g.gemm_share, gg)) return e;
            k_group_gemm<GCfg64><<<gg, GCfg64::NT, 0, st>>>(g);
            break;
        }
        case kGReduce: k_group_reduce<<<grid, 256, 0, st>>>(g); break;
        case kGRecon: k_group_recon<<<grid, 256, 0, st>>>(g); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t group_gemm_slots(int slots[3]) {
    int dev = 0, cus = 0, occ[3] = {0, 0, 0};
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[0], k_group_gemm<GCfg256>, GCfg256::NT, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[1], k_group_gemm<GCfg128>, GCfg128::NT, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[2], k_group_gemm<GCfg64>, GCfg64::NT, 0);
    for (int c = 0; c < 3; ++c) slots[c] = std::max(cus, 1) * std::max(occ[c], 1);
    return e;
}

hipError_t launch_fill_pattern(double* p, int64_t n, uint32_t seed, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_fill_pattern<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, n, seed);
    return hipGetLastError();
}

}  // namespace oza
