// k_igemm.hip — step 3: INT8 x INT8 -> INT32 GEMMs on gfx950 matrix cores, with fused epilogues.
//
// Operands come from step 2 as fragment-tiled INT8 planes (A [s][Mp][Kp], B transposed
// [s][Np][Kp]), so every MFMA operand fragment is one contiguous 1 KiB block.  A workgroup owns a
// BM x BN output tile; each of its waves owns a (BM/NWM) x (BN/NWN) sub-tile made of 32x32
// v_mfma_i32_32x32x32_i8 fragments with exact INT32 accumulation (mainloop in igemm_core.hpp).
//
//   Scheme II (reference: rocblas_gemm_strided_batched_ex, src/hip/pipeline.hip:237-245):
//     one launch covers every (tile, modulus); the epilogue reduces C_s to the CRT digit
//     c_s = C_s mod m_s and stores it as one byte (layout below), or stores
//     the raw INT32 C_s planes in the reference's [s][M][N] layout.
//   Scheme I (reference: rocblas_gemm_ex loop, pipeline.hip:110-135): one launch covers every
//     (group g, tile); the pairs of group g = s + t accumulate into one INT32 fragment set and
//     T_g is stored as a blocked INT32 plane; step 4 (k_recon.hip) sums the groups in FP64 in
//     the reference's ascending order.  The planar form for the stage entry point walks all
//     pairs per tile and writes [g-2][M][N].
//
// Digit layout (Scheme II, C_tc workspace): planes [s] of Mp x Np bytes in 32 x 32 element
// blocks, row-major over blocks, each block in the MFMA 32x32 output order (see store_digits);
// Scheme I's T planes use the same blocking with INT32 elements (store_i32_blocked).
#include <algorithm>

#include "igemm_core.hpp"
#include "recon_core.hpp"

namespace oza {
namespace {

using namespace core;

// ------------------------------------------------------------------------------------------
// Scheme II
// ------------------------------------------------------------------------------------------
struct S2Args {
    const int8_t* A;
    const int8_t* B;
    int64_t M, N, Mp, Np, Kp;
    int S, KT, tiles_m, tiles_n;
    int tm0;               // first tile row: the launch covers tile rows [tm0, tm0 + tiles_m)
    void* out;             // digits or int32 planes
    const CrtSet* crt;     // device copy (m, inv, rcp used)
    unsigned* queue;       // 8 work counters, zero at launch (dynamic order), or null (static order)
};

// Persistent (one workgroup per CU).  Work items are taken in logical order l = 0..total-1
// (modulus-major; inside a modulus, tiles in groups of GM tile rows), split into 8 contiguous
// runs, one per XCD (workgroup b runs on XCD b % 8).
//   static order (queue == null): workgroup b takes w = b, b + gridDim.x, ... through xcd_remap;
//   dynamic order (queue != null, gridDim.x a multiple of 8, total >= 8): the workgroups of XCD v
//     take the positions of run v in the order they ask for them (counter queue[v]), so that the
//     tiles in flight on one XCD stay consecutive and share their A and B panels in L2 even when
//     CUs drift apart.  The ask for the next position is made as a tile starts; its reply is
//     passed on through LDS after the tile's mainloop.
// After a tile's mainloop, the next tile's first stages are issued before the epilogue, so the
// epilogue's ALU work hides their load latency.
template <class C, bool DIGITS>
__global__ void __launch_bounds__(C::NT) k_igemm_s2(S2Args p) {
    __shared__ __attribute__((aligned(1024))) int8_t smem[C::LDS_BYTES];
    __shared__ int64_t s_next;  // dynamic order: next position of this workgroup's run
    const int64_t tiles = (int64_t)p.tiles_m * p.tiles_n, total = tiles * p.S;
    const bool dyn = p.queue != nullptr;
    const int xcd = blockIdx.x & 7;
    const int64_t run = total >> 3, run0 = (int64_t)xcd * run;
    const int64_t nrun = xcd == 7 ? total - 7 * run : run;  // items of this XCD's run
    const int64_t nwork = dyn ? nrun : total;
    int64_t w;  // dynamic: position in the run; static: work index
    if (dyn) {
        if (threadIdx.x == 0) s_next = atomicAdd(&p.queue[xcd], 1u);
        __syncthreads();
        w = s_next;
        __syncthreads();  // every wave has read s_next before it is written again
    } else {
        w = blockIdx.x;
    }
    if (w >= nwork) return;  // whole workgroup leaves together
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / C::NWN, wn = wave % C::NWN;
    struct Item {
        int s, tm, tn;
        const int8_t* a;
        const int8_t* b;
    };
    auto decode = [&](int64_t wi) {
        Item t;
        const int64_t l = dyn ? run0 + wi : xcd_remap(wi, total);
        t.s = (int)(l / tiles);
        grouped_tile<C>(l - (int64_t)t.s * tiles, p.tiles_m, p.tiles_n, t.tm, t.tn);
        t.tm += p.tm0;
        // tile (tm, tn) starts at block-row tm*BM/32 of the fragment-tiled planes (BM*Kp bytes)
        t.a = p.A + (int64_t)t.s * p.Mp * p.Kp + (int64_t)t.tm * C::BM * p.Kp;
        t.b = p.B + (int64_t)t.s * p.Np * p.Kp + (int64_t)t.tn * C::BN * p.Kp;
        return t;
    };
    Item cur = decode(w);
    issue_prologue<C>(smem, 1, p.KT, p.Kp, [&](int, const int8_t*& pa, const int8_t*& pb) {
        pa = cur.a;
        pb = cur.b;
    });
    for (;;) {
        unsigned ticket = 0;
        if (dyn && threadIdx.x == 0) ticket = atomicAdd(&p.queue[xcd], 1u);
        v16i acc[C::FM][C::FN];
        zero_acc<C>(acc);
        mainloop<C, false>(
            acc, smem, 1, p.KT, p.Kp,
            [&](int, const int8_t*& pa, const int8_t*& pb) {
                pa = cur.a;
                pb = cur.b;
            },
            [](int) {});
        if (dyn) {
            // Every wave read the previous s_next before this tile's mainloop barriers.
            if (threadIdx.x == 0) s_next = ticket;
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): the write lands before the barrier
        }
        block_sync_lds();  // every wave is done with the ring
        const int64_t w2 = dyn ? s_next : w + gridDim.x;
        Item nxt = cur;
        if (w2 < nwork) {
            nxt = decode(w2);
            issue_prologue<C>(smem, 1, p.KT, p.Kp, [&](int, const int8_t*& pa, const int8_t*& pb) {
                pa = nxt.a;
                pb = nxt.b;
            });
        }
        const int s = cur.s, tm = cur.tm, tn = cur.tn;
        if (DIGITS) {
            uint8_t* plane = reinterpret_cast<uint8_t*>(p.out) + (int64_t)s * p.Mp * p.Np;
            store_digits<C, false>(acc, plane, p.Np, (int64_t)tm * C::BM + wm * C::WTM, (int64_t)tn * C::BN + wn * C::WTN,
                            p.M, p.N, p.crt->m[s], p.crt->rcp[s], lane);
        } else {
            int32_t* o = reinterpret_cast<int32_t*>(p.out) + (int64_t)s * p.M * p.N;
#pragma unroll
            for (int fm = 0; fm < C::FM; ++fm)
#pragma unroll
                for (int fn = 0; fn < C::FN; ++fn) {
                    const int64_t col = (int64_t)tn * C::BN + wn * C::WTN + fn * 32 + (lane & 31);
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const int64_t row = (int64_t)tm * C::BM + wm * C::WTM + fm * 32 + frag_row(j, lane);
                        if (row < p.M && col < p.N) o[row * p.N + col] = acc[fm][fn][j];
                    }
                }
        }
        if (w2 >= nwork) break;
        // The next mainloop's first wait (vmcnt(0) with two stages) also covers these stores.
        static_assert(C::STAGES == 2, "the epilogue's stores are drained by the first stage wait");
        w = w2;
        cur = nxt;
    }
}

// ------------------------------------------------------------------------------------------
// Scheme I
// ------------------------------------------------------------------------------------------
struct S1Args {
    const int8_t* A;
    const int8_t* B;
    int64_t M, N, Mp, Np, Kp;
    int S, KT, tiles_m, tiles_n, beta;
    int tm0;                 // blocked form: the launch covers tile rows [tm0, tm0 + tiles_m)
    int32_t* T;              // group products: int32 [S][M][N], or blocked [S][Mp][Np]
    unsigned* queue;         // 8 work counters, zero at launch (dynamic order), or null (static order)
};

// Scheme I group products in the reference's planar layout (the stage entry point): one
// workgroup per tile walks every pair (s, t) in the reference's order -- groups g = s + t = 2..S+1
// ascending, s ascending inside a group (src/hip/pipeline.hip:110-135) -- and stores each
// finished T_g to int32 [g-2][M][N].
template <class C>
__global__ void __launch_bounds__(C::NT) k_igemm_s1(S1Args p) {
    __shared__ __attribute__((aligned(1024))) int8_t smem[C::LDS_BYTES];
    const int64_t tiles = (int64_t)p.tiles_m * p.tiles_n;
    int tm, tn;
    grouped_tile<C>(xcd_remap(blockIdx.x, tiles), p.tiles_m, p.tiles_n, tm, tn);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / C::NWN, wn = wave % C::NWN;
    const int64_t planeA = p.Mp * p.Kp, planeB = p.Np * p.Kp;
    const int8_t* abase = p.A + (int64_t)tm * C::BM * p.Kp;
    const int8_t* bbase = p.B + (int64_t)tn * C::BN * p.Kp;
    const int S = p.S;
    const int npairs = S * (S + 1) / 2;
    auto decode = [](int q, int& g, int& sidx) {  // q -> group g (2-based) and s (1-based)
        g = 2;
        while (q >= g - 1) {
            q -= g - 1;
            ++g;
        }
        sidx = q + 1;
    };
    v16i acc[C::FM][C::FN];
    zero_acc<C>(acc);
    mainloop<C>(
        acc, smem, npairs, p.KT, p.Kp,
        [&](int q, const int8_t*& pa, const int8_t*& pb) {
            int g, sidx;
            decode(q, g, sidx);
            pa = abase + (int64_t)(sidx - 1) * planeA;
            pb = bbase + (int64_t)(g - sidx - 1) * planeB;
        },
        [&](int q) {
            int g, sidx;
            decode(q, g, sidx);
            if (sidx != g - 1) return;  // group continues
            int32_t* o = p.T + (int64_t)(g - 2) * p.M * p.N;
#pragma unroll
            for (int fm = 0; fm < C::FM; ++fm)
#pragma unroll
                for (int fn = 0; fn < C::FN; ++fn) {
                    const int64_t col = (int64_t)tn * C::BN + wn * C::WTN + fn * 32 + (lane & 31);
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const int64_t row = (int64_t)tm * C::BM + wm * C::WTM + fm * 32 + frag_row(j, lane);
                        if (row < p.M && col < p.N) o[row * p.N + col] = acc[fm][fn][j];
                    }
                }
            zero_acc<C>(acc);
        });
}

// Scheme I group products with the 256 x 256 tile of the Scheme II GEMM: work item = (group g,
// tile); the pairs s = 1..g-1 (t = g - s) of the group accumulate in INT32 through one pass of
// the pipelined mainloop, and T_g is stored as a blocked INT32 plane (store_i32_blocked).  The
// FP64 sum over groups is step 4 (k_recon_s1_blocked).
// Work order: items differ in length (g - 1 pairs), so the XCDs split the tiles, not the items:
// XCD x (workgroups b with b % 8 == x) owns the tiles [x*T/8, (x+1)*T/8) of the grouped tile order
// in every group, and walks its items largest group first, then tile -- every XCD gets the same
// mix of work and its CUs share A and B panels in L2.  The workgroups of XCD x take the positions
// of its sequence in the order they ask for them (counter queue[x]).  Without counters, or with a
// grid that is not a multiple of 8 (fewer items than CUs), workgroup b walks the whole item list,
// largest group first, at b, b + gridDim.x, ...
template <class C>
__global__ void __launch_bounds__(C::NT) k_igemm_s1g(S1Args p) {
    __shared__ __attribute__((aligned(1024))) int8_t smem[C::LDS_BYTES];
    __shared__ int64_t s_next;  // dynamic order: next position of this XCD's sequence
    const int64_t tiles = (int64_t)p.tiles_m * p.tiles_n;
    const bool dyn = (gridDim.x & 7) == 0 && p.queue != nullptr;
    const int xcd = blockIdx.x & 7;
    const int64_t t0 = dyn ? tiles * xcd / 8 : 0, ntile = dyn ? tiles * (xcd + 1) / 8 - t0 : tiles;
    const int64_t nwork = ntile * p.S;
    int64_t w;
    if (dyn) {
        if (threadIdx.x == 0) s_next = atomicAdd(&p.queue[xcd], 1u);
        __syncthreads();
        w = s_next;
        __syncthreads();  // every wave has read s_next before it is written again
    } else {
        w = blockIdx.x;
    }
    if (w >= nwork) return;  // whole workgroup leaves together
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / C::NWN, wn = wave % C::NWN;
    const int64_t planeA = p.Mp * p.Kp, planeB = p.Np * p.Kp;
    struct Item {
        int g, tm, tn;
        const int8_t* abase;
        const int8_t* bbase;
    };
    auto decode = [&](int64_t k) {
        Item t;
        const int gi = (int)(k / ntile);
        t.g = p.S + 1 - gi;  // S + 1 down to 2
        grouped_tile<C>(t0 + (k - (int64_t)gi * ntile), p.tiles_m, p.tiles_n, t.tm, t.tn);
        t.tm += p.tm0;
        t.abase = p.A + (int64_t)t.tm * C::BM * p.Kp;
        t.bbase = p.B + (int64_t)t.tn * C::BN * p.Kp;
        return t;
    };
    Item cur = decode(w);
    auto pairs_of = [&](const Item& t) {
        return [&t, planeA, planeB](int pr, const int8_t*& pa, const int8_t*& pb) {
            pa = t.abase + (int64_t)pr * planeA;             // s = pr + 1
            pb = t.bbase + (int64_t)(t.g - pr - 2) * planeB;  // t = g - s
        };
    };
    issue_prologue<C>(smem, cur.g - 1, p.KT, p.Kp, pairs_of(cur));
    for (;;) {
        unsigned ticket = 0;
        if (dyn && threadIdx.x == 0) ticket = atomicAdd(&p.queue[xcd], 1u);
        v16i acc[C::FM][C::FN];
        zero_acc<C>(acc);
        mainloop<C, false>(acc, smem, cur.g - 1, p.KT, p.Kp, pairs_of(cur), [](int) {});
        if (dyn) {
            // Every wave read the previous s_next before this item's mainloop barriers.
            if (threadIdx.x == 0) s_next = ticket;
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): the write lands before the barrier
        }
        block_sync_lds();  // every wave is done with the ring
        const int64_t w2 = dyn ? s_next : w + gridDim.x;
        Item nxt = cur;
        if (w2 < nwork) {
            nxt = decode(w2);
            issue_prologue<C>(smem, nxt.g - 1, p.KT, p.Kp, pairs_of(nxt));
        }
        store_i32_blocked<C>(acc, p.T + (int64_t)(cur.g - 2) * p.Mp * p.Np, p.Np,
                             (int64_t)cur.tm * C::BM + wm * C::WTM, (int64_t)cur.tn * C::BN + wn * C::WTN, lane);
        if (w2 >= nwork) break;
        static_assert(C::STAGES == 2, "the epilogue's stores are drained by the first stage wait");
        w = w2;
        cur = nxt;
    }
}

// One resident workgroup per CU for the persistent GEMMs (the 128 KiB LDS ring allows no more).
hipError_t persistent_grid(int64_t total, unsigned& grid) {
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    grid = (unsigned)std::min<int64_t>(total, std::max(cus, 1));
    return e;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------------------
using CfgS2 = IgemmCfgS2;
using CfgS1 = IgemmCfgS1;

hipError_t launch_igemm_s2(const Dims& d, const int8_t* A_sl, const int8_t* B_sl, S2Out mode, void* out,
                           const CrtSet* crt_dev, hipStream_t st, int tm0, int tms, unsigned* queue) {
    S2Args p;
    p.A = A_sl;
    p.B = B_sl;
    p.M = d.M;
    p.N = d.N;
    p.Mp = d.Mp;
    p.Np = d.Np;
    p.Kp = d.Kp;
    p.S = d.S;
    p.KT = (int)(d.Kp / CfgS2::BK);
    const int tiles_m = (int)((d.M + CfgS2::BM - 1) / CfgS2::BM);
    if (tm0 < 0 || tm0 > tiles_m) return hipErrorInvalidValue;
    p.tm0 = tm0;
    p.tiles_m = tms < 0 ? tiles_m - tm0 : std::min(tms, tiles_m - tm0);
    p.tiles_n = (int)((d.N + CfgS2::BN - 1) / CfgS2::BN);
    p.out = out;
    p.crt = crt_dev;
    const int64_t total = (int64_t)p.tiles_m * p.tiles_n * d.S;
    if (total == 0) return hipSuccess;
    unsigned grid = 0;
    if (hipError_t e = persistent_grid(total, grid)) return e;
    p.queue = queue && grid % 8 == 0 && total >= 8 ? queue : nullptr;  // counters zeroed by the caller
    if (mode == S2Out::Digits)
        k_igemm_s2<CfgS2, true><<<grid, CfgS2::NT, 0, st>>>(p);
    else
        k_igemm_s2<CfgS2, false><<<grid, CfgS2::NT, 0, st>>>(p);
    return hipGetLastError();
}

static S1Args make_s1(const Dims& d, int beta, const int8_t* A_sl, const int8_t* B_sl) {
    S1Args p;
    p.A = A_sl;
    p.B = B_sl;
    p.M = d.M;
    p.N = d.N;
    p.Mp = d.Mp;
    p.Np = d.Np;
    p.Kp = d.Kp;
    p.S = d.S;
    p.KT = (int)(d.Kp / CfgS1::BK);
    p.tiles_m = (int)((d.M + CfgS1::BM - 1) / CfgS1::BM);
    p.tiles_n = (int)((d.N + CfgS1::BN - 1) / CfgS1::BN);
    p.beta = beta;
    p.T = nullptr;
    p.queue = nullptr;
    p.tm0 = 0;
    return p;
}

hipError_t launch_igemm_s1_blocked(const Dims& d, const int8_t* A_sl, const int8_t* B_sl, int32_t* T, hipStream_t st,
                                   unsigned* queue, int tm0, int tms) {
    S1Args p = make_s1(d, 7, A_sl, B_sl);
    p.KT = (int)(d.Kp / CfgS2::BK);
    const int tiles_m = (int)((d.M + CfgS2::BM - 1) / CfgS2::BM);
    if (tm0 < 0 || tm0 > tiles_m) return hipErrorInvalidValue;
    p.tm0 = tm0;
    p.tiles_m = tms < 0 ? tiles_m - tm0 : std::min(tms, tiles_m - tm0);
    p.tiles_n = (int)((d.N + CfgS2::BN - 1) / CfgS2::BN);
    p.T = T;
    const int64_t total = (int64_t)p.tiles_m * p.tiles_n * d.S;
    if (total == 0) return hipSuccess;
    unsigned grid = 0;
    if (hipError_t e = persistent_grid(total, grid)) return e;
    p.queue = queue && grid % 8 == 0 ? queue : nullptr;  // counters zeroed by the caller
    k_igemm_s1g<CfgS2><<<grid, CfgS2::NT, 0, st>>>(p);
    return hipGetLastError();
}

hipError_t launch_igemm_s1_groups(const Dims& d, const int8_t* A_sl, const int8_t* B_sl, int32_t* T,
                                  hipStream_t st) {
    S1Args p = make_s1(d, 7, A_sl, B_sl);
    p.T = T;
    const int64_t grid = (int64_t)p.tiles_m * p.tiles_n;
    if (grid == 0) return hipSuccess;
    k_igemm_s1<CfgS1><<<(unsigned)grid, CfgS1::NT, 0, st>>>(p);
    return hipGetLastError();
}

}  // namespace oza
