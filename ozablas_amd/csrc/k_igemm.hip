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
// Work order of the persistent GEMMs (one workgroup per CU).  Work items are split into 8
// contiguous runs, one per XCD (workgroup b runs on XCD b % 8):
//   static order (queue == null): workgroup b takes positions b, b + gridDim.x, ... of the whole
//     list (through xcd_remap for Scheme II);
//   dynamic order (queue != null, gridDim.x a multiple of 8): the workgroups of XCD v take the
//     positions of run v in the order they ask for them (counter queue[v]), so that the items in
//     flight on one XCD stay neighbours and share their operand panels in L2 even when CUs drift
//     apart.
// A workgroup knows its next item while it works on the current one (the chain issues the next
// item's first stages during the current item's last iterations).  With the dynamic order, thread 0
// asks for the item after next when the current item's mainloop ends; the answer comes back during
// the epilogue and is passed to every wave through LDS (Claims below).
// ------------------------------------------------------------------------------------------
struct Claims {
    int64_t w, w1;  // positions of the current and the next item
};

// First two positions of a workgroup.  Dynamic order: the first position is the workgroup's index
// among those of its XCD (blockIdx.x / 8), so that the first round runs the first items of the
// XCD's run side by side -- neighbours, which share operand panels (and, for Scheme I, the pair
// sequence of one group) -- and every later position comes from the XCD's counter, past that first
// round (publish_claim adds the same offset).  Two claims from the counter up front would hand the
// first claimants every other item of the first two rounds and leave the later ones with nothing
// when the run is short.  Positions only grow, so when the first is past the end so is the second.
__device__ __forceinline__ Claims first_claims(unsigned* queue, int xcd, unsigned* s_pos) {
    Claims c;
    if (queue) {
        if (threadIdx.x == 0) s_pos[1] = (gridDim.x >> 3) + atomicAdd(&queue[xcd], 1u);
        __syncthreads();
        c.w = blockIdx.x >> 3;
        c.w1 = s_pos[1];
    } else {
        c.w = blockIdx.x;
        c.w1 = c.w + gridDim.x;
    }
    return c;
}

// Position of the item after next, from the ticket thread 0 took (dynamic order).  Every wave read
// s_pos[1] at the previous publication and has passed the current item's barriers since; the
// barrier here makes the new value visible.
__device__ __forceinline__ int64_t publish_claim(unsigned ticket, unsigned* s_pos) {
    if (threadIdx.x == 0) s_pos[1] = (gridDim.x >> 3) + ticket;
    wait_lgkmcnt0();
    block_sync_lds();
    return s_pos[1];
}

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
    const CrtSet* crt;     // device copy (m, rcp used)
    unsigned* queue;       // 8 work counters, zero at launch (dynamic order), or null (static order)
    int split, kts;        // K parts and K steps per part (SPLIT only)
};

// Work items (modulus, tile) in logical order l = 0..total-1: modulus-major; inside a modulus,
// tiles in groups of GM tile rows.  SPLIT: items (part j, modulus, tile), part-major, part j over
// K steps [j kts, min((j + 1) kts, KT)), its digits into plane set j of the output.
template <class C, bool DIGITS, bool SPLIT = false>
__global__ void __launch_bounds__(C::NT) k_igemm_s2(S2Args p) {
    __shared__ __attribute__((aligned(1024))) int8_t smem[C::LDS_BYTES];
    __shared__ unsigned s_pos[2];
    const int64_t tiles = (int64_t)p.tiles_m * p.tiles_n, total = tiles * p.S;
    const int64_t items = SPLIT ? total * p.split : total;
    const bool dyn = p.queue != nullptr;
    const int xcd = blockIdx.x & 7;
    const int64_t run = items >> 3, run0 = (int64_t)xcd * run;
    const int64_t nwork = dyn ? (xcd == 7 ? items - 7 * run : run) : items;  // positions to take
    Claims c = first_claims(p.queue, xcd, s_pos);
    if (c.w >= nwork) return;  // whole workgroup leaves together
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / C::NWN, wn = wave % C::NWN;
    struct Item {
        int s, tm, tn, j, k0, KT;
        const int8_t* a;
        const int8_t* b;
        uint32_t m;
        double rcp;
    };
    auto decode = [&](int64_t wi) {
        Item t;
        int64_t l = dyn ? run0 + wi : xcd_remap(wi, items);
        t.j = 0;
        t.k0 = 0;
        t.KT = p.KT;
        if (SPLIT) {
            t.j = (int)(l / total);
            l -= (int64_t)t.j * total;
            t.k0 = t.j * p.kts;
            t.KT = min(p.kts, p.KT - t.k0);
        }
        t.s = (int)(l / tiles);
        grouped_tile<C>(l - (int64_t)t.s * tiles, p.tiles_m, p.tiles_n, t.tm, t.tn);
        t.tm += p.tm0;
        // tile (tm, tn) starts at block-row tm*BM/32 of the fragment-tiled planes (BM*Kp bytes)
        t.a = p.A + (int64_t)t.s * p.Mp * p.Kp + (int64_t)t.tm * C::BM * p.Kp;
        t.b = p.B + (int64_t)t.s * p.Np * p.Kp + (int64_t)t.tn * C::BN * p.Kp;
        // scalar loads: counted by lgkmcnt, so the epilogue does not wait for the direct loads
        t.m = as_constant(p.crt)->m[t.s];
        t.rcp = as_constant(p.crt)->rcp[t.s];
        return t;
    };
    Item cur = decode(c.w);
    auto own_of = [&](const Item& t) {
        return [&t, &p](int i) { return StageSrc{t.a, t.b, p.Kp, (t.k0 + i) * C::BK}; };
    };
    chain_start<C>(smem, 0, cur.KT, own_of(cur));
    int base = 0;  // ring slot of the current item's stage 0
    for (;;) {
        const bool more = c.w1 < nwork;
        const Item nxt = more ? decode(c.w1) : cur;
        // The next item's stage j; past its stages, or with no next item, a refetch of this item's
        // last stage.
        auto tail = [&](int j) {
            const bool real = more && j < nxt.KT;
            return StageSrc{real ? nxt.a : cur.a, real ? nxt.b : cur.b, p.Kp,
                            (real ? nxt.k0 + j : cur.k0 + cur.KT - 1) * C::BK};
        };
        v16i acc[C::FM][C::FN];
        zero_acc<C>(acc);
        mainloop_chain<C>(acc, smem, base, cur.KT, own_of(cur), tail);
        // The item after next (dynamic order).  Only thread 0's ticket is used, so the other lanes
        // leave it unset: a value merged from both paths would be copied as soon as the atomic
        // returns, and the wait for it would hold the epilogue behind the next item's loads.
        unsigned ticket;
        if (dyn && more && threadIdx.x == 0) ticket = atomicAdd(&p.queue[xcd], 1u);
        const int s = cur.s, tm = cur.tm, tn = cur.tn;
        if (DIGITS) {
            uint8_t* plane = reinterpret_cast<uint8_t*>(p.out) + ((int64_t)cur.j * p.S + s) * p.Mp * p.Np;
            store_digits<C, false>(acc, plane, p.Np, (int64_t)tm * C::BM + wm * C::WTM, (int64_t)tn * C::BN + wn * C::WTN,
                                   p.M, p.N, cur.m, cur.rcp, lane);
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
        if (!more) break;
        // the epilogue's stores are drained by the next chain start's vmcnt(0)
        const int64_t w2 = dyn ? publish_claim(ticket, s_pos) : c.w1 + gridDim.x;
        base = (base + cur.KT) & 1;
        c.w = c.w1;
        c.w1 = w2;
        cur = nxt;
    }
    wait_vmcnt<0>();  // no direct load in flight into LDS when the workgroup leaves
}

// Split-K digit reduction of a single call: 4 KiB of the S digit planes per workgroup (Mp * Np is a
// multiple of 65536, so a workgroup's bytes lie in one plane), partial sets after set 0.
__global__ void __launch_bounds__(256) k_digits_reduce(uint8_t* dg, int64_t plane, int S, int split,
                                                       const CrtSet* crt) {
    const int64_t off = (int64_t)blockIdx.x * 4096 + threadIdx.x * 16;
    const int s = (int)(off / plane);
    const int64_t stride = (int64_t)S * plane;
    reduce_digits16(dg, dg + stride, stride, split - 1, crt->m[s], off);
}

// ------------------------------------------------------------------------------------------
// Scheme I
// ------------------------------------------------------------------------------------------
struct S1Args {
    const int8_t* A;
    const int8_t* B;
    int64_t M, N, Mp, Np, Kp;
    int S, KT, tiles_m, tiles_n, beta;
    uint64_t kt_div;         // ceil(2^32 / KT): i / KT = (i * kt_div) >> 32 for the stage indices here
    int tm0;                 // blocked form: the launch covers tile rows [tm0, tm0 + tiles_m)
    int32_t* T;              // group products: int32 [S][M][N], or blocked [S][Mp][Np]
    unsigned* queue;         // 8 work counters, zero at launch (dynamic order), or null (static order)
};

// Stage i of the item of group g (pairs s = 1..g-1, t = g - s, in the reference's ascending-s
// order, KT stages each) of the tile whose operand tiles start at abase / bbase.
// i / KT by a multiply with ceil(2^32 / KT) (kt_div): exact while i * KT < 2^32 (the error
// i * (ceil(2^32/KT) KT - 2^32) / (KT 2^32) stays below 1/KT); here i < 21 KT and KT <= 2148
// (K <= 137438 in 64-byte steps), so i * KT < 10^8.  A division by a runtime KT would put a long
// scalar sequence at the head of every iteration.
__device__ __forceinline__ StageSrc s1_stage(const int8_t* abase, const int8_t* bbase, int64_t planeA,
                                             int64_t planeB, int g, int KT, uint64_t kt_div, int64_t Kp, int BK, int i) {
    const int pr = (int)(((uint64_t)(uint32_t)i * kt_div) >> 32), kt = i - pr * KT;
    return StageSrc{abase + (int64_t)pr * planeA,             // s = pr + 1
                    bbase + (int64_t)(g - pr - 2) * planeB,   // t = g - s
                    Kp, kt * BK};
}

// Scheme I group products in the reference's planar layout (the stage entry point): one
// workgroup per tile; its items are the groups g = s + t = 2..S+1 in ascending order
// (src/hip/pipeline.hip:110-135), chained through the ring; each finished T_g is stored to
// int32 [g-2][M][N].
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
    auto own_of = [&](int g) {
        return [=](int i) { return s1_stage(abase, bbase, planeA, planeB, g, p.KT, p.kt_div, p.Kp, C::BK, i); };
    };
    chain_start<C>(smem, 0, p.KT, own_of(2));
    int base = 0;
    for (int g = 2; g <= p.S + 1; ++g) {
        const int nit = (g - 1) * p.KT;
        const bool more = g < p.S + 1;
        auto tail = [&](int j) {  // the next group's stage j, or a refetch of this group's last stage
            return more ? s1_stage(abase, bbase, planeA, planeB, g + 1, p.KT, p.kt_div, p.Kp, C::BK, j)
                        : s1_stage(abase, bbase, planeA, planeB, g, p.KT, p.kt_div, p.Kp, C::BK, nit - 1);
        };
        v16i acc[C::FM][C::FN];
        zero_acc<C>(acc);
        mainloop_chain<C>(acc, smem, base, nit, own_of(g), tail);
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
        base = (base + nit) & 1;
    }
    wait_vmcnt<0>();  // no direct load in flight into LDS when the workgroup leaves
}

// Scheme I group products with the 256 x 256 tile of the Scheme II GEMM: work item = (group g,
// tile); the pairs s = 1..g-1 (t = g - s) of the group accumulate in INT32 through one pass of
// the chained mainloop, and T_g is stored as a blocked INT32 plane (store_i32_blocked).  The
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
    __shared__ unsigned s_pos[2];
    const int64_t tiles = (int64_t)p.tiles_m * p.tiles_n;
    const bool dyn = (gridDim.x & 7) == 0 && p.queue != nullptr;
    const int xcd = blockIdx.x & 7;
    const int64_t t0 = dyn ? tiles * xcd / 8 : 0, ntile = dyn ? tiles * (xcd + 1) / 8 - t0 : tiles;
    const int64_t nwork = ntile * p.S;
    Claims c = first_claims(dyn ? p.queue : nullptr, xcd, s_pos);
    if (c.w >= nwork) return;  // whole workgroup leaves together
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / C::NWN, wn = wave % C::NWN;
    const int64_t planeA = p.Mp * p.Kp, planeB = p.Np * p.Kp;
    struct Item {
        int g, tm, tn, nit;
        const int8_t* abase;
        const int8_t* bbase;
    };
    auto decode = [&](int64_t k) {
        Item t;
        const int gi = (int)(k / ntile);
        t.g = p.S + 1 - gi;  // S + 1 down to 2
        t.nit = (t.g - 1) * p.KT;
        grouped_tile<C>(t0 + (k - (int64_t)gi * ntile), p.tiles_m, p.tiles_n, t.tm, t.tn);
        t.tm += p.tm0;
        t.abase = p.A + (int64_t)t.tm * C::BM * p.Kp;
        t.bbase = p.B + (int64_t)t.tn * C::BN * p.Kp;
        return t;
    };
    Item cur = decode(c.w);
    auto own_of = [&](const Item& t) {
        return [&t, &p, planeA, planeB](int i) {
            return s1_stage(t.abase, t.bbase, planeA, planeB, t.g, p.KT, p.kt_div, p.Kp, C::BK, i);
        };
    };
    chain_start<C>(smem, 0, cur.nit, own_of(cur));
    int base = 0;
    for (;;) {
        const bool more = c.w1 < nwork;
        const Item nxt = more ? decode(c.w1) : cur;
        auto tail = [&](int j) {  // the next item's stage j, or a refetch of this item's last stage
            const bool real = more && j < nxt.nit;
            return real ? s1_stage(nxt.abase, nxt.bbase, planeA, planeB, nxt.g, p.KT, p.kt_div, p.Kp, C::BK, j)
                        : s1_stage(cur.abase, cur.bbase, planeA, planeB, cur.g, p.KT, p.kt_div, p.Kp, C::BK, cur.nit - 1);
        };
        v16i acc[C::FM][C::FN];
        zero_acc<C>(acc);
        mainloop_chain<C>(acc, smem, base, cur.nit, own_of(cur), tail);
        unsigned ticket;  // the item after next (dynamic order); thread 0's only, as in k_igemm_s2
        if (dyn && more && threadIdx.x == 0) ticket = atomicAdd(&p.queue[xcd], 1u);
        store_i32_blocked<C>(acc, p.T + (int64_t)(cur.g - 2) * p.Mp * p.Np, p.Np,
                             (int64_t)cur.tm * C::BM + wm * C::WTM, (int64_t)cur.tn * C::BN + wn * C::WTN, lane);
        if (!more) break;
        const int64_t w2 = dyn ? publish_claim(ticket, s_pos) : c.w1 + gridDim.x;
        base = (base + cur.nit) & 1;
        c.w = c.w1;
        c.w1 = w2;
        cur = nxt;
    }
    wait_vmcnt<0>();  // no direct load in flight into LDS when the workgroup leaves
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
                           const CrtSet* crt_dev, hipStream_t st, int tm0, int tms, unsigned* queue, int split) {
    if (split < 1 || (split > 1 && mode != S2Out::Digits)) return hipErrorInvalidValue;
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
    p.kts = (p.KT + split - 1) / split;
    p.split = (p.KT + p.kts - 1) / p.kts;  // non-empty parts only
    const int64_t total = (int64_t)p.tiles_m * p.tiles_n * d.S * p.split;
    if (total == 0) return hipSuccess;
    unsigned grid = 0;
    if (hipError_t e = persistent_grid(total, grid)) return e;
    // A workgroup of the dynamic order claims two positions before it starts, so with fewer than two
    // items per workgroup the first claimants would take them all and the others none: below that,
    // the static order (every workgroup its own positions) is the balanced one.
    p.queue = queue && grid % 8 == 0 && total >= 2 * (int64_t)grid ? queue : nullptr;  // counters zeroed by the caller
    if (mode == S2Out::Digits && p.split > 1)
        k_igemm_s2<CfgS2, true, true><<<grid, CfgS2::NT, 0, st>>>(p);
    else if (mode == S2Out::Digits)
        k_igemm_s2<CfgS2, true><<<grid, CfgS2::NT, 0, st>>>(p);
    else
        k_igemm_s2<CfgS2, false><<<grid, CfgS2::NT, 0, st>>>(p);
    return hipGetLastError();
}

hipError_t launch_digits_reduce(const Dims& d, uint8_t* digits, int split, const CrtSet* crt_dev, hipStream_t st) {
    const int KT = (int)(d.Kp / CfgS2::BK);
    const int kts = (KT + split - 1) / split;
    split = (KT + kts - 1) / kts;  // the parts launch_igemm_s2 ran
    if (split < 2 || d.M == 0 || d.N == 0) return hipSuccess;
    const int64_t plane = d.Mp * d.Np;
    k_digits_reduce<<<(unsigned)(d.S * plane / 4096), 256, 0, st>>>(digits, plane, d.S, split, crt_dev);
    return hipGetLastError();
}

// ceil(2^32 / KT), the multiplier of s1_stage's division (KT >= 1).
static uint64_t kt_divisor(int KT) { return ((1ull << 32) + (uint64_t)KT - 1) / (uint64_t)KT; }

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
    p.kt_div = kt_divisor(p.KT);
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
    p.kt_div = kt_divisor(p.KT);
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
    // two positions claimed up front: dynamic order only with at least two items per workgroup
    p.queue = queue && grid % 8 == 0 && total >= 2 * (int64_t)grid ? queue : nullptr;  // counters zeroed by the caller
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
