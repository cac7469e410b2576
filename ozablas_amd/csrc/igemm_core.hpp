// igemm_core.hpp — the INT8 MFMA GEMM mainloop shared by the single-problem and grouped kernels.
//
// A workgroup owns a BM x BN output tile; each wave owns a (BM/NWM) x (BN/NWN) sub-tile made of
// 32x32 v_mfma_i32_32x32x32_i8 fragments with exact INT32 accumulation.  K advances in BK-byte
// stages through an LDS ring filled by direct global->LDS loads.
#pragma once

#include "oza_device.hpp"
#include "oza_internal.hpp"
#include "k_igemm_cfg.hpp"

namespace oza {
namespace core {


// ------------------------------------------------------------------------------------------
// LDS ring filled by direct global->LDS loads (global_load_lds_dwordx4).  The operand planes are
// fragment-tiled (frag_off16): every 1 KiB block is one 32x32 MFMA operand fragment in lane order,
// so a wave-wide direct load copies one block contiguously from HBM/L2 into LDS and a wave reads a
// fragment back with one contiguous, bank-conflict-free ds_read_b128.  In LDS a stage holds A blocks
// [kb][BM/32] followed by B blocks [kb][BN/32]; wave w always loads blocks w, w + NWAVES, ... of a
// stage, so two waves never write the same LDS bytes.
// ------------------------------------------------------------------------------------------
#define OZA_LDS_PTR(p) ((__attribute__((address_space(3))) void*)(p))

// Waits until at most N vector-memory operations of this wave are outstanding (gfx9 s_waitcnt
// encoding: vmcnt bits [3:0] and [15:14]; expcnt and lgkmcnt left at their no-wait maxima).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Waits for this wave's LDS reads and writes (lgkmcnt(0); vmcnt and expcnt left at their maxima).
__device__ __forceinline__ void wait_lgkmcnt0() { __builtin_amdgcn_s_waitcnt(0xC07F); }

__device__ __forceinline__ void block_sync_lds() {
    // s_barrier without the release/acquire fence of __syncthreads(): the fence would wait for
    // every outstanding load, including the prefetch of later stages.  The empty asm statements
    // keep the compiler from moving LDS accesses across the barrier.
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

template <class C>
__device__ __forceinline__ void issue_stage(int8_t* stage, const int8_t* __restrict__ a, const int8_t* __restrict__ b,
                                            int64_t ldk, int k0, int wave, int lane) {
    const int64_t kblocks = ldk >> 5;  // fragment blocks per block-row of a plane
#pragma unroll
    for (int i = 0; i < C::ABLK / C::NWAVES; ++i) {
        const int blk = i * C::NWAVES + wave;  // = kb * (BM/32) + rb
        const int kb = blk / (C::BM / 32), rb = blk % (C::BM / 32);
        __builtin_amdgcn_global_load_lds(a + ((rb * kblocks + (k0 >> 5) + kb) << 10) + lane * 16,
                                         OZA_LDS_PTR(stage + blk * 1024), 16, 0, 0);
    }
    int8_t* sb = stage + C::BM * C::BK;
#pragma unroll
    for (int i = 0; i < C::BBLK / C::NWAVES; ++i) {
        const int blk = i * C::NWAVES + wave;
        const int kb = blk / (C::BN / 32), rb = blk % (C::BN / 32);
        __builtin_amdgcn_global_load_lds(b + ((rb * kblocks + (k0 >> 5) + kb) << 10) + lane * 16,
                                         OZA_LDS_PTR(sb + blk * 1024), 16, 0, 0);
    }
}

// One K stage of a work item: its operand tiles (fragment-tiled planes of ldk K-bytes per row)
// and the stage's first K byte.
struct StageSrc {
    const int8_t* a;
    const int8_t* b;
    int64_t ldk;
    int k0;
};

template <class C>
struct Frags {
    v4i a[C::FM];
    v4i b[C::FN];
};

template <class C>
__device__ __forceinline__ void load_frags(Frags<C>& f, const int8_t* stage, int kk, int wm, int wn, int lane) {
    const int8_t* sb = stage + C::BM * C::BK;
#pragma unroll
    for (int fm = 0; fm < C::FM; ++fm)
        f.a[fm] = *reinterpret_cast<const v4i*>(stage + ((kk * (C::BM / 32) + wm * C::FM + fm) << 10) + lane * 16);
#pragma unroll
    for (int fn = 0; fn < C::FN; ++fn)
        f.b[fn] = *reinterpret_cast<const v4i*>(sb + ((kk * (C::BN / 32) + wn * C::FN + fn) << 10) + lane * 16);
}

template <class C>
__device__ __forceinline__ void mma_frags(v16i (&acc)[C::FM][C::FN], const Frags<C>& f) {
#pragma unroll
    for (int fm = 0; fm < C::FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < C::FN; ++fn)
            acc[fm][fn] = __builtin_amdgcn_mfma_i32_32x32x32_i8(f.a[fm], f.b[fn], acc[fm][fn], 0, 0, 0);
}

// Scheduling hint for one k-step: interleave the next step's LDS fragment reads between the
// current step's MFMAs so that LDS latency hides behind matrix-core work.
template <class C>
__device__ __forceinline__ void interleave_hint() {
    constexpr int NMFMA = C::FM * C::FN;
    constexpr int NREAD = C::FM + C::FN;
#pragma unroll
    for (int i = 0; i < NMFMA; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                 // 1 MFMA
        if (i < NREAD) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // 1 DS read
    }
}

// Last k-step of an iteration: also spread the direct loads of a later stage between the MFMAs.
template <class C>
__device__ __forceinline__ void interleave_hint_last() {
    constexpr int NMFMA = C::FM * C::FN;
    constexpr int NREAD = C::FM + C::FN;
    constexpr int NLOAD = C::LPS;
#pragma unroll
    for (int i = 0; i < NMFMA; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);                 // 1 MFMA
        if (i < NREAD) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // 1 DS read
        if (i < NLOAD) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);  // 1 VMEM (direct load)
    }
}

// ------------------------------------------------------------------------------------------
// Mainloop of one work item in a chain of items that share one 2-slot LDS ring.  A workgroup's
// items (tiles of one or several problems, moduli, or Scheme I groups) run back to back: item
// stage i sits in slot (base + i) & 1, and the next item's stage j is the chain's stage nit + j,
// so the next item's first two stages are issued during this item's last two iterations and land
// while its last MFMAs and its epilogue run.
//
// Ring protocol (tests/test_ring_model.py checks these rules on a transcription of the loops):
//  R1 a wave reads a stage only after every wave waited vmcnt(0) for its loads of that stage and
//     then passed a workgroup barrier;
//  R2 a wave loads into a slot only after every wave waited lgkmcnt(0) for its reads of that slot
//     and then passed a barrier;
//  R3 a wave loads into a slot only after a vmcnt(0) that follows its previous load into the same
//     slot: the ring never depends on two direct loads into one LDS address landing in issue order;
//  R4 the caller waits vmcnt(0) after its last item, so no load is in flight into LDS that a later
//     workgroup may be given.
//
// Per iteration `it` (k-steps 0..KSTEPS-1 of stage it):
//   k-steps 0..KSTEPS-2: MFMAs on the fragments in registers while the next k-step's fragments are
//     read from LDS;
//   then (it < nit-1) vmcnt(0) -- stage it+1, issued one iteration ago --, lgkmcnt(0), barrier; the
//     first fragments of stage it+1 are read and stage it+2 is issued into slot it, which every
//     wave has finished reading at that barrier, both interleaved with the last k-step's MFMAs;
//   the last iteration waits for no load: lgkmcnt(0), barrier, and chain stage nit+1 is issued into
//     its slot before the last k-step's MFMAs.
//   Wave priority: the section after each stage barrier (the direct loads and the last k-step's
//     MFMAs) runs at priority 1, everything else at 0.  Same-box A/B: Scheme II step 3 -4.5 % and
//     Scheme I -3 % against no priorities; priority 1 over all MFMA sections gave -2.4 % (Scheme
//     II), over the first k-step only +4 %.
//
//   On entry this item's stage 0 has been issued, and its stage 1 too when nit >= 2 (chain_start
//   for a chain's first item; the previous item's last two iterations otherwise).  With nit == 1 no
//   earlier iteration can take chain stage 1, so it is issued after the first barrier.
//   own(i) gives this item's stage i, tail(j) the next item's stage j -- or, when there is no such
//   stage, any valid source (a refetch: nothing reads it, and the next chain start's vmcnt(0)
//   completes it before its slot is loaded again).  Both are evaluated and one is selected, so no
//   branch separates the loads from the MFMAs they are interleaved with.
// ------------------------------------------------------------------------------------------
template <class C, class Own>
__device__ __forceinline__ void chain_start(int8_t* smem, int base, int nit, Own own) {
    static_assert(C::STAGES == 2, "the chain keeps two slots");
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const StageSrc s0 = own(0);
    issue_stage<C>(smem + (base & 1) * C::STAGE_BYTES, s0.a, s0.b, s0.ldk, s0.k0, wave, lane);
    if (nit >= 2) {
        const StageSrc s1 = own(1);
        issue_stage<C>(smem + ((base + 1) & 1) * C::STAGE_BYTES, s1.a, s1.b, s1.ldk, s1.k0, wave, lane);
    }
}

template <class C, class Own, class Tail>
__device__ __forceinline__ void mainloop_chain(v16i (&acc)[C::FM][C::FN], int8_t* smem, int base, int nit, Own own,
                                               Tail tail) {
    static_assert(C::STAGES == 2, "the chain keeps two slots");
    constexpr int KSTEPS = C::BK / 32;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave / C::NWN, wn = wave % C::NWN;
    auto slot = [&](int i) { return smem + ((base + i) & 1) * C::STAGE_BYTES; };
    auto issue = [&](int i) {  // chain stage i: this item's stage i, or the next item's stage i - nit
        const StageSrc o = own(i), t = tail(i - nit);
        const bool mine = i < nit;  // selects, not a branch: the loads stay among the MFMAs
        issue_stage<C>(slot(i), mine ? o.a : t.a, mine ? o.b : t.b, mine ? o.ldk : t.ldk, mine ? o.k0 : t.k0, wave,
                       lane);
    };
    auto lds_barrier = [] {
        wait_lgkmcnt0();  // this wave's LDS reads are done
        block_sync_lds();
    };
    Frags<C> f[2];
    wait_vmcnt<0>();
    lds_barrier();
    if (nit == 1) issue(1);
    load_frags<C>(f[0], slot(0), 0, wm, wn, lane);
    for (int it = 0; it + 1 < nit; ++it) {
        const int8_t* cur = slot(it);
#pragma unroll
        for (int kk = 0; kk + 1 < KSTEPS; ++kk) {
            load_frags<C>(f[(kk + 1) & 1], cur, kk + 1, wm, wn, lane);
            mma_frags<C>(acc, f[kk & 1]);
            interleave_hint<C>();
        }
        wait_vmcnt<0>();  // stage it+1, issued one iteration ago
        lds_barrier();
        load_frags<C>(f[KSTEPS & 1], slot(it + 1), 0, wm, wn, lane);
        __builtin_amdgcn_s_setprio(1);
        issue(it + 2);
        mma_frags<C>(acc, f[(KSTEPS - 1) & 1]);
        interleave_hint_last<C>();
        __builtin_amdgcn_s_setprio(0);
        if (KSTEPS % 2 == 1) {  // keep the fragments of the next iteration in f[0]
#pragma unroll
            for (int i = 0; i < C::FM; ++i) f[0].a[i] = f[1].a[i];
#pragma unroll
            for (int i = 0; i < C::FN; ++i) f[0].b[i] = f[1].b[i];
        }
    }
    {  // last iteration: its slot is freed at the barrier for chain stage nit + 1
        const int8_t* cur = slot(nit - 1);
#pragma unroll
        for (int kk = 0; kk + 1 < KSTEPS; ++kk) {
            load_frags<C>(f[(kk + 1) & 1], cur, kk + 1, wm, wn, lane);
            mma_frags<C>(acc, f[kk & 1]);
            interleave_hint<C>();
        }
        lds_barrier();
        __builtin_amdgcn_s_setprio(1);
        issue(nit + 1);
        mma_frags<C>(acc, f[(KSTEPS - 1) & 1]);
        __builtin_amdgcn_s_setprio(0);
    }
}

template <class C>
__device__ __forceinline__ void zero_acc(v16i (&acc)[C::FM][C::FN]) {
#pragma unroll
    for (int fm = 0; fm < C::FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < C::FN; ++fn)
#pragma unroll
            for (int j = 0; j < 16; ++j) acc[fm][fn][j] = 0;
}

// Row of accumulator element j of lane l inside a 32x32 fragment.
__device__ __forceinline__ int frag_row(int j, int lane) { return 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3); }

// Work index -> (s, tm, tn).  Consecutive workgroups are dealt round-robin to the 8 XCDs; the
// remap gives each XCD a contiguous run of work items so that neighbouring tiles (which share
// A rows and B columns) meet in the same L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t id, int64_t total) {
    const int64_t per = total / 8;
    if (per == 0) return id;
    if (id >= per * 8) return id;
    return (id % 8) * per + id / 8;
}

// Tile index t in [0, tiles_m*tiles_n) -> (tm, tn), GM tile-rows at a time.
template <class C>
__device__ __forceinline__ void grouped_tile(int64_t t, int tiles_m, int tiles_n, int& tm, int& tn) {
    const int64_t per_group = (int64_t)C::GM * tiles_n;
    const int g = (int)(t / per_group);
    const int first = g * C::GM;
    const int gsz = min(tiles_m - first, C::GM);
    const int r = (int)(t - g * per_group);
    tm = first + r % gsz;
    tn = r / gsz;
}


// CRT-digit epilogue of one tile: c = C_s mod m_s per accumulator element, one 16-byte store per
// fragment and lane.  Digit planes are blocked in 32 x 32 element blocks: block
// (R, Q) of a plane with Np columns starts at byte (R * (Np/32) + Q) * 1024, and inside it lane
// l's 16 bytes j = 0..15 sit at l*16 + j, holding element (row 8*(j/4) + 4*(l/32) + j%4,
// column l%32) of the block, i.e. the MFMA 32x32 output layout.  The layout does not depend on
// the tile shape, so every GEMM configuration feeds the same reconstruction.  With SKIP_PAD,
// blocks that hold no element of the M x N product (tile padding) are neither computed nor stored
// (the reconstruction never reads them); the grouped GEMM, whose small problems pad much, uses it,
// while the single-problem GEMM keeps the branch-free epilogue (0.3 % faster at 8192^3).  NT
// makes the stores nontemporal (store_unit16).
template <class C, bool SKIP_PAD, bool NT = false>
__device__ __forceinline__ void store_digits(const v16i (&acc)[C::FM][C::FN], uint8_t* plane, int64_t Np,
                                             int64_t row0, int64_t col0, int64_t M, int64_t N, uint32_t mod,
                                             double rcp, int lane) {
    const double m = (double)mod;
    const int64_t qn = Np >> 5;
    // row0 and col0 are the same for the whole wave
    const int64_t fm_end = (M - (int64_t)__builtin_amdgcn_readfirstlane((int)(row0 >> 5)) * 32 + 31) >> 5;
    const int64_t fn_end = (N - (int64_t)__builtin_amdgcn_readfirstlane((int)(col0 >> 5)) * 32 + 31) >> 5;
#pragma unroll
    for (int fm = 0; fm < C::FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < C::FN; ++fn) {
            if (SKIP_PAD && (fm >= fm_end || fn >= fn_end)) continue;
            int d[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) d[j] = crt_digit(acc[fm][fn][j], m, rcp);
            const int64_t blk = ((row0 >> 5) + fm) * qn + (col0 >> 5) + fn;
            store_global16<NT>(plane + (blk << 10) + lane * 16, pack4u(d[0], d[1], d[2], d[3]),
                           pack4u(d[4], d[5], d[6], d[7]), pack4u(d[8], d[9], d[10], d[11]),
                           pack4u(d[12], d[13], d[14], d[15]));
        }
}

// Split-K digit reduction of 16 digits: dg[off..off+15] = (dg + sum_{j < parts} part_j) mod m, with
// part_j at dgp + j * stride + off.  Every digit is below m, so one conditional subtraction per
// part keeps the running sum reduced.  off is a multiple of 16.
__device__ __forceinline__ void reduce_digits16(uint8_t* __restrict__ dg, const uint8_t* __restrict__ dgp,
                                                int64_t stride, int parts, uint32_t m, int64_t off) {
    uint4* dst = reinterpret_cast<uint4*>(dg + off);
    const uint4 d0 = *dst;
    uint32_t w[4] = {d0.x, d0.y, d0.z, d0.w};
    uint32_t r[16];
#pragma unroll
    for (int b = 0; b < 16; ++b) r[b] = (w[b >> 2] >> (8 * (b & 3))) & 0xffu;
    for (int j = 0; j < parts; ++j) {
        const uint4 v = *reinterpret_cast<const uint4*>(dgp + (int64_t)j * stride + off);
        const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            const uint32_t x = r[b] + ((u[b >> 2] >> (8 * (b & 3))) & 0xffu);
            r[b] = x >= m ? x - m : x;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) w[q] = r[4 * q] | (r[4 * q + 1] << 8) | (r[4 * q + 2] << 16) | (r[4 * q + 3] << 24);
    *dst = make_uint4(w[0], w[1], w[2], w[3]);
}

// INT32 epilogue of one tile into a blocked plane: 32 x 32 element blocks of 4 KiB, block (R, Q)
// at element (R * (Np/32) + Q) * 1024, lane l's 16 values j at l*16 + j (the layout of
// store_digits with 4-byte elements).
template <class C>
__device__ __forceinline__ void store_i32_blocked(const v16i (&acc)[C::FM][C::FN], int32_t* plane, int64_t Np,
                                                  int64_t row0, int64_t col0, int lane) {
    const int64_t qn = Np >> 5;
#pragma unroll
    for (int fm = 0; fm < C::FM; ++fm)
#pragma unroll
        for (int fn = 0; fn < C::FN; ++fn) {
            const int64_t blk = ((row0 >> 5) + fm) * qn + (col0 >> 5) + fn;
            int4* o = reinterpret_cast<int4*>(plane + (blk << 10) + lane * 16);
            const v16i& a = acc[fm][fn];
            o[0] = make_int4(a[0], a[1], a[2], a[3]);
            o[1] = make_int4(a[4], a[5], a[6], a[7]);
            o[2] = make_int4(a[8], a[9], a[10], a[11]);
            o[3] = make_int4(a[12], a[13], a[14], a[15]);
        }
}

}  // namespace core
}  // namespace oza
