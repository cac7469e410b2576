// stats_core.hpp — step-1 block bodies (256 threads) shared by the single-problem and grouped
// kernels.  Partial maxima of the exponent key (exp_key) meet in the key arrays through 32-bit
// atomic max; the keys start at zero and are turned into shifts afterwards (exp_from_key).
#pragma once

#include "oza_device.hpp"
#include "oza_internal.hpp"

namespace oza {
namespace core {

constexpr int kStatThreads = 256;
constexpr int kRowChunk = 4096;  // A: columns per wave
constexpr int kColRows = 128;    // B: rows per block
constexpr int kStatUnroll = 8;   // loads in flight per thread

struct StatShape {
    int64_t chunksA;  // column chunks per row of A
    int64_t nblkA;    // blocks covering A (4 waves each)
    int64_t colblkB;  // column blocks of B
    int64_t nblkB;    // blocks covering B
};

inline StatShape stat_shape(int64_t M, int64_t N, int64_t K) {
    StatShape h;
    h.chunksA = (K + kRowChunk - 1) / kRowChunk;
    h.nblkA = (M * h.chunksA + 3) / 4;
    h.colblkB = (N + 2 * kStatThreads - 1) / (2 * kStatThreads);
    h.nblkB = h.colblkB * ((K + kColRows - 1) / kColRows);
    return h;
}

// Row maxima of a K-contiguous operand (element (r, k) at A[r * lda + k]): one wave per (row,
// 4096-column chunk); block b covers waves 4b .. 4b+3.  VEC needs lda even and A 16-byte aligned.
template <bool VEC>
__device__ __forceinline__ void stats_A_block(const double* __restrict__ A, int64_t M, int64_t K, int64_t lda,
                                              int64_t chunksA, uint32_t* __restrict__ keyA, int64_t blk) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t w = blk * (kStatThreads / 64) + wave;
    const int64_t row = w / chunksA;
    const int64_t chunk = w - row * chunksA;
    if (row >= M) return;  // whole wave leaves together
    const double* p = A + row * lda;
    const int64_t c0 = chunk * kRowChunk;
    const int64_t c1 = min(K, c0 + kRowChunk);
    // kStatUnroll loads per thread are issued before the first use.  Indices past the end are
    // clamped to the last element of the range, which is read again: harmless for a maximum, and
    // it keeps every load unconditional.
    uint32_t key = 0;
    if (VEC) {
        for (int64_t c = c0 + 2 * lane; c < c1; c += 128 * kStatUnroll) {
            double2 v[kStatUnroll];
#pragma unroll
            for (int u = 0; u < kStatUnroll; ++u)
                v[u] = load_stream2(p + min(c + 128 * u, c1 - 2));
#pragma unroll
            for (int u = 0; u < kStatUnroll; ++u) key = max(key, max(exp_key(v[u].x), exp_key(v[u].y)));
        }
    } else {
        for (int64_t c = c0 + lane; c < c1; c += 64 * kStatUnroll) {
            double v[kStatUnroll];
#pragma unroll
            for (int u = 0; u < kStatUnroll; ++u) v[u] = load_stream(p + min(c + 64 * u, c1 - 1));
#pragma unroll
            for (int u = 0; u < kStatUnroll; ++u) key = max(key, exp_key(v[u]));
        }
    }
    key = wave_max_u32(key);
    if (lane == 0) atomicMax(keyA + row, key);
}

// Column maxima of an operand stored K x N (element (k, n) at B[k * ldb + n]): a block covers 512
// columns x rowsB rows (kColRows unless the caller asks for shorter blocks); thread = two adjacent
// columns (one 16-byte load per row when VEC, which needs ldb even and B 16-byte aligned; two
// 8-byte loads otherwise).
template <bool VEC>
__device__ __forceinline__ void stats_B_block(const double* __restrict__ B, int64_t K, int64_t N, int64_t ldb,
                                              int64_t colblkB, uint32_t* __restrict__ keyB, int64_t b,
                                              int64_t rowsB = kColRows) {
    const int64_t cb = b % colblkB;
    const int64_t rb = b / colblkB;
    const int64_t r0 = rb * rowsB;
    const int64_t r1 = min(K, r0 + rowsB);
    const int64_t c = cb * (2 * kStatThreads) + 2 * threadIdx.x;
    if (c >= N) return;
    // Loads as in stats_A_block: rows past r1 are clamped to r1 - 1 (read again), and a missing
    // second column reads the first one again (its key is not stored).
    uint32_t k0 = 0, k1 = 0;
    if (VEC && c + 1 < N) {
        for (int64_t r = r0; r < r1; r += kStatUnroll) {
            double2 v[kStatUnroll];
#pragma unroll
            for (int u = 0; u < kStatUnroll; ++u)
                v[u] = load_stream2(B + min(r + u, r1 - 1) * ldb + c);
#pragma unroll
            for (int u = 0; u < kStatUnroll; ++u) {
                k0 = max(k0, exp_key(v[u].x));
                k1 = max(k1, exp_key(v[u].y));
            }
        }
    } else {
        const bool two = c + 1 < N;
        const int64_t c2 = two ? c + 1 : c;
        for (int64_t r = r0; r < r1; r += kStatUnroll) {
            double v0[kStatUnroll], v1[kStatUnroll];
#pragma unroll
            for (int u = 0; u < kStatUnroll; ++u) {
                const int64_t rr = min(r + u, r1 - 1);
                v0[u] = B[rr * ldb + c];
                v1[u] = B[rr * ldb + c2];
            }
#pragma unroll
            for (int u = 0; u < kStatUnroll; ++u) {
                k0 = max(k0, exp_key(v0[u]));
                k1 = max(k1, exp_key(v1[u]));
            }
        }
    }
    atomicMax(keyB + c, k0);
    if (c + 1 < N) atomicMax(keyB + c + 1, k1);
}

// One operand of step 1 as the kernels see it: `rows` logical rows (M for A, N for B) of K values,
// element (r, k) at p[r * ld + k] (kcontig) or p[k * ld + r].  keys[rows] receive the row maxima.
struct StatOp {
    const double* p;
    int64_t rows, ld;
    int32_t kcontig, vec;
    int64_t chunks, colblk, nblk;
    int64_t rowsB;  // rows per block of the column reader
    uint32_t* keys;
};

// Blocks the column reader aims for in a single-problem launch.  With kColRows-row blocks a small
// operand gives few blocks (1024 x 1024: 16), whose threads each wait for kColRows / kStatUnroll =
// 16 rounds of loads in turn (15 us for the launch); shorter blocks, down to two rounds, spread the
// same reads over more CUs at the cost of more atomic maxima per column.
constexpr int64_t kStatTargetBlocksB = 512;

inline StatOp make_stat_op(const double* p, int64_t rows, int64_t K, int64_t ld, bool kcontig, uint32_t* keys) {
    StatOp o;
    o.p = p;
    o.rows = rows;
    o.ld = ld;
    o.kcontig = kcontig;
    // 16-byte loads: pairs never straddle a row end (K even for K-contiguous rows; the column
    // reader checks its pair itself) and every pair is 16-byte aligned.
    o.vec = (ld % 2 == 0) && ((reinterpret_cast<uintptr_t>(p) & 15) == 0) && (!kcontig || K % 2 == 0);
    const StatShape h = stat_shape(rows, rows, K);
    o.chunks = h.chunksA;
    o.colblk = h.colblkB;
    o.rowsB = kColRows;
    while (!kcontig && o.rowsB > 2 * kStatUnroll && o.colblk * ((K + o.rowsB - 1) / o.rowsB) < kStatTargetBlocksB)
        o.rowsB >>= 1;
    o.nblk = kcontig ? h.nblkA : o.colblk * ((K + o.rowsB - 1) / o.rowsB);
    o.keys = keys;
    return o;
}

__device__ __forceinline__ void stats_op_block(const StatOp& o, int64_t K, int64_t blk) {
    if (o.kcontig) {
        if (o.vec) stats_A_block<true>(o.p, o.rows, K, o.ld, o.chunks, o.keys, blk);
        else stats_A_block<false>(o.p, o.rows, K, o.ld, o.chunks, o.keys, blk);
    } else {
        if (o.vec) stats_B_block<true>(o.p, K, o.rows, o.ld, o.colblk, o.keys, blk, o.rowsB);
        else stats_B_block<false>(o.p, K, o.rows, o.ld, o.colblk, o.keys, blk, o.rowsB);
    }
}

// Key -> shift (offset - E) in place.
__device__ __forceinline__ void key_to_shift(uint32_t* p, int32_t offset) {
    const int e = exp_from_key(*p);
    *reinterpret_cast<int32_t*>(p) = offset - e;
}

}  // namespace core
}  // namespace oza
