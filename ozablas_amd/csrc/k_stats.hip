// k_stats.hip — step 1: row exponents of A and column exponents of B in one launch.
//
// Reference semantics: src/pipeline/step1_statistics.hpp:23-99 (shift = offset - E, E the
// largest frexp exponent of the row/column, at least -1023).  The reference runs one block per
// row of A and one thread per column of B with a serial loop over all K rows; here both
// operands are read once with 16-byte coalesced loads by a 2-D grid, every wave reduces with
// DPP, and partial maxima meet in the output array through 32-bit atomic max on an exponent
// key (see exp_key).  A finalize launch converts keys to shifts in place.
#include "oza_device.hpp"
#include "oza_internal.hpp"

namespace oza {
namespace {

constexpr int kStatThreads = 256;
constexpr int kRowChunk = 4096;  // A: columns per wave
constexpr int kColRows = 128;    // B: rows per block

__global__ void __launch_bounds__(kStatThreads) k_keys_zero(uint32_t* a, int64_t na, uint32_t* b, int64_t nb) {
    int64_t i = (int64_t)blockIdx.x * kStatThreads + threadIdx.x;
    if (i < na) a[i] = 0;
    else if (i - na < nb) b[i - na] = 0;
}

template <bool VEC>
__global__ void __launch_bounds__(kStatThreads)
    k_stats(const double* __restrict__ A, const double* __restrict__ B, int64_t M, int64_t N, int64_t K,
            int64_t nblkA, int64_t chunksA, int64_t colblkB, uint32_t* __restrict__ keyA,
            uint32_t* __restrict__ keyB) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if ((int64_t)blockIdx.x < nblkA) {
        // ---- A: one wave per (row, column chunk)
        const int64_t w = (int64_t)blockIdx.x * (kStatThreads / 64) + wave;
        const int64_t row = w / chunksA;
        const int64_t chunk = w - row * chunksA;
        if (row >= M) return;  // whole wave leaves together
        const double* p = A + row * K;
        const int64_t c0 = chunk * kRowChunk;
        const int64_t c1 = min(K, c0 + kRowChunk);
        uint32_t key = 0;
        if (VEC) {
            for (int64_t c = c0 + 2 * lane; c < c1; c += 128) {
                const double2 v = *reinterpret_cast<const double2*>(p + c);
                key = max(key, max(exp_key(v.x), exp_key(v.y)));
            }
        } else {
            for (int64_t c = c0 + lane; c < c1; c += 64) key = max(key, exp_key(p[c]));
        }
        key = wave_max_u32(key);
        if (lane == 0) atomicMax(keyA + row, key);
    } else {
        // ---- B: a block covers (VEC ? 512 : 256) columns x kColRows rows
        const int64_t b = (int64_t)blockIdx.x - nblkA;
        const int64_t cb = b % colblkB;
        const int64_t rb = b / colblkB;
        const int64_t r0 = rb * kColRows;
        const int64_t r1 = min(K, r0 + kColRows);
        if (VEC) {
            const int64_t c = cb * (2 * kStatThreads) + 2 * threadIdx.x;
            if (c >= N) return;
            uint32_t k0 = 0, k1 = 0;
            for (int64_t r = r0; r < r1; ++r) {
                const double2 v = *reinterpret_cast<const double2*>(B + r * N + c);
                k0 = max(k0, exp_key(v.x));
                k1 = max(k1, exp_key(v.y));
            }
            atomicMax(keyB + c, k0);
            atomicMax(keyB + c + 1, k1);
        } else {
            const int64_t c = cb * kStatThreads + threadIdx.x;
            if (c >= N) return;
            uint32_t k0 = 0;
            for (int64_t r = r0; r < r1; ++r) k0 = max(k0, exp_key(B[r * N + c]));
            atomicMax(keyB + c, k0);
        }
    }
}

__global__ void __launch_bounds__(kStatThreads)
    k_keys_finalize(uint32_t* a, int64_t na, uint32_t* b, int64_t nb, int32_t offset) {
    int64_t i = (int64_t)blockIdx.x * kStatThreads + threadIdx.x;
    uint32_t* p = nullptr;
    if (i < na) p = a + i;
    else if (i - na < nb) p = b + (i - na);
    if (p) {
        const int e = exp_from_key(*p);
        *reinterpret_cast<int32_t*>(p) = offset - e;
    }
}

}  // namespace

hipError_t launch_stats(const Dims& d, const double* A, const double* B, int32_t offset, int32_t* shift_A,
                        int32_t* shift_B, hipStream_t st) {
    uint32_t* ka = reinterpret_cast<uint32_t*>(shift_A);
    uint32_t* kb = reinterpret_cast<uint32_t*>(shift_B);
    const int64_t tot = d.M + d.N;
    const unsigned gz = (unsigned)((tot + kStatThreads - 1) / kStatThreads);
    k_keys_zero<<<gz, kStatThreads, 0, st>>>(ka, d.M, kb, d.N);

    // Both operands go through one launch, vectorized when both allow 16-byte loads.
    const bool vec = (d.K % 2 == 0) && (d.N % 2 == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0) &&
                     ((reinterpret_cast<uintptr_t>(B) & 15) == 0);
    const int64_t chunksA = (d.K + kRowChunk - 1) / kRowChunk;
    const int64_t nblkA = (d.M * chunksA + 3) / 4;
    const int64_t colw = vec ? 2 * kStatThreads : kStatThreads;
    const int64_t colblkB = (d.N + colw - 1) / colw;
    const int64_t nblkB = colblkB * ((d.K + kColRows - 1) / kColRows);
    const unsigned grid = (unsigned)(nblkA + nblkB);
    if (vec)
        k_stats<true><<<grid, kStatThreads, 0, st>>>(A, B, d.M, d.N, d.K, nblkA, chunksA, colblkB, ka, kb);
    else
        k_stats<false><<<grid, kStatThreads, 0, st>>>(A, B, d.M, d.N, d.K, nblkA, chunksA, colblkB, ka, kb);
    k_keys_finalize<<<gz, kStatThreads, 0, st>>>(ka, d.M, kb, d.N, offset);
    return hipGetLastError();
}

}  // namespace oza
