// k_stats.hip — step 1: row exponents of A and column exponents of B in one launch.
//
// Reference semantics: src/pipeline/step1_statistics.hpp:23-99 (shift = offset - E, E the
// largest frexp exponent of the row/column, at least -1023).  The reference runs one block per
// row of A and one thread per column of B with a serial loop over all K rows; here both
// operands are read once with 16-byte coalesced loads by a 2-D grid, every wave reduces with
// DPP, and partial maxima meet in the output array through 32-bit atomic max on an exponent
// key (see exp_key).  A finalize launch converts keys to shifts in place.
#include "stats_core.hpp"

namespace oza {
namespace {

using namespace core;

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
    if ((int64_t)blockIdx.x < nblkA)
        stats_A_block<VEC>(A, M, K, chunksA, keyA, blockIdx.x);
    else
        stats_B_block<VEC>(B, K, N, colblkB, keyB, (int64_t)blockIdx.x - nblkA);
}

__global__ void __launch_bounds__(kStatThreads)
    k_keys_finalize(uint32_t* a, int64_t na, uint32_t* b, int64_t nb, int32_t offset) {
    int64_t i = (int64_t)blockIdx.x * kStatThreads + threadIdx.x;
    if (i < na) key_to_shift(a + i, offset);
    else if (i - na < nb) key_to_shift(b + (i - na), offset);
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
    const StatShape h = stat_shape(d.M, d.N, d.K, A, B);
    const unsigned grid = (unsigned)(h.nblkA + h.nblkB);
    if (h.vec)
        k_stats<true><<<grid, kStatThreads, 0, st>>>(A, B, d.M, d.N, d.K, h.nblkA, h.chunksA, h.colblkB, ka, kb);
    else
        k_stats<false><<<grid, kStatThreads, 0, st>>>(A, B, d.M, d.N, d.K, h.nblkA, h.chunksA, h.colblkB, ka, kb);
    k_keys_finalize<<<gz, kStatThreads, 0, st>>>(ka, d.M, kb, d.N, offset);
    return hipGetLastError();
}

}  // namespace oza
