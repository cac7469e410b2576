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

// Zeroes the keys of both operands and nc further words (the GEMM's work counters, or none).
__global__ void __launch_bounds__(kStatThreads)
    k_keys_zero(uint32_t* a, int64_t na, uint32_t* b, int64_t nb, uint32_t* c, int64_t nc) {
    int64_t i = (int64_t)blockIdx.x * kStatThreads + threadIdx.x;
    if (i < na) a[i] = 0;
    else if (i - na < nb) b[i - na] = 0;
    else if (i - na - nb < nc) c[i - na - nb] = 0;
}

// Blocks [0, a.nblk) reduce operand a, the rest operand b.
__global__ void __launch_bounds__(kStatThreads) k_stats(StatOp a, StatOp b, int64_t K) {
    if ((int64_t)blockIdx.x < a.nblk) stats_op_block(a, K, blockIdx.x);
    else stats_op_block(b, K, (int64_t)blockIdx.x - a.nblk);
}

__global__ void __launch_bounds__(kStatThreads)
    k_keys_finalize(uint32_t* a, int64_t na, uint32_t* b, int64_t nb, int32_t offset) {
    int64_t i = (int64_t)blockIdx.x * kStatThreads + threadIdx.x;
    if (i < na) key_to_shift(a + i, offset);
    else if (i - na < nb) key_to_shift(b + (i - na), offset);
}

}  // namespace

hipError_t launch_stats(const Dims& d, const Operand& A, const Operand& B, int32_t offset, int32_t* shift_A,
                        int32_t* shift_B, hipStream_t st, uint32_t* zero, int64_t nzero) {
    uint32_t* ka = reinterpret_cast<uint32_t*>(shift_A);
    uint32_t* kb = reinterpret_cast<uint32_t*>(shift_B);
    const int64_t tot = d.M + d.N;
    const unsigned gz = (unsigned)((tot + (zero ? nzero : 0) + kStatThreads - 1) / kStatThreads);
    k_keys_zero<<<gz, kStatThreads, 0, st>>>(ka, d.M, kb, d.N, zero, zero ? nzero : 0);
    // Both operands go through one launch; each is read along its contiguous direction.
    const StatOp a = make_stat_op(A.p, d.M, d.K, A.ld, A.kcontig, ka);
    const StatOp b = make_stat_op(B.p, d.N, d.K, B.ld, B.kcontig, kb);
    k_stats<<<(unsigned)(a.nblk + b.nblk), kStatThreads, 0, st>>>(a, b, d.K);
    k_keys_finalize<<<(unsigned)((tot + kStatThreads - 1) / kStatThreads), kStatThreads, 0, st>>>(ka, d.M, kb, d.N,
                                                                                                  offset);
    return hipGetLastError();
}

}  // namespace oza
