// k_slice.hip — step 2: split FP64 operands into INT8 planes laid out for the MFMA GEMM.
//
// Scheme II (reference: src/pipeline/step2_slicing.hpp:29-38, 110-206): x = ldexp(a, shift),
// t = rint(x) (round-half-even, which is what both the reference's llrint fast path and its
// rint/fmod heavy path round to), residue r_s = symmetric residue of t mod m_s.  Any exact
// method yields the same bytes; sym_residue uses two FMAs per modulus.
//
// Scheme I (reference: step2_slicing.hpp:20-25, 44-100): the reference's exact sequence of
// FP64 operations, sigma = 0.75 * 2^(E+1-beta+53-beta*s), a_s = (a + sigma) - sigma,
// slice = int(ldexp(a_s * 2^-(E+1-beta), beta*s)), a -= a_s.  Contraction is off in this file
// so every operation rounds where the reference's does.
//
// Output layout: A planes (Mp x Kp) and B planes (Np x Kp, B transposed so that both MFMA
// operands are K-contiguous) in the fragment-tiled order of frag_off16.  Padding is zero.  Each
// thread produces 16 consecutive K values of one row/column, i.e. one 16-byte unit per plane, and
// a wave's 64 units form contiguous runs of the tiled layout.
#pragma clang fp contract(off)

#include "slice_core.hpp"

namespace oza {
namespace {

using namespace core;

// Blocks [0, a.nblk) slice operand a into its planes, the rest operand b.
template <bool WIDE>
__global__ void __launch_bounds__(kSliceThreads)
    k_slice_s2(SliceOp a, SliceOp b, int64_t K, int64_t Kp, int S, ResidueConsts rc) {
    if ((int64_t)blockIdx.x < a.nblk) slice_s2_op_block<WIDE>(a, K, Kp, S, rc, blockIdx.x);
    else slice_s2_op_block<WIDE>(b, K, Kp, S, rc, (int64_t)blockIdx.x - a.nblk);
}

__global__ void __launch_bounds__(kSliceThreads)
    k_slice_s1(SliceOp a, SliceOp b, int64_t K, int64_t Kp, int S, int beta) {
    if ((int64_t)blockIdx.x < a.nblk) slice_s1_op_block(a, K, Kp, S, beta, blockIdx.x);
    else slice_s1_op_block(b, K, Kp, S, beta, (int64_t)blockIdx.x - a.nblk);
}

// ---------------------------------------------------------------- unpack to planar layout
__global__ void k_unpack_A(const int8_t* __restrict__ in, int64_t M, int64_t K, int64_t Mp, int64_t Kp, int S,
                           int8_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t tot = (int64_t)S * M * K;
    if (i >= tot) return;
    const int64_t s = i / (M * K);
    const int64_t r = (i / K) % M;
    const int64_t k = i % K;
    out[i] = in[s * Mp * Kp + frag_off16(r, k >> 4, Kp) + (k & 15)];
}

__global__ void k_unpack_B(const int8_t* __restrict__ in, int64_t K, int64_t N, int64_t Np, int64_t Kp, int S,
                           int8_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t tot = (int64_t)S * K * N;
    if (i >= tot) return;
    const int64_t s = i / (K * N);
    const int64_t k = (i / N) % K;
    const int64_t n = i % N;
    out[i] = in[s * Np * Kp + frag_off16(n, k >> 4, Kp) + (k & 15)];
}

}  // namespace

hipError_t launch_slice_s2(const Dims& d, const Operand& A, const Operand& B, const int32_t* shift_A,
                           const int32_t* shift_B, int8_t* A_sl, int8_t* B_sl, hipStream_t st) {
    const ResidueConsts rc = make_residue_consts(d.S);
    // |t| <= 2^kparam.  Up to kparam = 72 the two-part split alone is exact (|v| < 2^48, see
    // sym_residue); above it the high part is first reduced mod m (WIDE, exact for |t| < 2^80, which
    // covers K_param's largest value, 75 at S = 20 and K <= 2: test_stage2_tiny_K_wide_residues).
    const bool wide = scheme2_kparam(d.S, (size_t)d.K) > 72;
    const SliceOp a = make_slice_op(A.p, d.M, d.Mp, A.ld, A.kcontig, d.Kp, shift_A, A_sl);
    const SliceOp b = make_slice_op(B.p, d.N, d.Np, B.ld, B.kcontig, d.Kp, shift_B, B_sl);
    const unsigned grid = (unsigned)(a.nblk + b.nblk);
    if (wide) k_slice_s2<true><<<grid, kSliceThreads, 0, st>>>(a, b, d.K, d.Kp, d.S, rc);
    else k_slice_s2<false><<<grid, kSliceThreads, 0, st>>>(a, b, d.K, d.Kp, d.S, rc);
    return hipGetLastError();
}

hipError_t launch_slice_s1(const Dims& d, int beta, const Operand& A, const Operand& B, const int32_t* shift_A,
                           const int32_t* shift_B, int8_t* A_sl, int8_t* B_sl, hipStream_t st) {
    const SliceOp a = make_slice_op(A.p, d.M, d.Mp, A.ld, A.kcontig, d.Kp, shift_A, A_sl);
    const SliceOp b = make_slice_op(B.p, d.N, d.Np, B.ld, B.kcontig, d.Kp, shift_B, B_sl);
    k_slice_s1<<<(unsigned)(a.nblk + b.nblk), kSliceThreads, 0, st>>>(a, b, d.K, d.Kp, d.S, beta);
    return hipGetLastError();
}

hipError_t launch_unpack_A(const Dims& d, const int8_t* A_sl, int8_t* out, hipStream_t st) {
    const int64_t tot = (int64_t)d.S * d.M * d.K;
    if (tot == 0) return hipSuccess;
    k_unpack_A<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(A_sl, d.M, d.K, d.Mp, d.Kp, d.S, out);
    return hipGetLastError();
}

hipError_t launch_unpack_B(const Dims& d, const int8_t* B_sl, int8_t* out, hipStream_t st) {
    const int64_t tot = (int64_t)d.S * d.K * d.N;
    if (tot == 0) return hipSuccess;
    k_unpack_B<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(B_sl, d.K, d.N, d.Np, d.Kp, d.S, out);
    return hipGetLastError();
}

}  // namespace oza
