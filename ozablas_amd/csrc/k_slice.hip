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

template <bool WIDE>
__global__ void __launch_bounds__(kSliceThreads)
    k_slice_s2_A(const double* __restrict__ A, const int32_t* __restrict__ shift, int64_t M, int64_t K,
                 int64_t Mp, int64_t Kp, int S, ResidueConsts rc, int8_t* __restrict__ out) {
    slice_s2_A_block<WIDE>(A, shift, M, K, Mp, Kp, S, rc, out, blockIdx.x, blockIdx.y);
}

template <bool WIDE>
__global__ void __launch_bounds__(kSliceThreads)
    k_slice_s2_B(const double* __restrict__ B, const int32_t* __restrict__ shift, int64_t K, int64_t N,
                 int64_t Np, int64_t Kp, int S, ResidueConsts rc, int8_t* __restrict__ out) {
    slice_s2_B_block<WIDE>(B, shift, K, N, Np, Kp, S, rc, out, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(kSliceThreads)
    k_slice_s1_A(const double* __restrict__ A, const int32_t* __restrict__ neg_exp, int64_t M, int64_t K,
                 int64_t Mp, int64_t Kp, int S, int beta, int8_t* __restrict__ out) {
    slice_s1_A_block(A, neg_exp, M, K, Mp, Kp, S, beta, out, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(kSliceThreads)
    k_slice_s1_B(const double* __restrict__ B, const int32_t* __restrict__ neg_exp, int64_t K, int64_t N,
                 int64_t Np, int64_t Kp, int S, int beta, int8_t* __restrict__ out) {
    slice_s1_B_block(B, neg_exp, K, N, Np, Kp, S, beta, out, blockIdx.x, blockIdx.y);
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

hipError_t launch_slice_s2(const Dims& d, const double* A, const double* B, const int32_t* shift_A,
                           const int32_t* shift_B, int8_t* A_sl, int8_t* B_sl, hipStream_t st) {
    const ResidueConsts rc = make_residue_consts(d.S);
    // |t| <= 2^kparam; the two-part split is exact up to kparam = 74 (see sym_residue).
    const bool wide = scheme2_kparam(d.S, (size_t)d.K) > 72;
    const dim3 gA((unsigned)(d.Mp / 32), (unsigned)(d.Kp / 128));
    const dim3 gB((unsigned)(d.Np / 64), (unsigned)(d.Kp / 128));
    if (wide) {
        k_slice_s2_A<true><<<gA, kSliceThreads, 0, st>>>(A, shift_A, d.M, d.K, d.Mp, d.Kp, d.S, rc, A_sl);
        k_slice_s2_B<true><<<gB, kSliceThreads, 0, st>>>(B, shift_B, d.K, d.N, d.Np, d.Kp, d.S, rc, B_sl);
    } else {
        k_slice_s2_A<false><<<gA, kSliceThreads, 0, st>>>(A, shift_A, d.M, d.K, d.Mp, d.Kp, d.S, rc, A_sl);
        k_slice_s2_B<false><<<gB, kSliceThreads, 0, st>>>(B, shift_B, d.K, d.N, d.Np, d.Kp, d.S, rc, B_sl);
    }
    return hipGetLastError();
}

hipError_t launch_slice_s1(const Dims& d, int beta, const double* A, const double* B, const int32_t* shift_A,
                           const int32_t* shift_B, int8_t* A_sl, int8_t* B_sl, hipStream_t st) {
    const dim3 gA((unsigned)(d.Mp / 32), (unsigned)(d.Kp / 128));
    const dim3 gB((unsigned)(d.Np / 64), (unsigned)(d.Kp / 128));
    k_slice_s1_A<<<gA, kSliceThreads, 0, st>>>(A, shift_A, d.M, d.K, d.Mp, d.Kp, d.S, beta, A_sl);
    k_slice_s1_B<<<gB, kSliceThreads, 0, st>>>(B, shift_B, d.K, d.N, d.Np, d.Kp, d.S, beta, B_sl);
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
