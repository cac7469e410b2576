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

#include "oza_device.hpp"
#include "oza_internal.hpp"

namespace oza {
namespace {

constexpr int kSliceThreads = 256;

struct ResidueConsts {
    double m[kMaxSlices];
    double rcp[kMaxSlices];
    double c32[kMaxSlices];  // 2^32 mod m
};

ResidueConsts make_residue_consts(int S) {
    ResidueConsts rc;
    for (int s = 0; s < kMaxSlices; ++s) {
        const uint32_t m = kModuli[s];
        rc.m[s] = (double)m;
        rc.rcp[s] = 1.0 / (double)m;
        rc.c32[s] = (double)((uint64_t(1) << 32) % m);
    }
    (void)S;
    return rc;
}

// Symmetric residues of 16 split integers for one modulus, packed into 16 bytes.
template <bool WIDE>
__device__ __forceinline__ uint4 residues16(const SplitInt (&t)[16], double m, double rcp, double c32) {
    int r[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        SplitInt tj = t[j];
        if (WIDE) {  // |t| up to 2^80: first reduce the high part modulo m (exact, |th| < 2^48)
            const double q = rint(tj.th * rcp);
            tj.th = fma(-q, m, tj.th);
        }
        r[j] = sym_residue(tj, m, rcp, c32);
    }
    return make_uint4(pack4(r[0], r[1], r[2], r[3]), pack4(r[4], r[5], r[6], r[7]),
                      pack4(r[8], r[9], r[10], r[11]), pack4(r[12], r[13], r[14], r[15]));
}

// ---------------------------------------------------------------- Scheme II, A (row shifts)
// Block = 32 rows x 128 K; thread = row (tid % 32), 16-byte K chunk (tid / 32).  The 64 lanes of a
// wave cover one whole 1 KiB fragment block.
template <bool WIDE>
__global__ void __launch_bounds__(kSliceThreads)
    k_slice_s2_A(const double* __restrict__ A, const int32_t* __restrict__ shift, int64_t M, int64_t K,
                 int64_t Mp, int64_t Kp, int S, ResidueConsts rc, int8_t* __restrict__ out) {
    const int64_t row = (int64_t)blockIdx.x * 32 + (threadIdx.x & 31);
    const int64_t c = (int64_t)blockIdx.y * 8 + (threadIdx.x >> 5);
    const int64_t k0 = c * 16;
    SplitInt t[16];
    if (row < M) {
        const int sh = shift[row];
        const double* p = A + row * K + k0;
        double a[16];
        if (k0 + 16 <= K && ((reinterpret_cast<uintptr_t>(p) & 15) == 0)) {
#pragma unroll
            for (int j = 0; j < 16; j += 2) {
                const double2 v = *reinterpret_cast<const double2*>(p + j);
                a[j] = v.x;
                a[j + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = (k0 + j < K) ? p[j] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) t[j] = split32(rint(ldexp(a[j], sh)));
    } else {
#pragma unroll
        for (int j = 0; j < 16; ++j) t[j] = {0.0, 0.0};
    }
    const int64_t plane = Mp * Kp;
    int8_t* o = out + frag_off16(row, c, Kp);
    for (int s = 0; s < S; ++s)
        *reinterpret_cast<uint4*>(o + s * plane) = residues16<WIDE>(t, rc.m[s], rc.rcp[s], rc.c32[s]);
}

// ---------------------------------------------------------------- Scheme II, B (column shifts)
// Block = 64 columns x 4 groups of 32 K rows; thread = one column, 32 consecutive k.
template <bool WIDE>
__global__ void __launch_bounds__(kSliceThreads)
    k_slice_s2_B(const double* __restrict__ B, const int32_t* __restrict__ shift, int64_t K, int64_t N,
                 int64_t Np, int64_t Kp, int S, ResidueConsts rc, int8_t* __restrict__ out) {
    const int64_t col = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int64_t kbase = (int64_t)blockIdx.y * 128 + (threadIdx.x >> 6) * 32;
    const int64_t plane = Np * Kp;
    const int sh = col < N ? shift[col] : 0;
#pragma unroll 1
    for (int half = 0; half < 2; ++half) {
        const int64_t k0 = kbase + half * 16;
        SplitInt t[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const double a = (col < N && k0 + j < K) ? B[(k0 + j) * N + col] : 0.0;
            t[j] = split32(rint(ldexp(a, sh)));
        }
        int8_t* o = out + frag_off16(col, k0 >> 4, Kp);
        for (int s = 0; s < S; ++s)
            *reinterpret_cast<uint4*>(o + s * plane) = residues16<WIDE>(t, rc.m[s], rc.rcp[s], rc.c32[s]);
    }
}

// ---------------------------------------------------------------- Scheme I
// Slices of 16 values sharing one exponent e (row of A or column of B).
__device__ __forceinline__ void slice_s1_16(double (&a)[16], int e, int beta, int S, int8_t* o, int64_t plane) {
    const double scale = ldexp(1.0, e + 1 - beta);
    const double inv_scale = ldexp(1.0, -(e + 1 - beta));
    for (int s = 0; s < S; ++s) {
        const double sigma = 0.75 * ldexp(scale, 53 - beta * s);
        int r[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            // (a + sigma) - sigma: contraction and reassociation are off in this file.
            const double tmp = a[j] + sigma;
            const double a_s = tmp - sigma;
            const double normalized = a_s * inv_scale;
            const double scaled = ldexp(normalized, beta * s);
            r[j] = (int)scaled;
            a[j] = a[j] - a_s;
        }
        *reinterpret_cast<uint4*>(o + s * plane) =
            make_uint4(pack4(r[0], r[1], r[2], r[3]), pack4(r[4], r[5], r[6], r[7]),
                       pack4(r[8], r[9], r[10], r[11]), pack4(r[12], r[13], r[14], r[15]));
    }
}

__global__ void __launch_bounds__(kSliceThreads)
    k_slice_s1_A(const double* __restrict__ A, const int32_t* __restrict__ neg_exp, int64_t M, int64_t K,
                 int64_t Mp, int64_t Kp, int S, int beta, int8_t* __restrict__ out) {
    const int64_t row = (int64_t)blockIdx.x * 32 + (threadIdx.x & 31);
    const int64_t c = (int64_t)blockIdx.y * 8 + (threadIdx.x >> 5);
    const int64_t k0 = c * 16;
    int8_t* o = out + frag_off16(row, c, Kp);
    if (row < M) {
        const int e = -neg_exp[row];
        const double* p = A + row * K + k0;
        double a[16];
        if (k0 + 16 <= K && ((reinterpret_cast<uintptr_t>(p) & 15) == 0)) {
#pragma unroll
            for (int j = 0; j < 16; j += 2) {
                const double2 v = *reinterpret_cast<const double2*>(p + j);
                a[j] = v.x;
                a[j + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = (k0 + j < K) ? p[j] : 0.0;
        }
        slice_s1_16(a, e, beta, S, o, Mp * Kp);
    } else {
        for (int s = 0; s < S; ++s) *reinterpret_cast<uint4*>(o + s * Mp * Kp) = make_uint4(0, 0, 0, 0);
    }
}

__global__ void __launch_bounds__(kSliceThreads)
    k_slice_s1_B(const double* __restrict__ B, const int32_t* __restrict__ neg_exp, int64_t K, int64_t N,
                 int64_t Np, int64_t Kp, int S, int beta, int8_t* __restrict__ out) {
    const int64_t col = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int64_t kbase = (int64_t)blockIdx.y * 128 + (threadIdx.x >> 6) * 32;
    const int64_t plane = Np * Kp;
#pragma unroll 1
    for (int half = 0; half < 2; ++half) {
        const int64_t k0 = kbase + half * 16;
        int8_t* o = out + frag_off16(col, k0 >> 4, Kp);
        if (col < N) {
            const int e = -neg_exp[col];
            double a[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = (k0 + j < K) ? B[(k0 + j) * N + col] : 0.0;
            slice_s1_16(a, e, beta, S, o, plane);
        } else {
            for (int s = 0; s < S; ++s) *reinterpret_cast<uint4*>(o + s * plane) = make_uint4(0, 0, 0, 0);
        }
    }
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
