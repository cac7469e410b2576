// k_util.hip — accuracy tooling that runs beside the Ozaki pipeline on the GPU.
//
// dd_gemm: every C entry as a double-double dot product (Ogita-Rump-Oishi Dot2: TwoProduct by
// FMA, TwoSum, k ascending), rounded once to FP64, plus |A||B| (fma accumulation, k ascending) for
// componentwise error bounds.  It is the full-matrix accuracy reference for sizes where an FP128
// CPU product is too slow.  Its error is at most u|x| + gamma_K^2 (|A||B|)_ij with u = 2^-53 and
// gamma_K = K u / (1 - K u) (Ogita, Rump, Oishi 2005, Prop. 4.5): of order K^2 2^-106 relative to
// |A||B| (2^-80 at K = 8192), far below the FP64 rounding the compared products carry.
//
// error_stats: max |d|, max relative (the reference's metric, utils/matrix_compare.hpp:47),
// max componentwise |d| / (|A||B|)_ij and the number of differing bit patterns.
//
// mfma_ceiling: the sustained INT8 MFMA rate and shader clock with no memory traffic (below).
#pragma clang fp contract(off)

#include "oza_internal.hpp"

namespace oza {
namespace {

constexpr int kT = 16;

__global__ void __launch_bounds__(kT* kT)
    k_dd_gemm(int64_t M, int64_t N, int64_t K, const double* __restrict__ A, const double* __restrict__ B,
              double* __restrict__ C, double* __restrict__ Cabs) {
    __shared__ double sa[kT][kT + 1];
    __shared__ double sb[kT][kT + 1];
    const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;
    const int64_t row = (int64_t)blockIdx.y * kT + ty;
    const int64_t col = (int64_t)blockIdx.x * kT + tx;
    double s = 0.0, c = 0.0, ab = 0.0;
    for (int64_t k0 = 0; k0 < K; k0 += kT) {
        const int64_t ka = k0 + tx, kb = k0 + ty;
        sa[ty][tx] = (row < M && ka < K) ? A[row * K + ka] : 0.0;
        sb[ty][tx] = (kb < K && col < N) ? B[kb * N + col] : 0.0;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kT; ++k) {
            const double a = sa[ty][k], b = sb[k][tx];
            const double p = a * b;
            const double ep = fma(a, b, -p);
            const double x = s + p;
            const double z = x - s;
            const double es = (s - (x - z)) + (p - z);
            s = x;
            c += ep + es;
            ab = fma(fabs(a), fabs(b), ab);
        }
        __syncthreads();
    }
    if (row < M && col < N) {
        C[row * N + col] = s + c;
        if (Cabs) Cabs[row * N + col] = ab;
    }
}

__device__ __forceinline__ void atomic_max_pos(double* p, double v) {
    // non-negative doubles order like their bit patterns
    atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

__global__ void __launch_bounds__(256)
    k_error_stats(int64_t n, const double* __restrict__ e, const double* __restrict__ a,
                  const double* __restrict__ absab, double* __restrict__ stats) {
    double mabs = 0.0, mrel = 0.0, mcw = 0.0;
    unsigned long long diff = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double ev = e[i], av = a[i];
        const double d = fabs(av - ev);
        mabs = fmax(mabs, d);
        const double ae = fabs(ev);
        mrel = fmax(mrel, ae > 0.0 ? d / ae : d);
        if (absab) {
            const double b = absab[i];
            mcw = fmax(mcw, b > 0.0 ? d / b : d);
        }
        diff += (__double_as_longlong(ev) != __double_as_longlong(av)) ? 1ull : 0ull;
    }
    __shared__ double r0[256], r1[256], r2[256];
    __shared__ unsigned long long r3[256];
    r0[threadIdx.x] = mabs;
    r1[threadIdx.x] = mrel;
    r2[threadIdx.x] = mcw;
    r3[threadIdx.x] = diff;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            r0[threadIdx.x] = fmax(r0[threadIdx.x], r0[threadIdx.x + o]);
            r1[threadIdx.x] = fmax(r1[threadIdx.x], r1[threadIdx.x + o]);
            r2[threadIdx.x] = fmax(r2[threadIdx.x], r2[threadIdx.x + o]);
            r3[threadIdx.x] += r3[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomic_max_pos(stats + 0, r0[0]);
        atomic_max_pos(stats + 1, r1[0]);
        atomic_max_pos(stats + 2, r2[0]);
        atomicAdd(reinterpret_cast<unsigned long long*>(stats + 3), r3[0]);
    }
}

// mfma_ceiling: the INT8 MFMA rate this device sustains with no memory traffic at all -- every
// workgroup (8 waves, one per CU) runs v_mfma_i32_32x32x32_i8 on operands held in registers, eight
// independent accumulators per wave as in the GEMM's k-step -- on random residues in [-125, 125] or
// on zeros.  On random data the chip's power limit lowers the clock below the nominal 2.4 GHz, so
// this rate, not the nominal peak, is what a GEMM on random residues can approach.  Thread 0 of
// each workgroup adds its shader-clock cycles (s_memtime) and 100 MHz real-time ticks
// (s_memrealtime) over the loop to acc[0] and acc[1]; acc[2] receives a checksum of the results.
__global__ void __launch_bounds__(512) k_mfma_ceiling(int iters, int zero, unsigned long long* __restrict__ acc) {
    typedef int v4i __attribute__((ext_vector_type(4)));
    typedef int v16i __attribute__((ext_vector_type(16)));
    v4i a[4][2], b[4][4];
    uint32_t x = (blockIdx.x * 512u + threadIdx.x) * 2654435761u + 12345u;
    auto next = [&]() {  // four residues in [-125, 125] (or zeros), one byte each
        uint32_t w = 0;
        for (int k = 0; k < 4; ++k) {
            x ^= x << 13;
            x ^= x >> 17;
            x ^= x << 5;
            w |= (uint32_t)(uint8_t)(int8_t)((int)(x % 251u) - 125) << (8 * k);
        }
        return zero ? 0u : w;
    };
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
#pragma unroll
        for (int i = 0; i < 2; ++i) a[kk][i] = v4i{(int)next(), (int)next(), (int)next(), (int)next()};
#pragma unroll
        for (int i = 0; i < 4; ++i) b[kk][i] = v4i{(int)next(), (int)next(), (int)next(), (int)next()};
    }
    v16i c[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) c[i][j] = v16i{};
    const unsigned long long t0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime();
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
#pragma unroll
            for (int fm = 0; fm < 2; ++fm)
#pragma unroll
                for (int fn = 0; fn < 4; ++fn)
                    c[fm][fn] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[kk][fm], b[kk][fn], c[fm][fn], 0, 0, 0);
    }
    const unsigned long long t1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime();
    int sum = 0;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int k = 0; k < 16; ++k) sum += c[i][j][k];
    if (threadIdx.x == 0) {
        atomicAdd(acc + 0, t1 - t0);
        atomicAdd(acc + 1, r1 - r0);
    }
    if (sum == 0x7fffffff) atomicAdd(acc + 2, 1ull);  // keeps the products live
}

__global__ void k_zero4(double* p) {
    if (threadIdx.x < 4) reinterpret_cast<unsigned long long*>(p)[threadIdx.x] = 0ull;
}

}  // namespace

hipError_t launch_mfma_ceiling(int iters, bool zero, unsigned grid, double* acc4, hipStream_t st) {
    k_zero4<<<1, 64, 0, st>>>(acc4);
    k_mfma_ceiling<<<grid, 512, 0, st>>>(iters, zero ? 1 : 0, reinterpret_cast<unsigned long long*>(acc4));
    return hipGetLastError();
}

hipError_t launch_dd_gemm(int64_t M, int64_t N, int64_t K, const double* A, const double* B, double* C_hi,
                          double* C_abs, hipStream_t st) {
    if (M == 0 || N == 0) return hipSuccess;
    const dim3 grid((unsigned)((N + kT - 1) / kT), (unsigned)((M + kT - 1) / kT));
    k_dd_gemm<<<grid, kT * kT, 0, st>>>(M, N, K, A, B, C_hi, C_abs);
    return hipGetLastError();
}

hipError_t launch_error_stats(int64_t n, const double* exact, const double* actual, const double* absab,
                              double* stats, hipStream_t st) {
    k_zero4<<<1, 64, 0, st>>>(stats);
    if (n > 0) {
        int64_t blocks = (n + 255) / 256;
        if (blocks > 4096) blocks = 4096;
        k_error_stats<<<(unsigned)blocks, 256, 0, st>>>(n, exact, actual, absab, stats);
    }
    return hipGetLastError();
}

}  // namespace oza
