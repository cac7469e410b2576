// k_recon.hip — step 4: FP64 reconstruction.
//
// Scheme II (reference: src/pipeline/step4_reconstruction.hpp:59-184, crt_math.hpp:38-80).
// The reference sums c_s * (M/m_s) modulo M, c_s = (C_s mod m_s) * (M/m_s)^-1 mod m_s, in u64
// (S <= 7) or four u64 limbs (S > 7), maps acc >= floor(M/2) to -(M - acc), converts limb by limb
// (to_double_256) and scales by 2^-(shift_A[i] + shift_B[j]).  The value it converts is the fixed
// integer C mod M, so any exact method gives the same double.  Here, with r_s = C_s mod m_s and the
// CRT basis Q_s = (M/m_s)^-1 mod m_s * (M/m_s): X = sum_s r_s Q_s accumulated without reduction in
// 16-bit limbs held in 32-bit registers (from blocked digits: one v_dot4_u32_u8 per four terms
// and byte of a limb; from INT32 planes: one v_mad_u32_u24 per limb and term), the quotient
// q = round(sum_s r_s Q_s/M) from an FP64 dot product, Y = X - q*M by signed limb arithmetic, an
// exact range check only when q was ambiguous, then the reference's limb-wise conversion.
//
// Scheme I group reconstruction (step4_reconstruction.hpp:21-49): C = 0, then
// C = fma(T_g, scale_g, C) for g = 2..S+1 ascending, from the GEMM's blocked INT32 planes
// (k_recon_s1_blocked) or from the reference's planar [S][M][N] planes (stage entry point).
#include <algorithm>

#include "recon_core.hpp"

namespace oza {
namespace {

using namespace core;

// ---------------------------------------------------------------- from GEMM digits
// Workgroup = one 32 x 32 element block of C at a time: blocks [0, nblk) of the block rows from
// R0 on, block b = (R0 + b / ceil(N/32), b % ceil(N/32)), walked with a grid stride.
template <int S, bool PLAIN>
__global__ void __launch_bounds__(256)
    k_recon_digits(const uint8_t* __restrict__ dg, const int32_t* __restrict__ shA, const int32_t* __restrict__ shB,
                   int64_t M, int64_t N, int64_t Mp, int64_t Np, const CrtSet* __restrict__ cs, OutC c, int64_t R0,
                   int64_t nblk) {
    const int64_t qn = (N + 31) >> 5;
    for (int64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const int64_t r = b / qn, Q = b - r * qn;
        recon_digits_block<S, PLAIN>(dg, Mp * Np, Np, shA, shB, M, N, cs, c.p, c.ldc, c.alpha, c.beta, R0 + r, Q);
    }
}

// ---------------------------------------------------------------- from INT32 planes [s][M][N]
template <int S>
__global__ void __launch_bounds__(256)
    k_recon_int32(const int32_t* __restrict__ Ct, const int32_t* __restrict__ shA, const int32_t* __restrict__ shB,
                  int64_t M, int64_t N, const CrtSet* __restrict__ cs, double* __restrict__ Cout) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t MN = M * N;
    if (idx >= MN) return;
    const int64_t row = idx / N, col = idx - row * N;
    uint32_t c[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
        c[s] = (uint32_t)crt_digit(Ct[s * MN + idx], (double)cs->m[s], cs->rcp[s]);
    const double f = crt_value<S>(c, cs);
    Cout[idx] = ldexp(f, -(shA[row] + shB[col]));
}

// ---------------------------------------------------------------- Scheme I from INT32 group planes
__global__ void __launch_bounds__(256)
    k_recon_s1(const int32_t* __restrict__ T, const int32_t* __restrict__ shA, const int32_t* __restrict__ shB,
               int64_t M, int64_t N, int S, int beta, double* __restrict__ Cout) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t MN = M * N;
    if (idx >= MN) return;
    const int64_t row = idx / N, col = idx - row * N;
    const double u = ldexp(1.0, -shA[row] + 1 - beta);
    const double v = ldexp(1.0, -shB[col] + 1 - beta);
    double c = 0.0;
    for (int g = 2; g <= S + 1; ++g) {
        const double scale = ldexp(u * v, -beta * g + 2 * beta);
        c = fma((double)T[(int64_t)(g - 2) * MN + idx], scale, c);
    }
    Cout[idx] = c;
}

// ---------------------------------------------------------------- Scheme I from blocked INT32 planes
// One 32 x 32 element block (R, Q); thread t reads int4 t of the block in every plane (4
// elements of one column).  C = 0, then C = fma(T_g, 2^(2 beta - beta g) u_i v_j, C) for
// g = 2..S+1 (step4_reconstruction.hpp:21-49, pipeline.hip:80, ascending g).  The planes are read
// four at a time ahead of their FMAs.
__device__ __forceinline__ void recon_s1_block(const int32_t* __restrict__ T, const int32_t* __restrict__ shA,
                                               const int32_t* __restrict__ shB, int64_t M, int64_t N, int64_t Mp,
                                               int64_t Np, int S, int beta, const OutC& c, int64_t R, int64_t Q) {
    const int t = threadIdx.x;
    const int lane = t >> 2, wd = t & 3;
    const int64_t col = Q * 32 + (lane & 31);
    const int64_t row0 = R * 32 + 8 * wd + 4 * (lane >> 5);
    if (col >= N || row0 >= M) return;
    const int32_t* base = T + ((R * (Np >> 5) + Q) << 10) + t * 4;
    const int64_t plane = Mp * Np;
    const double v = ldexp(1.0, -shB[col] + 1 - beta);
    double u[4], acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int b = 0; b < 4; ++b) u[b] = row0 + b < M ? ldexp(1.0, -shA[row0 + b] + 1 - beta) : 0.0;
    for (int g0 = 2; g0 <= S + 1; g0 += 4) {
        int4 x[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)  // planes past S + 1 read the last one again and are not used
            x[i] = *reinterpret_cast<const int4*>(base + (int64_t)(min(g0 + i, S + 1) - 2) * plane);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int g = g0 + i;
            if (g > S + 1) break;
            const int xs[4] = {x[i].x, x[i].y, x[i].z, x[i].w};
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[b] = fma((double)xs[b], ldexp(u[b] * v, -beta * g + 2 * beta), acc[b]);
        }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b)
        if (row0 + b < M) store_c(c.p, (row0 + b) * c.ldc + col, acc[b], c.alpha, c.beta);
}

// Blocks [0, nblk) of the block rows from R0 on, walked with a grid stride (as k_recon_digits).
__global__ void __launch_bounds__(256)
    k_recon_s1_blocked(const int32_t* __restrict__ T, const int32_t* __restrict__ shA,
                       const int32_t* __restrict__ shB, int64_t M, int64_t N, int64_t Mp, int64_t Np, int S,
                       int beta, OutC c, int64_t R0, int64_t nblk) {
    const int64_t qn = (N + 31) >> 5;
    for (int64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        const int64_t r = b / qn, Q = b - r * qn;
        recon_s1_block(T, shA, shB, M, N, Mp, Np, S, beta, c, R0 + r, Q);
    }
}

// C = beta * C (BLAS quick path when alpha == 0 or K == 0); beta == 0 writes zeros without
// reading C.
__global__ void __launch_bounds__(256) k_scale_c(OutC c, int64_t M, int64_t N) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= M * N) return;
    const int64_t row = idx / N, col = idx - row * N;
    double* p = c.p + row * c.ldc + col;
    *p = c.beta == 0.0 ? 0.0 : c.beta * *p;
}

// Byte idx of the blocked digit planes -> (modulus s, row, col) of the M x N product.
__device__ __forceinline__ void digit_position(int64_t idx, int64_t plane, int64_t Np, int64_t& s, int64_t& row,
                                               int64_t& col) {
    s = idx / plane;
    const int64_t r = idx - s * plane;
    const int64_t blk = r >> 10;
    const int within = (int)(r & 1023);
    const int lane = within >> 4, j = within & 15;
    const int64_t R = blk / (Np >> 5), Q = blk - R * (Np >> 5);
    row = R * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
    col = Q * 32 + (lane & 31);
}

// Blocked digit planes -> planar [s][M][N] bytes (for inspection and parity tests).
__global__ void __launch_bounds__(256)
    k_digits_unpack(const uint8_t* __restrict__ dg, int64_t M, int64_t N, int64_t Mp, int64_t Np, int S,
                    uint8_t* __restrict__ out) {
    const int64_t plane = Mp * Np;
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= plane * S) return;
    int64_t s, row, col;
    digit_position(idx, plane, Np, s, row, col);
    if (row < M && col < N) out[(s * M + row) * N + col] = dg[idx];
}

// Planar [s][M][N] digits -> blocked digit planes, reduced mod m_s (the inverse of
// k_digits_unpack for digits in [0, m_s); padding positions get 0).
__global__ void __launch_bounds__(256)
    k_digits_pack(const uint8_t* __restrict__ in, int64_t M, int64_t N, int64_t Mp, int64_t Np, int S,
                  const CrtSet* __restrict__ cs, uint8_t* __restrict__ dg) {
    const int64_t plane = Mp * Np;
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= plane * S) return;
    int64_t s, row, col;
    digit_position(idx, plane, Np, s, row, col);
    dg[idx] = row < M && col < N ? (uint8_t)(in[(s * M + row) * N + col] % cs->m[s]) : 0;
}

template <int S>
hipError_t recon_digits_S(const Dims& d, const uint8_t* dg, const int32_t* shA, const int32_t* shB,
                          const CrtSet* cs, const OutC& c, hipStream_t st, int64_t R0, int64_t nR, unsigned grid) {
    const int64_t nblk = nR * ((d.N + 31) >> 5);
    if (nblk == 0) return hipSuccess;
    if (grid == 0 || grid > nblk) grid = (unsigned)nblk;
    if (c.alpha == 1.0 && c.beta == 0.0)
        k_recon_digits<S, true><<<grid, 256, 0, st>>>(dg, shA, shB, d.M, d.N, d.Mp, d.Np, cs, c, R0, nblk);
    else
        k_recon_digits<S, false><<<grid, 256, 0, st>>>(dg, shA, shB, d.M, d.N, d.Mp, d.Np, cs, c, R0, nblk);
    return hipGetLastError();
}

template <int S>
hipError_t recon_int32_S(const Dims& d, const int32_t* Ct, const int32_t* shA, const int32_t* shB,
                         const CrtSet* cs, double* C, hipStream_t st) {
    const int64_t n = d.M * d.N;
    k_recon_int32<S><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(Ct, shA, shB, d.M, d.N, cs, C);
    return hipGetLastError();
}

#define OZA_S_CASES(F, ...)                                                                          \
    switch (d.S) {                                                                                   \
        case 1: return F<1>(__VA_ARGS__);   case 2: return F<2>(__VA_ARGS__);                        \
        case 3: return F<3>(__VA_ARGS__);   case 4: return F<4>(__VA_ARGS__);                        \
        case 5: return F<5>(__VA_ARGS__);   case 6: return F<6>(__VA_ARGS__);                        \
        case 7: return F<7>(__VA_ARGS__);   case 8: return F<8>(__VA_ARGS__);                        \
        case 9: return F<9>(__VA_ARGS__);   case 10: return F<10>(__VA_ARGS__);                      \
        case 11: return F<11>(__VA_ARGS__); case 12: return F<12>(__VA_ARGS__);                      \
        case 13: return F<13>(__VA_ARGS__); case 14: return F<14>(__VA_ARGS__);                      \
        case 15: return F<15>(__VA_ARGS__); case 16: return F<16>(__VA_ARGS__);                      \
        case 17: return F<17>(__VA_ARGS__); case 18: return F<18>(__VA_ARGS__);                      \
        case 19: return F<19>(__VA_ARGS__); case 20: return F<20>(__VA_ARGS__);                      \
        default: return hipErrorInvalidValue;                                                        \
    }

}  // namespace

hipError_t launch_recon_s2_digits(const Dims& d, const uint8_t* digits, const int32_t* shift_A,
                                  const int32_t* shift_B, const CrtSet* crt_dev, const OutC& c, hipStream_t st,
                                  int64_t row0, int64_t rows, bool narrow) {
    if (d.M == 0 || d.N == 0) return hipSuccess;
    if (row0 < 0 || row0 > d.M || (row0 & 31)) return hipErrorInvalidValue;
    const int64_t end = rows < 0 ? d.M : std::min<int64_t>(d.M, row0 + rows);
    const int64_t R0 = row0 >> 5, nR = (end - row0 + 31) >> 5;
    unsigned grid = 0;  // one workgroup per block
    if (narrow) {
        int dev = 0, cus = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        grid = (unsigned)std::max(cus, 1);
    }
    OZA_S_CASES(recon_digits_S, d, digits, shift_A, shift_B, crt_dev, c, st, R0, nR, grid)
}

hipError_t launch_recon_s2_int32(const Dims& d, const int32_t* C_tc, const int32_t* shift_A,
                                 const int32_t* shift_B, const CrtSet* crt_dev, double* C, hipStream_t st) {
    if (d.M == 0 || d.N == 0) return hipSuccess;
    OZA_S_CASES(recon_int32_S, d, C_tc, shift_A, shift_B, crt_dev, C, st)
}

hipError_t launch_scale_c(const OutC& c, int64_t M, int64_t N, hipStream_t st) {
    const int64_t n = M * N;
    if (n == 0) return hipSuccess;
    k_scale_c<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(c, M, N);
    return hipGetLastError();
}

hipError_t launch_digits_unpack(const Dims& d, const uint8_t* digits, uint8_t* out, hipStream_t st) {
    const int64_t n = d.Mp * d.Np * d.S;
    if (n == 0) return hipSuccess;
    k_digits_unpack<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(digits, d.M, d.N, d.Mp, d.Np, d.S, out);
    return hipGetLastError();
}

hipError_t launch_digits_pack(const Dims& d, const uint8_t* planar, const CrtSet* crt_dev, uint8_t* digits,
                              hipStream_t st) {
    const int64_t n = d.Mp * d.Np * d.S;
    if (n == 0) return hipSuccess;
    k_digits_pack<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(planar, d.M, d.N, d.Mp, d.Np, d.S, crt_dev, digits);
    return hipGetLastError();
}

hipError_t launch_recon_s1_blocked(const Dims& d, int beta, const int32_t* T, const int32_t* shift_A,
                                   const int32_t* shift_B, const OutC& c, hipStream_t st, int64_t row0, int64_t rows,
                                   bool narrow) {
    if (d.M == 0 || d.N == 0) return hipSuccess;
    if (row0 < 0 || row0 > d.M || (row0 & 31)) return hipErrorInvalidValue;
    const int64_t end = rows < 0 ? d.M : std::min<int64_t>(d.M, row0 + rows);
    const int64_t nblk = ((end - row0 + 31) >> 5) * ((d.N + 31) >> 5);
    if (nblk == 0) return hipSuccess;
    unsigned grid = (unsigned)nblk;
    if (narrow) {
        int dev = 0, cus = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        grid = (unsigned)std::min<int64_t>(nblk, std::max(cus, 1));
    }
    k_recon_s1_blocked<<<grid, 256, 0, st>>>(T, shift_A, shift_B, d.M, d.N, d.Mp, d.Np, d.S, beta, c, row0 >> 5,
                                             nblk);
    return hipGetLastError();
}

hipError_t launch_recon_s1_groups(const Dims& d, int beta, const int32_t* T, const int32_t* shift_A,
                                  const int32_t* shift_B, double* C, hipStream_t st) {
    const int64_t n = d.M * d.N;
    if (n == 0) return hipSuccess;
    k_recon_s1<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(T, shift_A, shift_B, d.M, d.N, d.S, beta, C);
    return hipGetLastError();
}

}  // namespace oza
