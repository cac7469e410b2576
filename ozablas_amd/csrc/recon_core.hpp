// recon_core.hpp — Scheme II step-4 arithmetic shared by the single-problem and grouped
// reconstruction kernels: exact CRT reconstruction of one element from its digits, and the
// block body that reconstructs a 32 x 32 block of C from the blocked digit planes.
#pragma once

#include "oza_device.hpp"
#include "oza_internal.hpp"

namespace oza {
namespace core {

template <int S>
struct CrtShape {
    static constexpr int BITS = bitlen_M(S);
    static constexpr int LP = (BITS + 15) / 16;               // limbs of M / m_s
    static constexpr int NL = (BITS + 5 + 8 + 15) / 16 + 1;   // accumulator limbs + sign spare
    static_assert(NL <= kLimbs, "limb budget");
};

template <int NL>
__device__ __forceinline__ void propagate(int32_t (&y)[NL]) {
#pragma unroll
    for (int l = 0; l < NL - 1; ++l) {
        const int32_t c = y[l] >> 16;  // arithmetic shift: floor division
        y[l] -= c * 65536;
        y[l + 1] += c;
    }
}

// X = sum_s c_s Q_s in LP unpropagated 16-bit limbs (each < S * 255 * 65535 < 2^28.4), one
// v_mad_u32_u24 per limb and term.
template <int S>
__device__ __forceinline__ void crt_limbs(const uint32_t (&c)[S], const CrtSet* __restrict__ cs,
                                          int32_t (&y)[CrtShape<S>::NL]) {
    constexpr int LP = CrtShape<S>::LP, NL = CrtShape<S>::NL;
#pragma unroll
    for (int l = 0; l < NL; ++l) y[l] = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
#pragma unroll
        for (int l = 0; l < LP; ++l) y[l] = (int32_t)__umul24(c[s], cs->Q16[s][l]) + y[l];
    }
}

// The same limbs from digits packed four moduli to a word (byte k of P[j] = c_{4j+k}, 0 past S):
// the sums against the low and high bytes of each 16-bit limb of Q_s by v_dot4_u32_u8 (four
// products and the add in one instruction; each sum < S * 255^2 < 2^21), then lo + 256 hi, which
// is crt_limbs' limb exactly.
template <int S>
__device__ __forceinline__ void crt_limbs_packed(const uint32_t (&P)[(S + 3) / 4], const CrtSet* __restrict__ cs,
                                                 int32_t (&y)[CrtShape<S>::NL]) {
    constexpr int LP = CrtShape<S>::LP, NL = CrtShape<S>::NL, NJ = (S + 3) / 4;
    uint32_t b8[2 * LP];  // byte-limb sums; one table row (contiguous) per group of four moduli
#pragma unroll
    for (int l = 0; l < 2 * LP; ++l) b8[l] = 0;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
#pragma unroll
        for (int l = 0; l < 2 * LP; ++l) b8[l] = __builtin_amdgcn_udot4(P[j], cs->Q8x4[j][l], b8[l], false);
    }
#pragma unroll
    for (int l = 0; l < LP; ++l) y[l] = (int32_t)(b8[2 * l] + (b8[2 * l + 1] << 8));
#pragma unroll
    for (int l = LP; l < NL; ++l) y[l] = 0;
}

// Signed reconstruction of one element from the limbs of X = sum_s c_s Q_s (crt_limbs or
// crt_limbs_packed), returned as the reference's double.
//   q = round(X/M), with X/M estimated from the top min(LP, 5) limbs: the limbs below contribute
//     < 2^-50 and the FP64 sum errs by < 2^-38 (X/M < 20 * 256), inside the 1e-9 margin of qthr;
//   Y = X - qM.  When X/M is clearly inside (q - 1/2, q + 1/2), Y has the sign of qf - q except
//     when |Y| is below the estimate's error; |Y| is formed by negating the limbs before the one
//     carry pass, and a negative top limb afterwards (the rare wrong guess) flips it back.
//     Otherwise the exact comparisons against floor(M/2) decide (the reference's mapping of
//     X mod M >= floor(M/2) to negative).
template <int S>
__device__ __forceinline__ double crt_finish(int32_t (&y)[CrtShape<S>::NL], const CrtSet* __restrict__ cs) {
    constexpr int LP = CrtShape<S>::LP, NL = CrtShape<S>::NL;
    constexpr int LQ = LP < 5 ? LP : 5;
    double qf = 0.0;
#pragma unroll
    for (int l = LP - 1; l >= LP - LQ; --l) qf = fma((double)y[l], cs->wl[l], qf);
    const double q = rint(qf);
    const int32_t qi = (int32_t)q;
#pragma unroll
    for (int l = 0; l < LP; ++l) y[l] += __mul24(-qi, (int32_t)cs->M16[l]);  // 0 <= qi < 2^13, M16 < 2^16
    bool neg;
    if (fabs(qf - q) > cs->qthr) {
        propagate<NL>(y);
        // Bring Y into [-(floor(M/2)+1), floor(M/2)-1] exactly.
        int32_t d[NL];
#pragma unroll
        for (int l = 0; l < NL; ++l) d[l] = y[l] - (int32_t)(l < kLimbs ? cs->Mhalf16[l] : 0);
        propagate<NL>(d);
        if (d[NL - 1] >= 0) {
#pragma unroll
            for (int l = 0; l < NL; ++l) y[l] -= (int32_t)cs->M16[l];
            propagate<NL>(y);
        } else {
#pragma unroll
            for (int l = 0; l < NL; ++l) d[l] = y[l] + (int32_t)cs->Mhalf16[l] + (l == 0 ? 1 : 0);
            propagate<NL>(d);
            if (d[NL - 1] < 0) {
#pragma unroll
                for (int l = 0; l < NL; ++l) y[l] += (int32_t)cs->M16[l];
                propagate<NL>(y);
            }
        }
        neg = y[NL - 1] < 0;
        if (neg) {
#pragma unroll
            for (int l = 0; l < NL; ++l) y[l] = -y[l];
            propagate<NL>(y);
        }
    } else {
        neg = qf < q;
        if (neg) {
#pragma unroll
            for (int l = 0; l < NL; ++l) y[l] = -y[l];
        }
        propagate<NL>(y);
        if (y[NL - 1] < 0) {
            neg = !neg;
#pragma unroll
            for (int l = 0; l < NL; ++l) y[l] = -y[l];
            propagate<NL>(y);
        }
    }
    // to_double_256 (crt_math.hpp:73-80) on the 32-bit words w of |Y| (16-bit limbs in pairs): the
    // 64-bit word d_i = w[2i+1] 2^32 + w[2i] converted with one rounding, as one fma of its exact
    // halves, already scaled by its power of two 2^(64 i) (exact); then the reference's additions in
    // its order (0 + RN(d0) is RN(d0); the word d3 is zero).
    uint32_t w[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const uint32_t lo = 2 * k < NL ? (uint32_t)y[2 * k] : 0u;
        const uint32_t hi = 2 * k + 1 < NL ? (uint32_t)y[2 * k + 1] : 0u;
        w[k] = lo | (hi << 16);
    }
    const double d0 = fma((double)w[1], 0x1p32, (double)w[0]);
    const double d1 = fma((double)w[3], 0x1p96, (double)w[2] * 0x1p64);
    const double d2 = fma((double)w[5], 0x1p160, (double)w[4] * 0x1p128);
    double f = d0 + d1;
    f = f + d2;
    return (neg && f != 0.0) ? -f : f;  // Y = 0 is +0 as in the reference
}

template <int S>
__device__ __forceinline__ double crt_value(const uint32_t (&c)[S], const CrtSet* __restrict__ cs) {
    int32_t y[CrtShape<S>::NL];
    crt_limbs<S>(c, cs, y);
    return crt_finish<S>(y, cs);
}

// C[row * ldc + col] = alpha * x + beta * C[...] as fma(alpha, x, beta * C); beta == 0 does not
// read C, and alpha == 1 with beta == 0 stores x itself.
__device__ __forceinline__ void store_c(double* __restrict__ C, int64_t idx, double x, double alpha, double beta) {
    if (beta == 0.0) C[idx] = alpha == 1.0 ? x : alpha * x;
    else C[idx] = fma(alpha, x, beta * C[idx]);
}

// One 32 x 32 element block (R, Q) of C from blocked digit planes (layout of store_digits:
// block (R, Q) at byte (R * Np/32 + Q) * 1024 of each plane, MFMA 32x32 order inside).
// 256 threads; thread t reads dword t of the block in every plane (4 elements of one column).
// PLAIN: alpha = 1, beta = 0 (C = X), compiled without the output rule.
template <int S, bool PLAIN = true>
__device__ __forceinline__ void recon_digits_block(const uint8_t* __restrict__ dg, int64_t plane, int64_t Np,
                                                   const int32_t* __restrict__ shA, const int32_t* __restrict__ shB,
                                                   int64_t M, int64_t N, const CrtSet* __restrict__ cs,
                                                   double* __restrict__ Cout, int64_t ldc, double alpha, double beta,
                                                   int64_t R, int64_t Q) {
    const int t = threadIdx.x;
    const int lane = t >> 2, wd = t & 3;
    const int64_t col = Q * 32 + (lane & 31);
    const int64_t row0 = R * 32 + 8 * wd + 4 * (lane >> 5);
    if (col >= N || row0 >= M) return;
    const uint8_t* base = dg + ((R * (Np >> 5) + Q) << 10) + t * 4;
    constexpr int NJ = (S + 3) / 4;
    uint32_t w[4 * NJ];
#pragma unroll
    for (int s = 0; s < 4 * NJ; ++s) w[s] = s < S ? *reinterpret_cast<const uint32_t*>(base + s * plane) : 0u;
    const int32_t sb = shB[col];
    // Not unrolled: the CRT tables are then re-read through the scalar cache for each element
    // instead of being held in (spilled) scalar registers across all four.
#pragma unroll 1
    for (int b = 0; b < 4; ++b) {
        const int64_t row = row0 + b;
        if (row < M) {
            // Element b's digits, four moduli to a word: byte b of w[4j..4j+3] (v_perm_b32 byte
            // selectors: 0-3 pick from the second source, 4-7 from the first, 0x0c gives zero).
            const uint32_t pick = 0x0c0c0000u | ((4u + b) << 8) | (uint32_t)b;
            uint32_t P[NJ];
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const uint32_t lo = __builtin_amdgcn_perm(w[4 * j + 1], w[4 * j], pick);
                const uint32_t hi = __builtin_amdgcn_perm(w[4 * j + 3], w[4 * j + 2], pick);
                P[j] = lo | (hi << 16);
            }
            int32_t y[CrtShape<S>::NL];
            crt_limbs_packed<S>(P, cs, y);
            const double f = crt_finish<S>(y, cs);
            const double x = ldexp(f, -(shA[row] + sb));
            if (PLAIN) Cout[row * ldc + col] = x;
            else store_c(Cout, row * ldc + col, x, alpha, beta);
        }
    }
}

}  // namespace core
}  // namespace oza
