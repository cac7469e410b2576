// k_recon.hip — step 4: FP64 reconstruction.
//
// Scheme II (reference: src/pipeline/step4_reconstruction.hpp:59-184, crt_math.hpp:38-80).
// The reference sums c_s * (M/m_s) modulo M in u64 (S <= 7) or four u64 limbs (S > 7), maps
// acc >= floor(M/2) to -(M - acc), converts limb by limb (to_double_256) and scales by
// 2^-(shift_A[i] + shift_B[j]).  The value it converts is a fixed integer, so any exact method
// gives the same double.  Here: X = sum_s c_s * (M/m_s) accumulated without reduction in 16-bit
// limbs held in 32-bit registers (one v_mad_u32_u24 per limb and term), the quotient
// q = round(sum_s c_s / m_s) from an FP64 dot product, Y = X - q*M by signed limb arithmetic,
// an exact range check only when q was ambiguous, then the reference's limb-wise conversion.
//
// Scheme I group reconstruction (step4_reconstruction.hpp:21-49) is fused into the GEMM; the
// planar-product form here serves the reference-layout stage entry point.
#include "oza_device.hpp"
#include "oza_internal.hpp"
#include "k_igemm_cfg.hpp"

namespace oza {
namespace {

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

// Signed reconstruction of the CRT digits of one element, returned as the reference's double.
template <int S>
__device__ __forceinline__ double crt_value(const uint32_t (&c)[S], const CrtSet* __restrict__ cs) {
    constexpr int LP = CrtShape<S>::LP, NL = CrtShape<S>::NL;
    int32_t y[NL];
#pragma unroll
    for (int l = 0; l < NL; ++l) y[l] = 0;
    double qf = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        qf = fma((double)c[s], cs->rcp[s], qf);
#pragma unroll
        for (int l = 0; l < LP; ++l) y[l] = (int32_t)__umul24(c[s], cs->part16[s][l]) + y[l];
    }
    const double q = rint(qf);
    const int32_t qi = (int32_t)q;
#pragma unroll
    for (int l = 0; l < LP; ++l) y[l] -= qi * (int32_t)cs->M16[l];
    propagate<NL>(y);
    if (fabs(qf - q) > cs->qthr) {
        // Reference mapping: X mod M >= floor(M/2) is negative.  Bring Y into
        // [-(floor(M/2)+1), floor(M/2)-1] exactly.
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
    }
    const bool neg = y[NL - 1] < 0;
    if (neg) {
#pragma unroll
        for (int l = 0; l < NL; ++l) y[l] = -y[l];
        propagate<NL>(y);
    }
    uint64_t d64[3] = {0, 0, 0};
#pragma unroll
    for (int l = 0; l < NL && l < 12; ++l) d64[l / 4] |= (uint64_t)(uint32_t)y[l] << (16 * (l % 4));
    // to_double_256 (crt_math.hpp:73-80): four roundings in this order; the top limb is zero.
    double f = 0.0;
    f += (double)d64[0];
    f += ldexp((double)d64[1], 64);
    f += ldexp((double)d64[2], 128);
    return neg ? -f : f;
}

// ---------------------------------------------------------------- from GEMM digits
// Thread = one dword (4 accumulator bytes) of a 16-byte lane group of the digit layout.
template <class C, int S>
__global__ void __launch_bounds__(256)
    k_recon_digits(const uint8_t* __restrict__ dg, const int32_t* __restrict__ shA, const int32_t* __restrict__ shB,
                   int64_t M, int64_t N, int tiles_m, int tiles_n, const CrtSet* __restrict__ cs,
                   double* __restrict__ Cout) {
    constexpr int64_t TB = (int64_t)C::BM * C::BN;
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t tiles = (int64_t)tiles_m * tiles_n;
    if (idx >= tiles * (TB / 4)) return;
    const int64_t tile = idx / (TB / 4);
    const int64_t r = idx - tile * (TB / 4);
    const int g16 = (int)(r >> 2), wd = (int)(r & 3);
    const int tm = (int)(tile / tiles_n), tn = (int)(tile % tiles_n);
    const int lane = g16 & 63;
    int rest = g16 >> 6;
    const int fn = rest % C::FN;
    rest /= C::FN;
    const int fm = rest % C::FM;
    const int wave = rest / C::FM;
    const int wm = wave / C::NWN, wn = wave % C::NWN;

    const int64_t col = (int64_t)tn * C::BN + wn * C::WTN + fn * 32 + (lane & 31);
    const int64_t row0 = (int64_t)tm * C::BM + wm * C::WTM + fm * 32 + 8 * wd + 4 * (lane >> 5);
    if (col >= N || row0 >= M) return;

    uint32_t w[S];
    const int64_t plane = tiles * TB;
    const uint8_t* base = dg + tile * TB + (int64_t)g16 * 16 + wd * 4;
#pragma unroll
    for (int s = 0; s < S; ++s) w[s] = *reinterpret_cast<const uint32_t*>(base + s * plane);
    const int32_t sb = shB[col];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int64_t row = row0 + b;
        if (row < M) {
            uint32_t c[S];
#pragma unroll
            for (int s = 0; s < S; ++s) c[s] = (w[s] >> (8 * b)) & 0xffu;
            const double f = crt_value<S>(c, cs);
            Cout[row * N + col] = ldexp(f, -(shA[row] + sb));
        }
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
        c[s] = (uint32_t)crt_digit(Ct[s * MN + idx], (double)cs->m[s], (double)cs->inv[s], cs->rcp[s],
                                   (int)cs->m[s]);
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

// Digit layout -> planar [s][M][N] bytes (for inspection and parity tests).
template <class C>
__global__ void __launch_bounds__(256)
    k_digits_unpack(const uint8_t* __restrict__ dg, int64_t M, int64_t N, int tiles_m, int tiles_n, int S,
                    uint8_t* __restrict__ out) {
    constexpr int64_t TB = (int64_t)C::BM * C::BN;
    const int64_t tiles = (int64_t)tiles_m * tiles_n;
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= tiles * TB * S) return;
    const int64_t s = idx / (tiles * TB);
    const int64_t rr = idx - s * tiles * TB;
    const int64_t tile = rr / TB;
    const int off = (int)(rr - tile * TB);
    const int j = off & 15, g16 = off >> 4;
    const int lane = g16 & 63;
    int rest = g16 >> 6;
    const int fn = rest % C::FN;
    rest /= C::FN;
    const int fm = rest % C::FM;
    const int wave = rest / C::FM;
    const int wm = wave / C::NWN, wn = wave % C::NWN;
    const int tm = (int)(tile / tiles_n), tn = (int)(tile % tiles_n);
    const int64_t row = (int64_t)tm * C::BM + wm * C::WTM + fm * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
    const int64_t col = (int64_t)tn * C::BN + wn * C::WTN + fn * 32 + (lane & 31);
    if (row < M && col < N) out[(s * M + row) * N + col] = dg[idx];
}

template <int S>
hipError_t recon_digits_S(const Dims& d, const uint8_t* dg, const int32_t* shA, const int32_t* shB,
                          const CrtSet* cs, double* C, hipStream_t st) {
    using Cfg = IgemmCfgS2;
    const int tm = (int)((d.M + Cfg::BM - 1) / Cfg::BM), tn = (int)((d.N + Cfg::BN - 1) / Cfg::BN);
    const int64_t n = (int64_t)tm * tn * (Cfg::BM * Cfg::BN / 4);
    k_recon_digits<Cfg, S><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(dg, shA, shB, d.M, d.N, tm, tn, cs, C);
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
                                  const int32_t* shift_B, const CrtSet* crt_dev, double* C, hipStream_t st) {
    if (d.M == 0 || d.N == 0) return hipSuccess;
    OZA_S_CASES(recon_digits_S, d, digits, shift_A, shift_B, crt_dev, C, st)
}

hipError_t launch_recon_s2_int32(const Dims& d, const int32_t* C_tc, const int32_t* shift_A,
                                 const int32_t* shift_B, const CrtSet* crt_dev, double* C, hipStream_t st) {
    if (d.M == 0 || d.N == 0) return hipSuccess;
    OZA_S_CASES(recon_int32_S, d, C_tc, shift_A, shift_B, crt_dev, C, st)
}

hipError_t launch_digits_unpack(const Dims& d, const uint8_t* digits, uint8_t* out, hipStream_t st) {
    using Cfg = IgemmCfgS2;
    const int tm = (int)((d.M + Cfg::BM - 1) / Cfg::BM), tn = (int)((d.N + Cfg::BN - 1) / Cfg::BN);
    const int64_t n = (int64_t)tm * tn * Cfg::BM * Cfg::BN * d.S;
    if (n == 0) return hipSuccess;
    k_digits_unpack<Cfg><<<(unsigned)((n + 255) / 256), 256, 0, st>>>(digits, d.M, d.N, tm, tn, d.S, out);
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
