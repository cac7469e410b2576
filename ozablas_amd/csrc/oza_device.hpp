// oza_device.hpp — device helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace oza {

// ---- exponent statistics -------------------------------------------------------------------
// The reference takes E = max(-1023, max_k e_k) where frexp(|a|) = f * 2^e, f in [1/2, 1),
// zeros and NaNs skipped (step1_statistics.hpp:33-42).  E depends only on the high word of
// |a|: for normal numbers e = biased_exponent - 1022; a subnormal gives -1022 when mantissa
// bit 51 is set and something <= -1023 otherwise; +-Inf has frexp exponent 0.  E is monotone
// in that high word, so a row maximum of the high word (after mapping NaN -> 0 and
// Inf -> 0x3FE00000, a word whose exponent is 0) gives the row's E.
__device__ __forceinline__ uint32_t exp_key(double x) {
    const uint64_t ab = (uint64_t)__double_as_longlong(x) & 0x7fffffffffffffffull;
    uint32_t hi = (uint32_t)(ab >> 32);
    if (ab > 0x7ff0000000000000ull) hi = 0;                // NaN
    else if (ab == 0x7ff0000000000000ull) hi = 0x3FE00000u;  // Inf: frexp exponent 0
    return hi;
}

__host__ __device__ __forceinline__ int exp_from_key(uint32_t key) {
    const int be = (int)(key >> 20);
    if (be == 0) return (key & 0x80000u) ? -1022 : -1023;
    return be - 1022;
}

// Wave-wide max of a 32-bit unsigned value: DPP within each 16-lane row (quad permutes,
// half-row and row mirrors), then two cross-row exchanges.  All 64 lanes must be active.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, false));  // row_mirror
    v = max(v, (uint32_t)__shfl_xor((int)v, 16));
    v = max(v, (uint32_t)__shfl_xor((int)v, 32));
    return v;
}

// ---- Scheme II residues ------------------------------------------------------------------------
// Symmetric residue of an integer-valued double t (|t| < 2^80) modulo an odd m < 256, in
// [-(m-1)/2, (m-1)/2].  t is split exactly as t = th * 2^32 + tl (tl in [0, 2^32)); then
// v = th * (2^32 mod m) + tl is exact (|v| < 2^57 ... in practice < 2^46) and congruent to t.
// rint(v / m) is the nearest integer to v/m: the exact quotient is at least 1/(2m) away from a
// half-integer and the product v * rcp(m) errs by far less, so r = v - m * rint(v * rcp) is the
// symmetric residue itself, computed exactly by one FMA.
struct SplitInt {
    double th, tl;
};
__device__ __forceinline__ SplitInt split32(double t) {
    const double th = floor(t * 0x1p-32);
    const double tl = fma(-th, 0x1p32, t);
    return {th, tl};
}
__device__ __forceinline__ int sym_residue(SplitInt t, double m, double rcp, double c32) {
    const double v = fma(t.th, c32, t.tl);
    const double q = rint(v * rcp);
    return (int)fma(-q, m, v);
}

// CRT digit of an INT32 product: the Euclidean residue c = v mod m in [0, m) -- the reference's
// `rem` (step4_reconstruction.hpp:92-94) -- for any int32 v and modulus 3 <= m < 512, with
// rcp = RN(1/m), in five FP64 operations.
//   x = fma(v, rcp, 2^31 + 2^-10).  |v/m| < 2^31/3, so 2^30 < x < 2^32 and x is rounded to a
//   multiple of at most 2^-21 (error <= 2^-22); |v rcp - v/m| <= |v|/m 2^-53 < 2^-23.  With
//   v = q m + c, x = (q + 2^31) + c/m + 2^-10 + e, |e| < 2^-21.
//   fract(x) = c/m + 2^-10 + e exactly (in (0, 1) since c <= m - 1 and m < 1024), and
//   fma(fract(x), m, 0.5) = c + 0.5 + m (2^-10 + e) lies strictly between c + 0.5 and c + 1 for
//   m < 512 (the product and sum are exact: fract(x) is a multiple of 2^-22 below 1), so the
//   truncating conversion returns c.
// The reference's multiplication by (M/m)^-1 mod m is folded into the CRT basis Q_s used by the
// reconstruction.
__device__ __forceinline__ int crt_digit(int v, double m, double rcp) {
    constexpr double kBias = 2147483648.0 + 0x1p-10;  // 2^31 + 2^-10, exact
    const double x = fma((double)v, rcp, kBias);
    return (int)fma(__builtin_amdgcn_fract(x), m, 0.5);
}

// ---- MFMA-fragment-tiled operand planes -------------------------------------------------------
// An operand plane (rows x Kp bytes, rows = M or N padded to 256, K contiguous) is stored as 1 KiB
// blocks of 32 rows x 32 K-bytes, block (rb, kb) at ((rb * Kp/32) + kb) * 1024.  Inside a block the
// 16-byte unit of lane l holds row l%32, K-bytes 16*(l/32) .. +15: exactly the register operand
// of lane l for v_mfma_i32_32x32x32_i8, so one wave moves a fragment with a single contiguous
// 1 KiB load (global -> LDS -> registers).  frag_off16 gives the byte offset of the 16-byte unit
// holding row `row`, K-bytes 16*c .. 16*c+15.
__host__ __device__ __forceinline__ int64_t frag_off16(int64_t row, int64_t c, int64_t Kp) {
    return (((row >> 5) * (Kp >> 5) + (c >> 1)) << 10) + (((row & 31) + ((c & 1) << 5)) << 4);
}

// ---- address spaces -------------------------------------------------------------------------------
// A 16-byte store through a global (address space 1) pointer.  Through a generic pointer the
// compiler emits a flat store, which may address LDS, and so first waits for every direct
// global->LDS load in flight (vmcnt(0)); a global store carries no such wait.
// NT: nontemporal (a wave instruction writes whole lines, and the data is not read back soon).
template <bool NT = false>
__device__ __forceinline__ void store_global16(void* p, uint32_t x, uint32_t y, uint32_t z, uint32_t w) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 v = {x, y, z, w};
    if (NT) __builtin_nontemporal_store(v, (__attribute__((address_space(1))) u32x4*)p);
    else *(__attribute__((address_space(1))) u32x4*)p = v;
}

// Read-only data in global memory seen as constant (address space 4): uniform reads become scalar
// loads, which are counted apart from the vector loads (lgkmcnt, not vmcnt), so waiting for them
// does not also wait for the direct loads in flight.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* as_constant(const T* p) {
    return (const __attribute__((address_space(4))) T*)p;
}

// Streaming loads of operands read once per pass (nontemporal: not kept in the caches).  Only for
// accesses whose wave covers whole cache lines in one instruction: a line read piecewise by
// several nontemporal loads is fetched from memory each time (slicing's 16-byte-per-lane row
// reads ran 25 % slower with them; the statistics' contiguous wave reads 11 % faster).
__device__ __forceinline__ double2 load_stream2(const double* p) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 v = __builtin_nontemporal_load(reinterpret_cast<const d2*>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ double load_stream(const double* p) { return __builtin_nontemporal_load(p); }

// 16-byte store of a plane unit, nontemporal with NT (every use writes a whole 1 KiB fragment block
// or half of one per wave instruction, so whole lines).  The grouped pipeline uses NT: its many
// small planes are not read again while still cached (grouped steps 2 and 3 -2.5 % and -5 %),
// whereas at 8192^3 it was neutral (slower slicing, slightly faster GEMM).
template <bool NT>
__device__ __forceinline__ void store_unit16(void* p, uint4 v) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 x = {v.x, v.y, v.z, v.w};
    if (NT) __builtin_nontemporal_store(x, reinterpret_cast<u32x4*>(p));
    else *reinterpret_cast<u32x4*>(p) = x;
}

// ---- byte packing ---------------------------------------------------------------------------------
// Four values already in [0, 255] into one word (byte 0 = a).
__device__ __forceinline__ uint32_t pack4u(int a, int b, int c, int d) {
    return (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)c << 16) | ((uint32_t)d << 24);
}

// Low bytes of four values (any sign) into one word (byte 0 = a): v_perm_b32 picks byte 0 of each
// source (selector 0-3: second source, 4-7: first source, 0x0c: zero).
__device__ __forceinline__ uint32_t pack4_lo(int a, int b, int c, int d) {
    const uint32_t ab = __builtin_amdgcn_perm((uint32_t)b, (uint32_t)a, 0x0c0c0400u);
    const uint32_t cd = __builtin_amdgcn_perm((uint32_t)d, (uint32_t)c, 0x04000c0cu);
    return ab | cd;
}

}  // namespace oza
