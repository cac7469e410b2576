// slice_core.hpp — step-2 block bodies (256 threads) shared by the single-problem and grouped
// kernels.  Operands with K-contiguous rows are read in blocks of 32 rows x 128 K, operands with
// row-contiguous storage in blocks of 64 rows x 128 K; see k_slice.hip for the arithmetic and the
// output layout.
#pragma once

#include "oza_device.hpp"
#include "oza_internal.hpp"

namespace oza {
namespace core {

constexpr int kSliceThreads = 256;

struct ResidueConsts {
    double m[kMaxSlices];
    double rcp[kMaxSlices];
    double c32[kMaxSlices];  // 2^32 mod m
};

inline ResidueConsts make_residue_consts(int S) {
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
    return make_uint4(pack4_lo(r[0], r[1], r[2], r[3]), pack4_lo(r[4], r[5], r[6], r[7]),
                      pack4_lo(r[8], r[9], r[10], r[11]), pack4_lo(r[12], r[13], r[14], r[15]));
}

// ---------------------------------------------------------------- Scheme II, K-contiguous rows
// Operand element (row, k) at A[row * lda + k] (the reference's A; also op(B) = B^T stored N x K).
// Block = 32 rows x 128 K; thread = row (tid % 32), 16-byte K chunk (tid / 32).  The 64 lanes of a
// wave cover one whole 1 KiB fragment block.
template <bool WIDE, bool NT = false>
__device__ __forceinline__ void slice_s2_A_block(const double* __restrict__ A, int64_t lda,
                                                 const int32_t* __restrict__ shift, int64_t M, int64_t K, int64_t Mp,
                                                 int64_t Kp, int S, const ResidueConsts& rc, int8_t* __restrict__ out,
                                                 int64_t bx, int64_t by) {
    const int64_t row = bx * 32 + (threadIdx.x & 31);
    const int64_t c = by * 8 + (threadIdx.x >> 5);
    const int64_t k0 = c * 16;
    SplitInt t[16];
    if (row < M) {
        const int sh = shift[row];
        const double* p = A + row * lda + k0;
        double a[16];
        if (k0 + 16 <= K && ((reinterpret_cast<uintptr_t>(p) & 15) == 0)) {
#pragma unroll
            for (int j = 0; j < 16; j += 2) {
                const double2 v = *reinterpret_cast<const double2*>(p + j);
                a[j] = v.x;
                a[j + 1] = v.y;
            }
        } else {  // ragged or unaligned: every load unconditional (clamped index), padding zeroed after
            const double* row_p = A + row * lda;
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = row_p[min(k0 + j, K - 1)];
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = k0 + j < K ? a[j] : 0.0;
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
        store_unit16<NT>(o + s * plane, residues16<WIDE>(t, rc.m[s], rc.rcp[s], rc.c32[s]));
}

// ---------------------------------------------------------------- Scheme II, steps 1 and 2 fused
// Short K-contiguous rows (Kp <= kFuseMaxKp): a block holds R whole rows (R = fuse_rows(Kp)) in
// registers, so a row's exponent maximum and its residues come from one read of the row instead
// of one read for step 1 and another for step 2.  Same results as stats_A_block + key_to_shift +
// slice_s2_A_block: the maximum is order-free, and every element goes through the same residue
// arithmetic.  Unit u = (row u % R, 16-byte K chunk u / R) of the block; thread t takes units
// t, t + 256, ... (at most 4), all of row t % R, so 8 consecutive lanes write one whole 128-byte
// line of a fragment block.  Rows from M to roundup(M, 32) get zero residues (as in
// slice_s2_A_block); shift[row] is written for rows < M.
constexpr int64_t kFuseMaxKp = 2048;
constexpr int kFuseThreads = 256;

// Rows per fused block: the largest power of two R <= 128 with R * Kp/16 <= 1024 units (R >= 8
// for Kp <= kFuseMaxKp).
__host__ __device__ inline int fuse_rows(int64_t Kp) {
    const int64_t units = Kp / 16;
    int r = 128;
    while (r > 8 && (int64_t)r * units > 4 * kFuseThreads) r >>= 1;
    return r;
}

// 16 values of row `row` from K-byte 16c on, zero past K (loads as in slice_s2_A_block).
__device__ __forceinline__ void load_row16(const double* __restrict__ A, int64_t lda, int64_t row, int64_t k0,
                                           int64_t K, double (&a)[16]) {
    const double* p = A + row * lda + k0;
    if (k0 + 16 <= K && ((reinterpret_cast<uintptr_t>(p) & 15) == 0)) {
#pragma unroll
        for (int j = 0; j < 16; j += 2) {
            const double2 v = *reinterpret_cast<const double2*>(p + j);
            a[j] = v.x;
            a[j + 1] = v.y;
        }
    } else {
        const double* row_p = A + row * lda;
#pragma unroll
        for (int j = 0; j < 16; ++j) a[j] = row_p[min(k0 + j, K - 1)];
#pragma unroll
        for (int j = 0; j < 16; ++j) a[j] = k0 + j < K ? a[j] : 0.0;
    }
}

// s_key: kFuseThreads words of LDS, s_shift: 128 words.
template <bool WIDE>
__device__ __forceinline__ void fused_s2_A_block(const double* __restrict__ A, int64_t lda, int64_t M, int64_t K,
                                                 int64_t Mp, int64_t Kp, int32_t kparam, int S,
                                                 const ResidueConsts& rc, int32_t* __restrict__ shift,
                                                 int8_t* __restrict__ out, int64_t blk, uint32_t* s_key,
                                                 int32_t* s_shift) {
    const int R = fuse_rows(Kp);
    const int64_t U = Kp / 16;  // units per row
    const int t = threadIdx.x;
    const int rl = t % R;
    const int64_t row = blk * R + rl;
    const int64_t rows_out = (M + 31) / 32 * 32;
    const int64_t nunits = (int64_t)R * U;
    double a[4][16];
    uint32_t key = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t u = t + (int64_t)kFuseThreads * i;
        if (u < nunits && row < M) {
            load_row16(A, lda, row, (u / R) * 16, K, a[i]);
#pragma unroll
            for (int j = 0; j < 16; ++j) key = max(key, exp_key(a[i][j]));
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) a[i][j] = 0.0;
        }
    }
    s_key[t] = key;
    __syncthreads();
    if (t < R) {
        uint32_t k = 0;
        for (int j = t; j < kFuseThreads; j += R) k = max(k, s_key[j]);
        const int32_t sh = kparam - exp_from_key(k);  // key_to_shift
        s_shift[t] = sh;
        if (row < M) shift[row] = sh;
    }
    __syncthreads();
    const int sh = s_shift[rl];
    if (row >= rows_out) return;  // past the rows the planes hold for this operand
    const int64_t plane = Mp * Kp;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t u = t + (int64_t)kFuseThreads * i;
        if (u >= nunits) break;
        SplitInt v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = split32(rint(ldexp(a[i][j], sh)));
        int8_t* o = out + frag_off16(row, u / R, Kp);
        for (int s = 0; s < S; ++s)
            store_unit16<true>(o + s * plane, residues16<WIDE>(v, rc.m[s], rc.rcp[s], rc.c32[s]));
    }
}

// ---------------------------------------------------------------- Scheme II, row-contiguous columns
// Operand element (col, k) at B[k * ldb + col] (the reference's B; also op(A) = A^T stored K x M).
// Block = 64 columns x 4 groups of 32 K rows; thread = one column, 32 consecutive k.
template <bool WIDE, bool NT = false>
__device__ __forceinline__ void slice_s2_B_block(const double* __restrict__ B, int64_t ldb,
                                                 const int32_t* __restrict__ shift, int64_t K, int64_t N, int64_t Np,
                                                 int64_t Kp, int S, const ResidueConsts& rc, int8_t* __restrict__ out,
                                                 int64_t bx, int64_t by) {
    const int64_t col = bx * 64 + (threadIdx.x & 63);
    const int64_t kbase = by * 128 + (threadIdx.x >> 6) * 32;
    const int64_t plane = Np * Kp;
    const int sh = col < N ? shift[col] : 0;
    const int64_t cc = min(col, N - 1);  // clamped: every load unconditional, padding zeroed after
#pragma unroll 1
    for (int half = 0; half < 2; ++half) {
        const int64_t k0 = kbase + half * 16;
        double a[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) a[j] = B[min(k0 + j, K - 1) * ldb + cc];
        SplitInt t[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) t[j] = split32(rint(ldexp(col < N && k0 + j < K ? a[j] : 0.0, sh)));
        int8_t* o = out + frag_off16(col, k0 >> 4, Kp);
        for (int s = 0; s < S; ++s)
            store_unit16<NT>(o + s * plane, residues16<WIDE>(t, rc.m[s], rc.rcp[s], rc.c32[s]));
    }
}

// ---------------------------------------------------------------- Scheme I
// Slices of 16 values sharing one exponent e (row of A or column of B).
__device__ __forceinline__ void slice_s1_16(double (&a)[16], int e, int beta, int S, int8_t* o, int64_t plane) {
#pragma clang fp contract(off)  // every operation rounds where the reference's does
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
            make_uint4(pack4_lo(r[0], r[1], r[2], r[3]), pack4_lo(r[4], r[5], r[6], r[7]),
                       pack4_lo(r[8], r[9], r[10], r[11]), pack4_lo(r[12], r[13], r[14], r[15]));
    }
}

__device__ __forceinline__ void slice_s1_A_block(const double* __restrict__ A, int64_t lda,
                                                 const int32_t* __restrict__ neg_exp, int64_t M, int64_t K, int64_t Mp,
                                                 int64_t Kp, int S, int beta, int8_t* __restrict__ out, int64_t bx,
                                                 int64_t by) {
    const int64_t row = bx * 32 + (threadIdx.x & 31);
    const int64_t c = by * 8 + (threadIdx.x >> 5);
    const int64_t k0 = c * 16;
    int8_t* o = out + frag_off16(row, c, Kp);
    if (row < M) {
        const int e = -neg_exp[row];
        const double* p = A + row * lda + k0;
        double a[16];
        if (k0 + 16 <= K && ((reinterpret_cast<uintptr_t>(p) & 15) == 0)) {
#pragma unroll
            for (int j = 0; j < 16; j += 2) {
                const double2 v = *reinterpret_cast<const double2*>(p + j);
                a[j] = v.x;
                a[j + 1] = v.y;
            }
        } else {  // ragged or unaligned: every load unconditional (clamped index), padding zeroed after
            const double* row_p = A + row * lda;
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = row_p[min(k0 + j, K - 1)];
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = k0 + j < K ? a[j] : 0.0;
        }
        slice_s1_16(a, e, beta, S, o, Mp * Kp);
    } else {
        for (int s = 0; s < S; ++s) *reinterpret_cast<uint4*>(o + s * Mp * Kp) = make_uint4(0, 0, 0, 0);
    }
}

__device__ __forceinline__ void slice_s1_B_block(const double* __restrict__ B, int64_t ldb,
                                                 const int32_t* __restrict__ neg_exp, int64_t K, int64_t N, int64_t Np,
                                                 int64_t Kp, int S, int beta, int8_t* __restrict__ out, int64_t bx,
                                                 int64_t by) {
    const int64_t col = bx * 64 + (threadIdx.x & 63);
    const int64_t kbase = by * 128 + (threadIdx.x >> 6) * 32;
    const int64_t plane = Np * Kp;
#pragma unroll 1
    for (int half = 0; half < 2; ++half) {
        const int64_t k0 = kbase + half * 16;
        int8_t* o = out + frag_off16(col, k0 >> 4, Kp);
        if (col < N) {
            const int e = -neg_exp[col];
            double a[16];  // every load unconditional (clamped index), padding zeroed after
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = B[min(k0 + j, K - 1) * ldb + col];
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = k0 + j < K ? a[j] : 0.0;
            slice_s1_16(a, e, beta, S, o, plane);
        } else {
            for (int s = 0; s < S; ++s) *reinterpret_cast<uint4*>(o + s * plane) = make_uint4(0, 0, 0, 0);
        }
    }
}

// One operand of step 2: logical rows (padded to rows_p in the plane) of K values, element (r, k)
// at p[r * ld + k] (kcontig: 32-row blocks) or p[k * ld + r] (64-row blocks); nblk blocks of
// 128 K each, over the blocks that hold a logical row.  The plane's blocks past the last logical
// row are left unwritten: they only meet product rows (or columns) past M (or N), whose digits
// nothing reads, while the K padding inside the written blocks is zero.
struct SliceOp {
    const double* p;
    int64_t rows, rows_p, ld;
    int32_t kcontig;
    const int32_t* shift;
    int8_t* out;
    int64_t nblk;
};

inline SliceOp make_slice_op(const double* p, int64_t rows, int64_t rows_p, int64_t ld, bool kcontig,
                             int64_t Kp, const int32_t* shift, int8_t* out) {
    SliceOp o;
    o.p = p;
    o.rows = rows;
    o.rows_p = rows_p;
    o.ld = ld;
    o.kcontig = kcontig;
    o.shift = shift;
    o.out = out;
    const int64_t rb = kcontig ? 32 : 64;
    o.nblk = ((rows + rb - 1) / rb) * (Kp / 128);
    return o;
}

template <bool WIDE>
__device__ __forceinline__ void slice_s2_op_block(const SliceOp& o, int64_t K, int64_t Kp, int S,
                                                  const ResidueConsts& rc, int64_t blk) {
    const int64_t kb = Kp / 128;
    const int64_t bx = blk / kb, by = blk - bx * kb;
    if (o.kcontig) slice_s2_A_block<WIDE>(o.p, o.ld, o.shift, o.rows, K, o.rows_p, Kp, S, rc, o.out, bx, by);
    else slice_s2_B_block<WIDE>(o.p, o.ld, o.shift, K, o.rows, o.rows_p, Kp, S, rc, o.out, bx, by);
}

__device__ __forceinline__ void slice_s1_op_block(const SliceOp& o, int64_t K, int64_t Kp, int S, int beta,
                                                  int64_t blk) {
    const int64_t kb = Kp / 128;
    const int64_t bx = blk / kb, by = blk - bx * kb;
    if (o.kcontig) slice_s1_A_block(o.p, o.ld, o.shift, o.rows, K, o.rows_p, Kp, S, beta, o.out, bx, by);
    else slice_s1_B_block(o.p, o.ld, o.shift, K, o.rows, o.rows_p, Kp, S, beta, o.out, bx, by);
}

}  // namespace core
}  // namespace oza
