/*
 * ozoracle.h — CPU restatement of the OzaBLAS Ozaki Scheme I / II pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP product is
 * compared against; it is never linked into, loaded by, or called from the
 * product path (ozablas_amd/).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * Every function cites the reference file:line it restates.  Matrices are
 * dense row-major; indices are 64-bit (the reference's 32-bit `int` index
 * overflow, step2_slicing.hpp:54, is deliberately not inherited).
 *
 * Pinning (what checks each function, so that a plausible slip fails a test):
 *   orc_generate_matrix    SHA-256 fixture + the reference generator compiled
 *                          from utils/matrix_gen.hpp (oracle/_ref)
 *   orc_crt_tables         closed forms (products, floor(M/2), M/m_i, inverse
 *                          by m_i-brute force) + reference tables (oracle/_ref)
 *   orc_scheme1_beta / orc_scheme2_kparam   closed forms + oracle/_ref
 *   orc_row/col_shifts     brute force via the frexp definition + the
 *                          reference's GPU kernels (oracle/_ref, GPU tests)
 *   orc_slice_s1_*         invariant sum(slice_s * 2^(E+1-beta-beta*s)) ==
 *                          truncated input, |slice| <= 2^(beta-1) + reference
 *                          GPU kernels
 *   orc_slice_s2_*         residue == rint(x) mod m (exact big-int check),
 *                          |r| <= (m-1)/2 + reference GPU kernels
 *   orc_igemm / groups     brute-force int64 GEMM on tiny inputs, linearity
 *   orc_recon_s2           CRT result == exact integer product when in range;
 *                          congruence X == C_s (mod m_s); reference GPU kernels
 *   orc_recon_s1           exact-power-of-two closed form + reference GPU kernels
 *   orc_gemm_fp128         closed-form products of small integers, compared
 *                          with oracle/_ref's compute_exact_gemm_fp128
 */
#ifndef OZORACLE_H
#define OZORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_SLICES 20

/* utils/matrix_gen.hpp:15-34 — a_ij = ((1 - U) - 0.5) * exp(phi * Z), mt19937(seed),
 * U from uniform_real_distribution<double>(0,1), Z from normal_distribution<double>. */
void orc_generate_matrix(int64_t rows, int64_t cols, double phi, int seed, double *out);

/* src/common/crt_tables.hpp:13-16 (the moduli list) and :19-112 (derived tables),
 * recomputed here from closed forms.  M, M_half, partial[i] are 4 x u64 limbs. */
void orc_moduli(uint8_t out[ORC_MAX_SLICES]);
int orc_crt_tables(int S, uint64_t M[4], uint64_t M_half[4],
                   uint64_t partial[ORC_MAX_SLICES][4], uint64_t inv[ORC_MAX_SLICES]);

/* src/common/crt_math.hpp:73-80 */
double orc_to_double_256(const uint64_t a[4]);

/* src/pipeline/step1_statistics.hpp:12-15 */
int orc_scheme1_beta(int64_t K);
/* src/pipeline/step1_statistics.hpp:17-20 with M_double from src/hip/pipeline.hip:205-207 */
int orc_scheme2_kparam(int S, int64_t K);

/* src/pipeline/step1_statistics.hpp:23-76 (rows of A) and :78-99 (columns of B) */
void orc_row_shifts(const double *A, int64_t rows, int64_t cols, int32_t offset, int32_t *shifts);
void orc_col_shifts(const double *B, int64_t rows, int64_t cols, int32_t offset, int32_t *shifts);

/* src/pipeline/step2_slicing.hpp:44-71 (row exponents, A) and :73-100 (column exponents, B).
 * out is planar [s][rows][cols]. */
void orc_slice_s1_rows(const double *A, const int32_t *neg_exp, int64_t rows, int64_t cols,
                       int S, int beta, int8_t *out);
void orc_slice_s1_cols(const double *B, const int32_t *neg_exp, int64_t rows, int64_t cols,
                       int S, int beta, int8_t *out);

/* src/pipeline/step2_slicing.hpp:29-38 */
int orc_symmetric_mod(int64_t a, int m);
/* src/pipeline/step2_slicing.hpp:110-157 (row shifts, A) and :163-206 (column shifts, B). */
void orc_slice_s2_rows(const double *A, const int32_t *shifts, int64_t rows, int64_t cols,
                       int S, int8_t *out);
void orc_slice_s2_cols(const double *B, const int32_t *shifts, int64_t rows, int64_t cols,
                       int S, int8_t *out);

/* Step 3.  src/hip/pipeline.hip:237-245 (Scheme II, C_s = A_s B_s) and :110-135
 * (Scheme I, T_g = sum_{s+t=g} A_{s-1} B_{t-1}).  Exact: accumulates in int64 and
 * returns the number of entries that do not fit int32 (0 for valid inputs). */
int64_t orc_igemm(const int8_t *A, const int8_t *B, int64_t M, int64_t N, int64_t K, int32_t *C);
int64_t orc_s2_products(const int8_t *A_sl, const int8_t *B_sl, int64_t M, int64_t N, int64_t K,
                        int S, int32_t *C_tc);
int64_t orc_s1_group_products(const int8_t *A_sl, const int8_t *B_sl, int64_t M, int64_t N,
                              int64_t K, int S, int32_t *T);

/* Step 4.  src/pipeline/step4_reconstruction.hpp:21-49 applied for g = 2..S+1 on a
 * zeroed C (src/hip/pipeline.hip:80,110-144). */
void orc_recon_s1(const int32_t *T, const int32_t *neg_exp_A, const int32_t *neg_exp_B,
                  int64_t M, int64_t N, int S, int beta, double *C);
/* src/pipeline/step4_reconstruction.hpp:59-110 (S<=7) and :116-184 (S>7). */
void orc_recon_s2(const int32_t *C_tc, const int32_t *shift_A, const int32_t *shift_B,
                  int64_t M, int64_t N, int S, double *C);
/* The Euclidean residue r_s = C_s mod m_s in [0, m_s), the reference's `rem`
 * (step4_reconstruction.hpp:92-94), one plane per modulus: the digit format of the GEMM epilogue. */
void orc_crt_digits(const int32_t *C_tc, int64_t MN, int S, uint8_t *digits);

/* Whole pipelines in the reference's stage order (src/hip/pipeline.hip:48-176, 182-275). */
void orc_s1_gemm(const double *A, const double *B, double *C, int64_t M, int64_t N, int64_t K, int S);
void orc_s2_gemm(const double *A, const double *B, double *C, int64_t M, int64_t N, int64_t K, int S);

/* utils/matrix_reference.hpp:22-51: FP128 dot products, one rounding to FP64. */
void orc_gemm_fp128(const double *A, const double *B, double *C, int64_t M, int64_t N, int64_t K);
/* The same exact-product sum in double-double arithmetic (TwoProd/TwoSum, Ogita-Rump-Oishi
 * Dot2), returned as hi + lo rounded to FP64. */
void orc_gemm_dd(const double *A, const double *B, double *C, int64_t M, int64_t N, int64_t K);
/* |A||B| in FP64 (for componentwise error bounds). */
void orc_gemm_abs(const double *A, const double *B, double *C, int64_t M, int64_t N, int64_t K);

/* utils/matrix_compare.hpp:24-65: out = {max_abs, max_rel, frobenius, relative_frobenius}. */
void orc_errors(const double *expected, const double *actual, int64_t n, double out[4]);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
