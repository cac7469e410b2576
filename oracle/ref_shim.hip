// ref_shim.hip — C entry points over the reference's own sources, compiled where they lie
// under the reference checkout (see oracle/Makefile, target `ref`).  TEST INFRASTRUCTURE
// ONLY: the built library (oracle/_ref/libozref.so) is the checker's checker; the product
// never loads it.
//
// Built from:
//   utils/matrix_gen.hpp, utils/matrix_reference.hpp, utils/matrix_compare.hpp
//   src/common/crt_tables.hpp, src/common/crt_math.hpp
//   src/pipeline/constants.hpp, step1_statistics.hpp, step2_slicing.hpp,
//   step4_reconstruction.hpp  (the reference's stage 1/2/4 GPU kernels)
// Not built: src/hip/pipeline.hip and src/core/* (they need rocBLAS for step 3; step 3 is
// exact integer arithmetic and is pinned by brute force instead).  The launch shapes below
// restate src/hip/pipeline.hip:39-42, 82-97, 139-141, 209-224, 252-259.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "matrix_gen.hpp"
#include "matrix_reference.hpp"
#include "matrix_compare.hpp"
#include "common/crt_tables.hpp"
#include "common/crt_math.hpp"
#include "pipeline/constants.hpp"
#include "pipeline/step1_statistics.hpp"
#include "pipeline/step2_slicing.hpp"
#include "pipeline/step4_reconstruction.hpp"

using namespace ozablas;

extern "C" {

// ---------------------------------------------------------------- host side
void ref_generate_matrix(int rows, int cols, double phi, int seed, double* out) {
    auto v = matrix_utils::generation::generate_ozaki_matrix<double>(rows, cols, phi, seed);
    std::memcpy(out, v.data(), sizeof(double) * v.size());
}

void ref_gemm_fp128(const double* A, const double* B, double* C, size_t M, size_t N, size_t K) {
    std::vector<double> a(A, A + M * K), b(B, B + K * N);
    auto c = matrix_utils::reference::compute_exact_gemm_fp128(a, b, M, N, K);
    std::memcpy(C, c.data(), sizeof(double) * M * N);
}

void ref_errors(const double* expected, const double* actual, size_t n, double out[4]) {
    std::vector<double> e(expected, expected + n), a(actual, actual + n);
    auto m = matrix_utils::compare::compute_errors(e, a);
    out[0] = m.max_abs_error;
    out[1] = m.max_relative_error;
    out[2] = m.frobenius_error;
    out[3] = m.relative_frobenius_error;
}

void ref_moduli(uint8_t out[20]) { std::memcpy(out, crt::moduli_all, sizeof(crt::moduli_all)); }

void ref_crt_tables(int S, uint64_t M[4], uint64_t M_half[4], uint64_t partial[20][4], uint64_t inv[20]) {
    std::memcpy(M, crt::M_prod_20[S - 1], 32);
    std::memcpy(M_half, crt::M_half_20[S - 1], 32);
    std::memcpy(partial, crt::partial_moduli_20[S - 1], sizeof(crt::partial_moduli_20[0]));
    std::memcpy(inv, crt::mod_inv_20[S - 1], sizeof(crt::mod_inv_20[0]));
}

double ref_to_double_256(const uint64_t a[4]) { return crt::to_double_256(crt::uint256_t(a)); }

int ref_scheme1_beta(int K) { return pipeline::calculate_scheme1_beta(K); }

int ref_scheme2_kparam(int S, int K) {
    crt::uint256_t M_prod(crt::M_prod_20[S - 1]);
    double M_double = crt::to_double_256(M_prod);
    return pipeline::calculate_scheme2_k_param(M_double, K);
}

// ---------------------------------------------------------------- device side
// All pointers are device pointers; every call runs on the null stream and synchronizes.
static dim3 block2d() { return dim3(16, 16); }
static dim3 grid2d(int rows, int cols) { return dim3((cols + 15) / 16, (rows + 15) / 16); }

int ref_gpu_init(void) {
    hipError_t e = hipSuccess;
    if (!e) e = hipMemcpyToSymbol(HIP_SYMBOL(OZA_c_moduli_all), crt::moduli_all, sizeof(crt::moduli_all));
    if (!e) e = hipMemcpyToSymbol(HIP_SYMBOL(OZA_c_M_prod_20), crt::M_prod_20, sizeof(crt::M_prod_20));
    if (!e) e = hipMemcpyToSymbol(HIP_SYMBOL(OZA_c_M_half_20), crt::M_half_20, sizeof(crt::M_half_20));
    if (!e) e = hipMemcpyToSymbol(HIP_SYMBOL(OZA_c_partial_moduli_20), crt::partial_moduli_20,
                                  sizeof(crt::partial_moduli_20));
    if (!e) e = hipMemcpyToSymbol(HIP_SYMBOL(OZA_c_mod_inv_20), crt::mod_inv_20, sizeof(crt::mod_inv_20));
    return (int)e;
}

static int finish() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    return (int)hipDeviceSynchronize();
}

int ref_gpu_row_shifts(const double* A, int rows, int cols, int offset, int32_t* shifts) {
    pipeline::compute_row_shifts_A<<<rows, 256>>>(A, rows, cols, offset, shifts);
    return finish();
}

int ref_gpu_col_shifts(const double* B, int rows, int cols, int offset, int32_t* shifts) {
    pipeline::compute_col_shifts_B<<<(cols + 255) / 256, 256>>>(B, rows, cols, offset, shifts);
    return finish();
}

int ref_gpu_slice_s1_A(const double* A, const int32_t* sh, int rows, int cols, int S, int beta, int8_t* out) {
    pipeline::slice_scheme1_A<<<grid2d(rows, cols), block2d()>>>(A, sh, rows, cols, S, beta, out);
    return finish();
}

int ref_gpu_slice_s1_B(const double* B, const int32_t* sh, int rows, int cols, int S, int beta, int8_t* out) {
    pipeline::slice_scheme1_B<<<grid2d(rows, cols), block2d()>>>(B, sh, rows, cols, S, beta, out);
    return finish();
}

int ref_gpu_slice_s2_A(const double* A, const int32_t* sh, int rows, int cols, int S, int8_t* out) {
    pipeline::slice_scheme2_A<<<grid2d(rows, cols), block2d()>>>(A, sh, rows, cols, S, out);
    return finish();
}

int ref_gpu_slice_s2_B(const double* B, const int32_t* sh, int rows, int cols, int S, int8_t* out) {
    pipeline::slice_scheme2_B<<<grid2d(rows, cols), block2d()>>>(B, sh, rows, cols, S, out);
    return finish();
}

int ref_gpu_recon_s1_group(const int32_t* T, const int32_t* shA, const int32_t* shB, int rows, int cols,
                           int g, int beta, double* C) {
    pipeline::reconstruct_scheme1_group<<<grid2d(rows, cols), block2d()>>>(T, shA, shB, rows, cols, g, beta, C);
    return finish();
}

int ref_gpu_recon_s2(const int32_t* Ctc, const int32_t* shA, const int32_t* shB, int rows, int cols, int S,
                     double* C) {
    if (S <= 7)
        pipeline::reconstruct_scheme2_leq7<<<grid2d(rows, cols), block2d()>>>(Ctc, shA, shB, rows, cols, S, C);
    else
        pipeline::reconstruct_scheme2_gt7<<<grid2d(rows, cols), block2d()>>>(Ctc, shA, shB, rows, cols, S, C);
    return finish();
}

}  // extern "C"
