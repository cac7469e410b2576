// blas.cpp — ozablas::gemm over the C ABI's oza_gemm.
#include "ozablas/blas.hpp"

#include <stdexcept>

#include "ozablas/core/executor.hpp"
#include "status.hpp"

namespace ozablas {

void gemm(const Executor& exec, int scheme, int slices, Layout layout, Op transA, Op transB, size_t M, size_t N,
          size_t K, double alpha, const double* A, size_t lda, const double* B, size_t ldb, double beta, double* C,
          size_t ldc) {
    const auto* hip = dynamic_cast<const HipExecutor*>(&exec);
    if (!hip) throw std::runtime_error("OzaBLAS: gemm needs a HipExecutor.");
    detail::check(oza_gemm(hip->context(), scheme, slices, (int)layout, (int)transA, (int)transB, M, N, K, alpha, A,
                           lda, B, ldb, beta, C, ldc, nullptr, 0));
}

}  // namespace ozablas
