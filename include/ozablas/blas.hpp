// ozablas/blas.hpp — BLAS-style front end (gfx950 build): C = alpha op(A) op(B) + beta C with
// row- or column-major storage, transposes and leading dimensions, emulated with Ozaki Scheme I
// or II.  No counterpart in the reference, whose entry points take dense row-major A, B, C only
// (include/ozablas/ozablas.hpp:20,31).  With alpha = 1 and beta = 0 the result is bit-identical to
// ozaki_scheme{1,2}_gemm on dense row-major copies of op(A) and op(B); otherwise each entry is
// fma(alpha, X, beta * C) of the emulated product X, and beta == 0 does not read C.
#pragma once

#include <cstddef>

namespace ozablas {

class Executor;

enum class Layout { RowMajor = 0, ColMajor = 1 };
enum class Op { N = 0, T = 1 };

// scheme: 1 (slices) or 2 (moduli); slices: 1..20.  A, B, C are device pointers.  Uses the
// executor's own workspace (kept between calls); asynchronous on the executor's stream.
void gemm(const Executor& exec, int scheme, int slices, Layout layout, Op transA, Op transB, size_t M, size_t N,
          size_t K, double alpha, const double* A, size_t lda, const double* B, size_t ldb, double beta, double* C,
          size_t ldc);

}  // namespace ozablas
