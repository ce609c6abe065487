// Tall-skinny FP64 products for the Davidson subspace algebra (see skinny.hip).
#pragma once

#include "common.hpp"

namespace adc2g {

// C = alpha A B + beta C; A is M x K (ld lda), B is K x N (ld ldb), N <= 64.
// Returns false when the shape is not supported (the caller falls back to
// the general GEMM).
bool skinny_nn(int64_t M, int N, int K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
               int64_t ldc, double alpha, double beta, hipStream_t st);

// C = alpha A^T B + beta C; A is K x M, B is K x N (K large); ws holds the
// per-block partials.
bool skinny_tn(int M, int N, int64_t K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
               int64_t ldc, double alpha, double beta, DevBuf<double>& ws, hipStream_t st);

}  // namespace adc2g
