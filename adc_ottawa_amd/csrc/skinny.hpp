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

// X = V Z, WY = W Z, Res = X diag(theta) - WY and, when dpr, the DPR step
// T = Res / guard(theta - hii), for nb <= 32 columns (V, W: n x c, Z: c x nb).
// Per-block partial sums of Res^2 go to part (grown here); returns the number
// of blocks (0 = shape not supported, nothing launched).
int ritz_fused(int64_t n, int c, int nb, const double* V, const double* W, int64_t ld, const double* Z, int ldz,
               const double* theta, const double* hii, bool dpr, double* X, double* WY, double* Res, double* T,
               DevBuf<double>& part, hipStream_t st);

}  // namespace adc2g
