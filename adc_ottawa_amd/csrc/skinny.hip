// Tall-skinny FP64 products of the Davidson subspace algebra (HBM-bound):
//   skinny_nn: C (M x N) = alpha A (M x K) B (K x N) + beta C, M = ndim rows,
//              K subspace columns, N <= 64 block columns.  One row of A per
//              lane, B staged in LDS (broadcast reads), accumulators in
//              registers.
//   skinny_tn: C (M x N) = alpha A^T B + beta C, A (K x M), B (K x N) with K =
//              ndim rows: each block reduces a K-chunk into a partial M x N
//              through LDS tiles of 16 rows, a second kernel sums the
//              partials in a fixed order (deterministic).
#include "skinny.hpp"

namespace adc2g {

namespace {

constexpr int NN_THREADS = 256;
constexpr int NN_KC = 128;  // K rows of B staged per pass

template <int NMAX>
__global__ void __launch_bounds__(NN_THREADS) skinny_nn_kernel(int64_t M, int N, int K, const double* __restrict__ A,
                                                               int64_t lda, const double* __restrict__ B,
                                                               int64_t ldb, double* __restrict__ C, int64_t ldc,
                                                               double alpha, double beta) {
  __shared__ double sB[NN_KC * NMAX];
  const int64_t i = blockIdx.x * (int64_t)NN_THREADS + threadIdx.x;
  const bool row = i < M;
  double acc[NMAX];
#pragma unroll
  for (int n = 0; n < NMAX; ++n) acc[n] = 0.0;
  for (int k0 = 0; k0 < K; k0 += NN_KC) {
    const int kc = min(NN_KC, K - k0);
    __syncthreads();
    for (int t = threadIdx.x; t < kc * NMAX; t += NN_THREADS) {
      const int k = t / NMAX, n = t % NMAX;
      sB[t] = (n < N) ? B[(int64_t)(k0 + k) + ldb * n] : 0.0;
    }
    __syncthreads();
    if (row) {
      const double* a = A + i + lda * k0;
#pragma unroll 4
      for (int k = 0; k < kc; ++k) {
        const double av = a[lda * k];
        const double* b = sB + k * NMAX;
#pragma unroll
        for (int n = 0; n < NMAX; ++n) acc[n] += av * b[n];
      }
    }
  }
  if (!row) return;
#pragma unroll
  for (int n = 0; n < NMAX; ++n) {
    if (n < N) {
      double* c = C + i + ldc * n;
      *c = alpha * acc[n] + (beta != 0.0 ? beta * *c : 0.0);
    }
  }
}

constexpr int TN_THREADS = 256;
constexpr int TN_RB = 16;  // rows per LDS tile

// partial[blk][m*N + n] = sum over the block's rows of A[k][m] B[k][n]
__global__ void __launch_bounds__(TN_THREADS) skinny_tn_kernel(int64_t K, int M, int N, const double* __restrict__ A,
                                                               int64_t lda, const double* __restrict__ B,
                                                               int64_t ldb, int64_t rows_per_block,
                                                               double* __restrict__ part) {
  extern __shared__ double sm[];
  double* sA = sm;                   // [TN_RB][M]
  double* sB = sm + TN_RB * M;       // [TN_RB][N]
  const int mb = (M + 3) / 4, nb = (N + 3) / 4;
  const int tiles = mb * nb;
  // each lane owns up to 4 tiles of 4x4 outputs
  double acc[4][16];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[q][e] = 0.0;
  const int64_t r0 = blockIdx.x * rows_per_block;
  const int64_t r1 = min(K, r0 + rows_per_block);
  for (int64_t k0 = r0; k0 < r1; k0 += TN_RB) {
    const int kr = (int)min((int64_t)TN_RB, r1 - k0);
    __syncthreads();
    for (int t = threadIdx.x; t < TN_RB * M; t += TN_THREADS) {
      const int k = t % TN_RB, m = t / TN_RB;
      sA[k * M + m] = (k < kr) ? A[(k0 + k) + lda * m] : 0.0;
    }
    for (int t = threadIdx.x; t < TN_RB * N; t += TN_THREADS) {
      const int k = t % TN_RB, n = t / TN_RB;
      sB[k * N + n] = (k < kr) ? B[(k0 + k) + ldb * n] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int tile = threadIdx.x + q * TN_THREADS;
      if (tile < tiles) {
      const int m0 = (tile / nb) * 4, n0 = (tile % nb) * 4;
#pragma unroll 4
      for (int k = 0; k < TN_RB; ++k) {
        double a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          a[u] = (m0 + u < M) ? sA[k * M + m0 + u] : 0.0;
          b[u] = (n0 + u < N) ? sB[k * N + n0 + u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int w = 0; w < 4; ++w) acc[q][u * 4 + w] += a[u] * b[w];
      }
      }
    }
  }
  double* p = part + (int64_t)blockIdx.x * M * N;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int tile = threadIdx.x + q * TN_THREADS;
    if (tile < tiles) {
      const int m0 = (tile / nb) * 4, n0 = (tile % nb) * 4;
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int w = 0; w < 4; ++w)
          if (m0 + u < M && n0 + w < N) p[(m0 + u) * N + n0 + w] = acc[q][u * 4 + w];
    }
  }
}

__global__ void skinny_tn_reduce(int nblk, int M, int N, const double* __restrict__ part, double* __restrict__ C,
                                 int64_t ldc, double alpha, double beta) {
  const int total = M * N;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * total + t];
    const int m = t / N, n = t % N;
    double* c = C + m + ldc * n;
    *c = alpha * s + (beta != 0.0 ? beta * *c : 0.0);
  }
}

}  // namespace

bool skinny_nn(int64_t M, int N, int K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
               int64_t ldc, double alpha, double beta, hipStream_t st) {
  if (N <= 0 || K <= 0 || M <= 0) return true;
  if (N > 64) return false;
  const int grid = (int)((M + NN_THREADS - 1) / NN_THREADS);
#define L(NM)                                                                                                      \
  hipLaunchKernelGGL((skinny_nn_kernel<NM>), dim3(grid), dim3(NN_THREADS), 0, st, M, N, K, A, lda, B, ldb, C, ldc, \
                     alpha, beta)
  if (N <= 8) L(8);
  else if (N <= 16) L(16);
  else if (N <= 24) L(24);
  else if (N <= 32) L(32);
  else if (N <= 48) L(48);
  else L(64);
#undef L
  HIP_CHECK(hipGetLastError());
  return true;
}

bool skinny_tn(int M, int N, int64_t K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
               int64_t ldc, double alpha, double beta, DevBuf<double>& ws, hipStream_t st) {
  if (M <= 0 || N <= 0) return true;
  const int tiles = ((M + 3) / 4) * ((N + 3) / 4);
  if (tiles > 4 * TN_THREADS) return false;
  const size_t lds = sizeof(double) * TN_RB * (size_t)(M + N);
  if (lds > 64 * 1024) return false;
  int nblk = 1024;
  int64_t rpb = (K + nblk - 1) / nblk;
  rpb = ((rpb + TN_RB - 1) / TN_RB) * TN_RB;
  nblk = (int)((K + rpb - 1) / rpb);
  ws.alloc((size_t)nblk * M * N);
  hipLaunchKernelGGL(skinny_tn_kernel, dim3(nblk), dim3(TN_THREADS), lds, st, K, M, N, A, lda, B, ldb, rpb, ws.get());
  HIP_CHECK(hipGetLastError());
  const int tot = M * N;
  hipLaunchKernelGGL(skinny_tn_reduce, dim3((tot + 255) / 256), dim3(256), 0, st, nblk, M, N, ws.get(), C, ldc, alpha,
                     beta);
  HIP_CHECK(hipGetLastError());
  return true;
}

}  // namespace adc2g
