// Tall-skinny FP64 products of the Davidson subspace algebra (HBM-bound):
//   skinny_nn: C (M x N) = alpha A (M x K) B (K x N) + beta C, M = ndim rows,
//              K subspace columns, N <= 64 block columns.  One row of A per
//              lane, B staged in LDS (broadcast reads), accumulators in
//              registers.
//   skinny_tn: C (M x N) = alpha A^T B + beta C, A (K x M), B (K x N) with K =
//              ndim rows: FP64 MFMA over K-splits (see tn_mfma_kernel), then a
//              fixed-order sum of the split partials (deterministic).
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

// Ritz vectors, W y, residuals and (optionally) the DPR step in one pass over
// the subspace: X = V Z, WY = W Z, Res = X diag(theta) - WY,
// T = Res / guard(theta - hii); one row per lane, Z staged in LDS, per-block
// partial sums of Res^2 per column into part[block][col].
template <int NMAX>
__global__ void __launch_bounds__(NN_THREADS) ritz_kernel(int64_t n, int c, int nb, const double* __restrict__ V,
                                                          const double* __restrict__ W, int64_t ld,
                                                          const double* __restrict__ Z, int ldz,
                                                          const double* __restrict__ theta,
                                                          const double* __restrict__ hii, int dpr,
                                                          double* __restrict__ X, double* __restrict__ WY,
                                                          double* __restrict__ Res, double* __restrict__ T,
                                                          double* __restrict__ part) {
  __shared__ double sZ[NN_KC * NMAX];
  __shared__ double red[NN_THREADS / 64][NMAX];
  const int64_t i = blockIdx.x * (int64_t)NN_THREADS + threadIdx.x;
  const bool row = i < n;
  double ax[NMAX], aw[NMAX];
#pragma unroll
  for (int q = 0; q < NMAX; ++q) ax[q] = aw[q] = 0.0;
  for (int k0 = 0; k0 < c; k0 += NN_KC) {
    const int kc = min(NN_KC, c - k0);
    __syncthreads();
    for (int t = threadIdx.x; t < kc * NMAX; t += NN_THREADS) {
      const int k = t / NMAX, q = t % NMAX;
      sZ[t] = (q < nb) ? Z[(k0 + k) + (int64_t)ldz * q] : 0.0;
    }
    __syncthreads();
    if (row) {
      const double* v = V + i + ld * k0;
      const double* w = W + i + ld * k0;
#pragma unroll 2
      for (int k = 0; k < kc; ++k) {
        const double vv = v[ld * k], ww = w[ld * k];
        const double* z = sZ + k * NMAX;
#pragma unroll
        for (int q = 0; q < NMAX; ++q) {
          ax[q] += vv * z[q];
          aw[q] += ww * z[q];
        }
      }
    }
  }
  const double h = row ? hii[i] : 0.0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NMAX; ++q) {
    double sq = 0.0;
    if (q < nb) {
      const double th = theta[q];
      const double r = th * ax[q] - aw[q];
      if (row) {
        X[i + ld * q] = ax[q];
        WY[i + ld * q] = aw[q];
        Res[i + ld * q] = r;
        if (dpr) {
          double den = th - h;
          if (fabs(den) < 1e-10) den = den < 0.0 ? -1e-10 : 1e-10;
          T[i + ld * q] = r / den;
        }
        sq = r * r;
      }
    }
    for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
    if (lane == 0) red[wv][q] = sq;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nb; q += NN_THREADS) {
    double s = 0.0;
    for (int w = 0; w < NN_THREADS / 64; ++w) s += red[w][q];
    part[(int64_t)blockIdx.x * nb + q] = s;
  }
}

// C = A^T B with K (= ndim) rows: each workgroup takes 64 columns of A (one
// 16-column tile per wave) and all N <= 16*NT columns of B over one K-split,
// streams K in chunks of 64 rows through LDS (each column run is 512
// contiguous bytes) and accumulates with FP64 MFMA 16x16x4.  The splits of one
// column block share an XCD (blockIdx -> XCD is round-robin), so B is re-read
// from that XCD's L2.  Partials go to part[split][m][n]; tn_reduce sums them
// in a fixed order.
constexpr int TN_KC = 64;
constexpr int TN_MB = 64;

typedef double dbl4 __attribute__((ext_vector_type(4)));

template <int NT>
__global__ void __launch_bounds__(256) tn_mfma_kernel(int64_t K, int M, int N, const double* __restrict__ A,
                                                      int64_t lda, const double* __restrict__ B, int64_t ldb,
                                                      int mblocks, int splits, int64_t rows_per_split,
                                                      double* __restrict__ part) {
  __shared__ double sA[TN_MB][TN_KC + 1];
  __shared__ double sB[16 * NT][TN_KC + 1];
  const int L = blockIdx.x;
  const int xcd = L & 7, local = L >> 3;
  const int mb = local % mblocks;
  const int split = (local / mblocks) * 8 + xcd;
  if (split >= splits) return;  // whole workgroup, before any barrier
  const int m0 = mb * TN_MB;
  const int64_t r0 = (int64_t)split * rows_per_split;
  const int64_t r1 = min(K, r0 + rows_per_split);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double ra[16], rb[4 * NT];
  auto load = [&](int64_t k0) {
    const int64_t k = k0 + lane;
    const bool ok = k < r1;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const int m = m0 + wv + 4 * c;
      ra[c] = (ok && m < M) ? A[k + lda * m] : 0.0;
    }
#pragma unroll
    for (int c = 0; c < 4 * NT; ++c) {
      const int n = wv + 4 * c;
      rb[c] = (ok && n < N) ? B[k + ldb * n] : 0.0;
    }
  };
  dbl4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
  load(r0);
  const int row = lane & 15, kg = lane >> 4;
  for (int64_t k0 = r0; k0 < r1; k0 += TN_KC) {
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 16; ++c) sA[wv + 4 * c][lane] = ra[c];
#pragma unroll
    for (int c = 0; c < 4 * NT; ++c) sB[wv + 4 * c][lane] = rb[c];
    __syncthreads();
    if (k0 + TN_KC < r1) load(k0 + TN_KC);
    if (m0 + 16 * wv >= M) continue;  // this wave's column tile is empty
#pragma unroll
    for (int ks = 0; ks < TN_KC / 4; ++ks) {
      const double a = sA[16 * wv + row][4 * ks + kg];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const double b = sB[16 * t + row][4 * ks + kg];
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
      }
    }
  }
  // D[i][j] of a 16x16 tile sits at lane j + 16 (i % 4), register i / 4
  double* p = part + (int64_t)split * M * N;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + 16 * wv + 4 * r + kg;
      const int n = 16 * t + row;
      if (m < M && n < N) p[(int64_t)m * N + n] = acc[t][r];
    }
  }
}

// C[m + ldc n] = alpha sum_s part[s][m][n] + beta C: 64 outputs per
// workgroup, splits dealt round-robin to the 16 waves, fixed summation order.
__global__ void __launch_bounds__(1024) tn_reduce_kernel(int splits, int M, int N, const double* __restrict__ part,
                                                         double* __restrict__ C, int64_t ldc, double alpha,
                                                         double beta) {
  __shared__ double red[16][64];
  const int total = M * N;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int t = blockIdx.x * 64 + lane;
  double s0 = 0.0, s1 = 0.0;
  if (t < total) {
    int b = wv;
    for (; b + 16 < splits; b += 32) {
      s0 += part[(int64_t)b * total + t];
      s1 += part[(int64_t)(b + 16) * total + t];
    }
    if (b < splits) s0 += part[(int64_t)b * total + t];
  }
  red[wv][lane] = s0 + s1;
  __syncthreads();
  if (wv == 0 && t < total) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < 16; ++w) s += red[w][lane];
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

int ritz_fused(int64_t n, int c, int nb, const double* V, const double* W, int64_t ld, const double* Z, int ldz,
               const double* theta, const double* hii, bool dpr, double* X, double* WY, double* Res, double* T,
               DevBuf<double>& part, hipStream_t st) {
  if (nb <= 0 || nb > 32 || c <= 0) return 0;
  const int grid = (int)((n + NN_THREADS - 1) / NN_THREADS);
  part.alloc((size_t)grid * nb);
#define L(NM)                                                                                                       \
  hipLaunchKernelGGL((ritz_kernel<NM>), dim3(grid), dim3(NN_THREADS), 0, st, n, c, nb, V, W, ld, Z, ldz, theta, hii, \
                     dpr ? 1 : 0, X, WY, Res, T, part.get())
  if (nb <= 8) L(8);
  else if (nb <= 16) L(16);
  else if (nb <= 24) L(24);
  else L(32);
#undef L
  HIP_CHECK(hipGetLastError());
  return grid;
}

bool skinny_tn(int M, int N, int64_t K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
               int64_t ldc, double alpha, double beta, DevBuf<double>& ws, hipStream_t st) {
  if (M <= 0 || N <= 0) return true;
  if (N > 64 || K < 4 * TN_KC) return false;
  const int mblocks = (M + TN_MB - 1) / TN_MB;
  int splits = (1024 + mblocks - 1) / mblocks;
  splits = (int)std::min<int64_t>(splits, K / (4 * TN_KC));
  splits = std::max(splits, 1);
  int64_t rps = (K + splits - 1) / splits;
  rps = ((rps + TN_KC - 1) / TN_KC) * TN_KC;
  splits = (int)((K + rps - 1) / rps);
  ws.alloc((size_t)splits * M * N);
  const int grid = ((splits + 7) / 8) * 8 * mblocks;
  const int nt = (N + 15) / 16;
#define L(NT)                                                                                                  \
  hipLaunchKernelGGL((tn_mfma_kernel<NT>), dim3(grid), dim3(256), 0, st, K, M, N, A, lda, B, ldb, mblocks, splits, \
                     rps, ws.get())
  if (nt == 1) L(1);
  else if (nt == 2) L(2);
  else if (nt == 3) L(3);
  else L(4);
#undef L
  HIP_CHECK(hipGetLastError());
  const int tot = M * N;
  hipLaunchKernelGGL(tn_reduce_kernel, dim3((tot + 63) / 64), dim3(1024), 0, st, splits, M, N, ws.get(), C, ldc,
                     alpha, beta);
  HIP_CHECK(hipGetLastError());
  return true;
}

}  // namespace adc2g
