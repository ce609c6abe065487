// Small dense FP64 linear algebra on one workgroup, and the streaming vector
// kernels of the Davidson solver.
//
// syev: Householder reduction of the full symmetric matrix to tridiagonal
// form (one reflection per column, trailing block updated by the rank-2 form
// A <- A - v q^T - q v^T with q = 2(p - (v.p) v), p = A v), implicit QL with a
// Wilkinson shift on the tridiagonal (control on one lane, each sweep's
// Givens rotations then applied to the eigenvector rows by all lanes), back
// transformation with the stored reflectors, and an ascending sort.
#include <cfloat>

#include "linalg.hpp"

namespace adc2g {

namespace {

constexpr int SYEV_THREADS = 512;
constexpr int SYEV_NMAX = 512;

__device__ double block_sum(double x, double* sh) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = x;
  __syncthreads();
  double t = 0.0;
  const int nw = blockDim.x >> 6;
  for (int i = 0; i < nw; ++i) t += sh[i];
  return t;
}

struct SyevShared {
  double d[SYEV_NMAX], e[SYEV_NMAX], rc[SYEV_NMAX], rs[SYEV_NMAX];
  double red[16];
  int perm[SYEV_NMAX];
  int lo, hi, go, info;
};

// Eigen-decomposition of the symmetric n x n matrix A (destroyed); eigenvalues
// ascending into w, eigenvectors into the columns of Z; Zw is n*n scratch.
__device__ void syev_dev(int n, double* A, int lda, double* w, double* Z, int ldz, double* Zw, int* info,
                         SyevShared& S) {
  const int tid = threadIdx.x, nt = blockDim.x;
  // ---- Householder tridiagonalisation (reflector k stored in column k below row k)
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    double part = 0.0;
    for (int i = k + 2 + tid; i < n; i += nt) {
      const double x = A[i + (int64_t)k * lda];
      part += x * x;
    }
    const double x0 = A[k + 1 + (int64_t)k * lda];
    const double tail = block_sum(part, S.red);
    if (tail == 0.0) {
      if (tid == 0) {
        S.e[k] = x0;
        A[k + 1 + (int64_t)k * lda] = 0.0;  // no reflection: zero reflector
      }
      __syncthreads();
      continue;
    }
    const double sig = tail + x0 * x0;
    const double alpha = -copysign(sqrt(sig), x0);
    const double v0 = x0 - alpha;
    const double inv = 1.0 / sqrt(tail + v0 * v0);
    __syncthreads();
    for (int i = k + 1 + tid; i < n; i += nt) {
      const double x = (i == k + 1) ? v0 : A[i + (int64_t)k * lda];
      A[i + (int64_t)k * lda] = x * inv;
    }
    if (tid == 0) S.e[k] = alpha;
    __syncthreads();
    const double* v = A + (int64_t)k * lda;  // v[i] for i in k+1..n-1
    // p = A22 v, one row per lane (coalesced over the column-major storage)
    for (int i = k + 1 + tid; i < n; i += nt) {
      double s = 0.0;
      for (int j = k + 1; j < n; ++j) s += A[i + (int64_t)j * lda] * v[j];
      Zw[i] = s;
    }
    __syncthreads();
    part = 0.0;
    for (int j = k + 1 + tid; j < n; j += nt) part += v[j] * Zw[j];
    const double Kv = block_sum(part, S.red);
    for (int j = k + 1 + tid; j < n; j += nt) Zw[n + j] = 2.0 * (Zw[j] - Kv * v[j]);  // q
    __syncthreads();
    const double* q = Zw + n;
    const int64_t mm = (int64_t)m * m;
    for (int64_t idx = tid; idx < mm; idx += nt) {
      const int i = k + 1 + (int)(idx % m), j = k + 1 + (int)(idx / m);
      A[i + (int64_t)j * lda] -= v[i] * q[j] + q[i] * v[j];
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += nt) S.d[i] = A[i + (int64_t)i * lda];
  if (tid == 0) {
    if (n >= 2) S.e[n - 2] = A[n - 1 + (int64_t)(n - 2) * lda];
    S.e[n - 1] = 0.0;
    S.info = 0;
  }
  for (int64_t idx = tid; idx < (int64_t)n * n; idx += nt) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    Zw[idx] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  // ---- implicit QL on the tridiagonal (d, e), rotations accumulated into Zw
  int ql_info = 0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    while (true) {
      if (tid == 0) {
        double* d = S.d;
        double* e = S.e;
        int m = l;
        for (; m < n - 1; ++m) {
          const double dd = fabs(d[m]) + fabs(d[m + 1]);
          if (fabs(e[m]) <= DBL_EPSILON * dd) break;
        }
        if (m == l) {
          S.go = 0;
        } else if (iter++ >= 60) {
          S.go = 0;
          S.info = 1;
        } else {
          double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
          double r = hypot(g, 1.0);
          g = d[m] - d[l] + e[l] / (g + copysign(r, g));
          double s = 1.0, c = 1.0, p = 0.0;
          int i = m - 1;
          bool under = false;
          for (; i >= l; --i) {
            const double f = s * e[i], b = c * e[i];
            r = hypot(f, g);
            e[i + 1] = r;
            if (r == 0.0) {
              d[i + 1] -= p;
              e[m] = 0.0;
              under = true;
              break;
            }
            s = f / r;
            c = g / r;
            g = d[i + 1] - p;
            r = (d[i] - g) * s + 2.0 * c * b;
            p = s * r;
            d[i + 1] = g + p;
            g = c * r - b;
            S.rc[i] = c;
            S.rs[i] = s;
          }
          S.lo = under ? i + 1 : l;
          S.hi = m - 1;
          if (!under) {
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
          }
          S.go = 1;
        }
      }
      __syncthreads();
      const int go = S.go, lo = S.lo, hi = S.hi;
      ql_info = S.info;
      __syncthreads();  // everyone has read the control words before lane 0 rewrites them
      if (!go) break;
      for (int r = tid; r < n; r += nt) {
        // the element of column i+1 is carried in a register down the sweep,
        // so the loads of column i do not wait on the previous store
        double carry = Zw[r + (int64_t)(hi + 1) * n];
        for (int i = hi; i >= lo; --i) {
          const double c = S.rc[i], s = S.rs[i];
          const double zi = Zw[r + (int64_t)i * n];
          Zw[r + (int64_t)(i + 1) * n] = s * zi + c * carry;
          carry = c * zi - s * carry;
        }
        Zw[r + (int64_t)lo * n] = carry;
      }
      __syncthreads();
    }
    if (ql_info) break;
  }
  __syncthreads();
  if (ql_info) {
    if (tid == 0) info[0] = 1;
    __syncthreads();
    return;
  }
  // ---- back transformation: Zw <- H_0 ... H_{n-3} Zw (one column per wave,
  // rows split over the lanes)
  {
    const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
    for (int j = wid; j < n; j += nw) {
      double* z = Zw + (int64_t)j * n;
      for (int k = n - 3; k >= 0; --k) {
        const double* v = A + (int64_t)k * lda;
        double t = 0.0;
        for (int i = k + 1 + lane; i < n; i += 64) t += v[i] * z[i];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        t *= 2.0;
        if (t != 0.0)
          for (int i = k + 1 + lane; i < n; i += 64) z[i] -= t * v[i];
      }
    }
  }
  __syncthreads();
  // ---- ascending order
  if (tid == 0) {
    for (int i = 0; i < n; ++i) S.perm[i] = i;
    for (int i = 0; i < n; ++i) {
      int best = i;
      for (int j = i + 1; j < n; ++j)
        if (S.d[S.perm[j]] < S.d[S.perm[best]]) best = j;
      const int t = S.perm[i];
      S.perm[i] = S.perm[best];
      S.perm[best] = t;
    }
    info[0] = 0;
  }
  __syncthreads();
  for (int j = tid; j < n; j += nt) w[j] = S.d[S.perm[j]];
  for (int64_t idx = tid; idx < (int64_t)n * n; idx += nt) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    Z[i + (int64_t)j * ldz] = Zw[i + (int64_t)S.perm[j] * n];
  }
  __syncthreads();
}

__global__ void __launch_bounds__(SYEV_THREADS) syev_kernel(int n, double* A, int lda, double* w, double* Z,
                                                             int ldz, double* work, int* info) {
  __shared__ SyevShared S;
  syev_dev(n, A, lda, w, Z, ldz, work, info, S);
}

// M = D U_kept Lambda^{-1/2} for the scaled Gram matrix D G D (D = diag(G)^{-1/2}).
// work: r*r (copy) + r*r (U) + r (lambda) + r*r (syev scratch).
__global__ void __launch_bounds__(SYEV_THREADS) orth_coeffs_kernel(int r, const double* G, int ldg, double* M,
                                                                    int ldm, double* work, int* kept, double tol) {
  __shared__ SyevShared S;
  __shared__ double dinv[SYEV_NMAX];
  __shared__ int sel[SYEV_NMAX];
  __shared__ int nsel;
  const int tid = threadIdx.x, nt = blockDim.x;
  double* Gc = work;
  double* U = work + (int64_t)r * r;
  double* lam = U + (int64_t)r * r;
  double* Zw = lam + r;
  for (int i = tid; i < r; i += nt) {
    const double g = G[i + (int64_t)i * ldg];
    dinv[i] = (g > 0.0) ? 1.0 / sqrt(g) : 0.0;
  }
  __syncthreads();
  for (int64_t idx = tid; idx < (int64_t)r * r; idx += nt) {
    const int i = (int)(idx % r), j = (int)(idx / r);
    Gc[idx] = dinv[i] * G[i + (int64_t)j * ldg] * dinv[j];
  }
  __syncthreads();
  __shared__ int sinfo;
  syev_dev(r, Gc, r, lam, U, r, Zw, &sinfo, S);
  __syncthreads();
  if (tid == 0) {
    int c = 0;
    if (sinfo == 0) {
      const double lmax = lam[r - 1];
      // largest eigenvalues first, so the best-conditioned directions come first
      for (int j = r - 1; j >= 0; --j)
        if (lam[j] > tol * lmax && lam[j] > 0.0) sel[c++] = j;
    }
    nsel = c;
    kept[0] = (sinfo == 0) ? c : -1;
  }
  __syncthreads();
  const int c = nsel;
  for (int64_t idx = tid; idx < (int64_t)r * c; idx += nt) {
    const int i = (int)(idx % r), t = (int)(idx / r);
    const int j = sel[t];
    M[i + (int64_t)t * ldm] = dinv[i] * U[i + (int64_t)j * r] / sqrt(lam[j]);
  }
}

__global__ void residual_kernel(int64_t n, int nc, const double* __restrict__ theta, const double* __restrict__ X,
                                const double* __restrict__ WY, double* __restrict__ Res, int64_t ld,
                                double* __restrict__ part) {
  __shared__ double red[16];
  for (int c = 0; c < nc; ++c) {
    const double th = theta[c];
    const double* x = X + ld * c;
    const double* wy = WY + ld * c;
    double* r = Res + ld * c;
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
      const double v = th * x[i] - wy[i];
      r[i] = v;
      s += v * v;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * nc + c] = s;
  }
}

__global__ void colsum_kernel(int nblk, int nc, const double* __restrict__ part, double* __restrict__ out) {
  for (int c = threadIdx.x; c < nc; c += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * nc + c];
    out[c] = s;
  }
}

__device__ __forceinline__ double guard_den(double den) {
  const double DEN_MIN = 1e-10;
  if (fabs(den) < DEN_MIN) return den < 0.0 ? -DEN_MIN : DEN_MIN;
  return den;
}

__global__ void dpr_kernel(int64_t n, int nc, const double* __restrict__ theta, const double* __restrict__ hii,
                           const double* __restrict__ Res, double* __restrict__ T, int64_t ld) {
  const int64_t total = n * nc;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = idx / n, i = idx - c * n;
    T[i + ld * c] = Res[i + ld * c] / guard_den(theta[c] - hii[i]);
  }
}

__global__ void olsen_dots_kernel(int64_t n, int nc, const double* __restrict__ theta, const double* __restrict__ hii,
                                  const double* __restrict__ X, const double* __restrict__ Res, int64_t ld,
                                  double* __restrict__ part) {
  __shared__ double red[16];
  for (int c = 0; c < nc; ++c) {
    const double th = theta[c];
    double sxz = 0.0, sxy = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
      const double cd = 1.0 / guard_den(hii[i] - th);
      const double x = X[i + ld * c];
      sxz += x * Res[i + ld * c] * cd;
      sxy += x * x * cd;
    }
    sxz = block_sum(sxz, red);
    sxy = block_sum(sxy, red);
    if (threadIdx.x == 0) {
      part[(int64_t)blockIdx.x * 2 * nc + 2 * c] = sxz;
      part[(int64_t)blockIdx.x * 2 * nc + 2 * c + 1] = sxy;
    }
  }
}

__global__ void olsen_step_kernel(int64_t n, int nc, const double* __restrict__ theta, const double* __restrict__ hii,
                                  const double* __restrict__ X, const double* __restrict__ Res,
                                  const double* __restrict__ sums, double* __restrict__ T, int64_t ld) {
  const int64_t total = n * nc;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = idx / n, i = idx - c * n;
    const double alpha = sums[2 * c] / sums[2 * c + 1];
    const double cd = 1.0 / guard_den(hii[i] - theta[c]);
    T[i + ld * c] = alpha * cd * X[i + ld * c] - cd * Res[i + ld * c];
  }
}

__global__ void unit_vectors_kernel(int nc, const int64_t* __restrict__ idx, double* __restrict__ V, int64_t ld) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nc) V[idx[c] + ld * c] = 1.0;
}

__global__ void symmetrize_kernel(int c0, int c1, double* R, int ld) {
  // rows c0..c1 of columns 0..c1 <- transpose of the freshly computed columns
  const int64_t w = c1 - c0;
  const int64_t total = w * c1;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int i = c0 + (int)(idx % w);  // row in the new range
    const int j = (int)(idx / w);       // column 0..c1
    if (j < c0 || j > i) R[i + (int64_t)j * ld] = R[j + (int64_t)i * ld];
  }
}

__global__ void cdot_kernel(int64_t n, const double* __restrict__ a, const double* __restrict__ b, int64_t ld,
                            double* __restrict__ part) {
  __shared__ double red[16];
  const double* ar = a;
  const double* ai = a + ld;
  const double* br = b;
  const double* bi = b + ld;
  double sr = 0.0, si = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    sr += ar[i] * br[i] + ai[i] * bi[i];
    si += ar[i] * bi[i] - ai[i] * br[i];
  }
  sr = block_sum(sr, red);
  si = block_sum(si, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sr;
    part[2 * blockIdx.x + 1] = si;
  }
}

}  // namespace

void launch_syev(int n, double* A, int lda, double* w, double* Z, int ldz, double* work, int* info,
                 hipStream_t st) {
  if (n > SYEV_NMAX) throw Error(ADC2G_EARG, "syev: n > " + std::to_string(SYEV_NMAX));
  hipLaunchKernelGGL(syev_kernel, dim3(1), dim3(SYEV_THREADS), 0, st, n, A, lda, w, Z, ldz, work, info);
  HIP_CHECK(hipGetLastError());
}

int launch_residual(int64_t n, int nc, const double* theta, const double* X, const double* WY, double* Res,
                    int64_t ld, double* part, hipStream_t st) {
  const int nb = reduce_blocks(n);
  hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(256), 0, st, n, nc, theta, X, WY, Res, ld, part);
  HIP_CHECK(hipGetLastError());
  return nb;
}

void launch_colsum(int nblk, int nc, const double* part, double* out, hipStream_t st) {
  hipLaunchKernelGGL(colsum_kernel, dim3(1), dim3(256), 0, st, nblk, nc, part, out);
  HIP_CHECK(hipGetLastError());
}

void launch_dpr(int64_t n, int nc, const double* theta, const double* hii, const double* Res, double* T,
                int64_t ld, hipStream_t st) {
  hipLaunchKernelGGL(dpr_kernel, dim3(grid_for(n * nc)), dim3(256), 0, st, n, nc, theta, hii, Res, T, ld);
  HIP_CHECK(hipGetLastError());
}

int launch_olsen_dots(int64_t n, int nc, const double* theta, const double* hii, const double* X,
                      const double* Res, int64_t ld, double* part, hipStream_t st) {
  const int nb = reduce_blocks(n);
  hipLaunchKernelGGL(olsen_dots_kernel, dim3(nb), dim3(256), 0, st, n, nc, theta, hii, X, Res, ld, part);
  HIP_CHECK(hipGetLastError());
  return nb;
}

void launch_olsen_step(int64_t n, int nc, const double* theta, const double* hii, const double* X,
                       const double* Res, const double* sums, double* T, int64_t ld, hipStream_t st) {
  hipLaunchKernelGGL(olsen_step_kernel, dim3(grid_for(n * nc)), dim3(256), 0, st, n, nc, theta, hii, X, Res, sums,
                     T, ld);
  HIP_CHECK(hipGetLastError());
}

void launch_unit_vectors(int nc, const int64_t* idx, double* V, int64_t ld, hipStream_t st) {
  hipLaunchKernelGGL(unit_vectors_kernel, dim3((nc + 255) / 256), dim3(256), 0, st, nc, idx, V, ld);
  HIP_CHECK(hipGetLastError());
}

void launch_orth_coeffs(int r, const double* G, int ldg, double* M, int ldm, double* work, int* kept, double tol,
                        hipStream_t st) {
  if (r > SYEV_NMAX) throw Error(ADC2G_EARG, "orth_coeffs: block too large");
  hipLaunchKernelGGL(orth_coeffs_kernel, dim3(1), dim3(SYEV_THREADS), 0, st, r, G, ldg, M, ldm, work, kept, tol);
  HIP_CHECK(hipGetLastError());
}

void launch_symmetrize_cols(int c0, int c1, double* R, int ld, hipStream_t st) {
  if (c1 <= c0) return;
  hipLaunchKernelGGL(symmetrize_kernel, dim3(grid_for((int64_t)(c1 - c0) * c1)), dim3(256), 0, st, c0, c1, R, ld);
  HIP_CHECK(hipGetLastError());
}

int launch_cdot_partial(int64_t n, const double* a, const double* b, int64_t ld, double* part, hipStream_t st) {
  const int nb = reduce_blocks(n);
  hipLaunchKernelGGL(cdot_kernel, dim3(nb), dim3(256), 0, st, n, a, b, ld, part);
  HIP_CHECK(hipGetLastError());
  return nb;
}

}  // namespace adc2g
