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
constexpr int SYEVX_ZCAP = 10240;  // doubles of LDS for the eigenvector block (80 KB)
constexpr int ORTH_ZCAP = 4096;    // the same for the Gram-matrix eigensolve (blocks up to 64)

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
  double d[SYEV_NMAX], e[SYEV_NMAX], rc[SYEV_NMAX], rs[SYEV_NMAX], qv[SYEV_NMAX];
  double red[16];
  int perm[SYEV_NMAX];
  int lo, hi, go, info;
};

// Householder reduction of the symmetric n x n matrix A to tridiagonal form.
// On exit S.d / S.e hold the diagonal and the off-diagonal (e[n-1] = 0) and
// column k of A below row k holds reflector k (unit 2-norm, H_k = I - 2 v v^T).
// Four barriers per column: wave 0 forms the reflector (v in S.rs), all waves
// form p = A22 v (partials in S.rc), wave 0 forms q = 2 (p - (v.p) v) (S.qv),
// all waves apply A22 <- A22 - v q^T - q v^T.  Zw is unused (kept for the
// callers' workspace layout).
__device__ void tridiag_dev(int n, double* A, int lda, double* Zw, SyevShared& S) {
  (void)Zw;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
  double* vv = S.rs;
  double* qv = S.qv;
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    double* col = A + (int64_t)k * lda;
    if (wv == 0) {
      double part = 0.0;
      for (int i = k + 2 + lane; i < n; i += 64) part += col[i] * col[i];
      for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
      const double x0 = col[k + 1];
      if (part == 0.0) {
        for (int i = k + 1 + lane; i < n; i += 64) vv[i] = 0.0;
        if (lane == 0) {
          S.e[k] = x0;
          col[k + 1] = 0.0;  // no reflection: zero reflector
          S.go = 0;
        }
      } else {
        const double alpha = -copysign(sqrt(part + x0 * x0), x0);
        const double v0 = x0 - alpha;
        const double inv = 1.0 / sqrt(part + v0 * v0);
        for (int i = k + 1 + lane; i < n; i += 64) {
          const double x = (i == k + 1 ? v0 : col[i]) * inv;
          vv[i] = x;
          col[i] = x;
        }
        if (lane == 0) {
          S.e[k] = alpha;
          S.go = 1;
        }
      }
    }
    __syncthreads();
    if (!S.go) {
      __syncthreads();  // everyone has read S.go before wave 0 rewrites it
      continue;
    }
    // p = A22 v: the m rows in chunks of 64 (one row per lane, coalesced over
    // the column-major storage), the columns in ns slices, one (chunk, slice)
    // per wave
    {
      const int rch = (m + 63) >> 6;
      const int ns = max(1, nw / rch);
      const int per = (m + ns - 1) / ns;
      for (int task = wv; task < rch * ns; task += nw) {
        const int rc = task % rch, sl = task / rch;
        const int i = k + 1 + rc * 64 + lane;
        const int j0 = k + 1 + sl * per, j1 = min(n, j0 + per);
        if (i < n) {
          double s = 0.0;
          for (int j = j0; j < j1; ++j) s += A[i + (int64_t)j * lda] * vv[j];
          S.rc[sl * m + (i - k - 1)] = s;
        }
      }
      __syncthreads();
      if (wv == 0) {
        double kv = 0.0;
        for (int i = k + 1 + lane; i < n; i += 64) {
          double s = 0.0;
          for (int sl = 0; sl < ns; ++sl) s += S.rc[sl * m + (i - k - 1)];
          qv[i] = s;  // p for now
          kv += s * vv[i];
        }
        for (int off = 32; off > 0; off >>= 1) kv += __shfl_xor(kv, off, 64);
        for (int i = k + 1 + lane; i < n; i += 64) qv[i] = 2.0 * (qv[i] - kv * vv[i]);
      }
    }
    __syncthreads();
    for (int j = k + 1 + wv; j < n; j += nw) {
      const double vj = vv[j], qj = qv[j];
      double* Aj = A + (int64_t)j * lda;
      for (int i = k + 1 + lane; i < n; i += 64) Aj[i] -= vv[i] * qj + qv[i] * vj;
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += nt) S.d[i] = A[i + (int64_t)i * lda];
  if (tid == 0) {
    if (n >= 2) S.e[n - 2] = A[n - 1 + (int64_t)(n - 2) * lda];
    S.e[n - 1] = 0.0;
  }
  __syncthreads();
}

// Eigen-decomposition of the symmetric n x n matrix A (destroyed); eigenvalues
// ascending into w, eigenvectors into the columns of Z; Zw is n*n scratch.
__device__ void syev_dev(int n, double* A, int lda, double* w, double* Z, int ldz, double* Zw, int* info,
                         SyevShared& S) {
  const int tid = threadIdx.x, nt = blockDim.x;
  tridiag_dev(n, A, lda, Zw, S);
  if (tid == 0) S.info = 0;
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

// ---- lowest k eigenpairs: bisection + inverse iteration on the tridiagonal

struct BisectShared {
  double lo[SYEV_NMAX], hi[SYEV_NMAX], mu[SYEV_NMAX];
  double xs[SYEV_THREADS];
  int cs[SYEV_THREADS];
  double red[3][16];
};

// Number of eigenvalues of the tridiagonal (d, e2 = e^2) below x (Sturm count).
__device__ int sturm_count(int n, const double* d, const double* e2, double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q < 0.0;
  for (int i = 1; i < n; ++i) {
    q = (d[i] - x) - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0;
  }
  return c;
}

__device__ __forceinline__ double unit_rand(uint64_t a, uint64_t b) {
  uint64_t z = a * 0x9E3779B97F4A7C15ull + b * 0xBF58476D1CE4E5B9ull + 0x94D049BB133111EBull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;  // [-1, 1)
}

// Lowest k eigenpairs of the symmetric n x n matrix A (destroyed), eigenvalues
// ascending in w[0..k), eigenvectors in Z[:, 0..k).  Householder reduction,
// then on the tridiagonal T:
//  * eigenvalue j by multisection: nt/k lanes evaluate Sturm counts at
//    interior points of its bracket each round, until the bracket is within
//    eps (2 |lambda| + ||T||);
//  * eigenvector j by inverse iteration on lane j: LU of T - mu_j I with
//    partial pivoting (shifts of coincident eigenvalues separated by
//    10 eps ||T||), three solves from a fixed pseudo-random start, then
//    two modified Gram-Schmidt sweeps over the k vectors (one wave);
//  * back-transformation of the k vectors with the stored reflectors.
// work: 2n + 6nk doubles.
// zl / zcap: LDS for the k eigenvectors ([i * k + j]) through inverse
// iteration, Gram-Schmidt and back-transformation when k * n <= zcap (their
// steps are serial chains, so the latency of each access is what counts);
// otherwise they run on Z in global memory.
__device__ void syevx_dev(int n, int k, double* A, int lda, double* w, double* Z, int ldz, double* work, int* info,
                          SyevShared& S, BisectShared& B, double* zl = nullptr, int zcap = 0) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
  tridiag_dev(n, A, lda, work, S);
  const double* d = S.d;
  double* e2 = S.rc;
  for (int i = tid; i < n; i += nt) e2[i] = S.e[i] * S.e[i];
  // Gershgorin interval, ||T|| and the largest e^2
  double glo = INFINITY, ghi = -INFINITY, em = 0.0;
  int bad = 0;
  for (int i = tid; i < n; i += nt) {
    const double r = (i > 0 ? fabs(S.e[i - 1]) : 0.0) + fabs(S.e[i]);
    glo = fmin(glo, d[i] - r);
    ghi = fmax(ghi, d[i] + r);
    em = fmax(em, S.e[i] * S.e[i]);
    if (!isfinite(d[i]) || !isfinite(S.e[i])) bad = 1;
  }
  for (int off = 32; off > 0; off >>= 1) {
    glo = fmin(glo, __shfl_xor(glo, off, 64));
    ghi = fmax(ghi, __shfl_xor(ghi, off, 64));
    em = fmax(em, __shfl_xor(em, off, 64));
  }
  if (lane == 0) {
    B.red[0][wv] = glo;
    B.red[1][wv] = ghi;
    B.red[2][wv] = em;
  }
  bad = __syncthreads_or(bad);
  if (bad) {
    if (tid == 0) info[0] = 1;
    return;
  }
  glo = B.red[0][0];
  ghi = B.red[1][0];
  em = B.red[2][0];
  for (int i = 1; i < nw; ++i) {
    glo = fmin(glo, B.red[0][i]);
    ghi = fmax(ghi, B.red[1][i]);
    em = fmax(em, B.red[2][i]);
  }
  const double tnorm = fmax(fmax(fabs(glo), fabs(ghi)), DBL_MIN);
  const double pivmin = DBL_MIN * fmax(1.0, em);
  const double fudge = 2.0 * tnorm * DBL_EPSILON * n + 2.0 * pivmin;
  glo -= fudge;
  ghi += fudge;
  for (int j = tid; j < k; j += nt) {
    B.lo[j] = glo;
    B.hi[j] = ghi;
  }
  __syncthreads();
  // ---- multisection
  const int G = max(1, nt / k);
  for (int round = 0; round < 96; ++round) {
    const int j = tid / G, g = tid - j * G;
    if (j < k) {
      const double lo = B.lo[j], hi = B.hi[j];
      const double x = lo + (hi - lo) * (double)(g + 1) / (double)(G + 1);
      B.xs[tid] = x;
      B.cs[tid] = sturm_count(n, d, e2, x, pivmin);
    }
    __syncthreads();
    int open = 0;
    for (int jj = tid; jj < k; jj += nt) {
      double lo = B.lo[jj], hi = B.hi[jj];
      for (int gg = 0; gg < G; ++gg) {
        const double x = B.xs[jj * G + gg];
        if (B.cs[jj * G + gg] <= jj) lo = fmax(lo, x);
        else hi = fmin(hi, x);
      }
      B.lo[jj] = lo;
      B.hi[jj] = hi;
      const double tol = DBL_EPSILON * (2.0 * fmax(fabs(lo), fabs(hi)) + tnorm);
      const double mid = 0.5 * (lo + hi);
      if (hi - lo > tol && mid > lo && mid < hi) open = 1;
    }
    if (!__syncthreads_or(open)) break;
  }
  for (int j = tid; j < k; j += nt) w[j] = 0.5 * (B.lo[j] + B.hi[j]);
  if (tid == 0) {
    const double pertol = 10.0 * DBL_EPSILON * tnorm;
    double prev = -INFINITY;
    for (int j = 0; j < k; ++j) {
      double m = 0.5 * (B.lo[j] + B.hi[j]);
      if (j > 0 && m - prev < pertol) m = prev + pertol;
      B.mu[j] = m;
      prev = m;
    }
  }
  __syncthreads();
  // ---- inverse iteration, one lane per eigenvalue; arrays [i * k + j]
  {
    double* U1 = work + 2 * n;
    double* U2 = U1 + (int64_t)n * k;
    double* U3 = U2 + (int64_t)n * k;
    double* ML = U3 + (int64_t)n * k;
    double* PV = ML + (int64_t)n * k;
    double* Y = (zl && (int64_t)n * k <= zcap) ? zl : PV + (int64_t)n * k;
    const double tiny = DBL_EPSILON * tnorm;
    for (int j = tid; j < k; j += nt) {
      const double mu = B.mu[j];
      double cu = d[0] - mu, cv = (n > 1) ? S.e[0] : 0.0;
      for (int i = 0; i + 1 < n; ++i) {
        const double a1 = d[i + 1] - mu, b1 = S.e[i], c1 = (i + 2 < n) ? S.e[i + 1] : 0.0;
        double m;
        if (fabs(cu) >= fabs(b1)) {
          m = (cu != 0.0) ? b1 / cu : 0.0;
          U1[(int64_t)i * k + j] = cu;
          U2[(int64_t)i * k + j] = cv;
          U3[(int64_t)i * k + j] = 0.0;
          PV[(int64_t)i * k + j] = 0.0;
          cu = a1 - m * cv;
          cv = c1;
        } else {
          m = cu / b1;
          U1[(int64_t)i * k + j] = b1;
          U2[(int64_t)i * k + j] = a1;
          U3[(int64_t)i * k + j] = c1;
          PV[(int64_t)i * k + j] = 1.0;
          const double nu = cv - m * a1;
          cv = -m * c1;
          cu = nu;
        }
        ML[(int64_t)i * k + j] = m;
      }
      U1[(int64_t)(n - 1) * k + j] = cu;
      // the last row has no super-diagonals; the back substitution reads
      // them (times zero), so they must not hold stale NaN or Inf
      U2[(int64_t)(n - 1) * k + j] = 0.0;
      U3[(int64_t)(n - 1) * k + j] = 0.0;
      for (int i = 0; i < n; ++i) {
        double& u = U1[(int64_t)i * k + j];
        if (fabs(u) < tiny) u = (u < 0.0) ? -tiny : tiny;
      }
      for (int i = 0; i < n; ++i) Y[(int64_t)i * k + j] = unit_rand(j + 1, i);
      for (int itr = 0; itr < 3; ++itr) {
        for (int i = 0; i + 1 < n; ++i) {
          double bi = Y[(int64_t)i * k + j], bn = Y[(int64_t)(i + 1) * k + j];
          if (PV[(int64_t)i * k + j] != 0.0) {
            const double t = bi;
            bi = bn;
            bn = t;
          }
          Y[(int64_t)i * k + j] = bi;
          Y[(int64_t)(i + 1) * k + j] = bn - ML[(int64_t)i * k + j] * bi;
        }
        double y2 = 0.0, y1 = 0.0, ymax = 0.0;
        for (int i = n - 1; i >= 0; --i) {
          const double y = (Y[(int64_t)i * k + j] - U2[(int64_t)i * k + j] * y1 - U3[(int64_t)i * k + j] * y2) /
                           U1[(int64_t)i * k + j];
          Y[(int64_t)i * k + j] = y;
          y2 = y1;
          y1 = y;
          ymax = fmax(ymax, fabs(y));
        }
        const double sc = (ymax > 0.0 && isfinite(ymax)) ? 1.0 / ymax : 0.0;
        for (int i = 0; i < n; ++i) Y[(int64_t)i * k + j] *= sc;
      }
      if (Y != zl)
        for (int i = 0; i < n; ++i) Z[i + (int64_t)j * ldz] = Y[(int64_t)i * k + j];
    }
  }
  __syncthreads();
  // element (i, j) of the eigenvector block: LDS or Z
  const bool in_lds = zl && (int64_t)n * k <= zcap;
  auto zat = [&](int i, int j) -> double& { return in_lds ? zl[(int64_t)i * k + j] : Z[i + (int64_t)j * ldz]; };
  // ---- two modified Gram-Schmidt sweeps (wave 0)
  if (wv == 0) {
    for (int j = 0; j < k; ++j) {
      for (int pass = 0; pass < 2; ++pass) {
        for (int i = 0; i < j; ++i) {
          double t = 0.0;
          for (int r = lane; r < n; r += 64) t += zat(r, i) * zat(r, j);
          for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
          for (int r = lane; r < n; r += 64) zat(r, j) -= t * zat(r, i);
        }
      }
      double t = 0.0;
      for (int r = lane; r < n; r += 64) t += zat(r, j) * zat(r, j);
      for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
      const double sc = (t > 0.0) ? 1.0 / sqrt(t) : 0.0;
      for (int r = lane; r < n; r += 64) zat(r, j) *= sc;
    }
  }
  __syncthreads();
  // ---- back transformation of the k vectors (one column per wave)
  for (int j = wv; j < k; j += nw) {
    for (int kk = n - 3; kk >= 0; --kk) {
      const double* v = A + (int64_t)kk * lda;
      double t = 0.0;
      for (int i = kk + 1 + lane; i < n; i += 64) t += v[i] * zat(i, j);
      for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
      t *= 2.0;
      if (t != 0.0)
        for (int i = kk + 1 + lane; i < n; i += 64) zat(i, j) -= t * v[i];
    }
  }
  __syncthreads();
  if (in_lds)
    for (int64_t idx = tid; idx < (int64_t)n * k; idx += nt) {
      const int i = (int)(idx % n), j = (int)(idx / n);
      Z[i + (int64_t)j * ldz] = zl[(int64_t)i * k + j];
    }
  if (tid == 0) info[0] = 0;
  __syncthreads();
}

__global__ void __launch_bounds__(SYEV_THREADS) syevx_kernel(int n, int k, double* A, int lda, double* w, double* Z,
                                                              int ldz, double* work, int* info) {
  __shared__ SyevShared S;
  __shared__ BisectShared B;
  __shared__ double zl[SYEVX_ZCAP];
  syevx_dev(n, k, A, lda, w, Z, ldz, work, info, S, B, zl, SYEVX_ZCAP);
}

// M = D U_kept Lambda^{-1/2} for the scaled Gram matrix D G D (D = diag(G)^{-1/2}).
// work: r*r (copy) + r*r (U) + r (lambda) + r*r (syev scratch).
__global__ void __launch_bounds__(SYEV_THREADS) orth_coeffs_kernel(int r, const double* G, int ldg, double* M,
                                                                    int ldm, double* work, int* kept, double tol,
                                                                    double* cond, const double* bnorm) {
  __shared__ SyevShared S;
  __shared__ BisectShared Bs;
  __shared__ double dinv[SYEV_NMAX];
  __shared__ int sel[SYEV_NMAX];
  __shared__ int nsel;
  const int tid = threadIdx.x, nt = blockDim.x;
  double* Gc = work;
  double* U = work + (int64_t)r * r;
  double* lam = U + (int64_t)r * r;
  double* Zw = lam + r;
  // a column that kept less than tol of its norm through the projection
  // that made it (bnorm: squared norms before it) is rounding noise of the
  // subspace: its row and column of the scaled Gram matrix are zeroed, so it
  // is dropped
  for (int i = tid; i < r; i += nt) {
    const double g = G[i + (int64_t)i * ldg];
    dinv[i] = (g > 0.0 && !(bnorm && g <= tol * tol * bnorm[i])) ? 1.0 / sqrt(g) : 0.0;
  }
  __syncthreads();
  for (int64_t idx = tid; idx < (int64_t)r * r; idx += nt) {
    const int i = (int)(idx % r), j = (int)(idx / r);
    Gc[idx] = dinv[i] * G[i + (int64_t)j * ldg] * dinv[j];
  }
  __syncthreads();
  __shared__ int sinfo;
  __shared__ double zl[ORTH_ZCAP];
  syevx_dev(r, r, Gc, r, lam, U, r, Zw, &sinfo, S, Bs, zl, ORTH_ZCAP);
  __syncthreads();
  if (tid == 0) {
    int c = 0;
    double lmin = 0.0;
    if (sinfo == 0) {
      const double lmax = lam[r - 1];
      // largest eigenvalues first, so the best-conditioned directions come first
      for (int j = r - 1; j >= 0; --j)
        if (lam[j] > tol * lmax && lam[j] > 0.0) {
          sel[c++] = j;
          lmin = lam[j];
        }
      if (cond) cond[0] = (c > 0) ? lmax / lmin : INFINITY;
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

// need[0] = 1 when some column kept less than DGKS_KEEP of its squared norm
// through the projection: before^2 = |S1[:,j]|^2 + |S2[:,j]|^2 + G[j,j],
// after^2 = G[j,j].  One classical Gram-Schmidt pass leaves the column
// orthogonal to the subspace up to about eps / sqrt(after^2 / before^2); with
// DGKS_KEEP = 1e-6 that is ~1e-13, far below what the Rayleigh-Ritz step of a
// 1e-7 residual needs (1e-7 / ||H||), so the second pass runs only when the
// cancellation is that severe (the textbook 1/2 re-orthogonalises much more
// often than FP64 requires here).
constexpr double DGKS_KEEP = 1e-6;
__global__ void dgks_check_kernel(int r, int nb1, const double* __restrict__ S1, int ld1, int nb2,
                                  const double* __restrict__ S2, int ld2, const double* __restrict__ G, int ldg,
                                  int* __restrict__ need, double* __restrict__ bnorm) {
  int bad = 0;
  for (int j = threadIdx.x; j < r; j += blockDim.x) {
    double b = 0.0;
    for (int i = 0; i < nb1; ++i) b += S1[i + (int64_t)j * ld1] * S1[i + (int64_t)j * ld1];
    for (int i = 0; i < nb2; ++i) b += S2[i + (int64_t)j * ld2] * S2[i + (int64_t)j * ld2];
    const double a = G[j + (int64_t)j * ldg];
    b += a;
    if (bnorm) bnorm[j] = b;
    if (!(a >= DGKS_KEEP * b)) bad = 1;
  }
  bad = __syncthreads_or(bad);
  if (threadIdx.x == 0) need[0] = bad;
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

// One workgroup per column; fixed summation order.
__global__ void colsum_kernel(int nblk, int nc, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[16];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[(int64_t)b * nc + c];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[c] = s;
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

void launch_syevx(int n, int k, double* A, int lda, double* w, double* Z, int ldz, double* work, int* info,
                  hipStream_t st) {
  if (n > SYEV_NMAX) throw Error(ADC2G_EARG, "syevx: n > " + std::to_string(SYEV_NMAX));
  if (k < 1 || k > n) throw Error(ADC2G_EARG, "syevx: need 1 <= k <= n");
  hipLaunchKernelGGL(syevx_kernel, dim3(1), dim3(SYEV_THREADS), 0, st, n, k, A, lda, w, Z, ldz, work, info);
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
  if (nc <= 0) return;
  hipLaunchKernelGGL(colsum_kernel, dim3(nc), dim3(256), 0, st, nblk, nc, part, out);
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

void launch_dgks_check(int r, int nb1, const double* S1, int ld1, int nb2, const double* S2, int ld2, const double* G,
                       int ldg, int* need, hipStream_t st, double* bnorm) {
  hipLaunchKernelGGL(dgks_check_kernel, dim3(1), dim3(256), 0, st, r, nb1, S1, ld1, nb2, S2, ld2, G, ldg, need,
                     bnorm);
  HIP_CHECK(hipGetLastError());
}

void launch_unit_vectors(int nc, const int64_t* idx, double* V, int64_t ld, hipStream_t st) {
  hipLaunchKernelGGL(unit_vectors_kernel, dim3((nc + 255) / 256), dim3(256), 0, st, nc, idx, V, ld);
  HIP_CHECK(hipGetLastError());
}

void launch_orth_coeffs(int r, const double* G, int ldg, double* M, int ldm, double* work, int* kept, double tol,
                        hipStream_t st, double* cond, const double* bnorm) {
  if (r > SYEV_NMAX) throw Error(ADC2G_EARG, "orth_coeffs: block too large");
  hipLaunchKernelGGL(orth_coeffs_kernel, dim3(1), dim3(SYEV_THREADS), 0, st, r, G, ldg, M, ldm, work, kept, tol,
                     cond, bnorm);
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
