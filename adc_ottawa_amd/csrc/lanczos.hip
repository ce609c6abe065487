// Block Lanczos pseudospectrum on the GPU sigma-vector: run_block_lanczos
// (source/eigen/block_lanczos.f90:947-1256) without re-orthogonalisation
// (orthotype 0, the default), with partial re-orthogonalisation
// (orthotype 1: pro, update_omega, pro_mgs, localro, :1430-1706) or with the
// modified partial re-orthogonalisation (orthotype 2: mpro, mpro_mgs, :1710-1890), with
// hxq_incore (:1325-1361) replaced by the sigma-vector of the block.  Per cycle j (B_0 = 0):
//   U = H Q_j - Q_{j-1} B_{j-1}^T,  A_j = Q_j^T U,  R = U - Q_j A_j,
//   R = Q_{j+1} B_j  (B_j upper triangular).
// All Lanczos vectors stay in HBM (L, ndim x nvec); the products with them
// are tall-skinny FP64 kernels; the b x b algebra runs on the host.
//
// The factorisation of R is a shifted Cholesky QR followed by two Cholesky QR
// passes (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa, SIAM J. Sci.
// Comput. 42 (2020) A477): G = R^T R + s I with s = 11 (n b + b (b+1)) u
// |R|_F^2, R_1 = chol(G), Q = R R_1^-1, then twice R_k = chol(Q^T Q),
// Q <- Q R_k^-1; B_j = R_3 R_2 R_1.  It is orthonormal to rounding for
// condition numbers up to about 1/u, where the reference uses Householder QR
// (dgeqrf/dorgqr).  The diagonal of B_j is positive, so Q_{j+1} and B_j
// differ from the reference's by column and row signs and T by a diagonal
// +-1 similarity, which leaves eigenvalues, Ritz vectors and |transition
// moments| unchanged.
//
// T is diagonalised on the host threads (Householder tridiagonalisation,
// implicit QL with Wilkinson shifts, eigenvectors accumulated), as dsbev does in the
// reference (diagmat_banded, :2006-2052); the Ritz vectors L Y are formed on
// the device in column chunks, normalised, dotted with travec and copied out
// on request (ritzvecs_incore, :2237-2300).
#include <algorithm>
#include <array>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "linalg.hpp"
#include "solvers.hpp"

namespace adc2g {

namespace {

constexpr int LZ_NVEC_MAX = 4096;  // host eigen-decomposition of T is O(nvec^3)
constexpr int LZ_CHUNK = 256;      // Ritz vectors formed per pass
constexpr int LZ_THREADS = 256;

// ---- host dense algebra (column-major)

// Upper Cholesky factor: G = R^T R.  False when a pivot is not positive.
bool cholesky_upper(int b, const double* G, double* R) {
  std::fill(R, R + (size_t)b * b, 0.0);
  for (int j = 0; j < b; ++j) {
    double s = G[j + (size_t)b * j];
    for (int k = 0; k < j; ++k) s -= R[k + (size_t)b * j] * R[k + (size_t)b * j];
    if (!(s > 0.0) || !std::isfinite(s)) return false;
    const double rjj = std::sqrt(s);
    R[j + (size_t)b * j] = rjj;
    for (int i = j + 1; i < b; ++i) {
      double t = G[j + (size_t)b * i];
      for (int k = 0; k < j; ++k) t -= R[k + (size_t)b * j] * R[k + (size_t)b * i];
      R[j + (size_t)b * i] = t / rjj;
    }
  }
  return true;
}

// X = R^-1 for upper triangular R (back substitution per column).
void upper_inverse(int b, const double* R, double* X) {
  std::fill(X, X + (size_t)b * b, 0.0);
  for (int c = 0; c < b; ++c) {
    X[c + (size_t)b * c] = 1.0 / R[c + (size_t)b * c];
    for (int i = c - 1; i >= 0; --i) {
      double t = 0.0;
      for (int k = i + 1; k <= c; ++k) t += R[i + (size_t)b * k] * X[k + (size_t)b * c];
      X[i + (size_t)b * c] = -t / R[i + (size_t)b * i];
    }
  }
}

// C = A B for upper triangular b x b A and B (C upper triangular).
void upper_mul(int b, const double* A, const double* B, double* C) {
  std::fill(C, C + (size_t)b * b, 0.0);
  for (int j = 0; j < b; ++j)
    for (int i = 0; i <= j; ++i) {
      double t = 0.0;
      for (int k = i; k <= j; ++k) t += A[i + (size_t)b * k] * B[k + (size_t)b * j];
      C[i + (size_t)b * j] = t;
    }
}

// Runs f(begin, end) over [0, count) split into contiguous ranges of at
// least `grain` on up to 16 host threads.
template <class F>
void par_for(int count, int grain, F f) {
  const int hw = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  const int nt = std::max(1, std::min(hw, (count + grain - 1) / std::max(1, grain)));
  if (nt <= 1) {
    f(0, count);
    return;
  }
  const int per = (count + nt - 1) / nt;
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) {
    const int b = t * per, e = std::min(count, b + per);
    if (b < e) th.emplace_back(f, b, e);
  }
  for (auto& x : th) x.join();
}

// Barrier for a fixed team of host threads (generation counting).
struct SpinBarrier {
  explicit SpinBarrier(int n) : n_(n) {}
  void wait() {
    if (n_ <= 1) return;
    const int g = gen_.load(std::memory_order_acquire);
    if (count_.fetch_add(1, std::memory_order_acq_rel) == n_ - 1) {
      count_.store(0, std::memory_order_relaxed);
      gen_.fetch_add(1, std::memory_order_release);
    } else {
      while (gen_.load(std::memory_order_acquire) == g) std::this_thread::yield();
    }
  }
  int n_;
  std::atomic<int> count_{0}, gen_{0};
};

}  // namespace

// Eigen-decomposition of the symmetric n x n matrix A (destroyed): eigenvalues
// ascending in w, eigenvectors in the columns of Z.
//  1. Householder reduction to tridiagonal form: reflector k (unit 2-norm)
//     kept in column k of A below row k+1; the trailing block updated as
//     A <- A - v q^T - q v^T with p = A v, q = 2 (p - (v.p) v).
//  2. Implicit QL with a Wilkinson shift on the tridiagonal; the Givens
//     rotations of every sweep are recorded, then applied to Z = I in one
//     pass per block of rows (the rows of Z transform independently, so the
//     blocks run on separate threads and stay in cache).
//  3. Back transformation Z <- H_0 ... H_{n-3} Z, columns on separate threads.
//  4. Ascending sort.
// False if QL does not converge.
bool host_syev(int n, double* A, double* w, double* Z) {
  std::vector<double> d(n), e(n, 0.0), v(n), p(n), q(n);
  // 1. one team of threads for all columns, four barriers per column: the
  // reflector (thread 0), p over each thread's rows, q, the rank-2 update
  // over each thread's columns.  Every thread sums v.p itself in the same
  // order, so the result does not depend on the number of threads.
  const int hw = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  const int nt = n >= 512 ? hw : 1;
  SpinBarrier bar(nt);
  int skip = 0;
  auto team = [&](int tid) {
    for (int k = 0; k + 2 < n; ++k) {
      double* col = A + (size_t)k * n;
      if (tid == 0) {
        double part = 0.0;
        for (int i = k + 2; i < n; ++i) part += col[i] * col[i];
        const double x0 = col[k + 1];
        if (part == 0.0) {
          e[k] = x0;
          col[k + 1] = 0.0;  // zero reflector: no reflection
          skip = 1;
        } else {
          const double alpha = -std::copysign(std::sqrt(part + x0 * x0), x0);
          const double v0 = x0 - alpha;
          const double inv = 1.0 / std::sqrt(part + v0 * v0);
          for (int i = k + 1; i < n; ++i) {
            v[i] = (i == k + 1 ? v0 : col[i]) * inv;
            col[i] = v[i];
          }
          e[k] = alpha;
          skip = 0;
        }
      }
      bar.wait();
      const int sk = skip;
      bar.wait();  // everyone has read skip before thread 0 rewrites it
      if (sk) continue;
      const int m = n - k - 1, per = (m + nt - 1) / nt;
      const int i0 = k + 1 + std::min(m, tid * per), i1 = k + 1 + std::min(m, (tid + 1) * per);
      // p_i = A(:, i) . v over the trailing rows (A is kept symmetric)
      for (int i = i0; i < i1; ++i) {
        const double* Ai = A + (size_t)i * n;
        double t = 0.0;
        for (int j = k + 1; j < n; ++j) t += Ai[j] * v[j];
        p[i] = t;
      }
      bar.wait();
      double kv = 0.0;
      for (int i = k + 1; i < n; ++i) kv += p[i] * v[i];
      for (int i = i0; i < i1; ++i) q[i] = 2.0 * (p[i] - kv * v[i]);
      bar.wait();
      for (int j = i0; j < i1; ++j) {
        const double vj = v[j], qj = q[j];
        double* Aj = A + (size_t)j * n;
        for (int i = k + 1; i < n; ++i) Aj[i] -= v[i] * qj + q[i] * vj;
      }
      bar.wait();
    }
  };
  {
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(team, t);
    team(0);
    for (auto& x : th) x.join();
  }
  for (int i = 0; i < n; ++i) d[i] = A[i + (size_t)i * n];
  if (n >= 2) e[n - 2] = A[n - 1 + (size_t)(n - 2) * n];
  e[n - 1] = 0.0;

  struct Sweep {
    int lo, hi;
    size_t off;  // (c, s) of rotation t = hi..lo at rot[off + 2 (hi - t)]
  };
  std::vector<Sweep> sweeps;
  std::vector<double> rot;
  std::vector<double> rc(n), rs(n);
  for (int l = 0; l < n; ++l) {
    for (int iter = 0;; ++iter) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= DBL_EPSILON * dd) break;
      }
      if (m == l) break;
      if (iter >= 60) return false;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, pp = 0.0;
      int i = m - 1;
      bool under = false;
      for (; i >= l; --i) {
        const double f = s * e[i], bb = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= pp;
          e[m] = 0.0;
          under = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - pp;
        r = (d[i] - g) * s + 2.0 * c * bb;
        pp = s * r;
        d[i + 1] = g + pp;
        g = c * r - bb;
        rc[i] = c;
        rs[i] = s;
      }
      const int lo = under ? i + 1 : l, hi = m - 1;
      if (!under) {
        d[l] -= pp;
        e[l] = g;
        e[m] = 0.0;
      }
      if (hi >= lo) {
        sweeps.push_back({lo, hi, rot.size()});
        for (int t = hi; t >= lo; --t) {
          rot.push_back(rc[t]);
          rot.push_back(rs[t]);
        }
      }
    }
  }
  // rotations (t, t+1), t = hi..lo, of every sweep in order, applied to the
  // columns of Z = I one block of RB rows at a time; the column t+1 of the
  // running product is carried down the sweep
  constexpr int RB = 32;
  const int nrb = (n + RB - 1) / RB;
  par_for(nrb, 1, [&](int b0, int b1) {
    double carry[RB];
    for (int rb = b0; rb < b1; ++rb) {
      const int r0 = rb * RB, nr = std::min(RB, n - r0);
      for (int j = 0; j < n; ++j)
        for (int r = 0; r < nr; ++r) Z[r0 + r + (size_t)n * j] = (r0 + r == j) ? 1.0 : 0.0;
      for (const Sweep& sw : sweeps) {
        const double* rr = rot.data() + sw.off;
        const double* zc = Z + (size_t)(sw.hi + 1) * n + r0;
        for (int r = 0; r < nr; ++r) carry[r] = zc[r];
        for (int t = sw.hi; t >= sw.lo; --t, rr += 2) {
          const double ct = rr[0], st = rr[1];
          double* zi = Z + (size_t)t * n + r0;
          double* zn = zi + n;
          for (int r = 0; r < nr; ++r) {
            const double a = zi[r], cr = carry[r];
            zn[r] = st * a + ct * cr;
            carry[r] = ct * a - st * cr;
          }
        }
        double* zl = Z + (size_t)sw.lo * n + r0;
        for (int r = 0; r < nr; ++r) zl[r] = carry[r];
      }
    }
  });
  // back transformation Z <- H_0 ... H_{n-3} Z
  par_for(n, 8, [&](int j0, int j1) {
    for (int j = j0; j < j1; ++j) {
      double* z = Z + (size_t)j * n;
      for (int k = n - 3; k >= 0; --k) {
        const double* vk = A + (size_t)k * n;
        double t = 0.0;
        for (int i = k + 1; i < n; ++i) t += vk[i] * z[i];
        t *= 2.0;
        if (t != 0.0)
          for (int i = k + 1; i < n; ++i) z[i] -= t * vk[i];
      }
    }
  });
  // ascending order
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return d[x] < d[y]; });
  std::vector<double> Zs((size_t)n * n);
  for (int j = 0; j < n; ++j) {
    w[j] = d[perm[j]];
    std::copy(Z + (size_t)perm[j] * n, Z + (size_t)(perm[j] + 1) * n, Zs.begin() + (size_t)j * n);
  }
  std::copy(Zs.begin(), Zs.end(), Z);
  return true;
}

namespace {

// ---- device kernels of the Ritz stage

// part[blk * 2 nc + 2 c] = partial sum of R[:, c]^2, part[... + 1] = partial
// sum of f . R[:, c] (f may be NULL: 0).  grid (nblk, nc).
__global__ void __launch_bounds__(LZ_THREADS) ritz_stats_kernel(int64_t n, int nc, const double* __restrict__ R,
                                                                int64_t ld, const double* __restrict__ f,
                                                                double* __restrict__ part) {
  __shared__ double red[2][LZ_THREADS / 64];
  const int c = blockIdx.y;
  const double* r = R + ld * c;
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)LZ_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * LZ_THREADS) {
    const double x = r[i];
    s0 += x * x;
    if (f) s1 += f[i] * x;
  }
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_down(s0, off, 64);
    s1 += __shfl_down(s1, off, 64);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wv] = s0;
    red[1][wv] = s1;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double t = 0.0;
    for (int i = 0; i < LZ_THREADS / 64; ++i) t += red[threadIdx.x][i];
    part[(int64_t)blockIdx.x * 2 * nc + 2 * c + threadIdx.x] = t;
  }
}

// R[:, c] *= scale[c]; grid (nblk, nc).
__global__ void __launch_bounds__(LZ_THREADS) scale_cols_kernel(int64_t n, double* __restrict__ R, int64_t ld,
                                                                const double* __restrict__ scale) {
  const int c = blockIdx.y;
  const double sc = scale[c];
  double* r = R + ld * c;
  for (int64_t i = blockIdx.x * (int64_t)LZ_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * LZ_THREADS)
    r[i] *= sc;
}

struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
  Timer() {
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
  }
  ~Timer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

}  // namespace

void lanczos_run(Handle& H, const adc2g_lanczos_params& p, const double* q0_h, double* eigval, double* tmat,
                 const double* travec, double* tranmom, double* ritz, adc2g_lanczos_stats* stats) {
  using clk = std::chrono::steady_clock;
  const auto t_start = clk::now();
  const int b = p.blocksize, m = p.ncycles;
  const int64_t n = H.d.ndim;
  ARG_CHECK(b >= 1 && m >= 1, "blocksize and ncycles must be >= 1");
  ARG_CHECK((int64_t)b * m <= n, "blocksize * ncycles must not exceed ndim");
  ARG_CHECK((int64_t)b * m <= LZ_NVEC_MAX, "blocksize * ncycles must not exceed " + std::to_string(LZ_NVEC_MAX));
  ARG_CHECK((travec == nullptr) == (tranmom == nullptr), "travec and tranmom must both be given or both be NULL");
  const Prec pr = prec_from_int(p.precision);
  const int nvec = b * m;
  hipStream_t st = H.st;

  ARG_CHECK(p.orthotype >= 0 && p.orthotype <= 2, "orthotype must be 0, 1 or 2");
  const bool pro = p.orthotype >= 1;  // the omega recurrence (PRO and MPRO)
  DevBuf<double> L, U, X2, small, Pw, part, sums, scale;
  L.alloc((size_t)n * nvec);
  U.alloc((size_t)n * b);
  X2.alloc((size_t)n * b);
  small.alloc((size_t)b * b);
  const int nblk = reduce_blocks(n);
  if (pro) {
    Pw.alloc((size_t)n * b);
    part.alloc((size_t)nblk * 2 * b);
    sums.alloc((size_t)2 * b);
    scale.alloc(b);
  }
  HIP_CHECK(hipMemcpyAsync(L.get(), q0_h, sizeof(double) * n * b, hipMemcpyHostToDevice, st));

  std::vector<double> G((size_t)b * b), A((size_t)b * b), Bprev((size_t)b * b, 0.0), Bt((size_t)b * b);
  std::vector<double> R1((size_t)b * b), R2((size_t)b * b), R3((size_t)b * b), Ri((size_t)b * b),
      W1((size_t)b * b);
  std::vector<double> T((size_t)nvec * nvec, 0.0);

  auto gram_to_host = [&](const double* X, std::vector<double>& out) {
    gemm_tn(H, b, b, n, X, n, X, n, small.get(), b, 1.0, 0.0);
    HIP_CHECK(hipMemcpyAsync(out.data(), small.get(), sizeof(double) * b * b, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
  };
  auto upload_small = [&](const std::vector<double>& h) {
    HIP_CHECK(hipMemcpyAsync(small.get(), h.data(), sizeof(double) * b * b, hipMemcpyHostToDevice, st));
  };

  // the initial block must be orthonormal (init_vec, :128-197, delivers it so)
  gram_to_host(L.get(), G);
  {
    double dev = 0.0;
    for (int i = 0; i < b; ++i)
      for (int j = 0; j < b; ++j) dev = std::max(dev, std::fabs(G[i + (size_t)b * j] - (i == j ? 1.0 : 0.0)));
    ARG_CHECK(dev <= 1e-8, "q0 must have orthonormal columns (max |q0^T q0 - I| = " + std::to_string(dev) + ")");
  }

  // partial re-orthogonalisation (orthotype 1): the omega recurrence on the
  // host, the block operations on the device
  const double epmach = DBL_EPSILON;  // machine_precision (:927-943) returns 2^-52
  const double om_eps = epmach * b * std::sqrt((double)n), orthlim = std::sqrt(epmach);
  std::vector<double> om((size_t)(m + 2) * (m + 2), 0.0), anorm(m + 2, 0.0), bnorm(m + 2, 0.0);
  std::vector<std::vector<double>> A_all, B_all;
  auto OM = [&](int r, int c) -> double& { return om[r + (size_t)(m + 2) * c]; };
  // singular values of a b x b matrix: {largest, smallest}
  auto svd_ends = [&](const std::vector<double>& M) {
    std::vector<double> MtM((size_t)b * b), w(b), Z((size_t)b * b);
    for (int c = 0; c < b; ++c)
      for (int r = 0; r < b; ++r) {
        double t = 0.0;
        for (int k = 0; k < b; ++k) t += M[k + (size_t)b * r] * M[k + (size_t)b * c];
        MtM[r + (size_t)b * c] = t;
      }
    if (!host_syev(b, MtM.data(), w.data(), Z.data())) throw Error(ADC2G_EEIG, "block Lanczos: SVD failed");
    return std::make_pair(std::sqrt(std::max(w[b - 1], 0.0)), std::sqrt(std::max(w[0], 0.0)));
  };
  // Q <- Q - S C with C = S^T Q (b columns each)
  auto project_out = [&](const double* S, double* Q) {
    gemm_tn(H, b, b, n, S, n, Q, n, small.get(), b, 1.0, 0.0);
    gemm_nn(H, n, b, b, S, n, small.get(), b, Q, n, -1.0, 1.0);
  };
  // columns of Q scaled to unit 2-norm
  auto normalise_cols = [&](double* Q) {
    hipLaunchKernelGGL(ritz_stats_kernel, dim3(nblk, b), dim3(LZ_THREADS), 0, st, n, b, Q, n, nullptr, part.get());
    HIP_CHECK(hipGetLastError());
    launch_colsum(nblk, 2 * b, part.get(), sums.get(), st);
    std::vector<double> h((size_t)2 * b), sc(b);
    HIP_CHECK(hipMemcpyAsync(h.data(), sums.get(), sizeof(double) * 2 * b, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    for (int c = 0; c < b; ++c) sc[c] = 1.0 / std::sqrt(h[2 * c]);
    HIP_CHECK(hipMemcpyAsync(scale.get(), sc.data(), sizeof(double) * b, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(scale_cols_kernel, dim3(nblk, b), dim3(LZ_THREADS), 0, st, n, Q, n, scale.get());
    HIP_CHECK(hipGetLastError());
  };
  int npro = 0;

  // modified partial re-orthogonalisation (orthotype 2; mpro, mpro_mgs,
  // lowerbound, upperbound, :1710-1890): Q_{j+1} against the intervals of
  // stored blocks around the estimates that reach sqrt(eps), bounded where
  // the estimates fall below eps^(3/4); the intervals, widened by one block
  // each side, are used once more in the next cycle.
  const double eta = std::pow(epmach, 0.75);
  bool l2nd = false;
  std::vector<std::array<int, 2>> ivs;
  double* Qn_cur = nullptr;  // Q_{j+1} of the current cycle
  auto lowerbound = [&](int j, int k) -> int {
    if (k == 1 || k == 2) return 1;
    int lb = k - 1;
    for (int i = k - 2; i <= 1; ++i) {  // do i=k-2,1: runs only for k = 3
      if (std::fabs(OM(j + 1, i)) > eta) lb = i;
      else break;
    }
    return lb;
  };
  auto upperbound = [&](int j, int k) -> int {
    int ub = k + 1;
    for (int i = k + 1; i <= j - 1; ++i) {
      if (std::fabs(OM(j + 1, i)) >= eta) ub = i;
      else break;
    }
    return ub;
  };
  auto mpro_mgs = [&]() {
    for (const auto& iv : ivs)
      for (int k = iv[0]; k <= iv[1]; ++k) project_out(L.get() + (size_t)n * b * (k - 1), Qn_cur);
    normalise_cols(Qn_cur);
  };
  auto mpro = [&](int jj) -> bool {
    if (l2nd) {
      ++npro;
      mpro_mgs();
      for (const auto& iv : ivs)
        for (int l = iv[0]; l <= iv[1]; ++l) OM(jj + 1, l) = OM(l, jj + 1) = om_eps;
      ivs.clear();
      l2nd = false;
      return true;
    }
    ivs.clear();
    bool lro = false;
    int k = 1;
    for (;;) {  // the goto-10 scan of mpro
      if (OM(jj + 1, k) >= orthlim) {
        lro = true;
        const int lo = lowerbound(jj, k), hi = upperbound(jj, k);
        ivs.push_back({lo, hi});
        k += hi;
      } else {
        ++k;
      }
      if (!(k < jj - 1)) break;
    }
    if (!lro) return false;
    ++npro;
    mpro_mgs();
    for (auto& iv : ivs) {
      for (int l = iv[0]; l <= iv[1]; ++l) OM(jj + 1, l) = OM(l, jj + 1) = om_eps;
      if (iv[0] > 1) --iv[0];
      ++iv[1];
    }
    l2nd = true;
    return true;
  };

  Timer tm;
  double t_sigma = 0.0;
  const double u = 0.5 * DBL_EPSILON;
  for (int j = 0; j < m; ++j) {
    const double* Qj = L.get() + (size_t)n * b * j;
    HIP_CHECK(hipEventRecord(tm.a, st));
    H.sigma(Qj, n, U.get(), n, b, pr, st);
    HIP_CHECK(hipEventRecord(tm.b, st));
    if (j > 0) {
      // U -= Q_{j-1} B_{j-1}^T
      for (int r = 0; r < b; ++r)
        for (int c = 0; c < b; ++c) Bt[r + (size_t)b * c] = Bprev[c + (size_t)b * r];
      upload_small(Bt);
      gemm_nn(H, n, b, b, pro ? Pw.get() : Qj - (size_t)n * b, n, small.get(), b, U.get(), n, -1.0, 1.0);
    }
    // A_j = Q_j^T U
    gemm_tn(H, b, b, n, Qj, n, U.get(), n, small.get(), b, 1.0, 0.0);
    HIP_CHECK(hipMemcpyAsync(A.data(), small.get(), sizeof(double) * b * b, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, tm.a, tm.b));
    t_sigma += ms * 1e-3;
    const size_t c0 = (size_t)j * b;
    for (int c = 0; c < b; ++c)
      for (int r = 0; r < b; ++r) {
        const double a = (r <= c) ? A[r + (size_t)b * c] : A[c + (size_t)b * r];
        T[(c0 + r) + (size_t)nvec * (c0 + c)] = a;
      }
    if (j + 1 == m) break;  // T leaves out the last cycle's B
    // R = U - Q_j A_j
    upload_small(A);
    gemm_nn(H, n, b, b, Qj, n, small.get(), b, U.get(), n, -1.0, 1.0);
    // shifted Cholesky QR: X2 = U R1^-1
    gram_to_host(U.get(), G);
    double fro2 = 0.0;
    for (int i = 0; i < b; ++i) fro2 += G[i + (size_t)b * i];
    const double shift = 11.0 * ((double)n * b + (double)b * (b + 1)) * u * fro2;
    for (int i = 0; i < b; ++i) G[i + (size_t)b * i] += shift;
    if (!cholesky_upper(b, G.data(), R1.data()))
      throw Error(ADC2G_EBREAK, "block Lanczos: residual block of cycle " + std::to_string(j + 1) +
                                    " is not finite or has no positive Gram matrix");
    upper_inverse(b, R1.data(), Ri.data());
    upload_small(Ri);
    gemm_nn(H, n, b, b, U.get(), n, small.get(), b, X2.get(), n, 1.0, 0.0);
    // two Cholesky QR passes: U = X2 R2^-1, Q_{j+1} = U R3^-1
    double* Qn = L.get() + (size_t)n * b * (j + 1);
    Qn_cur = Qn;
    for (int pass = 0; pass < 2; ++pass) {
      const double* src = pass == 0 ? X2.get() : U.get();
      double* dst = pass == 0 ? U.get() : Qn;
      std::vector<double>& Rk = pass == 0 ? R2 : R3;
      gram_to_host(src, G);
      if (!cholesky_upper(b, G.data(), Rk.data()))
        throw Error(ADC2G_EBREAK, "block Lanczos: the residual block of cycle " + std::to_string(j + 1) +
                                      " lost rank (the Krylov space is exhausted)");
      upper_inverse(b, Rk.data(), Ri.data());
      upload_small(Ri);
      gemm_nn(H, n, b, b, src, n, small.get(), b, dst, n, 1.0, 0.0);
    }
    // B_j = R3 R2 R1
    upper_mul(b, R2.data(), R1.data(), W1.data());
    upper_mul(b, R3.data(), W1.data(), Bprev.data());
    if (pro) {
      // the working Q_j (the stored block keeps its value for the Ritz vectors)
      HIP_CHECK(hipMemcpyAsync(Pw.get(), Qj, sizeof(double) * n * b, hipMemcpyDeviceToDevice, st));
      A_all.push_back(A);
      B_all.push_back(Bprev);
      const int jj = j + 1;  // the reference's 1-based cycle
      bool lro = false;
      if (jj > 1 && jj < m) {
        // update_omega (:1530-1639)
        if (jj == 2) {
          anorm[1] = svd_ends(A_all[0]).first;
          bnorm[1] = svd_ends(B_all[0]).first;
          const auto s2 = svd_ends(B_all[1]);
          bnorm[2] = s2.first;
          OM(1, 1) = OM(2, 2) = om_eps;
          OM(1, 2) = OM(2, 1) = om_eps / s2.second;
        }
        anorm[jj] = svd_ends(A_all[jj - 1]).first;
        const auto sb = svd_ends(B_all[jj - 1]);
        bnorm[jj] = sb.first;
        const double inv = 1.0 / sb.second;
        OM(jj, jj + 1) = OM(jj + 1, jj) = OM(jj + 1, jj + 1) = om_eps;
        for (int k = 1; k < jj; ++k) {
          double x = bnorm[k + 1] * OM(jj, k + 1);
          if (k > 1) x += bnorm[k] * OM(jj, k - 1);
          x += bnorm[jj] * OM(jj - 1, k) + (anorm[jj] + anorm[k]) * OM(jj, k);
          OM(jj + 1, k) = OM(k, jj + 1) = inv * x;
        }
        if (p.orthotype == 2) {
          lro = mpro(jj);
        } else {
          for (int k = 1; k < jj; ++k)
            if (OM(jj + 1, k) >= orthlim) lro = true;
        }
        if (lro && p.orthotype == 1) {
          // pro_mgs (:1643-1706): Q_j and Q_{j+1} against the stored blocks 1..j-1,
          // Q_j normalised, then localro
          ++npro;
          for (int k = 0; k + 1 < jj; ++k) {
            const double* Lk = L.get() + (size_t)n * b * k;
            project_out(Lk, Pw.get());
            project_out(Lk, Qn);
          }
          normalise_cols(Pw.get());
          for (int k = 1; k <= jj; ++k) OM(jj + 1, k) = OM(jj, k) = om_eps;
        }
      }
      if (p.orthotype == 1) {
        // localro (:1430-1470): Q_{j+1} against Q_j, normalised
        project_out(Pw.get(), Qn);
        normalise_cols(Qn);
      }
    }
    for (int c = 0; c < b; ++c)
      for (int r = 0; r < b; ++r) {
        const double v = Bprev[r + (size_t)b * c];
        T[(c0 + b + r) + (size_t)nvec * (c0 + c)] = v;
        T[(c0 + c) + (size_t)nvec * (c0 + b + r)] = v;
      }
  }
  if (tmat) std::copy(T.begin(), T.end(), tmat);

  // eigenpairs of T (host)
  const auto t_eig = clk::now();
  std::vector<double> Y((size_t)nvec * nvec);
  if (!host_syev(nvec, T.data(), eigval, Y.data()))
    throw Error(ADC2G_EEIG, "block Lanczos: eigen-decomposition of T did not converge");
  const double time_tdiag = std::chrono::duration<double>(clk::now() - t_eig).count();

  // Ritz vectors, their norms and transition moments
  const auto t_ritz = clk::now();
  if (ritz || travec) {
    DevBuf<double> Yd, Rc, f, part, sums, scale;
    Yd.alloc((size_t)nvec * nvec);
    HIP_CHECK(hipMemcpyAsync(Yd.get(), Y.data(), sizeof(double) * Y.size(), hipMemcpyHostToDevice, st));
    const int cw = std::min(nvec, LZ_CHUNK);
    Rc.alloc((size_t)n * cw);
    if (travec) {
      f.alloc((size_t)n);
      HIP_CHECK(hipMemcpyAsync(f.get(), travec, sizeof(double) * n, hipMemcpyHostToDevice, st));
    }
    const int nblk = reduce_blocks(n);
    part.alloc((size_t)nblk * 2 * cw);
    sums.alloc((size_t)2 * cw);
    scale.alloc((size_t)cw);
    std::vector<double> hs((size_t)2 * cw), hsc((size_t)cw);
    for (int c0 = 0; c0 < nvec; c0 += cw) {
      const int nc = std::min(cw, nvec - c0);
      gemm_nn(H, n, nc, nvec, L.get(), n, Yd.get() + (size_t)nvec * c0, nvec, Rc.get(), n, 1.0, 0.0);
      hipLaunchKernelGGL(ritz_stats_kernel, dim3(nblk, nc), dim3(LZ_THREADS), 0, st, n, nc, Rc.get(), n,
                         travec ? f.get() : nullptr, part.get());
      HIP_CHECK(hipGetLastError());
      launch_colsum(nblk, 2 * nc, part.get(), sums.get(), st);
      HIP_CHECK(hipMemcpyAsync(hs.data(), sums.get(), sizeof(double) * 2 * nc, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      for (int c = 0; c < nc; ++c) {
        const double nrm = std::sqrt(hs[2 * c]);
        hsc[c] = 1.0 / nrm;
        if (tranmom) tranmom[c0 + c] = hs[2 * c + 1] / nrm;
      }
      if (ritz) {
        HIP_CHECK(hipMemcpyAsync(scale.get(), hsc.data(), sizeof(double) * nc, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(scale_cols_kernel, dim3(nblk, nc), dim3(LZ_THREADS), 0, st, n, Rc.get(), n, scale.get());
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipMemcpyAsync(ritz + (size_t)n * c0, Rc.get(), sizeof(double) * n * nc, hipMemcpyDeviceToHost,
                                 st));
        HIP_CHECK(hipStreamSynchronize(st));
      }
    }
  }
  const double time_ritz = std::chrono::duration<double>(clk::now() - t_ritz).count();

  double orth_loss = -1.0;
  if (p.check_orth) {
    DevBuf<double> Gd;
    Gd.alloc((size_t)nvec * nvec);
    gemm_tn(H, nvec, nvec, n, L.get(), n, L.get(), n, Gd.get(), nvec, 1.0, 0.0);
    std::vector<double> Gh((size_t)nvec * nvec);
    HIP_CHECK(hipMemcpyAsync(Gh.data(), Gd.get(), sizeof(double) * Gh.size(), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    orth_loss = 0.0;
    for (int j = 0; j < nvec; ++j)
      for (int i = 0; i < nvec; ++i)
        orth_loss = std::max(orth_loss, std::fabs(Gh[i + (size_t)nvec * j] - (i == j ? 1.0 : 0.0)));
  }
  HIP_CHECK(hipStreamSynchronize(st));
  if (stats) {
    stats->nvec = nvec;
    stats->nsigma = nvec;
    stats->time_sigma = t_sigma;
    stats->time_tdiag = time_tdiag;
    stats->time_ritz = time_ritz;
    stats->time_total = std::chrono::duration<double>(clk::now() - t_start).count();
    stats->orth_loss = orth_loss;
    stats->npro = npro;
    stats->reserved = 0;
  }
}

}  // namespace adc2g
