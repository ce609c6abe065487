// Coupled (Schur-complement) preconditioner for the block Davidson.
//
// The reference's correction steps (dpr / olsen, source/eigen/block_davidson.f90:1288-1419)
// divide the residual by theta - H_jj element by element, which ignores the
// 1h1p<->2h2p coupling C.  In ADC(2)-s the 2h2p block is diagonal (D), so
//
//     M(theta) = [[H11 - theta, C], [C^T, D - sigma]]
//
// -- the shifted matrix with the 2h2p diagonal shifted by a fixed sigma
// instead of theta -- is inverted exactly through the 1h1p Schur complement:
//
//     S(theta) = H11 - C P C^T - theta = A0 - theta,   P = (D - sigma)^-1,
//     t1 = S^-1 (r1 - C P r2),   t2 = P (r2 - C^T t1).
//
// A0 is n1 x n1.  It is formed once per sigma from sigma-vector halves of the
// n1 unit vectors of the 1h1p space (the 2h2p rows of H e_j give C^T e_j, the
// 1h1p rows of H [e_j; -P C^T e_j] give A0 e_j) and eigen-decomposed
// (rocSOLVER dsyevd), so every Ritz value theta costs only
// S^-1 = Q (Lambda - theta)^-1 Q^T.  Applying the step takes one down and one
// up half sigma-vector per column.  For roots well below every 2h2p diagonal
// element (the collective 2h2p states of these systems) the step is close to
// the exact shift-and-invert of H, and the Davidson iteration count drops by
// an order of magnitude.
//
// sigma is chosen by Newton steps on the partitioned eigenvalue problem
// (Lowdin): for the eigenpairs (lambda_i, q_i) of A0(sigma), a root of
// lambda_i(A0(w)) = w is estimated as w_i = sigma + (lambda_i - sigma) /
// (1 + |P C^T q_i|^2), and sigma moves to the mean of the wanted w_i (kept
// below min D by a margin).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>

#include "coupled.hpp"

#include <type_traits>
#include "linalg.hpp"
#include "solvers.hpp"

namespace adc2g {

namespace {

constexpr int CP_THREADS = 256;

// dst[n1 + j, c] = sign * src[n1 + j, c] / (hii[n1 + j] - sigma) for the 2h2p
// rows; the 1h1p rows of dst are zeroed when zero1 (else left alone).
__global__ void cp_scale2_kernel(int64_t n1, int64_t n, int nb, const double* __restrict__ hii, double sigma,
                                 double sign, const double* __restrict__ src, int64_t lds, double* __restrict__ dst,
                                 int64_t ldd, int zero1) {
  const int c = blockIdx.y;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (i < n1) {
      if (zero1) dst[i + ldd * c] = 0.0;
    } else {
      dst[i + ldd * c] = sign * src[i + lds * c] / (hii[i] - sigma);
    }
  }
}

// T[n1 + j, c] = (R[n1 + j, c] - U[n1 + j, c]) / (hii[n1 + j] - sigma)
// (T may be R: each element is read before it is written, by the same thread)
__global__ void cp_final2_kernel(int64_t n1, int64_t n, const double* __restrict__ hii, double sigma,
                                 const double* R, int64_t ldr, const double* __restrict__ U, int64_t ldu,
                                 double* T, int64_t ldt) {
  const int c = blockIdx.y;
  for (int64_t i = n1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    T[i + ldt * c] = (R[i + ldr * c] - U[i + ldu * c]) / (hii[i] - sigma);
}

// B[i + n1 c] = R[i + ldr c] - Wd[i + ldw c], i < n1
__global__ void cp_rhs1_kernel(int n1, const double* __restrict__ R, int64_t ldr, const double* __restrict__ Wd,
                               int64_t ldw, double* __restrict__ B) {
  const int c = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n1) B[i + (int64_t)n1 * c] = R[i + ldr * c] - Wd[i + ldw * c];
}

// Cm[j + n1 c] /= guard(lam[j] - theta[c])
__global__ void cp_shift_div_kernel(int n1, const double* __restrict__ lam, const double* __restrict__ theta,
                                    double* __restrict__ Cm) {
  const int c = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n1) return;
  double den = lam[j] - theta[c];
  if (fabs(den) < 1e-10) den = den < 0.0 ? -1e-10 : 1e-10;
  Cm[j + (int64_t)n1 * c] /= den;
}

// A = (A + A^T) / 2 in place (n x n, column-major)
__global__ void cp_symmetrize_kernel(int n, double* __restrict__ A) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * n) return;
  const int i = (int)(t % n), j = (int)(t / n);
  if (i >= j) return;
  const double m = 0.5 * (A[i + (int64_t)n * j] + A[j + (int64_t)n * i]);
  A[i + (int64_t)n * j] = m;
  A[j + (int64_t)n * i] = m;
}

// part[blk + nblk c] = sum over the 2h2p rows of (U[.,c] / (hii - sigma))^2
__global__ void cp_pnorm_kernel(int64_t n1, int64_t n, const double* __restrict__ hii, double sigma,
                                const double* __restrict__ U, int64_t ldu, double* __restrict__ part) {
  const int c = blockIdx.y;
  double s = 0.0;
  for (int64_t i = n1 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double y = U[i + ldu * c] / (hii[i] - sigma);
    s += y * y;
  }
  __shared__ double red[CP_THREADS / 64];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < CP_THREADS / 64; ++k) t += red[k];
    part[blockIdx.x + (int64_t)gridDim.x * c] = t;
  }
}

// Pair tensors of the 2h2p vectors P C^T e_j for the 1h1p unit vectors
// j = j0 .. j0+nb-1 (j = a'*o + k'), virtual-major layout Z[c][vec][k][(b,l)]
// (the layout the INT down contraction reads).  With the up map
//   G(c,b,k,l) = [k = k'] (c a'|b l) - [c = a'] (k k'|b l)
// (C^T e in the pair space) and CSF -> pair packing after pair -> CSF
// projection being the symmetrisation 2G + 2G^(bc,lk) - G^(bc) - G^(lk), the
// packed tensor of P C^T e is
//   Z = p (2G(c,b,k,l) + 2G(b,c,l,k) - G(b,c,k,l) - G(c,b,l,k)),
//   p = 1 / (e_c + e_b - e_k - e_l - sigma),
// so neither the up contraction nor the CSF vector is formed.  V, O: the FP64
// integral blocks V[a][c][(b,l)] = (ac|bl), O[k][m][(b,l)] = (km|bl).
template <class T>
__global__ void __launch_bounds__(CP_THREADS) cp_unit_pairs_kernel(int o, int v, int nb, int j0,
                                                                   const double* __restrict__ V,
                                                                   const double* __restrict__ O,
                                                                   const double* __restrict__ eps, double sigma,
                                                                   T* __restrict__ Z) {
  const int64_t n1 = (int64_t)o * v;
  const int row = blockIdx.x;  // (c, vec, k)
  const int k = row % o, vec = (row / o) % nb, c = row / (o * nb);
  const int j = j0 + vec, ap = j / o, kp = j % o;
  const double ec = eps[o + c], ek = eps[k];
  T* zr = Z + (int64_t)row * n1;
  for (int64_t bl = threadIdx.x; bl < n1; bl += CP_THREADS) {
    const int b = (int)(bl / o), l = (int)(bl % o);
    double g1 = 0.0, g2 = 0.0, g3 = 0.0, g4 = 0.0;
    if (k == kp) {
      g1 += V[((int64_t)c * v + ap) * n1 + bl];                  // (c a'|b l)
      g3 += V[((int64_t)b * v + ap) * n1 + (int64_t)c * o + l];  // (b a'|c l)
    }
    if (c == ap) {
      g1 -= O[((int64_t)k * o + kp) * n1 + bl];                  // (k k'|b l)
      g4 -= O[((int64_t)l * o + kp) * n1 + (int64_t)b * o + k];  // (l k'|b k)
    }
    if (l == kp) {
      g2 += V[((int64_t)b * v + ap) * n1 + (int64_t)c * o + k];  // (b a'|c k)
      g4 += V[((int64_t)c * v + ap) * n1 + (int64_t)b * o + k];  // (c a'|b k)
    }
    if (b == ap) {
      g2 -= O[((int64_t)l * o + kp) * n1 + (int64_t)c * o + k];  // (l k'|c k)
      g3 -= O[((int64_t)k * o + kp) * n1 + (int64_t)c * o + l];  // (k k'|c l)
    }
    const double p = 1.0 / (ec + eps[o + b] - ek - eps[l] - sigma);
    zr[bl] = (T)(p * (2.0 * g1 + 2.0 * g2 - g3 - g4));
  }
}

// The same tensor with one workgroup per (c, vec): every V value its o rows
// read -- R[x] = V[c][a'][x] for x = (b,l) and S[b][m] = V[b][a'][c o + m] --
// is staged in LDS by coalesced loads (the per-row kernel above gathers the
// l = k' and b-strided terms one element per cache line), then the o rows
// of n1 elements are written in order.  Each thread walks its (b, l) by a
// fixed stride (no integer division per element), and the BF16/FP32 tensors
// are formed in FP32 arithmetic (their rounding is coarser than its error).
// LDS: (2 n1 + nbas) doubles.
template <class T>
__global__ void __launch_bounds__(CP_THREADS) cp_unit_pairs_lds_kernel(int o, int v, int nb, int j0,
                                                                       const double* __restrict__ V,
                                                                       const double* __restrict__ O,
                                                                       const double* __restrict__ eps, double sigma,
                                                                       T* __restrict__ Z) {
  typedef typename std::conditional<std::is_same<T, double>::value, double, float>::type C;
  extern __shared__ double sm[];
  const int n1 = o * v;
  double* R = sm;
  double* S = sm + n1;
  double* ev = S + n1;
  const int vec = blockIdx.x % nb, c = blockIdx.x / nb;
  const int j = j0 + vec, ap = j / o, kp = j % o;
  const double* vr = V + ((int64_t)c * v + ap) * n1;
  for (int x = threadIdx.x; x < n1; x += CP_THREADS) R[x] = vr[x];
  for (int t = threadIdx.x; t < n1; t += CP_THREADS) {
    const int b = t / o, m = t - b * o;
    S[t] = V[((int64_t)b * v + ap) * n1 + (int64_t)c * o + m];
  }
  for (int t = threadIdx.x; t < o + v; t += CP_THREADS) ev[t] = eps[t];
  __syncthreads();
  const int qs = CP_THREADS / o, rs = CP_THREADS - qs * o;  // the stride CP_THREADS as (b, l) steps
  const int b0 = threadIdx.x / o, l0 = threadIdx.x - b0 * o;
  for (int k = 0; k < o; ++k) {
    T* zr = Z + ((int64_t)(c * nb + vec) * o + k) * n1;
    const C base = (C)(ev[o + c] - ev[k] - sigma);
    int b = b0, l = l0;
    for (int bl = threadIdx.x; bl < n1; bl += CP_THREADS) {
      double g1 = 0.0, g2 = 0.0, g3 = 0.0, g4 = 0.0;
      if (k == kp) {
        g1 += R[bl];          // (c a'|b l)
        g3 += S[b * o + l];   // (b a'|c l)
      }
      if (c == ap) {
        g1 -= O[((int64_t)k * o + kp) * n1 + bl];                  // (k k'|b l)
        g4 -= O[((int64_t)l * o + kp) * n1 + (int64_t)b * o + k];  // (l k'|b k)
      }
      if (l == kp) {
        g2 += S[b * o + k];   // (b a'|c k)
        g4 += R[b * o + k];   // (c a'|b k)
      }
      if (b == ap) {
        g2 -= O[((int64_t)l * o + kp) * n1 + (int64_t)c * o + k];  // (l k'|c k)
        g3 -= O[((int64_t)k * o + kp) * n1 + (int64_t)c * o + l];  // (k k'|c l)
      }
      const C num = (C)(2.0 * g1 + 2.0 * g2 - g3 - g4);
      zr[bl] = (T)(num / (base + (C)(ev[o + b] - ev[l])));
      l += rs;
      b += qs;
      if (l >= o) {
        l -= o;
        ++b;
      }
    }
  }
}

// A[:, j0 + c] = H11[:, j0 + c] - M[:, c] (n1 x nb; H11 symmetric)
__global__ void cp_a0_cols_kernel(int n1, int nb, int j0, const double* __restrict__ H11,
                                  const double* __restrict__ M, double* __restrict__ A) {
  const int c = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n1) A[i + (int64_t)n1 * (j0 + c)] = H11[i + (int64_t)n1 * (j0 + c)] - M[i + (int64_t)n1 * c];
}

inline int stream_grid(int64_t n) {
  int64_t g = (n + CP_THREADS - 1) / CP_THREADS;
  return (int)std::min<int64_t>(std::max<int64_t>(g, 1), 1024);
}

#define ROCBLAS_CHECK(expr)                                                                      \
  do {                                                                                           \
    rocblas_status s_ = (expr);                                                                  \
    if (s_ != rocblas_status_success)                                                            \
      throw ::adc2g::Error(ADC2G_EHIP, std::string(#expr) + " failed: rocblas status " + std::to_string((int)s_)); \
  } while (0)

}  // namespace

CoupledPrecon::~CoupledPrecon() {
  if (rb) (void)rocblas_destroy_handle((rocblas_handle)rb);
}

void CoupledPrecon::init(Handle& Hd, Prec p) {
  H = &Hd;
  prec = p;
  n1 = (int)Hd.d.n1;
  n = Hd.d.ndim;
  ARG_CHECK(!Hd.parent, "the coupled preconditioner works on the full operator only");
  // the 2h2p diagonal minimum: min over a,b,k,l of e_a + e_b - e_k - e_l
  std::vector<double> e(Hd.d.o + Hd.d.v);
  HIP_CHECK(hipMemcpyAsync(e.data(), Hd.eps.get(), sizeof(double) * e.size(), hipMemcpyDeviceToHost, Hd.st));
  HIP_CHECK(hipStreamSynchronize(Hd.st));
  const double eo_max = *std::max_element(e.begin(), e.begin() + Hd.d.o);
  const double ev_min = *std::min_element(e.begin() + Hd.d.o, e.end());
  dmin = 2.0 * ev_min - 2.0 * eo_max;
  if (Hd.adcx) {
    // ADC(2)-x: the step is that of the matrix with the 2h2p block replaced
    // by its diagonal (hii: D plus the first-order block's diagonal), whose
    // minimum bounds the shift
    std::vector<double> h2(n - n1);
    HIP_CHECK(hipMemcpyAsync(h2.data(), Hd.hii.get() + n1, sizeof(double) * h2.size(), hipMemcpyDeviceToHost,
                             Hd.st));
    HIP_CHECK(hipStreamSynchronize(Hd.st));
    dmin = *std::min_element(h2.begin(), h2.end());
  }
  nc = std::min(64, std::max(1, Hd.chunk));
  Xb.alloc((size_t)n * nc);
  Wb.alloc((size_t)n * nc);
  Qm.alloc((size_t)n1 * n1);
  lam.alloc(n1);
  offd.alloc(n1);
  Bm.alloc((size_t)n1 * nc);
  Cm.alloc((size_t)n1 * nc);
  th.alloc(nc);
  part.alloc((size_t)1024 * nc);
  nrm.alloc(nc);
  info.alloc(1);
  idx.alloc(n1);
  std::vector<int64_t> iota(n1);
  for (int i = 0; i < n1; ++i) iota[i] = i;
  HIP_CHECK(hipMemcpyAsync(idx.get(), iota.data(), sizeof(int64_t) * n1, hipMemcpyHostToDevice, Hd.st));
  if (!rb) {
    rocblas_handle h = nullptr;
    ROCBLAS_CHECK(rocblas_create_handle(&h));
    rb = h;
  }
  ROCBLAS_CHECK(rocblas_set_stream((rocblas_handle)rb, Hd.st));
  HIP_CHECK(hipStreamSynchronize(Hd.st));
}

// A0 = H11 - C P C^T into Qm (column-major, symmetrised).
void CoupledPrecon::form_schur(double s, Prec bp) {
  using clk = std::chrono::steady_clock;
  Handle& Hd = *H;
  hipStream_t st = Hd.st;
  const auto t0 = clk::now();
  sigma = s;
  const dim3 blk(CP_THREADS);
  if (Hd.form == Handle::FORM_INT && !Hd.adcx) {
    // pair tensors of P C^T e_j straight from the integral blocks, then the
    // down contractions: A0 e_j = H11 e_j - C P C^T e_j (P from orbital
    // energies; with ADC(2)-x the half sigma-vectors below use hii)
    Hd.ensure_lowprec(bp);
    const size_t es = bp == Prec::F64 ? 8 : (bp == Prec::F32 ? 4 : 2);
    Hd.ZG.alloc((size_t)n1 * n1 * nc * es);
    const double* V = (const double*)Hd.is[0].V.get();
    const double* O = (const double*)Hd.is[0].O.get();
    const int o = Hd.d.o, v = Hd.d.v;
    for (int j0 = 0; j0 < n1; j0 += nc) {
      const int nb = std::min(nc, n1 - j0);
      // staged variant when its LDS fits 64 KB (o = 16, v = 200: 52 KB)
      const size_t lds = sizeof(double) * ((size_t)2 * n1 + o + v);
      if (lds <= 65536) {
        const dim3 grid((unsigned)v * nb);
        switch (bp) {
          case Prec::F64:
            hipLaunchKernelGGL(cp_unit_pairs_lds_kernel<double>, grid, blk, lds, st, o, v, nb, j0, V, O, Hd.eps.get(),
                               sigma, (double*)Hd.ZG.get());
            break;
          case Prec::F32:
            hipLaunchKernelGGL(cp_unit_pairs_lds_kernel<float>, grid, blk, lds, st, o, v, nb, j0, V, O, Hd.eps.get(),
                               sigma, (float*)Hd.ZG.get());
            break;
          case Prec::BF16:
            hipLaunchKernelGGL(cp_unit_pairs_lds_kernel<__bf16>, grid, blk, lds, st, o, v, nb, j0, V, O,
                               Hd.eps.get(), sigma, (__bf16*)Hd.ZG.get());
            break;
        }
      } else {
        const dim3 grid((unsigned)v * nb * o);
        switch (bp) {
          case Prec::F64:
            hipLaunchKernelGGL(cp_unit_pairs_kernel<double>, grid, blk, 0, st, o, v, nb, j0, V, O, Hd.eps.get(), sigma,
                               (double*)Hd.ZG.get());
            break;
          case Prec::F32:
            hipLaunchKernelGGL(cp_unit_pairs_kernel<float>, grid, blk, 0, st, o, v, nb, j0, V, O, Hd.eps.get(), sigma,
                               (float*)Hd.ZG.get());
            break;
          case Prec::BF16:
            hipLaunchKernelGGL(cp_unit_pairs_kernel<__bf16>, grid, blk, 0, st, o, v, nb, j0, V, O, Hd.eps.get(),
                               sigma, (__bf16*)Hd.ZG.get());
            break;
        }
      }
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemsetAsync(Bm.get(), 0, sizeof(double) * n1 * nb, st));
      Hd.int_down(Hd.ZG.get(), nb, bp, Bm.get(), n1, st);
      hipLaunchKernelGGL(cp_a0_cols_kernel, dim3((n1 + 255) / 256, nb), dim3(256), 0, st, n1, nb, j0, Hd.H11.get(),
                         Bm.get(), Qm.get());
      HIP_CHECK(hipGetLastError());
    }
  } else {
    for (int j0 = 0; j0 < n1; j0 += nc) {
      const int nb = std::min(nc, n1 - j0);
      // X = [e_j; 0] -> 2h2p rows of H X = C^T e_j
      HIP_CHECK(hipMemsetAsync(Xb.get(), 0, sizeof(double) * n * nb, st));
      launch_unit_vectors(nb, idx.get() + j0, Xb.get(), n, st);
      Hd.sigma_part(Xb.get(), n, Wb.get(), n, nb, bp, Handle::SIGMA_UP, st);
      // X = [e_j; -P C^T e_j] -> 1h1p rows of H X = H11 e_j - C P C^T e_j = A0 e_j
      hipLaunchKernelGGL(cp_scale2_kernel, dim3(stream_grid(n), nb), blk, 0, st, (int64_t)n1, n, nb, Hd.hii.get(),
                         sigma, -1.0, Wb.get(), n, Xb.get(), n, 0);
      HIP_CHECK(hipGetLastError());
      Hd.sigma_part(Xb.get(), n, Wb.get(), n, nb, bp, Handle::SIGMA_DOWN, st);
      HIP_CHECK(hipMemcpy2DAsync(Qm.get() + (size_t)n1 * j0, sizeof(double) * n1, Wb.get(), sizeof(double) * n,
                                 sizeof(double) * n1, nb, hipMemcpyDeviceToDevice, st));
    }
  }
  const int64_t nn = (int64_t)n1 * n1;
  hipLaunchKernelGGL(cp_symmetrize_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, n1, Qm.get());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(st));
  t_build += std::chrono::duration<double>(clk::now() - t0).count();
}

void CoupledPrecon::build(double s, Prec bp) {
  using clk = std::chrono::steady_clock;
  form_schur(s, bp);
  hipStream_t st = H->st;
  const auto t1 = clk::now();
  ROCBLAS_CHECK(rocsolver_dsyevd((rocblas_handle)rb, rocblas_evect_original, rocblas_fill_upper, n1, Qm.get(), n1,
                                 lam.get(), offd.get(), info.get()));
  int inf = 0;
  HIP_CHECK(hipMemcpyAsync(&inf, info.get(), sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  if (inf != 0) throw Error(ADC2G_EEIG, "eigen-decomposition of the 1h1p Schur complement failed");
  const auto t2 = clk::now();
  t_eig += std::chrono::duration<double>(t2 - t1).count();
  nsigma_half += 2 * n1;
  ++builds;
}

// Root estimates w_i = sigma + (lambda_i - sigma) / (1 + |P C^T q_i|^2) of the
// k lowest eigenpairs of A0 (one Newton step on lambda_i(A0(w)) = w).
std::vector<double> CoupledPrecon::newton_roots(int k) {
  Handle& Hd = *H;
  hipStream_t st = Hd.st;
  k = std::min(k, n1);
  std::vector<double> lh(n1), w;
  HIP_CHECK(hipMemcpyAsync(lh.data(), lam.get(), sizeof(double) * n1, hipMemcpyDeviceToHost, st));
  for (int j0 = 0; j0 < k; j0 += nc) {
    const int nb = std::min(nc, k - j0);
    HIP_CHECK(hipMemsetAsync(Xb.get(), 0, sizeof(double) * n * nb, st));
    HIP_CHECK(hipMemcpy2DAsync(Xb.get(), sizeof(double) * n, Qm.get() + (size_t)n1 * j0, sizeof(double) * n1,
                               sizeof(double) * n1, nb, hipMemcpyDeviceToDevice, st));
    Hd.sigma_part(Xb.get(), n, Wb.get(), n, nb, prec, Handle::SIGMA_UP, st);
    nsigma_half += nb;
    const int g = stream_grid(n - n1);
    hipLaunchKernelGGL(cp_pnorm_kernel, dim3(g, nb), dim3(CP_THREADS), 0, st, (int64_t)n1, n, Hd.hii.get(), sigma,
                       Wb.get(), n, part.get());
    HIP_CHECK(hipGetLastError());
    launch_colsum(g, nb, part.get(), nrm.get(), st);
    std::vector<double> y(nb);
    HIP_CHECK(hipMemcpyAsync(y.data(), nrm.get(), sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    for (int i = 0; i < nb; ++i) w.push_back(sigma + (lh[j0 + i] - sigma) / (1.0 + y[i]));
  }
  return w;
}

// sigma for the k wanted roots.  The first Schur complement is formed just
// below min D.  Its inertia counts the roots of H below sigma, and Newton
// estimates of the k lowest roots from its eigenpairs move sigma to their
// mean.  Next to the pole at min D the first estimates are poor (the steep
// (D - sigma)^-1 damps the step), so the first step is doubled; from below
// the roots the estimates are close.  Once at most 4k roots lie below sigma,
// one more Newton step gives the final sigma.  Every A0 is formed with BF16
// contractions (FP32 accumulation, FP64 sum with H11): a preconditioner needs
// no more, and at o=16, v=200 the FP32 one converged in 21 iterations against
// 23 for BF16 while costing twice as much to form.
void CoupledPrecon::choose_sigma(int k, int max_search) {
  const double margin = 0.02 * std::max(1.0, std::fabs(dmin));
  double s = dmin - margin;
  std::vector<double> lh(n1);
  for (int b = 0; b < max_search; ++b) {
    build(s, Prec::BF16);
    HIP_CHECK(hipMemcpy(lh.data(), lam.get(), sizeof(double) * n1, hipMemcpyDeviceToHost));
    int below = 0;
    for (double x : lh) below += x < sigma ? 1 : 0;
    std::vector<double> w = newton_roots(k);
    double m = 0.0;
    for (double x : w) m += x;
    m /= (double)w.size();
    if (verbose)
      std::fprintf(stderr, "coupled: search sigma %.6f roots below %d, Newton estimates %.6f .. %.6f mean %.6f "
                   "(build %.3f s, eig %.3f s so far)\n", sigma, below, w.front(), w.back(), m, t_build, t_eig);
    if (below <= 4 * k) {
      s = std::min(m, dmin - margin);
      break;
    }
    s = std::min(b == 0 ? 2.0 * m - sigma : m, dmin - margin);
  }
  build(s, Prec::BF16);
  if (verbose) std::fprintf(stderr, "coupled: sigma %.6f (%d Schur complements, build %.3f s, eig %.3f s)\n", sigma,
                            builds, t_build, t_eig);
}

// During the iteration: when every active wanted Ritz value (an upper bound of
// its root) lies below sigma by more than their spread, the wanted roots sit
// well below the shift; sigma moves to their mean and A0 is formed again.
bool CoupledPrecon::track(const double* theta, int nact) {
  if (nact <= 0 || retracks >= 2) return false;
  double lo = theta[0], hi = theta[0], m = 0.0;
  for (int i = 0; i < nact; ++i) {
    lo = std::min(lo, theta[i]);
    hi = std::max(hi, theta[i]);
    m += theta[i];
  }
  m /= nact;
  if (!(hi + (hi - lo) < sigma)) return false;
  ++retracks;
  const double margin = 0.02 * std::max(1.0, std::fabs(dmin));
  build(std::min(m, dmin - margin), prec);
  if (verbose) std::fprintf(stderr, "coupled: Ritz values %.6f .. %.6f below the shift: sigma %.6f\n", lo, hi, sigma);
  return true;
}

// T[:, c] = M(theta_c)^-1 R[:, c] for nb <= nc columns (theta on the host).
// T may be R (with ldt = ldr): R's 1h1p rows are consumed into Bm before T's
// are written, and its 2h2p rows element by element in cp_final2_kernel.
void CoupledPrecon::apply(int nb, const double* theta_h, const double* R, int64_t ldr, double* T, int64_t ldt) {
  Handle& Hd = *H;
  hipStream_t st = Hd.st;
  ARG_CHECK(nb >= 1 && nb <= nc, "coupled step: block too large");
  const dim3 blk(CP_THREADS);
  // X = [0; P r2] -> 1h1p rows: C P r2
  hipLaunchKernelGGL(cp_scale2_kernel, dim3(stream_grid(n), nb), blk, 0, st, (int64_t)n1, n, nb, Hd.hii.get(), sigma,
                     1.0, R, ldr, Xb.get(), n, 1);
  HIP_CHECK(hipGetLastError());
  Hd.sigma_part(Xb.get(), n, Wb.get(), n, nb, prec, Handle::SIGMA_DOWN, st);
  hipLaunchKernelGGL(cp_rhs1_kernel, dim3((n1 + 255) / 256, nb), dim3(256), 0, st, n1, R, ldr, Wb.get(), n, Bm.get());
  HIP_CHECK(hipGetLastError());
  // t1 = Q (Lambda - theta)^-1 Q^T b, written into the 1h1p rows of X (2h2p rows zero)
  HIP_CHECK(hipMemcpyAsync(th.get(), theta_h, sizeof(double) * nb, hipMemcpyHostToDevice, st));
  gemm_tn(Hd, n1, nb, n1, Qm.get(), n1, Bm.get(), n1, Cm.get(), n1, 1.0, 0.0);
  hipLaunchKernelGGL(cp_shift_div_kernel, dim3((n1 + 255) / 256, nb), dim3(256), 0, st, n1, lam.get(), th.get(),
                     Cm.get());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemsetAsync(Xb.get(), 0, sizeof(double) * n * nb, st));
  gemm_nn(Hd, n1, nb, n1, Qm.get(), n1, Cm.get(), n1, Xb.get(), n, 1.0, 0.0);
  HIP_CHECK(hipMemcpy2DAsync(T, sizeof(double) * ldt, Xb.get(), sizeof(double) * n, sizeof(double) * n1, nb,
                             hipMemcpyDeviceToDevice, st));
  // t2 = P (r2 - C^T t1)
  Hd.sigma_part(Xb.get(), n, Wb.get(), n, nb, prec, Handle::SIGMA_UP, st);
  hipLaunchKernelGGL(cp_final2_kernel, dim3(stream_grid(n - n1), nb), blk, 0, st, (int64_t)n1, n, Hd.hii.get(), sigma,
                     R, ldr, Wb.get(), n, T, ldt);
  HIP_CHECK(hipGetLastError());
  nsigma_half += 2 * nb;
}

}  // namespace adc2g
