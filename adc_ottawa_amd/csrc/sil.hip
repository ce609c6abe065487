// Short-iterative-Lanczos wavepacket propagation on the GPU.
//
// Follows propagate_sillib (source/propagation/fvecprop.f90:365-564) with
// SILStep (source/propagation/sillib.f90:108-393) for the path the ADC driver
// takes (Restart = .true., StdForm = .true., Relax = .false.; PreFactor = -i):
//  * psi0 = F/|F| (init_wavepacket, fvecprop.f90:338-357), records at
//    i*tout for i = 1..int(tfinal/intperiod), intperiod = tout/2, with the
//    t/2 identity a0(2t) = sum_D psi_D(t)^2 |F|^2 (fvecprop.f90:521);
//  * each step: alpha_0 = <psi|dtpsi>, first Krylov vector and error |beta dt|
//    (sillib.f90:198-230); then Krylov vectors K_n = (-iH K_{n-1} - alpha
//    K_{n-1} - cbeta K_{n-2}) / cbeta_n with Err = |Err beta dt / n| until
//    order == IntOrder, Err <= TolError or beta < 1e-20 (:234-323); the next
//    diagonal element estimated as the previous one (:327); stepsize reduced
//    by (TolError/Err)^(1/order) when not converged (:331-334); eigenpairs of
//    the (order+1) tridiagonal (DSTEQR, :350-361) and
//    psi <- sum_p [sum_q exp(-i lambda_q dt) U_0q U_pq] K_p (:369-390).
// Complex vectors are planar on the device (re column, im column), so -iH v
// is one two-column sigma-vector: (Hv_im, -Hv_re).  Inner products, norms and
// updates run on the GPU in FP64; the recursion's scalars drive the control
// flow on the host, as the reference's single-threaded loop does.
#include <chrono>
#include <cmath>
#include <complex>
#include <functional>

#include "gemm.hpp"
#include "linalg.hpp"
#include "solvers.hpp"

namespace adc2g {

namespace {

typedef std::complex<double> cplx;

__device__ double block_sum2(double x, double* sh) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = x;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
  return t;
}

// out = (H v)_im, -(H v)_re : turn the two-column sigma of (re, im) into -i H v
__global__ void minus_i_kernel(int64_t n, const double* __restrict__ Hv, double* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double hr = Hv[i], hi = Hv[n + i];
    out[i] = hi;
    out[n + i] = -hr;
  }
}

// w <- w - a x - b z  (complex scalars, planar vectors of length n)
__global__ void caxpy2_kernel(int64_t n, double* __restrict__ w, double ar, double ai, const double* __restrict__ x,
                              double br, double bi, const double* __restrict__ z) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double wr = w[i], wi = w[n + i];
    const double xr = x[i], xi = x[n + i];
    wr -= ar * xr - ai * xi;
    wi -= ar * xi + ai * xr;
    if (z) {
      const double zr = z[i], zi = z[n + i];
      wr -= br * zr - bi * zi;
      wi -= br * zi + bi * zr;
    }
    w[i] = wr;
    w[n + i] = wi;
  }
}

// x <- c x
__global__ void cscale_kernel(int64_t n, double* __restrict__ x, double cr, double ci) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double xr = x[i], xi = x[n + i];
    x[i] = cr * xr - ci * xi;
    x[n + i] = cr * xi + ci * xr;
  }
}

// out = sum_p coef_p K_p (coef complex, np <= 32)
__global__ void combine_kernel(int64_t n, int np, const double* __restrict__ coef, const double* __restrict__ K,
                               int64_t stride, double* __restrict__ out) {
  __shared__ double c[64];
  if (threadIdx.x < 2 * np) c[threadIdx.x] = coef[threadIdx.x];
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double sr = 0.0, si = 0.0;
    for (int p = 0; p < np; ++p) {
      const double kr = K[p * stride + i], ki = K[p * stride + n + i];
      sr += c[2 * p] * kr - c[2 * p + 1] * ki;
      si += c[2 * p] * ki + c[2 * p + 1] * kr;
    }
    out[i] = sr;
    out[n + i] = si;
  }
}

// partial sums of psi^2 (not conjugated): re^2 - im^2, 2 re im
__global__ void square_sum_kernel(int64_t n, const double* __restrict__ x, double* __restrict__ part) {
  __shared__ double red[16];
  double sr = 0.0, si = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double a = x[i], b = x[n + i];
    sr += a * a - b * b;
    si += 2.0 * a * b;
  }
  sr = block_sum2(sr, red);
  si = block_sum2(si, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sr;
    part[2 * blockIdx.x + 1] = si;
  }
}

// part[blk] = partial sum_k a_k b_k (complex, unconjugated; planar re/im)
__global__ void bilinear_kernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                double* __restrict__ part) {
  __shared__ double red[16];
  double sr = 0.0, si = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double a = x[i], b = x[n + i], c = y[i], d = y[n + i];
    sr += a * c - b * d;
    si += a * d + b * c;
  }
  sr = block_sum2(sr, red);
  si = block_sum2(si, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sr;
    part[2 * blockIdx.x + 1] = si;
  }
}

// coef_p = sum_q exp(-i lam_q dt) U[0,q] U[p,q]   (one thread per p)
__global__ void expcoef_kernel(int np, const double* __restrict__ lam, const double* __restrict__ U, int ldu,
                               double dt, double* __restrict__ coef) {
  const int p = threadIdx.x;
  if (p >= np) return;
  double sr = 0.0, si = 0.0;
  for (int q = 0; q < np; ++q) {
    const double w = U[0 + q * ldu] * U[p + q * ldu];
    sr += cos(lam[q] * dt) * w;
    si += -sin(lam[q] * dt) * w;
  }
  coef[2 * p] = sr;
  coef[2 * p + 1] = si;
}

struct Sil {
  Handle& H;
  int64_t n;
  int kdim;
  Prec prec;
  DevBuf<double> K, psi, hv, tmp, part, sums, tri, U, lam, coef, work;
  DevBuf<int> info;
  int64_t nsigma = 0;
  int nb = 0;

  // the operator -i H v when set (laser propagation: ndim + 1 rows), else
  // the handle's sigma-vector
  std::function<void(const double*, double*)> op;

  Sil(Handle& h, int kd, Prec p, int64_t nrows = 0) : H(h), kdim(kd), prec(p) {
    n = nrows > 0 ? nrows : H.d.ndim;
    K.alloc((size_t)(kdim + 1) * 2 * n);
    psi.alloc(2 * n);
    hv.alloc(2 * n);
    tmp.alloc(2 * n);
    nb = reduce_blocks(n);
    part.alloc(2 * (size_t)nb);
    sums.alloc(4);
    tri.alloc((size_t)(kdim + 1) * (kdim + 1));
    U.alloc((size_t)(kdim + 1) * (kdim + 1));
    lam.alloc(kdim + 1);
    coef.alloc(2 * (kdim + 1));
    work.alloc((size_t)(kdim + 1) * (kdim + 1) + 64);
    info.alloc(2);
  }

  double* Kp(int p) { return K.get() + (int64_t)p * 2 * n; }

  // out = -i H v
  void func(const double* v, double* out) {
    if (op) {
      op(v, out);
      ++nsigma;
      return;
    }
    H.sigma(v, n, hv.get(), n, 2, prec, H.st);
    hipLaunchKernelGGL(minus_i_kernel, dim3(grid_for(n)), dim3(256), 0, H.st, n, hv.get(), out);
    HIP_CHECK(hipGetLastError());
    ++nsigma;
  }

  cplx cdot(const double* a, const double* b) {
    const int nbk = launch_cdot_partial(n, a, b, n, part.get(), H.st);
    launch_colsum(nbk, 2, part.get(), sums.get(), H.st);
    double h[2];
    HIP_CHECK(hipMemcpyAsync(h, sums.get(), sizeof(h), hipMemcpyDeviceToHost, H.st));
    HIP_CHECK(hipStreamSynchronize(H.st));
    return cplx(h[0], h[1]);
  }

  cplx bilinear(const double* a, const double* b) {
    hipLaunchKernelGGL(bilinear_kernel, dim3(nb), dim3(256), 0, H.st, n, a, b, part.get());
    HIP_CHECK(hipGetLastError());
    launch_colsum(nb, 2, part.get(), sums.get(), H.st);
    double h[2];
    HIP_CHECK(hipMemcpyAsync(h, sums.get(), sizeof(h), hipMemcpyDeviceToHost, H.st));
    HIP_CHECK(hipStreamSynchronize(H.st));
    return cplx(h[0], h[1]);
  }

  // out = H v (complex, planar): the i * matxvec_treal of the reference
  void hmul(const double* v, double* out) {
    H.sigma(v, n, out, n, 2, prec, H.st);
    ++nsigma;
  }

  cplx square_sum(const double* a) {
    hipLaunchKernelGGL(square_sum_kernel, dim3(nb), dim3(256), 0, H.st, n, a, part.get());
    HIP_CHECK(hipGetLastError());
    launch_colsum(nb, 2, part.get(), sums.get(), H.st);
    double h[2];
    HIP_CHECK(hipMemcpyAsync(h, sums.get(), sizeof(h), hipMemcpyDeviceToHost, H.st));
    HIP_CHECK(hipStreamSynchronize(H.st));
    return cplx(h[0], h[1]);
  }

  void axpy2(double* w, cplx a, const double* x, cplx b, const double* z) {
    hipLaunchKernelGGL(caxpy2_kernel, dim3(grid_for(n)), dim3(256), 0, H.st, n, w, a.real(), a.imag(), x, b.real(),
                       b.imag(), z);
    HIP_CHECK(hipGetLastError());
  }

  void scale(double* x, cplx c) {
    hipLaunchKernelGGL(cscale_kernel, dim3(grid_for(n)), dim3(256), 0, H.st, n, x, c.real(), c.imag());
    HIP_CHECK(hipGetLastError());
  }

  // exp(A dt) e_0 for the complex (m x m, column-major) matrix A: scaling
  // and squaring with a Taylor series (terms until below 1e-17 of the sum)
  static std::vector<cplx> expm_col0(const std::vector<cplx>& A, int m, double dt) {
    std::vector<cplx> B(A.size());
    double nrm = 0.0;  // 1-norm
    for (int q = 0; q < m; ++q) {
      double c = 0.0;
      for (int p = 0; p < m; ++p) {
        B[p + (size_t)q * m] = A[p + (size_t)q * m] * dt;
        c += std::abs(B[p + (size_t)q * m]);
      }
      nrm = std::max(nrm, c);
    }
    int sq = 0;
    while (nrm > 0.5) {
      nrm *= 0.5;
      ++sq;
    }
    const double sc = std::ldexp(1.0, -sq);
    for (auto& x : B) x *= sc;
    auto mul = [m](const std::vector<cplx>& X, const std::vector<cplx>& Y) {
      std::vector<cplx> Z((size_t)m * m, cplx(0, 0));
      for (int q = 0; q < m; ++q)
        for (int k = 0; k < m; ++k) {
          const cplx y = Y[k + (size_t)q * m];
          if (y == cplx(0, 0)) continue;
          for (int p = 0; p < m; ++p) Z[p + (size_t)q * m] += X[p + (size_t)k * m] * y;
        }
      return Z;
    };
    std::vector<cplx> E((size_t)m * m, cplx(0, 0)), T((size_t)m * m, cplx(0, 0));
    for (int p = 0; p < m; ++p) E[p + (size_t)p * m] = T[p + (size_t)p * m] = cplx(1, 0);
    for (int k = 1; k < 40; ++k) {
      T = mul(T, B);
      double tn = 0.0, en = 0.0;
      for (size_t i = 0; i < T.size(); ++i) {
        T[i] /= (double)k;
        E[i] += T[i];
        tn = std::max(tn, std::abs(T[i]));
        en = std::max(en, std::abs(E[i]));
      }
      if (tn <= 1e-17 * en) break;
    }
    for (int i = 0; i < sq; ++i) E = mul(E, E);
    return std::vector<cplx>(E.begin(), E.begin() + m);
  }

  // One CSILStep (Lanczos-Arnoldi, improved error criterion) on psi =
  // Kp(0) with dtpsi = -i H psi in Kp(1): the Krylov vectors orthogonalised
  // against all earlier ones (modified Gram-Schmidt) give the upper Hessenberg
  // matrix Hs of -i H; after each new order its last column is estimated
  // from the previous one (sub-diagonal mirrored, diagonal repeated) and the
  // coefficients c = exp(Hs dt) e_0 compared with those of the previous
  // order; the order grows until |c - c_old| <= tolerror (at least 2, at most
  // kdim), else the step shrinks by (tolerror / former error)^(1/order) with
  // the former error |err * dt * beta_k / k| accumulated over the orders.
  // psi <- |psi| sum_p c_p K_p.  The exponential is a scaled Taylor series of
  // Hs (the reference diagonalises Hs; the same exp(Hs dt) e_0).
  double csil_step(double intperiod, double tolerror, int& order_out) {
    const double RelativeMinStep = 1.0e-9, Tiny = 1.0e-18;
    const int MinOrder = 2;
    const cplx pref(0.0, -1.0);
    if (kdim < MinOrder) throw Error(ADC2G_ESIL_ORDER, "Illegal integration order");
    double stepsize = intperiod;
    const double minstep = std::fabs(RelativeMinStep * intperiod);
    const double norm = std::sqrt(cdot(Kp(0), Kp(0)).real());
    scale(Kp(0), cplx(1.0 / norm, 0.0));
    scale(Kp(1), cplx(1.0 / norm, 0.0));
    const int ld = kdim + 1;
    std::vector<cplx> Hs((size_t)ld * ld, cplx(0, 0)), cold(ld, cplx(0, 0)), c1;
    double former = 1.0, error = 0.0;
    int order = 0;
    auto hs = [&](int p, int q) -> cplx& { return Hs[p + (size_t)q * ld]; };
    auto coeffs = [&](int m, double dt) {
      std::vector<cplx> A((size_t)m * m);
      for (int q = 0; q < m; ++q)
        for (int p = 0; p < m; ++p) A[p + (size_t)q * m] = hs(p, q);
      return expm_col0(A, m, dt);
    };
    while (true) {
      if (std::fabs(stepsize) < std::fabs(minstep)) throw Error(ADC2G_ESIL_UNDERFLOW, "Stepsize underflow");
      const int next = order + 1;
      double* w = Kp(next);
      for (int p = 0; p <= order; ++p) {
        const cplx c = cdot(Kp(p), w);
        hs(p, order) = c;
        axpy2(w, c, Kp(p), cplx(0, 0), nullptr);
      }
      double sub = std::sqrt(cdot(w, w).real());
      hs(next, order) = pref * sub;
      cplx cinv(0, 0);
      if (sub < Tiny) {
        sub = 0.0;
      } else {
        cinv = 1.0 / hs(next, order);
      }
      scale(w, cinv);
      order = next;
      former = std::fabs(former * stepsize * sub / (double)order);
      hs(order - 1, order) = hs(order, order - 1);
      hs(order, order) = hs(order - 1, order - 1);
      c1 = coeffs(order + 1, stepsize);
      error = 0.0;
      if (order > 1) {
        for (int p = 0; p <= order; ++p) error += std::norm(cold[p] - c1[p]);
        error = std::sqrt(error);
      }
      for (int p = 0; p <= order; ++p) cold[p] = c1[p];
      if ((error > tolerror && order < kdim) || order < MinOrder) {
        func(Kp(order), Kp(order + 1));
        continue;
      }
      break;
    }
    if (!(error <= tolerror)) {
      stepsize = std::min(stepsize, stepsize * std::pow(tolerror / former, 1.0 / (double)order));
      c1 = coeffs(order + 1, stepsize);
    }
    std::vector<double> ch(2 * (order + 1));
    for (int p = 0; p <= order; ++p) {
      ch[2 * p] = norm * c1[p].real();
      ch[2 * p + 1] = norm * c1[p].imag();
    }
    HIP_CHECK(hipMemcpyAsync(coef.get(), ch.data(), sizeof(double) * ch.size(), hipMemcpyHostToDevice, H.st));
    hipLaunchKernelGGL(combine_kernel, dim3(grid_for(n)), dim3(256), 0, H.st, n, order + 1, coef.get(), K.get(), 2 * n,
                       tmp.get());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(Kp(0), tmp.get(), sizeof(double) * 2 * n, hipMemcpyDeviceToDevice, H.st));
    HIP_CHECK(hipStreamSynchronize(H.st));
    order_out = order;
    return stepsize;
  }

  // One SILStep on psi (dtpsi = -iH psi already in Kp(1)); returns stepsize.
  double step(double intperiod, double tolerror, int& order_out) {
    const double RelativeMinStep = 1.0e-10;
    const cplx pref(0.0, -1.0);
    if (kdim < 2) throw Error(ADC2G_ESIL_ORDER, "IntOrder is too small");
    double stepsize = intperiod;
    const double minstep = std::fabs(RelativeMinStep * intperiod);
    std::vector<double> diag(kdim + 2, 0.0), offd(kdim + 2, 0.0);
    // Kp(0) = psi, Kp(1) = dtpsi
    cplx alpha = cdot(Kp(0), Kp(1));
    diag[0] = (alpha / pref).real();
    axpy2(Kp(1), alpha, Kp(0), cplx(0, 0), nullptr);
    double beta = std::sqrt(cdot(Kp(1), Kp(1)).real());
    offd[1] = beta;
    cplx cbeta = pref * beta;
    cplx cinv(0, 0);
    if (beta < 1.0e-20) {
      beta = 0.0;
    } else {
      cinv = 1.0 / cbeta;
    }
    scale(Kp(1), cinv);
    double error = std::fabs(beta * stepsize);
    int order = 1;
    while (true) {
      if (std::fabs(stepsize) < std::fabs(minstep)) throw Error(ADC2G_ESIL_UNDERFLOW, "Stepsize underflow");
      ++order;
      func(Kp(order - 1), Kp(order));
      alpha = cdot(Kp(order - 1), Kp(order));
      diag[order - 1] = (alpha / pref).real();
      axpy2(Kp(order), alpha, Kp(order - 1), cbeta, Kp(order - 2));
      beta = std::sqrt(cdot(Kp(order), Kp(order)).real());
      offd[order] = beta;
      cbeta = pref * beta;
      if (beta < 1.0e-20) {
        offd[order] = 0.0;
        beta = 0.0;
        cinv = cplx(0, 0);
      } else {
        cinv = 1.0 / cbeta;
      }
      scale(Kp(order), cinv);
      error = std::fabs(error * beta * stepsize / (double)order);
      if (!(order < kdim && error > tolerror && beta > 1.0e-20)) break;
    }
    diag[order] = diag[order - 1];
    if (error > tolerror) stepsize = stepsize * std::pow(tolerror / error, 1.0 / (double)order);
    // tridiagonal eigenpairs on the device (DSTEQR('I') of the reference)
    const int np = order + 1;
    std::vector<double> T((size_t)np * np, 0.0);
    for (int i = 0; i < np; ++i) {
      T[i + (size_t)i * np] = diag[i];
      if (i + 1 < np) {
        T[i + 1 + (size_t)i * np] = offd[i + 1];
        T[i + (size_t)(i + 1) * np] = offd[i + 1];
      }
    }
    HIP_CHECK(hipMemcpyAsync(tri.get(), T.data(), sizeof(double) * np * np, hipMemcpyHostToDevice, H.st));
    launch_syev(np, tri.get(), np, lam.get(), U.get(), np, work.get(), info.get(), H.st);
    hipLaunchKernelGGL(expcoef_kernel, dim3(1), dim3(64), 0, H.st, np, lam.get(), U.get(), np, stepsize, coef.get());
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(combine_kernel, dim3(grid_for(n)), dim3(256), 0, H.st, n, np, coef.get(), K.get(), 2 * n,
                       tmp.get());
    HIP_CHECK(hipGetLastError());
    int inf = 0;
    HIP_CHECK(hipMemcpyAsync(&inf, info.get(), sizeof(int), hipMemcpyDeviceToHost, H.st));
    HIP_CHECK(hipMemcpyAsync(Kp(0), tmp.get(), sizeof(double) * 2 * n, hipMemcpyDeviceToDevice, H.st));
    HIP_CHECK(hipStreamSynchronize(H.st));
    if (inf != 0) throw Error(ADC2G_ESIL_DIAG, "Diagonalisation failed");
    order_out = order;
    return stepsize;
  }
};

}  // namespace

void sil_run(Handle& H, const adc2g_sil_params& p, const double* fvec, double* auto0, int64_t nrec_cap,
             int64_t* nrec_out, adc2g_sil_stats* stats, int autoord, double* auto1, double* auto2) {
  const auto t0 = std::chrono::steady_clock::now();
  ARG_CHECK(p.tout > 0 && p.tfinal >= 0, "tout must be > 0 and tfinal >= 0");
  ARG_CHECK(p.kdim >= 1 && p.kdim <= 30, "krydim must be in 1..30");
  ARG_CHECK(p.autotol > 0, "autotol must be > 0");
  if (p.kdim < 2) throw Error(ADC2G_ESIL_ORDER, "IntOrder is too small");
  const double intperiod = 0.5 * p.tout;
  const int64_t nrec = (int64_t)(p.tfinal / intperiod);
  ARG_CHECK(nrec_cap >= nrec + 1, "auto0 holds fewer than nrec+1 records");
  ARG_CHECK(autoord >= 0 && autoord <= 2, "autoord must be 0, 1 or 2");
  ARG_CHECK((autoord < 1 || auto1) && (autoord < 2 || auto2), "missing output array for the requested order");
  Sil S(H, p.kdim, prec_from_int(p.precision));
  DevBuf<double> hpsi, h2psi;
  if (autoord >= 1) hpsi.alloc(2 * (size_t)S.n);
  if (autoord >= 2) h2psi.alloc(2 * (size_t)S.n);
  const int64_t n = S.n;
  // psi0 = F / |F| (complex, imaginary part zero)
  std::vector<double> f(fvec, fvec + n);
  double norm2 = 0.0;
  for (int64_t i = 0; i < n; ++i) norm2 += f[i] * f[i];
  ARG_CHECK(norm2 > 0.0, "initial vector is zero");
  const double norm0 = std::sqrt(norm2);
  for (auto& x : f) x /= norm0;
  HIP_CHECK(hipMemcpyAsync(S.Kp(0), f.data(), sizeof(double) * n, hipMemcpyHostToDevice, H.st));
  HIP_CHECK(hipMemsetAsync(S.Kp(0) + n, 0, sizeof(double) * n, H.st));
  HIP_CHECK(hipStreamSynchronize(H.st));
  // t = 0: dot_product(psi0, psi0) * |F|^2
  const cplx a00 = S.cdot(S.Kp(0), S.Kp(0)) * norm2;
  auto0[0] = a00.real();
  auto0[1] = a00.imag();
  // first and second order at t = 0: dot_product(psi0, H^k psi0) |F|^2
  // (propagate_sillib, fvecprop.f90:456-471)
  if (autoord >= 1) {
    S.hmul(S.Kp(0), hpsi.get());
    const cplx a = S.cdot(S.Kp(0), hpsi.get()) * norm2;
    auto1[0] = a.real();
    auto1[1] = a.imag();
  }
  if (autoord >= 2) {
    S.hmul(hpsi.get(), h2psi.get());
    const cplx a = S.cdot(S.Kp(0), h2psi.get()) * norm2;
    auto2[0] = a.real();
    auto2[1] = a.imag();
  }
  const double tiny = 1e-9;
  int64_t nsteps = 0;
  int maxorder = 0;
  for (int64_t i = 1; i <= nrec; ++i) {
    double inttime = 0.0;
    while (true) {
      const double stepsize = intperiod - inttime;
      S.func(S.Kp(0), S.Kp(1));
      int order = 0;
      const double truestep = S.step(stepsize, p.autotol, order);
      ++nsteps;
      maxorder = std::max(maxorder, order);
      inttime += truestep;
      if (!(std::fabs(intperiod - inttime) > std::fabs(tiny * intperiod))) break;
    }
    const cplx a = S.square_sum(S.Kp(0)) * norm2;
    auto0[2 * i] = a.real();
    auto0[2 * i + 1] = a.imag();
    // orders 1 and 2 with the t/2 trick: psi^T H^k psi |F|^2 (fvecprop.f90:526-541)
    if (autoord >= 1) {
      S.hmul(S.Kp(0), hpsi.get());
      const cplx a1 = S.bilinear(S.Kp(0), hpsi.get()) * norm2;
      auto1[2 * i] = a1.real();
      auto1[2 * i + 1] = a1.imag();
    }
    if (autoord >= 2) {
      S.hmul(hpsi.get(), h2psi.get());
      const cplx a2 = S.bilinear(S.Kp(0), h2psi.get()) * norm2;
      auto2[2 * i] = a2.real();
      auto2[2 * i + 1] = a2.imag();
    }
    if (p.verbose && (i % 100 == 0 || i == nrec)) {
      std::fprintf(stderr, "sil: record %lld of %lld (t = %.2f a.u.), %lld steps, %lld sigma-vectors\n", (long long)i,
                   (long long)nrec, i * intperiod, (long long)nsteps, (long long)S.nsigma);
      std::fflush(stderr);
    }
  }
  if (nrec_out) *nrec_out = nrec + 1;
  if (stats) {
    stats->nsteps = nsteps;
    stats->nsigma = S.nsigma;
    stats->max_order = maxorder;
    stats->reserved = 0;
    stats->time_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
}

// propagate_wavepacket_sil (source/propagation/proplib_adc2.f90:486-630):
// integration period tout; in each period SIL steps from the current time
// until the period is covered (|intperiod - inttime| <= 1e-9 intperiod), each
// with -i H(t) at its start time; a period in which |dpsi/dt| <= 1e-10 is
// skipped (time advances by the whole period); the norm is recorded at the
// start of every period and at the end.
void sil_laser_run(Handle& H, const adc2g_laser_params& p, int npulse, const adc2g_pulse* pulses, double* psi_io,
                   double* norms, double* flux, double* trdens, int64_t nrec_cap, int64_t* nrec_out,
                   adc2g_sil_stats* stats) {
  const auto t0 = std::chrono::steady_clock::now();
  ARG_CHECK(p.tout > 0 && p.tfinal >= 0, "tout must be > 0 and tfinal >= 0");
  ARG_CHECK(p.kdim >= 1 && p.kdim <= 30, "krydim must be in 1..30");
  ARG_CHECK(p.proptol > 0, "proptol must be > 0");
  if (p.kdim < 2) throw Error(ADC2G_ESIL_ORDER, "IntOrder is too small");
  laser_check_pulses(npulse, pulses);
  ARG_CHECK(!flux || (H.cap.set && H.flux_theta.set), "the flux needs a CAP and the Theta operator");
  ARG_CHECK(!trdens || !H.stored, "the transition density needs the operator built from integrals or a view of it");
  const double intperiod = p.tout;
  const int64_t nrec = (int64_t)(p.tfinal / intperiod);
  ARG_CHECK(nrec_cap >= nrec + 1, "norms holds fewer than nrec+1 values");
  const Prec prec = prec_from_int(p.precision);
  const int64_t n = H.d.ndim + 1;
  Sil S(H, p.kdim, prec, n);
  DevBuf<double> hv;
  hv.alloc(2 * (size_t)n);
  double time = 0.0;
  // with a CAP the operator is not Hermitian: Lanczos-Arnoldi steps
  const bool csil = H.cap.set;
  S.op = [&](const double* v, double* out) {
    double E[3];
    laser_efield(npulse, pulses, time, E);
    laser_apply(H, E, v, hv.get(), out, prec, H.st);
  };
  std::vector<double> planar(2 * n);
  for (int64_t i = 0; i < n; ++i) {
    planar[i] = psi_io[2 * i];
    planar[n + i] = psi_io[2 * i + 1];
  }
  HIP_CHECK(hipMemcpyAsync(S.Kp(0), planar.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, H.st));
  const double tiny = 1e-9, tinier = 1e-10;
  int64_t nsteps = 0;
  int maxorder = 0;
  // the TD-NTO input (tdadc2_nto, nto.f90:382-547): the occupied-virtual
  // transition density of psi_IS / |psi_IS|, complex (re, im) per element
  const int64_t n1 = H.parent ? H.parent->d.n1 : H.d.n1;  // nvirt x nocc of the full orbital space
  DevBuf<double> tdev, tpart;
  std::vector<double> thost;
  if (trdens) {
    tdev.alloc(2 * (size_t)n1);
    tpart.alloc(2 * (size_t)reduce_blocks(H.d.ndim) + 2);
    thost.resize(2 * (size_t)n1);
  }
  auto record_trdens = [&](int64_t rec) {
    const double* x = S.Kp(0);
    const int nbk = launch_cdot_partial(H.d.ndim, x, x, n, tpart.get(), H.st);
    launch_colsum(nbk, 2, tpart.get(), tpart.get() + 2 * (size_t)reduce_blocks(H.d.ndim), H.st);
    double nrm[2];
    HIP_CHECK(hipMemcpyAsync(nrm, tpart.get() + 2 * (size_t)reduce_blocks(H.d.ndim), sizeof(nrm),
                             hipMemcpyDeviceToHost, H.st));
    trdens_vo_device(H, 2, x, n, tdev.get());
    HIP_CHECK(hipMemcpyAsync(thost.data(), tdev.get(), sizeof(double) * 2 * n1, hipMemcpyDeviceToHost, H.st));
    HIP_CHECK(hipStreamSynchronize(H.st));
    const double inv = nrm[0] > 0.0 ? 1.0 / std::sqrt(nrm[0]) : 0.0;
    double* out = trdens + rec * 2 * n1;
    for (int64_t r = 0; r < n1; ++r) {
      out[2 * r] = thost[r] * inv;
      out[2 * r + 1] = thost[n1 + r] * inv;
    }
  };
  for (int64_t i = 1; i <= nrec; ++i) {
    double inttime = 0.0;
    while (true) {
      const double stepsize = intperiod - inttime;
      S.func(S.Kp(0), S.Kp(1));
      if (inttime == 0.0) {
        norms[i - 1] = std::sqrt(S.cdot(S.Kp(0), S.Kp(0)).real());
        if (flux) flux[i - 1] = laser_flux(H, S.Kp(0), S.Kp(1));
        if (trdens) record_trdens(i - 1);
      }
      if (std::sqrt(S.cdot(S.Kp(1), S.Kp(1)).real()) > tinier) {
        int order = 0;
        const double truestep = csil ? S.csil_step(stepsize, p.proptol, order) : S.step(stepsize, p.proptol, order);
        ++nsteps;
        maxorder = std::max(maxorder, order);
        time += truestep;
        inttime += truestep;
        if (std::fabs(intperiod - inttime) > std::fabs(tiny * intperiod)) continue;
      } else {
        time += intperiod;
      }
      break;
    }
    if (p.verbose && (i % 100 == 0 || i == nrec)) {
      std::fprintf(stderr, "sil laser: period %lld of %lld (t = %.2f a.u.), %lld steps\n", (long long)i,
                   (long long)nrec, time, (long long)nsteps);
      std::fflush(stderr);
    }
  }
  norms[nrec] = std::sqrt(S.cdot(S.Kp(0), S.Kp(0)).real());
  if (flux) {  // the final flux value (proplib_adc2.f90:816-819)
    S.func(S.Kp(0), S.Kp(1));
    flux[nrec] = laser_flux(H, S.Kp(0), S.Kp(1));
  }
  if (trdens) record_trdens(nrec);
  HIP_CHECK(hipMemcpyAsync(planar.data(), S.Kp(0), sizeof(double) * 2 * n, hipMemcpyDeviceToHost, H.st));
  HIP_CHECK(hipStreamSynchronize(H.st));
  for (int64_t i = 0; i < n; ++i) {
    psi_io[2 * i] = planar[i];
    psi_io[2 * i + 1] = planar[n + i];
  }
  if (nrec_out) *nrec_out = nrec + 1;
  if (stats) {
    stats->nsteps = nsteps;
    stats->nsigma = S.nsigma;
    stats->max_order = maxorder;
    stats->reserved = 0;
    stats->time_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
}

// calc_fdstates_tdfd (source/propagation/fdstates_tdfd.f90:28-627): from
// psi0 = fvec / |fvec|, SIL periods of tout (calc_filterstates, :315-461);
// after period i (and at t = 0) the filter states gain
//   buffer_j += fac dt g(t) Re(exp(i E_j t) psi(t)),   E_j = E_1 + (j-1) dE,
// fac 1 at the first and last record and 2 between, g(t) = cos(pi t /
// (2 tfinal))^iwfunc (update_filterstates, :465-520; windowfunc, :524-535),
// and the states are v_i = sum_j fbas2eig(isel(i), j) buffer_j (:556-627).
// The buffers are never formed: v_i gains a_i(t) Re psi - b_i(t) Im psi with
// a_i + i b_i = fac dt g sum_j f2e[i, j] exp(i E_j t), coefficients known
// before the propagation.  The planar psi(t) of up to 16 records sit in a
// ring of device columns and one GEMM per 16 records adds them to V.
void sil_fdstates_run(Handle& H, const adc2g_sil_params& p, const double* fvec, int nfbas, const double ebound[2],
                      int iwfunc, int nsel, const double* f2e, double* vecs, double* energies,
                      adc2g_sil_stats* stats) {
  const auto t0 = std::chrono::steady_clock::now();
  ARG_CHECK(p.tout > 0 && p.tfinal >= 0, "tout must be > 0 and tfinal >= 0");
  ARG_CHECK(p.kdim >= 2 && p.kdim <= 30, "krydim must be in 2..30");
  ARG_CHECK(p.autotol > 0, "autotol must be > 0");
  ARG_CHECK(nfbas >= 2 && nsel >= 1 && iwfunc >= 0, "need nfbas >= 2, nsel >= 1 and iwfunc >= 0");
  ARG_CHECK(std::isfinite(ebound[0]) && std::isfinite(ebound[1]) && ebound[1] > ebound[0],
            "the energy window must be finite with E_max > E_min");
  const double dt = p.tout;
  const int64_t nrec = (int64_t)(p.tfinal / dt);
  const int64_t nr = nrec + 1;
  Sil S(H, p.kdim, prec_from_int(p.precision));
  const int64_t n = S.n;
  // coefficients of every record: rows 2 r (Re psi) and 2 r + 1 (Im psi), column i
  std::vector<double> coef((size_t)2 * nr * nsel);
  const double de = (ebound[1] - ebound[0]) / (double)(nfbas - 1);
  const double pi = 3.14159265358979323846;
  for (int64_t r = 0; r < nr; ++r) {
    const double t = (double)r * dt;
    const double fac = (r == 0 || r == nrec) ? 1.0 : 2.0;
    const double g = p.tfinal > 0.0 ? std::pow(std::cos(pi * t / (2.0 * p.tfinal)), iwfunc) : 1.0;
    for (int i = 0; i < nsel; ++i) {
      double a = 0.0, b = 0.0;
      for (int j = 0; j < nfbas; ++j) {
        const double ej = ebound[0] + (double)j * de;
        a += f2e[i + (size_t)nsel * j] * std::cos(ej * t);
        b += f2e[i + (size_t)nsel * j] * std::sin(ej * t);
      }
      coef[(size_t)2 * r + (size_t)2 * nr * i] = fac * dt * g * a;
      coef[(size_t)2 * r + 1 + (size_t)2 * nr * i] = -fac * dt * g * b;
    }
  }
  constexpr int RING = 16;
  DevBuf<double> Q, V, C;
  Q.alloc((size_t)2 * n * RING);
  V.alloc((size_t)n * nsel);
  C.alloc(coef.size());
  HIP_CHECK(hipMemcpyAsync(C.get(), coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, H.st));
  int64_t r0 = 0;  // first record not yet added to V
  auto flush = [&](int64_t rend) {
    GemmArgs g;
    g.M = (int)n; g.N = nsel; g.K = (int)(2 * (rend - r0));
    g.A = Q.get() + (size_t)2 * n * (r0 % RING); g.sam = 1; g.sak = n;
    g.B = C.get() + 2 * r0; g.sbk = 1; g.sbn = 2 * nr;
    g.map.C = V.get(); g.map.scm = 1; g.map.scn = n; g.map.beta = (r0 == 0) ? 0.0 : 1.0;
    gemm(g, Prec::F64, H.gws, H.st);
    r0 = rend;
  };
  auto record = [&](int64_t r) {
    if (r > r0 && r % RING == 0) flush(r);
    HIP_CHECK(hipMemcpyAsync(Q.get() + (size_t)2 * n * (r % RING), S.Kp(0), sizeof(double) * 2 * n,
                             hipMemcpyDeviceToDevice, H.st));
  };
  // psi0 = F / |F| (init_wavepacket, :291-311)
  std::vector<double> f(fvec, fvec + n);
  double norm2 = 0.0;
  for (int64_t i = 0; i < n; ++i) norm2 += f[i] * f[i];
  ARG_CHECK(norm2 > 0.0, "initial vector is zero");
  const double inv = 1.0 / std::sqrt(norm2);
  for (auto& x : f) x *= inv;
  HIP_CHECK(hipMemcpyAsync(S.Kp(0), f.data(), sizeof(double) * n, hipMemcpyHostToDevice, H.st));
  HIP_CHECK(hipMemsetAsync(S.Kp(0) + n, 0, sizeof(double) * n, H.st));
  record(0);
  const double tiny = 1e-9;
  int64_t nsteps = 0;
  int maxorder = 0;
  for (int64_t i = 1; i <= nrec; ++i) {
    double inttime = 0.0;
    while (true) {
      const double stepsize = dt - inttime;
      S.func(S.Kp(0), S.Kp(1));
      int order = 0;
      const double truestep = S.step(stepsize, p.autotol, order);
      ++nsteps;
      maxorder = std::max(maxorder, order);
      inttime += truestep;
      if (!(std::fabs(dt - inttime) > std::fabs(tiny * dt))) break;
    }
    record(i);
    if (p.verbose && (i % 100 == 0 || i == nrec)) {
      std::fprintf(stderr, "fd states: period %lld of %lld, %lld steps\n", (long long)i, (long long)nrec,
                   (long long)nsteps);
      std::fflush(stderr);
    }
  }
  flush(nr);
  fd_finish(H, V.get(), nsel, prec_from_int(p.precision), vecs, energies);
  if (stats) {
    stats->nsteps = nsteps;
    stats->nsigma = S.nsigma;
    stats->max_order = maxorder;
    stats->reserved = 0;
    stats->time_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
}

void matxvec_treal_host(Handle& H, const double* v1, double* v2, Prec p) {
  const int64_t n = H.d.ndim;
  std::vector<double> planar(2 * n);
  for (int64_t i = 0; i < n; ++i) {
    planar[i] = v1[2 * i];
    planar[n + i] = v1[2 * i + 1];
  }
  DevBuf<double> dv, dh;
  dv.alloc(2 * n);
  dh.alloc(2 * n);
  HIP_CHECK(hipMemcpyAsync(dv.get(), planar.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice, H.st));
  H.sigma(dv.get(), n, dh.get(), n, 2, p, H.st);
  HIP_CHECK(hipMemcpyAsync(planar.data(), dh.get(), sizeof(double) * 2 * n, hipMemcpyDeviceToHost, H.st));
  HIP_CHECK(hipStreamSynchronize(H.st));
  // -i (Hre + i Him) = Him - i Hre
  for (int64_t i = 0; i < n; ++i) {
    v2[2 * i] = planar[n + i];
    v2[2 * i + 1] = -planar[i];
  }
}

}  // namespace adc2g
