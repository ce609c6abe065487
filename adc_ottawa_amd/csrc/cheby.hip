// Chebyshev order-domain autocorrelation function of the ADC(2)-s matrix
// (chebyshev_auto, source/propagation/chebyspec.f90:245-319), with the matrix
// scaled to [-1, 1] by the spectral bounds as in matxvec_chebyshev_incore
// (source/propagation/tdselib.f90:1316-1398):
//   H_norm = (2/dE) H - (2 Emin/dE + 1) I,  dE = Emax - Emin,
//   q_1 = H_norm q_0,  q_k = 2 H_norm q_{k-1} - q_{k-2}   (chebyshev_recursion, tdselib.f90:1228-1257),
//   C_k = q_0.q_k for k = 0..K and, for k > K/2, C_2k = 2 q_k.q_k - C_0 and
//   C_{2k-1} = 2 q_{k-1}.q_k - C_1 (K = order, 2K+1 values).
// One sigma-vector per order; the scaling, the recursion and the three dot
// products of the order are one streaming kernel; the dot products stay on
// the device until the end (no host round trip per order).
#include <algorithm>
#include <cmath>
#include <vector>

#include "gemm.hpp"
#include "linalg.hpp"
#include "solvers.hpp"

namespace adc2g {

namespace {

constexpr int CB_THREADS = 256;

// qk = c (a w + b qkm1) - d qkm2; part[blk*3 + {0,1,2}] = partial q0.qk, qk.qk, qkm1.qk
__global__ void __launch_bounds__(CB_THREADS) cheby_step_kernel(int64_t n, double a, double b, double c, double d,
                                                                const double* __restrict__ w,
                                                                const double* __restrict__ qkm1,
                                                                const double* __restrict__ qkm2,
                                                                const double* __restrict__ q0,
                                                                double* __restrict__ qk, double* __restrict__ part) {
  __shared__ double red[3][CB_THREADS / 64];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  const bool three = (d != 0.0);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double p = qkm1[i];
    double x = c * (a * w[i] + b * p);
    if (three) x -= d * qkm2[i];
    qk[i] = x;
    s0 += q0[i] * x;
    s1 += x * x;
    s2 += p * x;
  }
  for (int off = 32; off > 0; off >>= 1) {
    s0 += __shfl_down(s0, off, 64);
    s1 += __shfl_down(s1, off, 64);
    s2 += __shfl_down(s2, off, 64);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wv] = s0;
    red[1][wv] = s1;
    red[2][wv] = s2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
    for (int i = 0; i < CB_THREADS / 64; ++i) t += red[threadIdx.x][i];
    part[(int64_t)blockIdx.x * 3 + threadIdx.x] = t;
  }
}

// V[:, i] *= s[i]
__global__ void fd_scale_kernel(int64_t n, int nsel, const double* __restrict__ s, double* __restrict__ V) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * nsel; t += (int64_t)gridDim.x * blockDim.x)
    V[t] *= s[t / n];
}

}  // namespace

void cheby_run(Handle& H, const double bounds[2], const double* q0_h, int order, Prec p, double* auto_out) {
  ARG_CHECK(order >= 1, "order must be >= 1");
  ARG_CHECK(std::isfinite(bounds[0]) && std::isfinite(bounds[1]) && bounds[1] > bounds[0],
            "bounds must be finite with Emax > Emin");
  const int64_t n = H.d.ndim;
  const double de = bounds[1] - bounds[0];
  const double a = 2.0 / de, b = -(2.0 * bounds[0] / de + 1.0);
  hipStream_t st = H.st;
  DevBuf<double> q0, qa, qb, qc, w, part, dots;
  q0.alloc(n);
  qa.alloc(n);
  qb.alloc(n);
  qc.alloc(n);
  w.alloc(n);
  const int nblk = reduce_blocks(n);
  part.alloc((size_t)nblk * 3);
  dots.alloc((size_t)3 * (order + 1));
  HIP_CHECK(hipMemcpyAsync(q0.get(), q0_h, sizeof(double) * n, hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemcpyAsync(qa.get(), q0_h, sizeof(double) * n, hipMemcpyHostToDevice, st));
  // order 0: the three dots of q0 with itself (C_0 from the first)
  hipLaunchKernelGGL(cheby_step_kernel, dim3(nblk), dim3(CB_THREADS), 0, st, n, 0.0, 1.0, 1.0, 0.0, qa.get(),
                     qa.get(), qa.get(), q0.get(), qc.get(), part.get());
  HIP_CHECK(hipGetLastError());
  launch_colsum(nblk, 3, part.get(), dots.get(), st);
  double* qkm2 = qc.get();  // unused at k = 1
  double* qkm1 = qa.get();
  double* qk = qb.get();
  for (int k = 1; k <= order; ++k) {
    H.sigma(qkm1, n, w.get(), n, 1, p, st);
    const double c = (k > 1) ? 2.0 : 1.0, d = (k > 1) ? 1.0 : 0.0;
    hipLaunchKernelGGL(cheby_step_kernel, dim3(nblk), dim3(CB_THREADS), 0, st, n, a, b, c, d, w.get(), qkm1, qkm2,
                       q0.get(), qk, part.get());
    HIP_CHECK(hipGetLastError());
    launch_colsum(nblk, 3, part.get(), dots.get() + 3 * k, st);
    double* t = qkm2;
    qkm2 = qkm1;
    qkm1 = qk;
    qk = t;
  }
  std::vector<double> h((size_t)3 * (order + 1));
  HIP_CHECK(hipMemcpyAsync(h.data(), dots.get(), sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  // the assignments in the reference's order (a doubled value C_{2k-1} with
  // 2k-1 <= K is overwritten by the direct one of a later order)
  auto_out[0] = h[0];
  for (int k = 1; k <= order; ++k) {
    auto_out[k] = h[3 * k];
    if (k > order / 2) {
      auto_out[2 * k] = 2.0 * h[3 * k + 1] - auto_out[0];
      auto_out[2 * k - 1] = 2.0 * h[3 * k + 2] - auto_out[1];
    }
  }
}

// Chebyshev filter-diagonalisation eigenstates (calc_eigenstates of
// fdstates_cfd, source/propagation/fdstates_cfd.f90:207-310): for the nsel
// selected states, v_i = sum_{k=0..K} coef[k, i] q_k with the same recursion,
// then v_i / |v_i| and E_i = v_i . H v_i (matxvec returns -H v, negated
// there).  The q_k go into a ring of KC device columns; every KC orders one
// GEMM adds Q (n x KC) . coef (KC x nsel) to V, so each q_k is written once
// and read once more.
void cheby_fdstates_run(Handle& H, const double bounds[2], const double* q0_h, int order, int nsel, const double* coef_h,
                        Prec p, double* vecs, double* energies) {
  ARG_CHECK(order >= 0 && nsel >= 1, "order must be >= 0 and nsel >= 1");
  ARG_CHECK(std::isfinite(bounds[0]) && std::isfinite(bounds[1]) && bounds[1] > bounds[0],
            "bounds must be finite with Emax > Emin");
  const int64_t n = H.d.ndim;
  const int nk = order + 1;
  const double de = bounds[1] - bounds[0];
  const double a = 2.0 / de, b = -(2.0 * bounds[0] / de + 1.0);
  hipStream_t st = H.st;
  constexpr int KC = 32;
  const int ring = std::min(KC, nk < 3 ? 3 : nk);
  DevBuf<double> Q, w, V, C, part;
  Q.alloc((size_t)n * ring);
  w.alloc((size_t)n);
  V.alloc((size_t)n * nsel);
  C.alloc((size_t)nk * nsel);
  part.alloc((size_t)reduce_blocks(n) * 3);
  HIP_CHECK(hipMemcpyAsync(C.get(), coef_h, sizeof(double) * nk * nsel, hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemcpyAsync(Q.get(), q0_h, sizeof(double) * n, hipMemcpyHostToDevice, st));
  auto slot = [&](int k) { return Q.get() + (size_t)(k % ring) * n; };
  int k0 = 0;  // first order not yet added to V
  auto flush = [&](int kend) {  // V += Q[k0..kend) coef[k0..kend, :]
    GemmArgs g;
    g.M = (int)n; g.N = nsel; g.K = kend - k0;
    g.A = slot(k0); g.sam = 1; g.sak = n;
    g.B = C.get() + k0; g.sbk = 1; g.sbn = nk;
    g.map.C = V.get(); g.map.scm = 1; g.map.scn = n; g.map.beta = (k0 == 0) ? 0.0 : 1.0;
    gemm(g, Prec::F64, H.gws, st);
    k0 = kend;
  };
  const int nblk = reduce_blocks(n);
  for (int k = 1; k <= order; ++k) {
    if (k % ring == k0 % ring && k > k0) flush(k);  // the slot of q_k still holds an unadded order
    double* qkm1 = slot(k - 1);
    double* qkm2 = (k > 1) ? slot(k - 2) : qkm1;
    H.sigma(qkm1, n, w.get(), n, 1, p, st);
    const double c = (k > 1) ? 2.0 : 1.0, d = (k > 1) ? 1.0 : 0.0;
    hipLaunchKernelGGL(cheby_step_kernel, dim3(nblk), dim3(CB_THREADS), 0, st, n, a, b, c, d, w.get(), qkm1, qkm2,
                       qkm1, slot(k), part.get());
    HIP_CHECK(hipGetLastError());
  }
  flush(nk);
  fd_finish(H, V.get(), nsel, p, vecs, energies);
}

// The states in V (ndim x nsel, device) normalised, and E_i = v_i . H v_i,
// as both filter-diagonalisation programs finish (fdstates_cfd.f90:286-299,
// fdstates_tdfd.f90:589-600): the norms and the Rayleigh quotients are the
// diagonals of two nsel x nsel FP64 GEMMs (V^T V, V^T H V), so only the
// result crosses PCIe.
void fd_finish(Handle& H, double* Vd, int nsel, Prec p, double* vecs, double* energies) {
  const int64_t n = H.d.ndim;
  hipStream_t st = H.st;
  DevBuf<double> G, W, sc;
  G.alloc((size_t)nsel * nsel);
  sc.alloc((size_t)nsel);
  auto gram = [&](const double* B) {  // G = V^T B
    GemmArgs g;
    g.M = nsel; g.N = nsel; g.K = (int)n;
    g.A = Vd; g.sam = n; g.sak = 1;
    g.B = B; g.sbk = 1; g.sbn = n;
    g.map.C = G.get(); g.map.scm = 1; g.map.scn = nsel;
    gemm(g, Prec::F64, H.gws, st);
    std::vector<double> h((size_t)nsel * nsel);
    HIP_CHECK(hipMemcpyAsync(h.data(), G.get(), sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    std::vector<double> dg(nsel);
    for (int i = 0; i < nsel; ++i) dg[i] = h[(size_t)i * nsel + i];
    return dg;
  };
  std::vector<double> nrm = gram(Vd);
  for (int i = 0; i < nsel; ++i) {
    ARG_CHECK(nrm[i] > 0.0, "a filter-diagonalisation state is zero");
    nrm[i] = 1.0 / std::sqrt(nrm[i]);
  }
  HIP_CHECK(hipMemcpyAsync(sc.get(), nrm.data(), sizeof(double) * nsel, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(fd_scale_kernel, dim3(grid_for(n * nsel)), dim3(256), 0, st, n, nsel, sc.get(), Vd);
  HIP_CHECK(hipGetLastError());
  W.alloc((size_t)n * nsel);
  H.sigma(Vd, n, W.get(), n, nsel, p, st);
  const std::vector<double> e = gram(W.get());
  for (int i = 0; i < nsel; ++i) energies[i] = e[i];
  if (vecs) {
    HIP_CHECK(hipMemcpyAsync(vecs, Vd, sizeof(double) * n * nsel, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
  }
}

// spectral_bounds (source/propagation/specbounds.f90): the lower bound is the
// lowest root of block Davidson (davstates 1, davtol 1e-8, dmain/maxsubdim 5/20
// for 'davlanc' (:85-160) and 2/10 for 'davpower' (:166-238), the final-space
// defaults precon_f = 2 (Olsen), ldfl_f = true, maxiter_f = 100 and a
// sub-diagonalisation guess with guessdim_f = 800; defaults.f90:84-95).  The
// upper bound is the highest eigenvalue of block Lanczos (lmain 2, ncycles
// 100, orthotype 1) from init_vec_singlesubspace (lancguess 9) for 'davlanc',
// or the power method for 'davpower'.  The caller of spectral_bounds widens
// them (chebyspec.f90:82-83 multiplies by 0.9 and 1.1); here they are
// returned as estimated.
namespace {

// part[blk] = {v1.v2} (mode 0) or {(e v1 - v2)^2, v2.v2} (mode 1), e = *ev
__global__ void __launch_bounds__(CB_THREADS) power_dots_kernel(int64_t n, int mode, const double* __restrict__ v1,
                                                                const double* __restrict__ v2,
                                                                const double* __restrict__ ev,
                                                                double* __restrict__ part) {
  __shared__ double red[2][CB_THREADS / 64];
  const double e = mode ? ev[0] : 0.0;
  double s0 = 0.0, s1 = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (mode == 0) {
      s0 += v1[i] * v2[i];
    } else {
      const double r = e * v1[i] - v2[i];
      s0 += r * r;
      s1 += v2[i] * v2[i];
    }
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
    for (int i = 0; i < CB_THREADS / 64; ++i) t += red[threadIdx.x][i];
    part[(int64_t)blockIdx.x * 2 + threadIdx.x] = t;
  }
}

// v1 = v2 / sqrt(*nrm2)
__global__ void power_next_kernel(int64_t n, const double* __restrict__ v2, const double* __restrict__ nrm2,
                                  double* __restrict__ v1) {
  const double inv = 1.0 / sqrt(nrm2[0]);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v1[i] = v2[i] * inv;
}

}  // namespace

// power_method (source/eigen/power.f90:22-100): guess from power_guess
// (:307-463, the highest eigenvector on the subdim ISs of highest H_jj),
// then power_iterations (:467-524): v2 = H v1, E = v1.v2, r = |E v1 - v2|,
// v2 /= |v2|; stop when r <= thrsh, else v1 = v2.  The eigenpair returned is
// (E, v1) of the converged iteration, as wreigenpair writes it.
void power_run(Handle& H, int subdim, int maxit, double thrsh, double* eval, double* vec, int* niter) {
  const int64_t n = H.d.ndim;
  ARG_CHECK(maxit >= 1 && thrsh >= 0.0, "maxit must be >= 1 and thrsh >= 0");
  hipStream_t st = H.st;
  DevBuf<double> v1, v2, part, sums;
  v1.alloc(n);
  v2.alloc(n);
  const int nblk = reduce_blocks(n);
  part.alloc((size_t)nblk * 2);
  sums.alloc(4);
  subdiag_guess(H, subdim > 0 ? subdim : 800, 1, v1.get(), true, true);
  double e = 0.0, r = INFINITY;
  int it = 0;
  for (it = 1; it <= maxit; ++it) {
    H.sigma(v1.get(), n, v2.get(), n, 1, Prec::F64, st);
    hipLaunchKernelGGL(power_dots_kernel, dim3(nblk), dim3(CB_THREADS), 0, st, n, 0, v1.get(), v2.get(), nullptr,
                       part.get());
    HIP_CHECK(hipGetLastError());
    launch_colsum(nblk, 2, part.get(), sums.get(), st);  // sums[0] = E
    hipLaunchKernelGGL(power_dots_kernel, dim3(nblk), dim3(CB_THREADS), 0, st, n, 1, v1.get(), v2.get(), sums.get(),
                       part.get());
    HIP_CHECK(hipGetLastError());
    launch_colsum(nblk, 2, part.get(), sums.get() + 2, st);  // |r|^2, |v2|^2
    double h[4];
    HIP_CHECK(hipMemcpyAsync(h, sums.get(), sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    e = h[0];
    r = std::sqrt(h[2]);
    if (r <= thrsh) break;
    hipLaunchKernelGGL(power_next_kernel, dim3(nblk), dim3(CB_THREADS), 0, st, n, v2.get(), sums.get() + 3, v1.get());
    HIP_CHECK(hipGetLastError());
  }
  if (eval) *eval = e;
  if (niter) *niter = std::min(it, maxit);
  if (vec) {
    HIP_CHECK(hipMemcpyAsync(vec, v1.get(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
  }
  if (!(r <= thrsh)) throw Error(ADC2G_ENOCONV, "power method: convergence not reached...");
}

// init_vec_doublesubspace (block_lanczos.f90:443-744): the first bs/2
// columns are the lowest eigenvectors of the subdim ISs of lowest H_jj, the
// others eigenvectors bs/2+1 .. bs (ascending order) of the subdim ISs of
// highest H_jj, as the reference's column loop takes them;
// init_vec_singlesubspace (:748-923): the lowest bs eigenvectors of the
// highest-H_jj block.
void lanczos_guess(Handle& H, int kind, int subdim, int bs, double* V) {
  ARG_CHECK(kind == 8 || kind == 9, "kind must be 8 (double subspace) or 9 (single subspace)");
  ARG_CHECK(bs >= 1 && subdim >= 1, "blocksize and subdim must be positive");
  if (kind == 9) {
    subdiag_cols(H, subdim, true, 0, bs, V);
    return;
  }
  ARG_CHECK(bs % 2 == 0, "the double-subspace guess needs an even blocksize");
  subdiag_cols(H, subdim, false, 0, bs / 2, V);
  subdiag_cols(H, subdim, true, bs / 2, bs / 2, V + (size_t)H.d.ndim * (bs / 2));
}

void spectral_bounds_run(Handle& H, int method, int ncycles, double bounds[2]) {
  ARG_CHECK(method >= 1 && method <= 3, "method must be 1 (davlanc), 2 (davpower) or 3 (lanczos)");
  const int64_t n = H.d.ndim;
  if (method == 3) {
    // spectral_bounds_lanczos (specbounds.f90:52-105): block Lanczos with
    // lancguess 8, lmain 2, ncycles 50 and the default orthotype 0; the
    // lowest and highest Lanczos state energies
    const int m = ncycles > 0 ? ncycles : 50;
    DevBuf<double> g;
    g.alloc((size_t)n * 2);
    lanczos_guess(H, 8, 800, 2, g.get());
    std::vector<double> q0((size_t)n * 2);
    HIP_CHECK(hipMemcpyAsync(q0.data(), g.get(), sizeof(double) * q0.size(), hipMemcpyDeviceToHost, H.st));
    HIP_CHECK(hipStreamSynchronize(H.st));
    adc2g_lanczos_params lp{};
    lp.blocksize = 2;
    lp.ncycles = m;
    lp.orthotype = 0;
    std::vector<double> ev((size_t)2 * m);
    lanczos_run(H, lp, q0.data(), ev.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
    bounds[0] = ev.front();
    bounds[1] = ev.back();
    return;
  }
  adc2g_dav_params dp{};
  dp.nstates = 1;
  dp.blocksize = method == 1 ? 5 : 2;   // davlanc: dmain_f 5, maxsubdim_f 20; davpower: 2, 10
  dp.maxvec = method == 1 ? 20 : 10;
  dp.maxiter = 100;
  dp.tol = 1e-8;
  dp.precon = 2;
  dp.deflate = 1;
  dp.precision = 0;
  dp.refine = 1;
  dp.guess = ADC2G_GUESS_SUBDIAG;
  dp.subdim = 800;
  double lo = 0.0;
  davidson_run(H, dp, &lo, nullptr, nullptr, nullptr);
  if (method == 2) {
    // upper bound from the power method (specbounds.f90:218-229)
    double hi = 0.0;
    power_run(H, 800, 350, 1e-8, &hi, nullptr, nullptr);
    bounds[0] = lo;
    bounds[1] = hi;
    return;
  }
  const int m = ncycles > 0 ? ncycles : 100;
  DevBuf<double> g;
  g.alloc((size_t)n * 2);
  lanczos_guess(H, 9, 800, 2, g.get());
  std::vector<double> q0((size_t)n * 2);
  HIP_CHECK(hipMemcpyAsync(q0.data(), g.get(), sizeof(double) * q0.size(), hipMemcpyDeviceToHost, H.st));
  HIP_CHECK(hipStreamSynchronize(H.st));
  adc2g_lanczos_params lp{};
  lp.blocksize = 2;
  lp.ncycles = m;
  lp.precision = 0;
  lp.check_orth = 0;
  lp.orthotype = 1;
  std::vector<double> ev((size_t)2 * m);
  lanczos_run(H, lp, q0.data(), ev.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
  bounds[0] = lo;
  bounds[1] = ev.back();
}

}  // namespace adc2g
