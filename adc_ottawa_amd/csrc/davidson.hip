// Block Davidson on the GPU.
//
// Semantics follow davdiag_block / run_block_davidson
// (source/eigen/block_davidson.f90:33-145, 690-774): block size and subspace
// defaults (:195-205), unit-vector guesses on the smallest |H_jj| (:463-532),
// Rayleigh-Ritz on the lowest blocksize_curr pairs, r = theta x - W y and
// ||r||_2 < tol (:1029-1085), the convergence count reset without deflation
// (:1077), deflation by moving converged roots to the end of the arrays
// (:1147-1222), the collapse rule currdim > maxvec_curr - blocksize_curr
// (:1109-1116), DPR/Olsen (:1288-1419) and projection against converged
// vectors (:1226-1284), eigenpairs returned sorted ascending (:1469-1532).
//
// Different by design (same subspaces in exact arithmetic, so the same Ritz
// values iteration by iteration):
//  * W is kept: only the new block gets a sigma-vector (the reference
//    recomputes W for all currdim columns each iteration, nmult += currdim;
//    that count is reported as nmult_ref).  A collapse keeps W Y.
//  * The new block is orthonormalised against V (and the locked vectors) by
//    block classical Gram-Schmidt -- a second pass only when the DGKS test
//    finds that a column lost more than half its squared norm -- and within
//    itself through the eigen-decomposition of its Gram matrix (repeated when
//    that matrix is ill-conditioned), instead of a Householder QR of the
//    whole subspace (:1423-1465); numerically dependent directions are
//    dropped.
//  * Rayleigh-Ritz computes only the bs_curr lowest eigenpairs of the
//    subspace matrix (the reference's dsyev computes all of them).
//  * Preconditioner denominators below 1e-10 in magnitude are clamped (the
//    reference divides unguarded).
// Mixed precision: with precision != FP64 the first phase uses low-precision
// sigma-vectors without convergence counting; when every residual is below
// lowprec_tol, or the largest one has not improved by 10% for 40 iterations,
// the sigma-vectors of the whole subspace are recomputed in FP64 and the
// iteration continues on them (refinement), so convergence is always judged
// on FP64 residuals.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <numeric>

#include "coupled.hpp"
#include "linalg.hpp"
#include "precon.hpp"
#include "skinny.hpp"
#include "solvers.hpp"

namespace adc2g {

void gemm_tn(Handle& H, int M, int N, int64_t K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
             int64_t ldc, double alpha, double beta) {
  if (K > 4096 && skinny_tn(M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, H.skinny_ws, H.st)) return;
  GemmArgs g;
  g.M = M; g.N = N; g.K = (int)K;
  g.A = A; g.sam = lda; g.sak = 1;
  g.B = B; g.sbk = 1; g.sbn = ldb;
  g.map.C = C; g.map.scm = 1; g.map.scn = ldc; g.map.alpha = alpha; g.map.beta = beta;
  gemm(g, Prec::F64, H.gws, H.st);
}

void gemm_nn(Handle& H, int64_t M, int N, int K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
             int64_t ldc, double alpha, double beta) {
  if (M > 4096 && skinny_nn(M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, H.st)) return;
  GemmArgs g;
  g.M = (int)M; g.N = N; g.K = K;
  g.A = A; g.sam = 1; g.sak = lda;
  g.B = B; g.sbk = 1; g.sbn = ldb;
  g.map.C = C; g.map.scm = 1; g.map.scn = ldc; g.map.alpha = alpha; g.map.beta = beta;
  gemm(g, Prec::F64, H.gws, H.st);
}

namespace {

// rows idx[0..nr) of the nc columns of W -> out[r + ldo c]
__global__ void gather_rows_kernel(int nr, int nc, const int64_t* __restrict__ idx, const double* __restrict__ W,
                                   int64_t ldw, double* __restrict__ out, int ldo) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)nr * nc) return;
  const int r = (int)(t % nr), c = (int)(t / nr);
  out[r + (int64_t)ldo * c] = W[idx[r] + ldw * c];
}

// V[idx[r] + ldv c] = U[r + ldu c] for the nc columns (V zeroed before)
__global__ void scatter_rows_kernel(int nr, int nc, const int64_t* __restrict__ idx, const double* __restrict__ U,
                                    int ldu, double* __restrict__ V, int64_t ldv) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)nr * nc) return;
  const int r = (int)(t % nr), c = (int)(t / nr);
  V[idx[r] + ldv * c] = U[r + (int64_t)ldu * c];
}

}  // namespace

// initvec_subdiag (source/eigen/block_davidson.f90:293-459): the subdim ISs
// of lowest H_jj (ascending; subdim grows while the next H_jj is within 1e-6
// of the last, so a degenerate group is not split), the stored matrix on
// them (diagonal H_jj, off-diagonal elements |h| > 1e-12 taken from the upper
// triangle i < j as the reference stores them), its lowest bs eigenvectors
// scattered into V (n x bs, ld n).  The columns come from FP64 sigma-vectors
// of unit vectors on the selected ISs.  With descending, the ISs of highest
// H_jj (init_vec_singlesubspace, source/eigen/block_lanczos.f90:748-923, which
// also keeps the lowest eigenvectors of that block); with top, the highest bs
// eigenvectors instead (power_guess, source/eigen/power.f90:307-463).
void subdiag_guess(Handle& H, int subdim, int bs, double* V, bool descending, bool top) {
  subdiag_cols(H, subdim, descending, top ? -1 : 0, bs, V);
}

// The same sub-diagonalisation with eigenvectors first .. first+count-1
// (ascending eigenvalue order; first = -1: the highest count) scattered into
// V (n x count, ld n).
void subdiag_cols(Handle& H, int subdim, bool descending, int first, int bs, double* V) {
  const int64_t n = H.d.ndim;
  hipStream_t st = H.st;
  std::vector<double> hii(n);
  HIP_CHECK(hipMemcpyAsync(hii.data(), H.hii.get(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  std::vector<int64_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  if (descending)
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return hii[a] > hii[b]; });
  else
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return hii[a] < hii[b]; });
  int64_t sd = std::min<int64_t>(subdim, n);
  while (sd < n && std::fabs(hii[order[sd - 1]] - hii[order[sd]]) < 1e-6) ++sd;
  ARG_CHECK(sd <= 4096, "the sub-diagonalisation guess is limited to 4096 ISs");
  ARG_CHECK(bs <= sd && first + bs <= sd, "blocksize exceeds the sub-diagonalisation dimension");
  const int m = (int)sd;
  DevBuf<int64_t> idx;
  idx.alloc(m);
  HIP_CHECK(hipMemcpyAsync(idx.get(), order.data(), sizeof(int64_t) * m, hipMemcpyHostToDevice, st));
  const int cb = std::min(m, 128);
  DevBuf<double> X, W, Hs;
  X.alloc((size_t)n * cb);
  W.alloc((size_t)n * cb);
  Hs.alloc((size_t)m * m);
  for (int c0 = 0; c0 < m; c0 += cb) {
    const int nc = std::min(cb, m - c0);
    HIP_CHECK(hipMemsetAsync(X.get(), 0, sizeof(double) * n * nc, st));
    launch_unit_vectors(nc, idx.get() + c0, X.get(), n, st);
    H.sigma(X.get(), n, W.get(), n, nc, Prec::F64, st);
    const int64_t tot = (int64_t)m * nc;
    hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, m, nc, idx.get(),
                       W.get(), n, Hs.get() + (size_t)m * c0, m);
    HIP_CHECK(hipGetLastError());
  }
  std::vector<double> hc((size_t)m * m);
  HIP_CHECK(hipMemcpyAsync(hc.data(), Hs.get(), sizeof(double) * hc.size(), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  // hc[p + m q] = H[order[p], order[q]]; the stored element of a pair is the
  // one with the smaller full-space index as its row
  std::vector<double> A((size_t)m * m);
  for (int q = 0; q < m; ++q)
    for (int pr = 0; pr < m; ++pr) {
      double h;
      if (pr == q) {
        h = hii[order[pr]];
      } else {
        h = order[pr] < order[q] ? hc[pr + (size_t)m * q] : hc[q + (size_t)m * pr];
        if (std::fabs(h) <= 1e-12) h = 0.0;
      }
      A[pr + (size_t)m * q] = h;
    }
  std::vector<double> w(m), U((size_t)m * m);
  if (!host_syev(m, A.data(), w.data(), U.data()))
    throw Error(ADC2G_EEIG, "the diagonalisation of the subspace Hamiltonian failed");
  DevBuf<double> Ud;
  Ud.alloc((size_t)m * bs);
  // eigenvectors first.. (the lowest bs for 0; the highest for -1, power_guess)
  const size_t f0 = first < 0 ? (size_t)(m - bs) : (size_t)first;
  HIP_CHECK(hipMemcpyAsync(Ud.get(), U.data() + f0 * m, sizeof(double) * m * bs, hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemsetAsync(V, 0, sizeof(double) * n * bs, st));
  const int64_t tot = (int64_t)m * bs;
  hipLaunchKernelGGL(scatter_rows_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, m, bs, idx.get(),
                     Ud.get(), m, V, n);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(st));
}

namespace {

struct Dav {
  Handle& H;
  int64_t n;
  int bs, mv;
  DevBuf<double> V, W, X, WY, Res, T, Rm, Rt, Zs, thd, part, nrm, small, work, M, S1, S2, Gm, cond, rpart, tmpcol, bn;
  DevBuf<int> info;
  DevBuf<int64_t> idx;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double t_sigma = 0;
  int nsigma = 0;
  int n_second_pass = 0;
  // the correction block shares the residual block's memory (a step that
  // reads each residual column before writing its correction, and needs the
  // residual no more afterwards: the coupled step without the network)
  bool t_alias = false;
  bool verbose = false;

  explicit Dav(Handle& h) : H(h) {}
  double* t() const { return t_alias ? Res.get() : T.get(); }
  ~Dav() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }

  void alloc() {
    n = H.d.ndim;
    V.alloc((size_t)n * mv);
    W.alloc((size_t)n * mv);
    X.alloc((size_t)n * bs);
    WY.alloc((size_t)n * bs);
    Res.alloc((size_t)n * bs);
    if (!t_alias) T.alloc((size_t)n * bs);
    Rm.alloc((size_t)mv * mv);
    Rt.alloc((size_t)mv * mv);
    Zs.alloc((size_t)mv * mv);
    thd.alloc(mv);
    part.alloc((size_t)1024 * 2 * bs);
    nrm.alloc(2 * bs);
    small.alloc((size_t)mv * std::max(mv, bs) + 16);
    work.alloc((size_t)4 * mv * mv + 4 * mv);
    M.alloc((size_t)bs * bs);
    S1.alloc((size_t)mv * bs);
    S2.alloc((size_t)bs * bs);
    Gm.alloc((size_t)bs * bs);
    bn.alloc(bs);
    cond.alloc(1);
    info.alloc(4);
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
  }

  void sigma(int c0, int nc, Prec p) {
    HIP_CHECK(hipEventRecord(e0, H.st));
    H.sigma(V.get() + n * c0, n, W.get() + n * c0, n, nc, p, H.st);
    HIP_CHECK(hipEventRecord(e1, H.st));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    t_sigma += ms * 1e-3;
    nsigma += nc;
  }

  // Rm[0:c1, c0:c1] = V[:, :c1]^T W[:, c0:c1], then mirror.
  void rayleigh_cols(int c0, int c1) {
    gemm_tn(H, c1, c1 - c0, n, V.get(), n, W.get() + n * c0, n, Rm.get() + (int64_t)mv * c0, mv, 1.0, 0.0);
    launch_symmetrize_cols(c0, c1, Rm.get(), mv, H.st);
  }

  // One block classical Gram-Schmidt pass of T[:, :r] against the orthonormal
  // columns of V1[:, :nb1] and V2[:, :nb2]; coefficients kept in S1 / S2.
  void cgs_pass(const double* V1, int nb1, const double* V2, int nb2, double* Tp, int r) {
    if (nb1 > 0) gemm_tn(H, nb1, r, n, V1, n, Tp, n, S1.get(), mv, 1.0, 0.0);
    if (nb2 > 0) gemm_tn(H, nb2, r, n, V2, n, Tp, n, S2.get(), bs, 1.0, 0.0);
    if (nb1 > 0) gemm_nn(H, n, r, nb1, V1, n, S1.get(), mv, Tp, n, -1.0, 1.0);
    if (nb2 > 0) gemm_nn(H, n, r, nb2, V2, n, S2.get(), bs, Tp, n, -1.0, 1.0);
  }

  // Orthonormalise Tp[:, :r] within itself into dst (dst != Tp) from its Gram
  // matrix Gm (r x r, ld r); returns the number of directions kept and the
  // condition number of the kept part of the Gram matrix.  With bnorm (the
  // squared column norms before the projection against the subspace), a
  // column that kept less than 1e-12 of its norm through the projection is
  // dropped: it is rounding noise of the subspace, not a new direction.
  int orth_from_gram(double* Tp, int r, double* dst, double* gram_cond, const double* bnorm = nullptr) {
    launch_orth_coeffs(r, Gm.get(), r, M.get(), r, work.get(), info.get(), 1e-12, H.st, cond.get(), bnorm);
    int kept = 0;
    HIP_CHECK(hipMemcpyAsync(&kept, info.get(), sizeof(int), hipMemcpyDeviceToHost, H.st));
    HIP_CHECK(hipMemcpyAsync(gram_cond, cond.get(), sizeof(double), hipMemcpyDeviceToHost, H.st));
    HIP_CHECK(hipStreamSynchronize(H.st));
    if (kept < 0) throw Error(ADC2G_EEIG, "Gram-matrix eigensolve failed in block orthonormalisation");
    if (kept > 0) gemm_nn(H, n, kept, r, Tp, n, M.get(), r, dst, n, 1.0, 0.0);
    return kept;
  }

  // New block T[:, :r] -> V[:, c0:c0+kept], orthonormal to V[:, :c0] and to
  // the locked vectors Xc[:, :ncv].  One block CGS pass, a second only when
  // the DGKS test asks for it; within the block, orthonormalisation through
  // the Gram matrix, repeated when that matrix was ill-conditioned.
  int append_block(int c0, int r, const double* Xc, int ncv) {
    cgs_pass(V.get(), c0, Xc, ncv, t(), r);
    gemm_tn(H, r, r, n, t(), n, t(), n, Gm.get(), r, 1.0, 0.0);
    const bool projected = c0 + ncv > 0;
    if (projected) {
      launch_dgks_check(r, c0, S1.get(), mv, ncv, S2.get(), bs, Gm.get(), r, info.get() + 1, H.st, bn.get());
      int need = 0;
      HIP_CHECK(hipMemcpyAsync(&need, info.get() + 1, sizeof(int), hipMemcpyDeviceToHost, H.st));
      HIP_CHECK(hipStreamSynchronize(H.st));
      if (need) {
        ++n_second_pass;
        cgs_pass(V.get(), c0, Xc, ncv, t(), r);
        gemm_tn(H, r, r, n, t(), n, t(), n, Gm.get(), r, 1.0, 0.0);
      }
    }
    double* slot = V.get() + n * c0;
    double gc = 0.0;
    if (verbose && projected) {
      std::vector<double> g((size_t)r * r), b(r);
      HIP_CHECK(hipMemcpyAsync(g.data(), Gm.get(), sizeof(double) * g.size(), hipMemcpyDeviceToHost, H.st));
      HIP_CHECK(hipMemcpyAsync(b.data(), bn.get(), sizeof(double) * r, hipMemcpyDeviceToHost, H.st));
      HIP_CHECK(hipStreamSynchronize(H.st));
      double mn = INFINITY;
      for (int j = 0; j < r; ++j) mn = std::min(mn, g[(size_t)j * r + j] / b[j]);
      std::printf("  expansion: %d columns, smallest |t_after|^2/|t_before|^2 %.3e\n", r, mn);
    }
    const int k1 = orth_from_gram(t(), r, slot, &gc, projected ? bn.get() : nullptr);
    if (verbose) std::printf("  expansion: kept %d of %d, Gram condition %.3e\n", k1, r, gc);
    if (k1 == 0) return 0;
    if (!(gc < 1e3)) {
      // the combination of nearly dependent columns amplifies their rounding
      // components along the subspace by up to sqrt(gc): project the
      // normalised block once more before orthonormalising it again (else V
      // loses orthogonality, and the Rayleigh matrix grows spurious Ritz
      // pairs of near-zero vectors that pass the residual test)
      if (projected) cgs_pass(V.get(), c0, Xc, ncv, slot, k1);
      gemm_tn(H, k1, k1, n, slot, n, slot, n, Gm.get(), k1, 1.0, 0.0);
      const int k2 = orth_from_gram(slot, k1, t(), &gc);
      if (k2 > 0) HIP_CHECK(hipMemcpyAsync(slot, t(), sizeof(double) * n * k2, hipMemcpyDeviceToDevice, H.st));
      return k2;
    }
    return k1;
  }

  void swap_cols(DevBuf<double>& B, int i, int j) {
    if (i == j) return;
    tmpcol.alloc(n);
    double* tmp = tmpcol.get();
    HIP_CHECK(hipMemcpyAsync(tmp, B.get() + n * i, sizeof(double) * n, hipMemcpyDeviceToDevice, H.st));
    HIP_CHECK(hipMemcpyAsync(B.get() + n * i, B.get() + n * j, sizeof(double) * n, hipMemcpyDeviceToDevice, H.st));
    HIP_CHECK(hipMemcpyAsync(B.get() + n * j, tmp, sizeof(double) * n, hipMemcpyDeviceToDevice, H.st));
  }
};

}  // namespace

void davidson_run(Handle& H, const adc2g_dav_params& p, double* eigval, double* eigvec, adc2g_dav_stats* stats,
                  DavTrace* trace, const double* guess) {
  const auto t0 = std::chrono::steady_clock::now();
  ARG_CHECK(p.nstates >= 1, "nstates must be >= 1");
  ARG_CHECK(p.tol > 0, "tol must be > 0");
  ARG_CHECK(p.precon >= ADC2G_PRE_DPR && p.precon <= ADC2G_PRE_LEARNED_COUPLED,
            "precon must be 1 (DPR), 2 (Olsen), 3 (learned), 4 (coupled) or 5 (learned on the coupled step)");
  const bool learned = p.precon == ADC2G_PRE_LEARNED || p.precon == ADC2G_PRE_LEARNED_COUPLED;
  const bool coupled = p.precon == ADC2G_PRE_COUPLED || p.precon == ADC2G_PRE_LEARNED_COUPLED;
  ARG_CHECK(!((learned || coupled) && (H.parent || H.stored)),
            "the learned and coupled preconditioners work on the full operator built from integrals only");
  if (learned && !precon_ready(H))
    throw Error(ADC2G_ESTATE, "learned preconditioner requested but no network is loaded");
  const int nstates = p.nstates;
  // davinitialise (block_davidson.f90:195-205)
  int bs = p.blocksize > 0 ? p.blocksize : std::max(nstates + 5, 2 * nstates);
  int mv = p.maxvec >= 0 ? p.maxvec : 4 * bs;
  mv = (int)std::floor((double)mv / bs) * bs;
  ARG_CHECK(bs >= nstates, "blocksize must be >= nstates");
  ARG_CHECK(mv >= 2 * bs, "maxvec must be at least twice the block size");
  ARG_CHECK((int64_t)mv <= H.d.ndim, "maxvec exceeds the matrix dimension");
  ARG_CHECK(mv <= 512, "maxvec > 512 is not supported by the subspace eigensolver");

  Dav D(H);
  D.bs = bs;
  D.mv = mv;
  D.t_alias = coupled && !learned && trace == nullptr;
  D.verbose = p.verbose != 0;
  D.alloc();
  const int64_t n = D.n;
  hipStream_t st = H.st;
  Prec cur = prec_from_int(p.precision);
  bool lowphase = (cur != Prec::F64);
  const bool refine = lowphase ? (p.refine != 0) : true;
  const double lowtol = p.lowprec_tol > 0 ? p.lowprec_tol : 1e-4;

  ARG_CHECK(p.guess >= ADC2G_GUESS_ONDIAG && p.guess <= ADC2G_GUESS_VECTORS, "guess must be 0, 1 or 2");
  ARG_CHECK((p.guess == ADC2G_GUESS_VECTORS) == (guess != nullptr), "initial block missing or unexpected");
  if (p.guess == ADC2G_GUESS_VECTORS) {
    // initvec_adc1: the caller's block (e.g. ADC(1) eigenvectors)
    HIP_CHECK(hipMemcpyAsync(D.V.get(), guess, sizeof(double) * n * bs, hipMemcpyHostToDevice, st));
    DevBuf<double> G;
    G.alloc((size_t)bs * bs);
    gemm_tn(H, bs, bs, n, D.V.get(), n, D.V.get(), n, G.get(), bs, 1.0, 0.0);
    std::vector<double> g((size_t)bs * bs);
    HIP_CHECK(hipMemcpyAsync(g.data(), G.get(), sizeof(double) * g.size(), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    double dev = 0.0;
    for (int j = 0; j < bs; ++j)
      for (int i = 0; i < bs; ++i) dev = std::max(dev, std::fabs(g[i + (size_t)bs * j] - (i == j ? 1.0 : 0.0)));
    ARG_CHECK(dev <= 1e-8, "the initial block must have orthonormal columns (max |G - I| = " + std::to_string(dev) +
                               ")");
  } else if (p.guess == ADC2G_GUESS_SUBDIAG) {
    subdiag_guess(H, p.subdim > 0 ? p.subdim : 800, bs, D.V.get(), false, false);
  } else {
    // initvec_ondiag: unit vectors on the bs smallest |H_jj| (stable order for ties)
    std::vector<double> hii(n);
    HIP_CHECK(hipMemcpyAsync(hii.data(), H.hii.get(), sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    std::vector<int64_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::partial_sort(order.begin(), order.begin() + bs, order.end(), [&](int64_t a, int64_t b) {
      const double fa = std::fabs(hii[a]), fb = std::fabs(hii[b]);
      return fa < fb || (fa == fb && a < b);
    });
    DevBuf<int64_t> didx;
    didx.alloc(bs);
    HIP_CHECK(hipMemcpyAsync(didx.get(), order.data(), sizeof(int64_t) * bs, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemsetAsync(D.V.get(), 0, sizeof(double) * n * bs, st));
    launch_unit_vectors(bs, didx.get(), D.V.get(), n, st);
    HIP_CHECK(hipStreamSynchronize(st));
  }
  // coupled (Schur-complement) step: A0 = H11 - C P C^T at the shift chosen
  // for the nstates wanted roots, eigen-decomposed once per run
  std::unique_ptr<CoupledPrecon> cpl;
  double t_precon = 0.0;
  if (coupled) {
    const auto tc = std::chrono::steady_clock::now();
    cpl.reset(new CoupledPrecon);
    cpl->init(H, Prec::F32);
    cpl->verbose = p.verbose;
    cpl->choose_sigma(nstates, 6);
    t_precon += std::chrono::duration<double>(std::chrono::steady_clock::now() - tc).count();
  }
  int currdim = bs, bs_curr = bs, mv_curr = mv, ns_curr = nstates;
  int nconv = 0, nconv_prev = 0;
  int64_t nmult_ref = 0;
  int niter_low = 0, nfallback = 0;
  D.sigma(0, bs, cur);
  D.rayleigh_cols(0, bs);

  std::vector<double> theta(bs), norm(bs), stored_norm(bs, 0.0);
  std::vector<double> prev_norm(bs, INFINITY);
  // learned step: per root, iterations without a 10 % residual drop, and the
  // sticky switch to the base step once two such iterations occurred
  std::vector<int> stall(bs, 0), sticky(bs, 0);
  double low_best = INFINITY;
  int low_stall = 0;
  bool converged = false;
  int it = 0;
  const int maxiter = p.maxiter > 0 ? p.maxiter : 100;
  for (it = 1; it <= maxiter; ++it) {
    nmult_ref += currdim;
    // Rayleigh-Ritz (diagrmat)
    HIP_CHECK(hipMemcpy2DAsync(D.Rt.get(), sizeof(double) * mv, D.Rm.get(), sizeof(double) * mv,
                               sizeof(double) * currdim, currdim, hipMemcpyDeviceToDevice, st));
    launch_syevx(currdim, bs_curr, D.Rt.get(), mv, D.thd.get(), D.Zs.get(), mv, D.work.get(), D.info.get(), st);
    // Ritz vectors and W y (calcritzvec, calcres), fused with the residuals
    // and, for DPR, the correction step
    const bool fused_dpr = (p.precon == 1);
    int nb = ritz_fused(n, currdim, bs_curr, D.V.get(), D.W.get(), n, D.Zs.get(), mv, D.thd.get(), H.hii.get(),
                        fused_dpr, D.X.get(), D.WY.get(), D.Res.get(), fused_dpr ? D.t() : nullptr, D.rpart, st);
    const bool have_t = fused_dpr && nb > 0;
    if (nb == 0) {
      gemm_nn(H, n, bs_curr, currdim, D.V.get(), n, D.Zs.get(), mv, D.X.get(), n, 1.0, 0.0);
      gemm_nn(H, n, bs_curr, currdim, D.W.get(), n, D.Zs.get(), mv, D.WY.get(), n, 1.0, 0.0);
      nb = launch_residual(n, bs_curr, D.thd.get(), D.X.get(), D.WY.get(), D.Res.get(), n, D.part.get(), st);
      launch_colsum(nb, bs_curr, D.part.get(), D.nrm.get(), st);
    } else {
      launch_colsum(nb, bs_curr, D.rpart.get(), D.nrm.get(), st);
    }
    int einfo = 0;
    HIP_CHECK(hipMemcpyAsync(&einfo, D.info.get(), sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(theta.data(), D.thd.get(), sizeof(double) * bs_curr, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(norm.data(), D.nrm.get(), sizeof(double) * bs_curr, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (einfo != 0) throw Error(ADC2G_EEIG, "Diagonalisation of the Rayleigh matrix failed");
    for (int i = 0; i < bs_curr; ++i) norm[i] = std::sqrt(norm[i]);
    if (lowphase) ++niter_low;

    if (p.verbose) {
      for (int i = 0; i < ns_curr; ++i)
        std::printf(i == 0 ? "%4d %c  %14.7f  %13.7e  %c\n" : "       %14.7f  %13.7e  %c\n",
                    i == 0 ? it : 0, lowphase ? 'L' : 'D', theta[i] * 27.2113845, norm[i],
                    norm[i] < p.tol ? 'y' : 'n');
      std::fflush(stdout);
    }
    if (trace) {
      // records of the active roots with the network's base step (the
      // coupled step when the run uses it, else DPR inside the sampler)
      const double* base = nullptr;
      if (coupled && ns_curr > 0) {
        for (int c0 = 0; c0 < ns_curr; c0 += cpl->nc) {
          const int nb = std::min(cpl->nc, ns_curr - c0);
          cpl->apply(nb, theta.data() + c0, D.Res.get() + n * c0, n, D.t() + n * c0, n);
        }
        base = D.t();
      }
      davtrace_iteration(H, *trace, it, bs_curr, ns_curr, theta.data(), norm.data(), D.X.get(), D.Res.get(), n,
                         base);
    }

    bool newly = false;
    if (lowphase) {
      double mx = 0;
      for (int i = 0; i < nstates; ++i) mx = std::max(mx, norm[i]);
      // the low-precision phase ends at lowprec_tol, or when the largest
      // residual has not improved by 10% over its best for 40 iterations
      // (the noise floor of the low-precision sigma)
      if (mx < low_best * 0.9) {
        low_best = mx;
        low_stall = 0;
      } else {
        ++low_stall;
      }
      if (refine && (mx < lowtol || low_stall >= 40)) {
        // refinement: the subspace is kept and its sigma-vectors W = H V are
        // recomputed in FP64, so the iteration continues where the
        // low-precision phase left it, now with exact Rayleigh matrices
        lowphase = false;
        cur = Prec::F64;
        D.sigma(0, currdim, cur);
        D.rayleigh_cols(0, currdim);
        std::fill(prev_norm.begin(), prev_norm.end(), INFINITY);
        continue;
      }
      if (!refine) {
        int c = 0;
        for (int i = 0; i < ns_curr; ++i)
          if (norm[i] < p.tol) ++c;
        if (c == nstates) {
          converged = true;
          break;
        }
      }
    } else {
      // calcres: convergence count
      if (!p.deflate) nconv = 0;
      for (int i = 0; i < ns_curr; ++i)
        if (norm[i] < p.tol) ++nconv;
      if (nconv == nstates) {
        converged = true;
        break;
      }
      // deflate (block_davidson.f90:1147-1222)
      newly = p.deflate && nconv > nconv_prev;
      if (newly) {
        std::vector<int> convmap;
        for (int i = 0; i < bs_curr; ++i)
          if (norm[i] < p.tol) convmap.push_back(i);
        for (int i = 1; i <= nconv - nconv_prev; ++i) {
          const int indx = bs - nconv_prev - i;
          const int src = convmap[i - 1];
          D.swap_cols(D.X, indx, src);
          D.swap_cols(D.Res, indx, src);
          D.swap_cols(D.WY, indx, src);
          if (have_t) D.swap_cols(D.T, indx, src);
          std::swap(theta[indx], theta[src]);
          std::swap(norm[indx], norm[src]);
          std::swap(prev_norm[indx], prev_norm[src]);
          std::swap(stall[indx], stall[src]);
          std::swap(sticky[indx], sticky[src]);
          stored_norm[indx] = norm[indx];
        }
        const int dd = nconv - nconv_prev;
        bs_curr -= dd;
        mv_curr -= dd;
        ns_curr -= dd;
        currdim -= dd;
      }
    }
    bool collapse = !(currdim <= mv_curr - bs_curr);
    if (newly) collapse = true;
    int ilast = currdim;
    if (collapse) {
      HIP_CHECK(hipMemcpyAsync(D.V.get(), D.X.get(), sizeof(double) * n * bs_curr, hipMemcpyDeviceToDevice, st));
      HIP_CHECK(hipMemcpyAsync(D.W.get(), D.WY.get(), sizeof(double) * n * bs_curr, hipMemcpyDeviceToDevice, st));
      D.rayleigh_cols(0, bs_curr);
      ilast = bs_curr;
    }
    // new directions (dpr / olsen / learned)
    HIP_CHECK(hipMemcpyAsync(D.thd.get(), theta.data(), sizeof(double) * bs_curr, hipMemcpyHostToDevice, st));
    if (p.precon == 1) {
      if (!have_t) launch_dpr(n, bs_curr, D.thd.get(), H.hii.get(), D.Res.get(), D.t(), n, st);
    } else if (p.precon == 2) {
      const int nb2 = launch_olsen_dots(n, bs_curr, D.thd.get(), H.hii.get(), D.X.get(), D.Res.get(), n,
                                        D.part.get(), st);
      launch_colsum(nb2, 2 * bs_curr, D.part.get(), D.nrm.get(), st);
      launch_olsen_step(n, bs_curr, D.thd.get(), H.hii.get(), D.X.get(), D.Res.get(), D.nrm.get(), D.t(), n, st);
    } else {
      const auto tc = std::chrono::steady_clock::now();
      if (coupled) {
        if (it % 10 == 0) cpl->track(theta.data(), ns_curr);
        for (int c0 = 0; c0 < bs_curr; c0 += cpl->nc) {
          const int nb = std::min(cpl->nc, bs_curr - c0);
          cpl->apply(nb, theta.data() + c0, D.Res.get() + n * c0, n, D.t() + n * c0, n);
        }
      }
      if (learned) {
        // learned step on the base step (DPR or coupled); a root that falls
        // short of the base step's usual progress keeps the base step from
        // then on: with DPR (slow, a few % per iteration) when its residual
        // missed a 10 % drop in two iterations, with the coupled step (close
        // to shift-and-invert, a factor of several per iteration) when it
        // missed halving once
        const double need = coupled ? 0.5 : 0.9;
        const int patience = coupled ? 1 : 2;
        std::vector<int> use_base(bs_curr, 0);
        for (int i = 0; i < bs_curr; ++i) {
          if (norm[i] > need * prev_norm[i]) ++stall[i]; else stall[i] = 0;
          if (stall[i] >= patience) sticky[i] = 1;
          if (sticky[i]) {
            use_base[i] = 1;
            ++nfallback;
          }
        }
        precon_learned_step(H, n, bs_curr, D.thd.get(), norm.data(), use_base.data(), D.Res.get(), D.t(), n,
                            coupled ? D.t() : nullptr);
      }
      HIP_CHECK(hipStreamSynchronize(st));
      t_precon += std::chrono::duration<double>(std::chrono::steady_clock::now() - tc).count();
    }
    for (int i = 0; i < bs_curr; ++i) prev_norm[i] = norm[i];
    // new block, orthogonal to the subspace and to the converged (deflated) vectors
    const bool lock = p.deflate && nconv > 0 && !lowphase;
    const int kept = D.append_block(ilast, bs_curr, lock ? D.X.get() + n * (bs - nconv) : nullptr, lock ? nconv : 0);
    if (kept == 0) throw Error(ADC2G_ENOCONV, "subspace expansion produced no new directions");
    D.sigma(ilast, kept, cur);
    D.rayleigh_cols(ilast, ilast + kept);
    currdim = ilast + kept;
    nconv_prev = nconv;
  }
  if (!converged) throw Error(ADC2G_ENOCONV, "Not all vectors have converged...");

  // wreigenpairs: deflated roots back after the active ones, ascending order
  std::vector<int> col(nstates);
  for (int i = 0; i < ns_curr; ++i) col[i] = i;
  for (int i = ns_curr; i < nstates; ++i) col[i] = bs - nconv_prev + (i - ns_curr);
  std::vector<double> ev(nstates), rn(nstates);
  for (int i = 0; i < nstates; ++i) {
    ev[i] = theta[col[i]];
    rn[i] = (i < ns_curr) ? norm[col[i]] : stored_norm[col[i]];
  }
  std::vector<int> order(nstates);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ev[a] < ev[b]; });
  double rmax = 0;
  for (int i = 0; i < nstates; ++i) {
    eigval[i] = ev[order[i]];
    rmax = std::max(rmax, rn[order[i]]);
  }
  if (eigvec) {
    for (int i = 0; i < nstates; ++i)
      HIP_CHECK(hipMemcpyAsync(eigvec + n * i, D.X.get() + n * col[order[i]], sizeof(double) * n,
                               hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
  }
  if (stats) {
    stats->niter = it;
    stats->nsigma = D.nsigma;
    stats->nmult_ref = nmult_ref;
    stats->niter_lowprec = niter_low;
    stats->nfallback = nfallback;
    stats->time_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    stats->time_sigma = D.t_sigma;
    stats->resnorm_max = rmax;
    stats->time_precon = t_precon;
    stats->nprecon_builds = cpl ? cpl->builds : 0;
    stats->precon_sigma = cpl ? cpl->sigma : 0.0;
  }
}

}  // namespace adc2g
