// Coupled (Schur-complement) Davidson preconditioner (see coupled.hip).
#pragma once

#include "handle.hpp"

namespace adc2g {

struct CoupledPrecon {
  Handle* H = nullptr;
  Prec prec = Prec::F32;  // precision of the half sigma-vectors (build and apply)
  int n1 = 0, nc = 0;     // 1h1p dimension; columns per half sigma-vector pass
  int64_t n = 0;
  double sigma = 0.0;     // shift of the 2h2p diagonal block
  double dmin = 0.0;      // smallest 2h2p diagonal element
  DevBuf<double> Xb, Wb;  // n x nc staging
  DevBuf<double> Qm, lam, offd;  // eigenvectors (n1 x n1, column-major) and eigenvalues of A0
  DevBuf<double> Bm, Cm, th, part, nrm;
  DevBuf<int> info;
  DevBuf<int64_t> idx;
  void* rb = nullptr;     // rocblas handle of the eigensolver
  int builds = 0;
  int verbose = 0;         // 1: one line per shift choice on stderr
  int64_t nsigma_half = 0;  // half sigma-vectors applied (build, Newton steps and corrections)
  double t_build = 0.0, t_eig = 0.0;

  CoupledPrecon() = default;
  CoupledPrecon(const CoupledPrecon&) = delete;
  CoupledPrecon& operator=(const CoupledPrecon&) = delete;
  ~CoupledPrecon();

  void init(Handle& H, Prec p);
  // A0 = H11 - C P(sigma) C^T into Qm (not decomposed), contractions in bp
  void form_schur(double sigma, Prec bp);
  // form_schur, then A0 = Q Lambda Q^T (Qm <- Q, lam <- Lambda)
  void build(double sigma, Prec bp);
  std::vector<double> newton_roots(int k);
  void choose_sigma(int k, int max_search);
  // rebuild at the mean of the active wanted Ritz values when they all lie
  // far below sigma (at most twice); true if A0 was formed again
  bool track(const double* theta, int nact);
  int retracks = 0;
  // T[:, c] = M(theta_c)^-1 R[:, c], c < nb <= nc (device R, T; host theta)
  void apply(int nb, const double* theta_h, const double* R, int64_t ldr, double* T, int64_t ldt);
};

}  // namespace adc2g
