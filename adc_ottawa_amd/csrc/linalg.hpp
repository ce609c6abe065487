// Small dense FP64 linear algebra and the Davidson/SIL vector kernels.
#pragma once

#include "common.hpp"

namespace adc2g {

// Symmetric eigensolver for small n (one workgroup): eigenvalues ascending in
// w[0..n), eigenvectors in the columns of Z (ld = ldz).  A (ld = lda) is
// destroyed.  work needs 4*n doubles.  info[0] = 0 on success.
void launch_syev(int n, double* A, int lda, double* w, double* Z, int ldz, double* work, int* info,
                 hipStream_t st);

// Lowest k eigenpairs of the symmetric n x n matrix A (destroyed; one
// workgroup): eigenvalues ascending in w[0..k), eigenvectors in Z[:, 0..k).
// work needs 2n + 6nk doubles.  info[0] = 0 on success, 1 on non-finite input.
void launch_syevx(int n, int k, double* A, int lda, double* w, double* Z, int ldz, double* work, int* info,
                  hipStream_t st);

// Res[:,i] = theta[i] X[:,i] - WY[:,i]; part[blk][i] = partial sums of Res^2
// (nblk = returned grid size), then colsum -> norms2[i].
int launch_residual(int64_t n, int nc, const double* theta, const double* X, const double* WY, double* Res,
                    int64_t ld, double* part, hipStream_t st);
// Column sums of part[nblk][nc] -> out[nc] (deterministic order).
void launch_colsum(int nblk, int nc, const double* part, double* out, hipStream_t st);

// DPR step: T[:,i] = Res[:,i] / guard(theta[i] - hii)
void launch_dpr(int64_t n, int nc, const double* theta, const double* hii, const double* Res, double* T,
                int64_t ld, hipStream_t st);
// Olsen dot products: part[blk][2i] = sum x r c, part[blk][2i+1] = sum x x c with c = 1/guard(hii-theta)
int launch_olsen_dots(int64_t n, int nc, const double* theta, const double* hii, const double* X,
                      const double* Res, int64_t ld, double* part, hipStream_t st);
// Olsen step: T = alpha c x - c r, alpha[i] = xz/xy (from the column sums, device array sums[2*nc])
void launch_olsen_step(int64_t n, int nc, const double* theta, const double* hii, const double* X,
                       const double* Res, const double* sums, double* T, int64_t ld, hipStream_t st);

// Unit vectors: V[idx[i] + ld*i] = 1 (V zeroed first).
void launch_unit_vectors(int nc, const int64_t* idx, double* V, int64_t ld, hipStream_t st);

// Block orthonormalisation helper: from the Gram matrix G (r x r, ld = ldg)
// of T, build M (r x r_kept, ld = ldm) with T M orthonormal; drops directions
// with relative eigenvalue below tol.  Returns r_kept through kept[0].
// Runs on one workgroup; scratch: work (see .hip).
// cond (optional, device) receives lambda_max / lambda_min of the kept directions.
// bnorm (optional, device): squared norms of the columns before a projection
// that made T; a column that kept less than tol of its norm through it is
// rounding noise and is dropped.
void launch_orth_coeffs(int r, const double* G, int ldg, double* M, int ldm, double* work, int* kept,
                        double tol, hipStream_t st, double* cond = nullptr, const double* bnorm = nullptr);

// Block Gram-Schmidt re-orthogonalisation test (DGKS form): after one
// projection T <- T - V1 S1 - V2 S2 of an orthonormal [V1 V2], need[0] = 1 if
// some column j kept less than 1e-6 of its squared norm (G = T^T T
// afterwards); above that, one pass leaves it orthogonal to ~1e-13.  bnorm
// (optional, device, r) receives the squared norms before the projection.
void launch_dgks_check(int r, int nb1, const double* S1, int ld1, int nb2, const double* S2, int ld2, const double* G,
                       int ldg, int* need, hipStream_t st, double* bnorm = nullptr);

// Copy the upper block of a symmetric matrix to the lower: R[j + ld*i] = R[i + ld*j]
// for i in [c0, c1), j < c0... (rows c0..c1 of the new columns)
void launch_symmetrize_cols(int c0, int c1, double* R, int ld, hipStream_t st);

// Complex helpers for SIL (planar storage: re column, im column).
// out[0] = sum(a_re*b_re + a_im*b_im), out[1] = sum(a_re*b_im - a_im*b_re) (i.e. <a|b>)
int launch_cdot_partial(int64_t n, const double* a, const double* b, int64_t ld, double* part, hipStream_t st);

// Host eigen-decomposition of the symmetric n x n matrix A (column-major,
// destroyed) on up to 16 threads: eigenvalues ascending in w, eigenvectors
// in the columns of Z (n x n).  False if the QL iteration does not converge.
bool host_syev(int n, double* A, double* w, double* Z);

inline int reduce_blocks(int64_t n) {
  int64_t b = (n + 1023) / 1024;
  if (b > 1024) b = 1024;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace adc2g
