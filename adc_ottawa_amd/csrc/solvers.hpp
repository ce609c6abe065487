// Solvers on top of the sigma-vector: block Davidson, SIL propagation and the
// learned preconditioner.
#pragma once

#include "handle.hpp"

namespace adc2g {

// FP64 products on the handle's stream: C = alpha A^T B + beta C (A: K x M,
// B: K x N) and C = alpha A B + beta C (A: M x K, B: K x N); the tall-skinny
// kernels when K (tn) or M (nn) is large and N <= 64, the general GEMM
// otherwise.  C must not alias A or B.
void gemm_tn(Handle& H, int M, int N, int64_t K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
             int64_t ldc, double alpha, double beta);
void gemm_nn(Handle& H, int64_t M, int N, int K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
             int64_t ldc, double alpha, double beta);

// v2 = -i H v1 for one complex vector (host arrays, interleaved re/im).
void matxvec_treal_host(Handle& H, const double* v1, double* v2, Prec p);

struct DavTrace;  // per-iteration log (harvest mode), may be NULL
// guess: the caller's initial block when p.guess == ADC2G_GUESS_VECTORS (else NULL)
void davidson_run(Handle& H, const adc2g_dav_params& p, double* eigval, double* eigvec, adc2g_dav_stats* stats,
                  DavTrace* trace, const double* guess = nullptr);

// autoord 1, 2 also fill auto1 / auto2 (nrec+1 complex values each)
void sil_run(Handle& H, const adc2g_sil_params& p, const double* fvec, double* auto0, int64_t nrec_cap,
             int64_t* nrec_out, adc2g_sil_stats* stats, int autoord = 0, double* auto1 = nullptr,
             double* auto2 = nullptr);

// Laser-driven propagation (laser.hip): E(t) of the pulses (host), the
// operator -i H(t) on planar complex vectors of ndim + 1 rows (v, out: re
// column then im column, leading dimension ndim + 1; hv: scratch of the same
// shape), and the SIL driver.
void laser_efield(int npulse, const adc2g_pulse* pulses, double t, double E[3]);
void laser_check_pulses(int npulse, const adc2g_pulse* pulses);
void laser_apply(Handle& H, const double E[3], const double* v, double* hv, double* out, Prec p, hipStream_t s);
void matxvec_treal_laser_host(Handle& H, double t, int npulse, const adc2g_pulse* pulses, const double* v1,
                              double* v2);
void sil_laser_run(Handle& H, const adc2g_laser_params& p, int npulse, const adc2g_pulse* pulses, double* psi,
                   double* norms, double* flux, double* trdens, int64_t nrec_cap, int64_t* nrec_out,
                   adc2g_sil_stats* stats);
// the CAP flux of adc2_flux_cap for planar psi and dpsi/dt (device, ndim + 1 rows)
double laser_flux(Handle& H, const double* psi, const double* dtpsi);

// Chebyshev order-domain autocorrelation C_0..C_2K (auto_out, host) of the
// host vector q0 with the spectrum scaled by bounds[0..1] = (Emin, Emax).
void cheby_run(Handle& H, const double bounds[2], const double* q0, int order, Prec p, double* auto_out);
// Chebyshev filter-diagonalisation eigenstates of the coefficients coef
// ((order + 1) x nsel column-major, host): vecs (ndim x nsel, host, or NULL)
// and energies (nsel)
void cheby_fdstates_run(Handle& H, const double bounds[2], const double* q0, int order, int nsel, const double* coef,
                        Prec p, double* vecs, double* energies);
// the filter-diagonalisation states in V (device, ndim x nsel) normalised,
// copied to vecs (host, or NULL), energies v . H v
void fd_finish(Handle& H, double* V, int nsel, Prec p, double* vecs, double* energies);
// time-domain filter-diagonalisation states (calc_fdstates_tdfd): SIL steps
// of tout from fvec / |fvec|, the filter states of nfbas energies on
// [ebound[0], ebound[1]] with the window cos(pi t / (2 tfinal))^iwfunc,
// combined by f2e (nsel x nfbas column-major, host)
void sil_fdstates_run(Handle& H, const adc2g_sil_params& p, const double* fvec, int nfbas, const double ebound[2],
                      int iwfunc, int nsel, const double* f2e, double* vecs, double* energies,
                      adc2g_sil_stats* stats);

// Block Lanczos pseudospectrum (orthotype 0) from the host block q0
// (ndim x blocksize); outputs as adc2g_block_lanczos documents them.
void lanczos_run(Handle& H, const adc2g_lanczos_params& p, const double* q0, double* eigval, double* tmat,
                 const double* travec, double* tranmom, double* ritz, adc2g_lanczos_stats* stats);

// Imaginary-time relaxation (XSIL) of p.nstates states; outputs as
// adc2g_relax documents them (guess may be NULL).
void relax_run(Handle& H, const adc2g_relax_params& p, const double* guess, double* eigval, double* eigvec,
               double* residual, int32_t* niter, adc2g_relax_stats* stats);

// The hmlt.dia / hmlt.off files of the matrix (upper triangle, |h_ij| > minc,
// records of maxbl elements); noffd and nbuf may be NULL.
void hmlt_write(Handle& H, const char* dia_path, const char* off_path, double minc, int maxbl, int64_t* noffd,
                int32_t* nbuf);

// MP2 energy, D2 diagnostic and ground-state density (rho: nbas x nbas
// column-major); any output may be NULL.
void mp2_run(Handle& H, double* e_mp2, double* d2, double* rho);
// ADC(2) F-vector of the nbas x nbas column-major operator matrix dpl (host) into f (host, ndim)
void fvector_run(Handle& H, const double* dpl, double* f);
// occupied-virtual transition density (v x o column-major per vector, at s*n1)
// of nvec vectors whose 1h1p rows are x[s*ldx + r], r < n1 (device), and of
// host vectors (ldx >= ndim)
void trdens_vo_device(Handle& H, int nvec, const double* x, int64_t ldx, double* out);
void trdens_vo_run(Handle& H, int nvec, const double* x_host, int64_t ldx, double* out_host);

// Sub-space diagonalisation guess: the lowest bs eigenvectors of the stored
// matrix on the subdim ISs of lowest (descending: highest) H_jj, into V
// (ndim x bs, device).
void subdiag_guess(Handle& H, int subdim, int bs, double* V, bool descending, bool top);
// eigenvectors first .. first+count-1 of that block (first = -1: the highest count)
void subdiag_cols(Handle& H, int subdim, bool descending, int first, int count, double* V);
// block Lanczos start block on the device (V: n x bs): kind 8
// init_vec_doublesubspace, kind 9 init_vec_singlesubspace
void lanczos_guess(Handle& H, int kind, int subdim, int bs, double* V);

// Power method for the highest eigenpair (power_method, source/eigen/power.f90);
// vec (host, ndim) and niter may be NULL.  Throws ADC2G_ENOCONV when the
// residual stays above thrsh after maxit iterations.
void power_run(Handle& H, int subdim, int maxit, double thrsh, double* eval, double* vec, int* niter);

// Spectral bounds of spectral_bounds (source/propagation/specbounds.f90):
// method 1 'davlanc' (:85-160), 2 'davpower' (:166-238).
void spectral_bounds_run(Handle& H, int method, int ncycles, double bounds[2]);

int precon_nparams();
void precon_harvest(Handle& H, const adc2g_dav_params& p, int64_t max_records, int64_t* added);
void precon_train(Handle& H, const Handle& data, int epochs, int batch, double lr, uint64_t seed, double* loss_out);
void precon_get(const Handle& H, float* params);
void precon_absorb(Handle& dst, const Handle& src, int64_t* nrec);
void precon_set(Handle& H, const float* params);
void precon_eval(Handle& H, int64_t n, const float* feat, const int32_t* cls, const float* dn, const float* en,
                 float* m_out, float* grad_out, double* loss_out);

}  // namespace adc2g
