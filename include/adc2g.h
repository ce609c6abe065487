/*
 * adc2g.h -- C ABI of the MI355X (gfx950) ADC(2)-s sigma-vector library.
 *
 * The library replaces the stored-matrix hot path of ADC_ottawa: the matrix
 * build and on-disk COO list, the sparse sigma-vector consumed by the block
 * Davidson solver and the complex sigma-vector consumed by the
 * short-iterative-Lanczos (SIL) propagator.  Everything lives behind an
 * opaque handle; device memory is owned by the library and every vector,
 * subspace and factor stays resident in HBM between calls.
 *
 * Conventions (those of the Fortran caller):
 *  - real(dp) = double, default INTEGER = int32_t, integer*8 = int64_t,
 *    complex(dp) = two interleaved doubles (re, im);
 *  - matrices are column-major (leading dimension = number of rows);
 *  - vectors use the reference's configuration order: the n1 = nocc*nvirt
 *    1h1p configurations (virtual outer, occupied inner), then the 2h2p
 *    blocks t1 (a=b,i=j), t2 (a<b,i=j), t3 (a=b,i<j), t4 and t5 (a<b,i<j,
 *    spin cases I and II), each in the loop order of
 *    source/adclib/select_fano.f90:539-718;
 *  - every function returns 0 on success or one of the ADC2G_E* codes; the
 *    message of the last failure on the calling thread is available from
 *    adc2g_last_error.
 *
 * Functions are not re-entrant per handle.  HIP streams are internal; the
 * *_device entry points run on the stream given (NULL = the handle's own
 * stream) and return without synchronising.
 */
#ifndef ADC2G_H
#define ADC2G_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes; 1..4 are the SIL codes of source/propagation/sillib.f90:84-89 */
#define ADC2G_OK 0
#define ADC2G_ESIL_ORDER 1      /* illegal integration order */
#define ADC2G_ESIL_UNDERFLOW 2  /* stepsize underflow */
#define ADC2G_ESIL_DIAG 3       /* diagonalisation failed */
#define ADC2G_ESIL_RESTART 4    /* stepsize too large for restart */
#define ADC2G_EARG 10           /* bad argument */
#define ADC2G_EHIP 11           /* HIP runtime failure */
#define ADC2G_ENOCONV 12        /* Davidson: not all vectors have converged */
#define ADC2G_EEIG 13           /* Rayleigh-Ritz eigensolve failed */
#define ADC2G_ENOMEM 14         /* device allocation failed */
#define ADC2G_ESTATE 15         /* call out of order (e.g. no trained network) */
#define ADC2G_EBREAK 16         /* block Lanczos: the residual block lost rank */

/* arithmetic of the 1h1p<->2h2p coupling contractions */
#define ADC2G_PREC_FP64 0       /* FP64 MFMA */
#define ADC2G_PREC_FP32 1       /* FP32 MFMA, FP32 accumulate */
#define ADC2G_PREC_BF16 2       /* BF16 MFMA, FP32 accumulate */

/* Davidson preconditioners (precon of source/eigen/block_davidson.f90:1121-1125) */
#define ADC2G_PRE_DPR 1         /* diagonal preconditioned residue */
#define ADC2G_PRE_OLSEN 2       /* Olsen */
#define ADC2G_PRE_LEARNED 3     /* MLP correction on top of DPR, DPR fallback */
#define ADC2G_PRE_COUPLED 4     /* exact inverse of H - theta with the 2h2p diagonal shifted by a fixed
                                   sigma, through the 1h1p Schur complement (no counterpart in the
                                   reference, whose steps are diagonal) */
#define ADC2G_PRE_LEARNED_COUPLED 5  /* MLP correction on top of the coupled step, coupled-step fallback */

/* Evaluation of the 1h1p<->2h2p coupling (same matrix either way) */
#define ADC2G_COUPLING_AUTO 0   /* INT when it needs fewer flops and its integrals fit, else DF */
#define ADC2G_COUPLING_DF 1     /* contractions with the factors B[P,p,q] (sums over naux) */
#define ADC2G_COUPLING_INT 2    /* contractions with (vv|ov), (oo|ov) assembled once from the factors */
#define ADC2G_COUPLING_STORED 3 /* (reported only) a stored matrix, adc2g_create_from_hmlt */

typedef struct adc2g_handle adc2g_handle;

typedef struct {
  int32_t device;        /* HIP device ordinal */
  int32_t build_phph;    /* 1: build the second-order ph-ph block (ADC(2)); 0: first order only */
  int32_t keep_lowprec;  /* 1: also keep FP32/BF16 copies of the coupling operands for mixed precision */
  int32_t coupling;      /* ADC2G_COUPLING_* */
} adc2g_opts;

/* Davidson parameters (the keys of rddiaginp, source/iomodules/rdinput.f90:1081-1288) */
typedef struct {
  int32_t nstates;       /* davstates */
  int32_t blocksize;     /* dmain; 0 -> max(nstates+5, 2*nstates) */
  int32_t maxvec;        /* maxsubdim; <0 -> 4*blocksize; rounded down to a multiple of blocksize */
  int32_t maxiter;       /* maxiter */
  double tol;            /* davtol: convergence when ||r||_2 < tol */
  int32_t precon;        /* ADC2G_PRE_* */
  int32_t deflate;       /* ldfl */
  int32_t precision;     /* ADC2G_PREC_* of the sigma-vectors in the first phase */
  int32_t refine;        /* mixed precision: 1 = finish with FP64 sigma-vectors (restarted refinement) */
  double lowprec_tol;    /* residual at which the low-precision phase hands over to FP64 */
  int32_t guess;         /* ADC2G_GUESS_*: initial vectors (davinitialise, block_davidson.f90:149-242) */
  int32_t verbose;       /* 1: print the iteration table to stdout */
  int32_t subdim;        /* guessdim: ISs of the sub-diagonalisation guess (<= 0: 800) */
  int32_t reserved;
} adc2g_dav_params;

/* Davidson initial vectors */
#define ADC2G_GUESS_ONDIAG 0   /* unit vectors on the smallest |H_jj| (initvec_ondiag, :463-532) */
#define ADC2G_GUESS_SUBDIAG 1  /* lowest eigenvectors of H on the subdim ISs of lowest H_jj
                                  (initvec_subdiag, :293-459) */
#define ADC2G_GUESS_VECTORS 2  /* caller's block, e.g. ADC(1) eigenvectors (initvec_adc1,
                                  :246-289): adc2g_davidson_vectors */

typedef struct {
  int32_t niter;         /* Davidson iterations */
  int32_t nsigma;        /* sigma-vectors actually computed */
  int64_t nmult_ref;     /* sigma-vectors the reference's hxvec_incore would have computed (sum of currdim) */
  int32_t niter_lowprec; /* iterations done with low-precision sigma-vectors */
  int32_t nfallback;     /* learned-preconditioner steps replaced by DPR */
  double time_total;     /* seconds, wall */
  double time_sigma;     /* seconds inside sigma-vectors */
  double resnorm_max;    /* max FP64 residual norm of the returned roots */
  double time_precon;    /* seconds in the coupled preconditioner (set-up and steps) */
  int32_t nprecon_builds;/* Schur complements formed by the coupled preconditioner */
  int32_t reserved;
  double precon_sigma;   /* shift of the 2h2p diagonal the coupled preconditioner used */
} adc2g_dav_stats;

typedef struct {
  double tout;           /* time between autocorrelation records (a.u.) */
  double tfinal;         /* final time; records at i*tout, i = 1..int(tfinal/(tout/2)) */
  int32_t kdim;          /* krydim: maximum Lanczos order */
  int32_t precision;     /* ADC2G_PREC_* of the sigma-vectors */
  double autotol;        /* SIL error tolerance */
  int32_t verbose;       /* 1: a progress line on stderr every 100 records */
  int32_t reserved;
} adc2g_sil_params;

typedef struct {
  int64_t nsteps;        /* SIL steps taken */
  int64_t nsigma;        /* complex sigma-vectors (each = 2 real ones) */
  int32_t max_order;
  int32_t reserved;
  double time_total;
} adc2g_sil_stats;

typedef struct {
  int32_t blocksize;     /* lmain: vectors per Lanczos block */
  int32_t ncycles;       /* iter: blocks generated; nvec = blocksize * ncycles */
  int32_t precision;     /* ADC2G_PREC_* of the sigma-vectors */
  int32_t check_orth;    /* 1: stats->orth_loss = max |L^T L - I| over the Lanczos vectors */
  int32_t orthotype;     /* 0: none (default); 1: partial re-orthogonalisation (pro);
                            2: modified partial re-orthogonalisation (mpro) */
  int32_t reserved;
} adc2g_lanczos_params;

typedef struct {
  int32_t nvec;          /* Lanczos states = blocksize * ncycles */
  int32_t nsigma;        /* sigma-vectors applied */
  int32_t npro;          /* re-orthogonalisations (orthotype 1 and 2) */
  int32_t reserved;
  double time_sigma;     /* s in the sigma-vectors (device events) */
  double time_tdiag;     /* s in the eigen-decomposition of T (host) */
  double time_ritz;      /* s forming the Ritz vectors and transition moments */
  double time_total;     /* s, whole call */
  double orth_loss;      /* max |L^T L - I| when check_orth, else -1 */
} adc2g_lanczos_stats;

typedef struct {
  int32_t nstates;       /* davstates: states relaxed one after another */
  int32_t krydim;        /* kdim: Lanczos vectors per timestep */
  int32_t maxdim;        /* maxsubdim: variational subspace, a multiple of krydim (<= 0: 2 krydim) */
  int32_t maxiter;       /* maxiter: timesteps per state */
  double step;           /* stepsize: imaginary timestep (a.u.) */
  double tol;            /* davtol: convergence when |H psi - E psi| <= tol */
  int32_t precision;     /* ADC2G_PREC_* of the sigma-vectors */
  int32_t verbose;       /* 1: one line per timestep on stderr (wrtable_1vec) */
  int32_t integrator;    /* integrator: 0 or 1 XSIL (default), 2 Bulirsch-Stoer (bs_rlx),
                            3 RKF45 (rkf45_rlx) */
  int32_t reserved;
  double toler;          /* siltol: error tolerance of integrators 2 and 3 */
} adc2g_relax_params;

typedef struct {
  int32_t nconv;         /* states converged */
  int32_t reserved;
  int64_t nsigma;        /* sigma-vectors applied (nmult) */
  double time_sigma;     /* s in the sigma-vectors (device events) */
  double time_total;     /* s, whole call */
} adc2g_relax_stats;

/* ---- lifetime ---------------------------------------------------------- */

/* Build the on-device ADC(2)-s operator from orbital energies eps[nocc+nvirt]
 * (occupied first) and density-fitting factors B[naux][nbas][nbas]
 * (C order: B[(P*nbas + p)*nbas + q], symmetric in p,q) so that
 * (pq|rs) = sum_P B[P,p,q] B[P,r,s].
 * Replaces write_fspace_adc2_1 (source/adclib/fspacetrial.f90:533-562) with
 * get_diag_adc2_save/get_offdiag_adc2_save_omp (source/adclib/get_matrix.f90:267-311,
 * 705-1186) and the integral globals of source/include/parameters.f90. */
int32_t adc2g_create(int32_t nocc, int32_t nvirt, int32_t naux, const double* eps,
                     const double* B, const adc2g_opts* opts, adc2g_handle** out);

/* Free everything the handle owns.  An operator with CVS views alive is not
 * destroyed (ADC2G_ESTATE): destroy the views first. */
int32_t adc2g_destroy(adc2g_handle* h);

/* A handle on a stored matrix: the reference's on-disk files hmlt.dia* and
 * hmlt.off* (the format adc2g_write_hmlt writes: wrtdg / wrtoffdg,
 * source/adclib/filetools.f90:24-44), loaded as its in-core Davidson loads
 * them (rdham_on / rdham_off, source/eigen/block_davidson.f90:588-686), with
 * the product of hxvec_incore (:844-883): H_jj x_j plus, for every stored
 * element (i, j, h_ij), h_ij x_j on row i and h_ij x_i on row j.  The
 * elements go to device memory as CSR with both entries (12 bytes each, two
 * per stored element).  The sigma-vector of every precision is computed in
 * FP64.  The solvers (Davidson with the DPR and Olsen steps and the ondiag,
 * subdiag and vector guesses, SIL, laser propagation, Chebyshev, block
 * Lanczos, power method, relaxation, spectral bounds), adc2g_dims (ndim; n1
 * and counts 0), adc2g_diagonal, adc2g_sigma, adc2g_sigma_device,
 * adc2g_matxvec_treal and adc2g_write_hmlt accept the handle; the entry points
 * that need the orbitals (views, ADC(2)-x, ph-ph block, Schur complement,
 * coupled and learned steps, MP2, F-vector, transition density, the
 * no-allocation sigma-vector) return ADC2G_EARG.  Files with records split
 * into gfortran sub-records (over 2 GiB) and matrices of 2^31 rows or more
 * are rejected. */
int32_t adc2g_create_from_hmlt(const char* dia_path, const char* off_path, int32_t device, adc2g_handle** out);

/* CVS view of an operator (core-valence separation; select_atom_is_cvs,
 * source/adclib/select_fano.f90:112-142, select_atom_d_cvs, :903-1027, with the
 * element functions of get_offdiag_adc2_save_cvs, source/adclib/get_matrix.f90:1323-1726).
 * The configurations are:
 *  - the 1h1p (a, i) with i one of the ncore core orbitals (1-based occupied
 *    indices in core[]), a outer, i inner;
 *  - no a=b,i=j or a!=b,i=j configurations;
 *  - the t3, t4 and t5 configurations in which exactly one of i < j is a core
 *    orbital, each in the order of the full space.
 * The CVS matrix is the principal sub-matrix of the full one on these
 * configurations.  The view's sigma-vector scatters into the full space,
 * applies the full sigma-vector and gathers.  Every solver entry point
 * (Davidson, relaxation, SIL, Chebyshev, block Lanczos) and adc2g_dims,
 * adc2g_diagonal, adc2g_sigma* and adc2g_write_hmlt accept the view; the ph-ph
 * block and the learned preconditioner belong to the full operator.  The
 * full operator must outlive the view. */
int32_t adc2g_create_cvs(adc2g_handle* full, int32_t ncore, const int32_t* core, adc2g_handle** out);

/* Frozen-core view of a full operator: every configuration none of whose
 * holes is one of the nfrozen frozen orbitals (1-based occupied indices in
 * frozen[]), each group (1h1p, t1..t5) in the order of the full space.
 * Replaces the lifrzcore selection of select_atom_is / select_atom_d
 * (source/adclib/select_fano.f90:27-40, 575, 610, 635-642, 665-672), which
 * restricts only the configurations: the matrix elements, with their sums
 * over all occupied orbitals, are those of the full matrix, so the view is
 * the principal sub-matrix on these configurations.  (The reference's MP2
 * with frozen core, source/adclib/mp2.f90:84-90, sums over the active
 * occupied orbitals: adc2g_mp2 of the view does so, and equals adc2g_mp2 of a
 * handle made without the frozen orbitals.)  The same solver entry points as
 * a CVS view accept it, and adc2g_mp2. */
int32_t adc2g_create_frozen_core(adc2g_handle* full, int32_t nfrozen, const int32_t* frozen, adc2g_handle** out);

/* ADC(2)-x: with enable = 1 the operator's sigma-vector (every solver entry
 * point, and views made afterwards) includes the first-order 2h2p-2h2p block
 * of get_offdiag_adc2ext_save (source/adclib/get_matrix.f90:2435-2893, element
 * functions C_1_1 .. C_4ii_4ii of source/adclib/adc_ph.f90:1102-1645), and the
 * diagonal gains its diagonal (get_diag_adc2ext_direct); enable = 0 returns to
 * ADC(2)-s.  Full operator only, before any view is made.  The block is
 * applied in FP64 in chunks of up to 8 vectors and needs the v^4 (vv|vv)
 * matrix in device memory (12.8 GB at v = 200), built on the first enable,
 * plus seven (ov)^2 work arrays per vector of a chunk.  With it the
 * coupled Davidson step is that of the matrix whose 2h2p block is replaced by
 * its diagonal (an approximation of M(theta), formed from half sigma-vectors).
 * The FP32 and BF16 sigma-vectors add the block in FP64 too; the
 * no-allocation sigma-vector rejects an ADC(2)-x handle (ADC2G_EARG). */
int32_t adc2g_set_adc2x(adc2g_handle* h, int32_t enable);

/* Sizes: ndim, n1, and the counts of t1..t5 (counts[0..4]); kpq(1..5,0) of
 * source/adclib/select_fano.f90 with kpq(5,0) repeated for t5. */
int32_t adc2g_dims(const adc2g_handle* h, int64_t* ndim, int64_t* n1, int64_t counts[5]);

/* Coupling form the handle evaluates (ADC2G_COUPLING_DF or _INT; no
 * counterpart in the reference, which stores the matrix elements). */
int32_t adc2g_coupling_form(const adc2g_handle* h, int32_t* form);

/* Diagonal H_jj of the full matrix in configuration order: the hmlt.dia record
 * of get_diag_adc2_save (source/adclib/get_matrix.f90:267-311). */
int32_t adc2g_diagonal(adc2g_handle* h, double* hii);

/* The dense ph-ph block (n1 x n1, column-major): first and second order, as
 * assembled by get_offdiag_adc2_save_omp (source/adclib/get_matrix.f90:848-884)
 * plus the diagonal of get_diag_adc2_save. */
int32_t adc2g_phph(adc2g_handle* h, double* h11);

/* The 1h1p Schur complement of the coupled preconditioner (ADC2G_PRE_COUPLED;
 * no counterpart in the reference): a0 (n1 x n1, column-major) receives
 * A0 = H11 - C (D - sigma)^-1 C^T, with C the 1h1p<->2h2p coupling block and
 * D the 2h2p diagonal, for a shift sigma below every 2h2p diagonal element.
 * The contractions with C run in the given precision (ADC2G_PREC_*), the sum
 * in FP64.  The number of eigenvalues of A0 below sigma equals the number of
 * eigenvalues of H below sigma (Sylvester's law of inertia). */
int32_t adc2g_coupled_schur(adc2g_handle* h, double sigma, int32_t precision, double* a0);

/* ---- sigma-vectors ----------------------------------------------------- */

/* W = H X for nvec columns, host arrays (ndim x nvec, ld = ndim).
 * The +H primitive of hxvec_incore (source/eigen/block_davidson.f90:844-883). */
int32_t adc2g_sigma(adc2g_handle* h, int32_t nvec, const double* X, double* W, int32_t precision);

/* Same with device pointers (ld = ndim) on the given HIP stream. */
int32_t adc2g_sigma_device(adc2g_handle* h, int32_t nvec, const double* dX, double* dW,
                           int32_t precision, void* stream);

/* The same as adc2g_sigma_device, allocating nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  Every workspace (pair tensors, packed
 * 1h1p blocks, GEMM split-K partials) lies in one static device array of
 * 256 MiB per GPU that the library holds until the process ends; the vectors
 * go through it in passes of as many columns as fit (at o=16, v=200: 2 FP64,
 * 4 FP32 or 9 BF16 vectors per pass).  ADC2G_ENOMEM when one vector's
 * workspace does not fit (n1^2 elements of the precision plus its partials,
 * e.g. FP64 beyond n1 of about 4000); ADC2G_EARG for a CVS view or a
 * precision whose operands are not resident yet (create the handle with
 * keep_lowprec = 1, or run one adc2g_sigma_device of that precision first).
 * Calls on one GPU share the array: calls that overlap on one GPU race. */
int32_t adc2g_sigma_device_noalloc(adc2g_handle* h, int32_t nvec, const double* dX, double* dW,
                                   int32_t precision, void* stream);

/* v2 = -i H v1 for one complex vector (interleaved re/im, length ndim):
 * matxvec_treal (source/propagation/tdselib.f90:323-437). */
int32_t adc2g_matxvec_treal(adc2g_handle* h, const double* v1, double* v2, int32_t precision);

/* ---- solvers ------------------------------------------------------------ */

/* Block Davidson for the lowest nstates eigenpairs: davdiag_block
 * (source/eigen/block_davidson.f90:33-145) via master_eig
 * (source/adclib/adc_common.f90:952-993).  eigval[nstates] ascending,
 * eigvec (ndim x nstates, may be NULL), stats may be NULL. */
int32_t adc2g_davidson(adc2g_handle* h, const adc2g_dav_params* p, double* eigval,
                       double* eigvec, adc2g_dav_stats* stats);

/* The same with the initial vectors given (p->guess = ADC2G_GUESS_VECTORS):
 * guess is a host block of ndim x blocksize orthonormal columns (rejected if
 * max |G^T G - I| > 1e-8), e.g. the ADC(1) eigenvectors that initvec_adc1
 * (source/eigen/block_davidson.f90:246-289) reads from SCRATCH/adc1_vecs,
 * zero on the 2h2p configurations; blocksize must be given (> 0). */
int32_t adc2g_davidson_vectors(adc2g_handle* h, const adc2g_dav_params* p, const double* guess,
                               double* eigval, double* eigvec, adc2g_dav_stats* stats);

/* SIL wavepacket propagation from fvec (real, ndim): propagate_sillib
 * (source/propagation/fvecprop.f90:365-564) with SILStep
 * (source/propagation/sillib.f90:108-393).  auto0 receives nrec+1 complex
 * values (interleaved), the zeroth-order autocorrelation at t = i*tout,
 * i = 0..nrec with nrec = int(tfinal/(tout/2)); nrec_out may be NULL. */
int32_t adc2g_sil_autocorr(adc2g_handle* h, const adc2g_sil_params* p, const double* fvec,
                           double* auto0, int64_t nrec_cap, int64_t* nrec_out,
                           adc2g_sil_stats* stats);

/* The same with the higher-order autocorrelation functions of autoord
 * (propagate_sillib, source/propagation/fvecprop.f90:421-541): for
 * autoord >= 1, auto1 receives psi(t/2)^T H psi(t/2) |F|^2 (dot_product(psi0, H psi0) |F|^2
 * at t = 0), and for autoord = 2, auto2 receives the same with H^2; one
 * (autoord 1) or two (autoord 2) more complex sigma-vectors per record.
 * auto1 / auto2 hold nrec+1 complex values each and may be NULL when not
 * requested; autoord = 0 is adc2g_sil_autocorr. */
int32_t adc2g_sil_autocorr_orders(adc2g_handle* h, const adc2g_sil_params* p, const double* fvec,
                                  int32_t autoord, double* auto0, double* auto1, double* auto2,
                                  int64_t nrec_cap, int64_t* nrec_out, adc2g_sil_stats* stats);

/* Time-domain filter-diagonalisation states (calc_fdstates_tdfd,
 * source/propagation/fdstates_tdfd.f90:28-627, driven by adc2_fdstates): SIL
 * periods of p->tout (p->kdim, p->autotol, p->precision; p->tfinal) from
 * fvec / |fvec|; at t = r tout, r = 0 .. int(tfinal / tout), the filter
 * state of energy E_j = ebound[0] + j (ebound[1] - ebound[0]) / (nfbas - 1)
 * gains fac tout g(t) Re(exp(i E_j t) psi(t)) (fac 1 at the first and last
 * record, 2 between; g = cos(pi t / (2 tfinal))^iwfunc), and state i is
 * sum_j fbas2eig[i + nsel j] buffer_j (fbas2eig: the selected rows
 * fbas2eig(isel(:), :) of the reference, nsel x nfbas column-major),
 * normalised, with energies[i] = v_i . H v_i.  vecs: ndim x nsel or NULL. */
int32_t adc2g_sil_fdstates(adc2g_handle* h, const adc2g_sil_params* p, const double* fvec, int32_t nfbas,
                           const double* ebound, int32_t iwfunc, int32_t nsel, const double* fbas2eig,
                           double* vecs, double* energies, adc2g_sil_stats* stats);

/* ---- Laser-driven propagation (propagate_laser_adc2,
 * source/propagation/proplib_adc2.f90:37-630, with matxvec_treal_laser,
 * source/propagation/tdselib.f90:569-720).  The wavefunction has ndim + 1
 * complex components: the intermediate states in the handle's configuration
 * order, then the MP2 ground state (statenumber = 0 starts from
 * (0, ..., 0, 1)).  Complex arrays are interleaved (re, im). */

/* One pulse of efield (tdselib.f90:850-897): E(t) = pulse_vec * strength *
 * envelope(t) * carrier(t).  carrier: 1 cos(freq (t - t0) - phase), 2 sin of
 * the same.  envelope: 1 cos^2 (envpar = tzero, fwhm: cos^2(pi (t - tzero) /
 * (2 fwhm)) for |t - tzero| < fwhm), 2 sin^2 ramp (envpar[0] = tau:
 * sin^2(pi t / (2 tau)) for 0 <= t <= tau, 1 after), 3 box (envpar = ti, tf:
 * 1 for ti <= t <= tf), 4 sin^2 (envpar = tzero, fwhm: sin^2(pi (t - tzero) /
 * (2 fwhm)) for tzero < t < tzero + 2 fwhm); 0 outside (tdselib.f90:904-1029). */
typedef struct {
  int32_t carrier;      /* ipulse: 1 cosine, 2 sine */
  int32_t envelope;     /* ienvelope: 1 cos^2, 2 sin^2 ramp, 3 box, 4 sin^2 */
  double pulse_vec[3];  /* polarisation (x, y, z) */
  double strength, freq, t0, phase;
  double envpar[2];
} adc2g_pulse;

typedef struct {
  double tout;        /* output interval = integration period (intperiod = tout) */
  double tfinal;      /* records at i * tout, i = 0 .. int(tfinal / tout) */
  int32_t kdim;       /* maximum Krylov order of a SIL step, 2..30 */
  int32_t precision;  /* ADC2G_PREC_* of the sigma-vectors */
  double proptol;     /* SIL error tolerance */
  int32_t verbose;
  int32_t reserved;
} adc2g_laser_params;

/* The intermediate-state representation of dipole component c (0 x, 1 y,
 * 2 z) that the reference reads from SCRATCH/dipole_<c> (opxvec_ext,
 * tdselib.f90:1031-1153: the diagonal, then each off-diagonal element once
 * and applied to both (i, j) and (j, i)): diag (ndim), nnz off-diagonal
 * elements (row, col, val) with 0-based row != col, the ground-state element
 * d00 and the ground-state column d0j (ndim).  Replaces any earlier set of c;
 * nnz = 0 and diag = d0j = NULL with d00 = 0 clears it. */
int32_t adc2g_set_dipole_is(adc2g_handle* h, int32_t component, const double* diag, int64_t nnz,
                            const int64_t* row, const int64_t* col, const double* val, double d00,
                            const double* d0j);

/* The complex absorbing potential: its intermediate-state representation W
 * (SCRATCH/cap, the same layout as the dipoles), ground-state element w00 and
 * column w0j.  With a CAP set, adc2g_matxvec_treal_laser adds -W v, the
 * -i (-i W) v of tdselib.f90:722-830, and adc2g_sil_laser integrates with the
 * short-iterative Lanczos-Arnoldi method (propagate_wavepacket_csil,
 * proplib_adc2.f90:659-830; CSILStep with the improved error criterion,
 * csillib.f90:120-520).  Clearing as for adc2g_set_dipole_is. */
int32_t adc2g_set_cap_is(adc2g_handle* h, const double* diag, int64_t nnz, const int64_t* row,
                         const int64_t* col, const double* val, double w00, const double* w0j);

/* The projector of the CAP term (iprojcap, tdselib.f90:735-830): the CAP
 * acts as P W P with P the projector onto the complement of nvec orthonormal
 * intermediate-state vectors (vecs: ndim x nvec, column-major: the selected
 * bound states the reference reads from its eigenvector file) and, with
 * ground = 1, of the ground state.  nvec = 0 and ground = 0 remove it. */
int32_t adc2g_set_cap_projector(adc2g_handle* h, int32_t nvec, const double* vecs, int32_t ground);

/* v2 = -i H(t) v1 (ndim + 1 complex components, interleaved): the -i H of
 * matxvec_treal on the intermediate states plus, for E(t) = sum of the pulses,
 * + i sum_c E_c [(D_c + d00_c) v1 + d0j_c v1_0] on the intermediate states and
 * + i sum_c E_c (d00_c v1_0 + d0j_c . v1) on the ground state
 * (matxvec_treal_laser, tdselib.f90:569-720), and the CAP term when a CAP is
 * set. */
int32_t adc2g_matxvec_treal_laser(adc2g_handle* h, double time, int32_t npulse, const adc2g_pulse* pulses,
                                  const double* v1, double* v2);

/* SIL propagation under the pulses (propagate_wavepacket_sil,
 * proplib_adc2.f90:486-630; with a CAP the Lanczos-Arnoldi steps of
 * propagate_wavepacket_csil, :659-830): psi (ndim + 1 complex, interleaved)
 * in and out; the periods i = 1 .. int(tfinal / tout) each integrated by
 * steps with H at the step's start time (Func(Time, ...)), a period
 * skipped when |dpsi/dt| <= 1e-10; norms[i] = |psi| at i * tout for
 * i = 0 .. nrec (the wrstepinfo values).  norms holds nrec_cap values.
 * flux (nrec_cap values, or NULL) receives the CAP flux of adc2_flux_cap
 * (flux.f90:179-411) at the same times; it needs a CAP and the Theta
 * operator (adc2g_set_flux_theta).  trdens (nrec_cap * 2 * nocc * nvirt
 * doubles, or NULL) receives at the same times the input of the TD-NTO
 * analysis of tdadc2_nto (nto.f90:382-547): the occupied-virtual transition
 * density (adc2g_trdens_vo) of psi_IS / |psi_IS|, complex (re, im) per
 * element, element (a, i) at a + nvirt * i of each record (zero when psi_IS
 * is zero).  The reference analyses records i >= 1 with |psi_IS| > 1e-4. */
int32_t adc2g_sil_laser(adc2g_handle* h, const adc2g_laser_params* p, int32_t npulse,
                        const adc2g_pulse* pulses, double* psi, double* norms, double* flux,
                        double* trdens, int64_t nrec_cap, int64_t* nrec_out, adc2g_sil_stats* stats);

/* The occupied-virtual block of the ADC(2) ground-to-excited-state
 * transition density of adc2_trdens_gs_virt_occ (density_matrix.f90:497-636)
 * for nvec vectors x (ndim rows each, leading dimension ldx >= ndim; only the
 * 1h1p rows enter) into trdens (nvirt x nocc column-major per vector, vector s
 * at s * nocc * nvirt): the trdens(nocc+1:nbas, 1:nocc) that adc2_nto_gs and
 * tdadc2_nto (nto.f90:76-260, 382-547) decompose into natural transition
 * orbitals.  Uses the MP2 density (as adc2g_mp2, so the (vv|ov) block must be
 * in memory) and keeps an n1 x n1 FP64 matrix on the device after the first
 * call (n1 = nocc * nvirt).  On a frozen-core view the view's 1h1p
 * amplitudes enter at their configurations, with the frozen-core MP2 density,
 * as the reference's loops over the frozen-core kpqf; on a CVS view the same
 * with the full MP2 density; trdens then covers all nocc x nvirt orbital
 * pairs.  Not available on a stored-matrix handle. */
int32_t adc2g_trdens_vo(adc2g_handle* h, int32_t nvec, const double* x, int64_t ldx, double* trdens);

/* The intermediate-state representation of the projector Theta onto the
 * absorbing region (SCRATCH/theta, the same layout as the dipoles), with
 * theta00 and theta0j, for the flux of adc2_flux_cap (flux.f90:179-411):
 *   flux = 2 Re <dpsi/dt| P Theta' P |psi> + 2 Re <psi| P W' P |psi>,
 * where, as the reference writes them, Theta' and W' add theta00 (w00) times
 * the operator's own IS output, (1 + theta00) Theta_IS u, rather than times u.
 * Clearing as for adc2g_set_dipole_is. */
int32_t adc2g_set_flux_theta(adc2g_handle* h, const double* diag, int64_t nnz, const int64_t* row,
                             const int64_t* col, const double* val, double theta00, const double* theta0j);

/* Chebyshev order-domain autocorrelation function (chebyshev_auto,
 * source/propagation/chebyspec.f90:245-319, with the scaled product of
 * matxvec_chebyshev_incore, source/propagation/tdselib.f90:1316-1398):
 * H_norm = (2/dE) H - (2 Emin/dE + 1), dE = Emax - Emin, bounds = {Emin, Emax}
 * (the reference estimates them with spectral_bounds and widens them by
 * 0.9 / 1.1, which assumes positive energies; here the caller passes bounds
 * that contain the spectrum); q_1 = H_norm q0, q_k = 2 H_norm q_{k-1} - q_{k-2};
 * auto_out[k] = q0.q_k for k <= order and, for k > order/2,
 * auto_out[2k] = 2 q_k.q_k - auto_out[0], auto_out[2k-1] = 2 q_{k-1}.q_k - auto_out[1]:
 * 2*order + 1 doubles.  q0: host vector of ndim.  precision: ADC2G_PREC_* of
 * the sigma-vectors (recursion and dot products in FP64). */
int32_t adc2g_cheby_auto(adc2g_handle* h, const double* bounds, const double* q0, int32_t order,
                         int32_t precision, double* auto_out);

/* Chebyshev filter-diagonalisation eigenstates (calc_eigenstates of
 * fdstates_cfd, source/propagation/fdstates_cfd.f90:207-310, called by
 * adc2_fdstates): with the recursion of adc2g_cheby_auto from q0 (ndim) and
 * the bounds, v_i = sum_{k=0..order} coef[k + (order + 1) i] q_k for the nsel
 * selected states (coef: the reference's k2eig(0:chebyord, isel(i)) columns),
 * normalised, and energies[i] = v_i . H v_i.  vecs: ndim x nsel column-major,
 * or NULL.  precision as for adc2g_cheby_auto (accumulation in FP64). */
int32_t adc2g_cheby_fdstates(adc2g_handle* h, const double* bounds, const double* q0, int32_t order,
                             int32_t nsel, const double* coef, int32_t precision, double* vecs,
                             double* energies);

/* Spectral bounds of the matrix as spectral_bounds
 * (source/propagation/specbounds.f90) estimates them.
 *  - method 3 'lanczos' (spectral_bounds_lanczos, :52-105): the lowest and
 *    highest eigenvalues of block Lanczos (block 2, ncycles cycles, 0 meaning
 *    the reference's 50, orthotype 0) started from the double-subspace guess
 *    (adc2g_lanczos_guess kind 8, 800 ISs).
 *  - methods 1 and 2: bounds[0] is the lowest root of block Davidson (1 root, tol 1e-8, Olsen,
 *    deflation, sub-diagonalisation guess on 800 ISs; block/subspace 5/20 for
 *    method 1, 2/10 for method 2).
 *  - bounds[1], method 1 'davlanc' (:85-160): the highest eigenvalue of block
 *    Lanczos (block 2, ncycles cycles, 0 meaning the reference's 100, partial
 *    re-orthogonalisation) started from the lowest eigenvectors of the 800
 *    ISs of highest H_jj (init_vec_singlesubspace,
 *    source/eigen/block_lanczos.f90:748-923).
 *  - bounds[1], method 2 'davpower' (:166-238): the power method
 *    (adc2g_power with its defaults).
 * The reference's Chebyshev driver then multiplies them by 0.9 and 1.1
 * (chebyspec.f90:82-83), which contains the spectrum only for positive
 * bounds; here they are returned unwidened.  ADC2G_ENOCONV when the
 * Davidson or the power iterations do not converge (the reference stops). */
int32_t adc2g_spectral_bounds(adc2g_handle* h, int32_t method, int32_t ncycles, double* bounds);

/* Start block of block Lanczos from sub-diagonalisations of the matrix, into
 * q0 (host, ndim x blocksize, column-major):
 *  - kind 8, init_vec_doublesubspace (source/eigen/block_lanczos.f90:443-744):
 *    columns 1..bs/2 the lowest eigenvectors of the subdim ISs of lowest
 *    H_jj; columns bs/2+1..bs eigenvectors bs/2+1..bs (ascending eigenvalue
 *    order, as the reference's column loop takes them) of the subdim ISs of
 *    highest H_jj; blocksize even;
 *  - kind 9, init_vec_singlesubspace (:748-923): the lowest blocksize
 *    eigenvectors of the subdim ISs of highest H_jj.
 * subdim 0 means the reference's 800; it grows while the next H_jj is within
 * 1e-6 of the last (at most 4096).  The sub-matrix elements come from
 * sigma-vectors of unit vectors, with |h| <= 1e-12 dropped as the stored
 * matrix drops them.  The two halves of kind 8 are orthonormal when the two
 * IS sets are disjoint (ndim >= 2 subdim) or when they are the whole space. */
int32_t adc2g_lanczos_guess(adc2g_handle* h, int32_t kind, int32_t subdim, int32_t blocksize, double* q0);

/* Highest eigenpair by the power method (power_method,
 * source/eigen/power.f90:22-100).  The guess is the highest eigenvector of the
 * stored matrix on the subdim (<= 0: 800, the reference's fixed value) ISs of
 * highest H_jj (power_guess, :307-463).  Up to maxit (<= 0: 350) iterations
 * of v2 = H v1, E = v1.v2, r = |E v1 - v2|, v1 = v2 / |v2|, until
 * r <= thrsh (<= 0: 1e-8) (power_iterations, :467-524).  eval and vec (host,
 * ndim, may be NULL) receive (E, v1); niter (may be NULL) the iterations.
 * ADC2G_ENOCONV if r stays above thrsh. */
int32_t adc2g_power(adc2g_handle* h, int32_t subdim, int32_t maxit, double thrsh, double* eval,
                    double* vec, int32_t* niter);

/* Block Lanczos pseudospectrum: run_block_lanczos
 * (source/eigen/block_lanczos.f90:947-1256) with hxq_incore (:1325-1361)
 * replaced by the sigma-vector of the block.  orthotype 0 (default) does no
 * re-orthogonalisation.  orthotype 1 adds partial re-orthogonalisation: pro
 * with the omega recurrence (:1474-1639), pro_mgs against the stored blocks
 * (:1643-1706) and localro (:1430-1470), with block Gram-Schmidt steps in
 * place of the vector-by-vector ones.  orthotype 2 is the modified partial
 * re-orthogonalisation (mpro, mpro_mgs, lowerbound, upperbound, :1710-1890):
 * the same omega recurrence, Q_{j+1} alone against the intervals of stored
 * blocks around the estimates that reach sqrt(eps), the widened intervals once
 * more in the next cycle, no localro.
 *   U = H Q_j - Q_{j-1} B_{j-1}^T,  A_j = Q_j^T U,  R = U - Q_j A_j = Q_{j+1} B_j
 * with B_j upper triangular.  The reference factors R with dgeqrf/dorgqr; here a
 * shifted Cholesky QR with two refinement passes gives the factor with a
 * positive diagonal, so Q_{j+1} and B_j differ from the reference's by the
 * signs of their columns and rows.  T is then D T_ref D for a diagonal D of
 * +-1: the same eigenvalues, Ritz vectors and |transition moments|.
 * q0: host block (ndim x blocksize) of orthonormal columns (the initial
 * vectors of init_vec, :128-197; rejected if max |q0^T q0 - I| > 1e-8).
 * Outputs (host; every one but eigval may be NULL):
 *   eigval[nvec]     eigenvalues of T ascending (dsbev, diagmat_banded :2006-2052);
 *   tmat[nvec*nvec]  T, block tridiagonal, diagonal blocks from the upper
 *                    triangle of A_j (the fill order of :1194-1217);
 *   tranmom[nvec]    travec . r_n for the normalised Ritz vectors r_n (the
 *                    products get_tranmom_1 forms, source/adclib/fspacetrial.f90:1076-1156;
 *                    travec and tranmom are both NULL or both given);
 *   ritz[ndim*nvec]  the normalised Ritz vectors r_n = L y_n / |L y_n|
 *                    (ritzvecs_incore :2237-2300), in the order of eigval.
 * The Lanczos vectors stay in HBM (8 ndim nvec bytes); the eigen-decomposition
 * of T runs on the host, as dsbev does in the reference.  Returns ADC2G_EBREAK
 * when a residual block is rank deficient (the Krylov space is exhausted). */
int32_t adc2g_block_lanczos(adc2g_handle* h, const adc2g_lanczos_params* p, const double* q0,
                            double* eigval, double* tmat, const double* travec, double* tranmom,
                            double* ritz, adc2g_lanczos_stats* stats);

/* Imaginary-time relaxation to the lowest eigenstates: relaxation
 * (source/propagation/relaxation.f90:36-133) with its default integrator,
 * XSIL (xsil_rlx, :139-309).  For each state s in turn: per timestep, krydim
 * Lanczos vectors from psi, whose first vector is orthogonalised against the
 * states already converged (get_lancvecs_current, :1207-1326).  psi is
 * propagated by exp(-H step) in the variational subspace of the last
 * maxdim/krydim timesteps (subspace_matrices, :1330-1381;
 * solve_subspace_tdse, :1385-1491; Lowdin canonical orthogonalisation that
 * drops overlap eigenvalues |sigma| <= 1e-6, canonical_ortho, :1495-1606;
 * subspace collapse, :1163-1203) and normalised.  It stops at
 * |H psi - E psi| <= tol (residual_1vec, :1610-1651).
 * guess: host ndim x nstates initial vectors, or NULL for unit vectors on the
 * nstates smallest |H_jj| (initvec_ondiag, :875-944).  eigval[nstates];
 * eigvec (ndim x nstates), residual[nstates], niter[nstates] and stats may
 * be NULL.  Returns ADC2G_ENOCONV if a state does not converge within maxiter
 * steps (the reference stops there; the outputs hold the states up to and
 * including that one), ADC2G_EBREAK if a Lanczos vector cannot be formed. */
int32_t adc2g_relax(adc2g_handle* h, const adc2g_relax_params* p, const double* guess,
                    double* eigval, double* eigvec, double* residual, int32_t* niter,
                    adc2g_relax_stats* stats);

/* ---- MP2 ground state ----------------------------------------------------- */

/* MP2 quantities of the operator's system (closed shell, no frozen core):
 *  - e_mp2: the correlation energy of mp2_ener (source/adclib/mp2.f90:68-154);
 *  - d2: the D2 diagnostic of mp2_d2 (:161-276);
 *  - rho: the nbas x nbas column-major ground-state density of rho_mp2
 *    (:301-413, rhogs2_uu/oo/ou :421-596), assembled as the reference's loops
 *    leave it: 2 + oo(i,i) on the occupied diagonal, oo(i,j) + oo(j,i) on the
 *    occupied off-diagonal, uu and ou assigned once and mirrored.
 * Any output may be NULL.  rho needs the (vv|ov) block in memory
 * (8 v^2 o v bytes).  On a frozen-core view: the frozen-core quantities of
 * lifrzcore (mp2.f90:84-90, 333-337; every occupied sum over the active
 * orbitals, the frozen rows of rho 2 on the diagonal only), rho over all
 * nbas orbitals.  Not available on a CVS view. */
int32_t adc2g_mp2(adc2g_handle* h, double* e_mp2, double* d2, double* rho);

/* ---- F-vector ---------------------------------------------------------------- */

/* The ADC(2) F-vector f_J = <Psi_J|D|Psi_0> of a one-electron operator with
 * the nbas x nbas column-major matrix dpl (the reference's dpl(p,q) in the MO
 * basis, used with the reference's index order), into f (ndim, the
 * configuration order of adc2g_sigma).  Replaces get_modifiedtm_adc2
 * (source/adclib/get_moment.f90:153-442) with the term functions of
 * source/adclib/dipole_ph.f90: 1h1p f = -sqrt(2) (dpl(a,k) + F0 + F25..F210 +
 * FA + FB + FC + F21..F24), 2h2p FI, FII, FIII, FIV1, FIV2 (closed shell,
 * C1 symmetry).  Needs the (vv|ov) block and eight n1 x n1 matrices in
 * device memory.  On a CVS view (adc2g_create_cvs) f has the view's ndim and
 * holds the elements of the CVS configurations, as get_modifiedtm_adc2 with
 * lcvs computes them (:372-373 skip the I and II blocks, which the CVS space
 * does not have). */
int32_t adc2g_fvector(adc2g_handle* h, const double* dpl, double* f);

/* ---- on-disk matrix ------------------------------------------------------ */

/* Write the matrix in the reference's on-disk format so that its readers
 * (rdham_on / rdham_off, source/eigen/block_davidson.f90:600-686) can load it:
 * dia_path gets the records (maxbl, nbuf) and H_jj (wrtdg,
 * source/adclib/filetools.f90:24-34); off_path gets nbuf records
 * (buff(maxbl), oi(maxbl), oj(maxbl), nlim) (wrtoffdg, :36-44) holding the
 * upper triangle i < j (1-based) with |h_ij| > minc (the reference keeps
 * |h| > minc = 1e-12), full records and a last, possibly empty one, as
 * get_offdiag_adc2_save_omp (source/adclib/get_matrix.f90:705-1186) writes
 * them.  maxbl <= 0 means 655360 (buf_size, source/include/parameters.f90:425).
 * The columns are formed by FP64 sigma-vectors of unit vectors, so the cost
 * is that of ndim sigma-vectors: meant for systems whose stored matrix the
 * reference itself could hold.  noffd (elements written) and nbuf may be NULL. */
int32_t adc2g_write_hmlt(adc2g_handle* h, const char* dia_path, const char* off_path, double minc,
                         int32_t maxbl, int64_t* noffd, int32_t* nbuf);

/* ---- learned preconditioner -------------------------------------------- */

/* Run FP64 Davidson on this handle's system with the base step of the
 * network (the reference DPR step, or the coupled step when p->precon is
 * ADC2G_PRE_COUPLED or ADC2G_PRE_LEARNED_COUPLED) and append (features, base
 * step, exact correction) records of every unconverged root and iteration to
 * the handle's training set (kept in HBM); at most max_records are kept per
 * call.  Returns the number added. */
int32_t adc2g_precon_harvest(adc2g_handle* h, const adc2g_dav_params* p, int64_t max_records,
                             int64_t* added);

/* Train the network on the records of `data` (a handle that harvested; may
 * be h itself) with Adam for `epochs` passes of mini-batches of `batch`
 * records; `loss_out[epochs]` (may be NULL) receives the mean loss per pass. */
int32_t adc2g_precon_train(adc2g_handle* h, const adc2g_handle* data, int32_t epochs,
                           int32_t batch, double lr, uint64_t seed, double* loss_out);

/* Append the training records harvested on src to those of dst (both stay
 * valid); *nrec (may be NULL) receives dst's record count afterwards. */
int32_t adc2g_precon_absorb(adc2g_handle* dst, const adc2g_handle* src, int64_t* nrec);

/* Copy the network to/from host memory (flat float32 array of
 * adc2g_precon_nparams() values in the order W1, E, b1, W2, b2, w3, b3). */
int32_t adc2g_precon_nparams(void);
int32_t adc2g_precon_get(const adc2g_handle* h, float* params);
int32_t adc2g_precon_set(adc2g_handle* h, const float* params);

/* Evaluate the network (forward) and its gradient of the mean loss
 * (backward) on host-supplied records, for checking against a CPU oracle.
 * feat[n*3] (transformed features), cls[n], dn[n], en[n]; m_out[n] and
 * grad_out[nparams] may be NULL. */
int32_t adc2g_precon_eval(adc2g_handle* h, int64_t n, const float* feat, const int32_t* cls,
                          const float* dn, const float* en, float* m_out, float* grad_out,
                          double* loss_out);

/* ---- diagnostics ---------------------------------------------------------- */

/* Copy the last error message (NUL-terminated, truncated to len). */
int32_t adc2g_last_error(char* buf, int32_t len);

/* Per-kernel device timing of sigma-vectors (no counterpart in the
 * reference; its timers are wall clocks around whole solver calls,
 * source/utilities/timingmod.f90:37-55).  enable = 1 starts recording HIP
 * events around the main coupling contractions (class 0: the sums over the
 * virtual index, or over naux in the DF form), the CSF -> pair-tensor packing
 * (class 1), the pair -> CSF projection fused with the 2h2p diagonal (class 2)
 * and the remaining coupling contractions (class 3) of every following
 * sigma-vector.  If out is not NULL it receives, for classes 0..3,
 * {milliseconds, launches, work} (12 doubles; work in algorithmic flops for
 * classes 0 and 3 and algorithmic HBM bytes for 1 and 2) of everything
 * recorded since the previous call, and the records are reset. */
int32_t adc2g_sigma_profile(adc2g_handle* h, int32_t enable, double* out);

#ifdef __cplusplus
}
#endif

#endif /* ADC2G_H */
