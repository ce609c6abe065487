!=======================================================================
! adc2g_mod: iso_c_binding interface to libadc2g (include/adc2g.h).
!
! The Fortran side of the drop-in.  A caller that today does
!   call write_fspace_adc2_1(ndim,kpq,noffdel,'i')   ! fspacetrial.f90:533
!   call master_eig(ndim,noffdel,'i')                 ! adc_common.f90:952
! builds an operator handle from its orbital energies and factors instead,
! runs adc2g_davidson, and writes the eigenpair file with
! adc2g_write_davstates in the record layout of wreigenpairs
! (source/eigen/block_davidson.f90:1515-1523), so readdavvc
! (source/adclib/adc_common.f90:999-1052) reads it unchanged.  Errors are
! turned into a message for error_control (source/iomodules/iomod.f90:94-112)
! through adc2g_errmsg.
!=======================================================================
module adc2g_mod

  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: ADC2G_OK = 0
  integer(c_int32_t), parameter :: ADC2G_PREC_FP64 = 0, ADC2G_PREC_FP32 = 1, ADC2G_PREC_BF16 = 2
  integer(c_int32_t), parameter :: ADC2G_PRE_DPR = 1, ADC2G_PRE_OLSEN = 2, ADC2G_PRE_LEARNED = 3, &
       ADC2G_PRE_COUPLED = 4, ADC2G_PRE_LEARNED_COUPLED = 5
  integer(c_int32_t), parameter :: ADC2G_COUPLING_AUTO = 0, ADC2G_COUPLING_DF = 1, ADC2G_COUPLING_INT = 2
  integer(c_int32_t), parameter :: ADC2G_GUESS_ONDIAG = 0, ADC2G_GUESS_SUBDIAG = 1, ADC2G_GUESS_VECTORS = 2
  real(c_double), parameter :: adc2g_eh2ev = 27.2113845d0       ! parameters.f90:192
  real(c_double), parameter :: adc2g_au2fs = 1.0d0/41.34137333656d0 ! fvecprop.f90:19

  type, bind(c) :: adc2g_opts
     integer(c_int32_t) :: device = 0
     integer(c_int32_t) :: build_phph = 1
     integer(c_int32_t) :: keep_lowprec = 0
     integer(c_int32_t) :: coupling = 0        ! ADC2G_COUPLING_*
  end type adc2g_opts

  type, bind(c) :: adc2g_dav_params
     integer(c_int32_t) :: nstates
     integer(c_int32_t) :: blocksize = 0
     integer(c_int32_t) :: maxvec = -1
     integer(c_int32_t) :: maxiter = 100
     real(c_double)     :: tol = 1.0d-7
     integer(c_int32_t) :: precon = 2
     integer(c_int32_t) :: deflate = 1
     integer(c_int32_t) :: precision = 0
     integer(c_int32_t) :: refine = 1
     real(c_double)     :: lowprec_tol = 1.0d-4
     integer(c_int32_t) :: guess = 0        ! ADC2G_GUESS_*
     integer(c_int32_t) :: verbose = 0
     integer(c_int32_t) :: subdim = 800     ! guessdim of the sub-diagonalisation guess
     integer(c_int32_t) :: reserved = 0
  end type adc2g_dav_params

  type, bind(c) :: adc2g_dav_stats
     integer(c_int32_t) :: niter, nsigma
     integer(c_int64_t) :: nmult_ref
     integer(c_int32_t) :: niter_lowprec, nfallback
     real(c_double)     :: time_total, time_sigma, resnorm_max
     real(c_double)     :: time_precon
     integer(c_int32_t) :: nprecon_builds, reserved
     real(c_double)     :: precon_sigma
  end type adc2g_dav_stats

  type, bind(c) :: adc2g_sil_params
     real(c_double)     :: tout, tfinal
     integer(c_int32_t) :: kdim = 10
     integer(c_int32_t) :: precision = 0
     real(c_double)     :: autotol = 1.0d-5
     integer(c_int32_t) :: verbose = 0
     integer(c_int32_t) :: reserved = 0
  end type adc2g_sil_params

  type, bind(c) :: adc2g_sil_stats
     integer(c_int64_t) :: nsteps, nsigma
     integer(c_int32_t) :: max_order, reserved
     real(c_double)     :: time_total
  end type adc2g_sil_stats

  ! laser pulse of efield (source/propagation/tdselib.f90:850-1029):
  ! carrier = ipulse, envelope = ienvelope, envpar as the reference's
  type, bind(c) :: adc2g_pulse
     integer(c_int32_t) :: carrier = 1, envelope = 1
     real(c_double)     :: pulse_vec(3) = 0.0d0
     real(c_double)     :: strength = 0.0d0, freq = 0.0d0, t0 = 0.0d0, phase = 0.0d0
     real(c_double)     :: envpar(2) = 0.0d0
  end type adc2g_pulse

  ! laser propagation (propagate_laser_adc2, proplib_adc2.f90:37-830)
  type, bind(c) :: adc2g_laser_params
     real(c_double)     :: tout, tfinal
     integer(c_int32_t) :: kdim = 20
     integer(c_int32_t) :: precision = 0
     real(c_double)     :: proptol = 1.0d-8
     integer(c_int32_t) :: verbose = 0
     integer(c_int32_t) :: reserved = 0
  end type adc2g_laser_params

  ! block Lanczos (lancdiag_block, source/eigen/block_lanczos.f90:28-81):
  ! blocksize = lmain, ncycles = iter of the lanczos section
  type, bind(c) :: adc2g_lanczos_params
     integer(c_int32_t) :: blocksize = 0
     integer(c_int32_t) :: ncycles = 0
     integer(c_int32_t) :: precision = 0
     integer(c_int32_t) :: check_orth = 0
     integer(c_int32_t) :: orthotype = 0      ! 0 none, 1 partial (pro), 2 modified partial (mpro)
     integer(c_int32_t) :: reserved = 0
  end type adc2g_lanczos_params

  ! relaxation (source/propagation/relaxation.f90:36-133, XSIL integrator):
  ! nstates = davstates, krydim = kdim, maxdim = maxsubdim, step = stepsize,
  ! tol = davtol of the relaxation input
  type, bind(c) :: adc2g_relax_params
     integer(c_int32_t) :: nstates = 0
     integer(c_int32_t) :: krydim = 10
     integer(c_int32_t) :: maxdim = -1
     integer(c_int32_t) :: maxiter = 100
     real(c_double)     :: step = 10.0d0
     real(c_double)     :: tol = 1.0d-7
     integer(c_int32_t) :: precision = 0
     integer(c_int32_t) :: verbose = 0
     integer(c_int32_t) :: integrator = 1      ! 1 XSIL, 2 Bulirsch-Stoer, 3 RKF45
     integer(c_int32_t) :: reserved = 0
     real(c_double)     :: toler = 1.0d-5      ! siltol: error tolerance of integrators 2 and 3
  end type adc2g_relax_params

  type, bind(c) :: adc2g_relax_stats
     integer(c_int32_t) :: nconv, reserved
     integer(c_int64_t) :: nsigma
     real(c_double)     :: time_sigma, time_total
  end type adc2g_relax_stats

  type, bind(c) :: adc2g_lanczos_stats
     integer(c_int32_t) :: nvec, nsigma, npro, reserved
     real(c_double)     :: time_sigma, time_tdiag, time_ritz, time_total, orth_loss
  end type adc2g_lanczos_stats

  interface
     ! a stored-matrix handle from the reference's hmlt.dia / hmlt.off files
     integer(c_int32_t) function adc2g_create_from_hmlt(dia_path, off_path, device, out) &
          bind(c, name='adc2g_create_from_hmlt')
       import :: c_int32_t, c_char, c_ptr
       character(kind=c_char), intent(in) :: dia_path(*), off_path(*)
       integer(c_int32_t), value :: device
       type(c_ptr), intent(out) :: out
     end function adc2g_create_from_hmlt
     integer(c_int32_t) function adc2g_create(nocc, nvirt, naux, eps, B, opts, h) bind(c, name='adc2g_create')
       import :: c_int32_t, c_double, c_ptr, adc2g_opts
       integer(c_int32_t), value :: nocc, nvirt, naux
       real(c_double), intent(in) :: eps(*), B(*)
       type(adc2g_opts), intent(in) :: opts
       type(c_ptr), intent(out) :: h
     end function adc2g_create

     integer(c_int32_t) function adc2g_destroy(h) bind(c, name='adc2g_destroy')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
     end function adc2g_destroy

     integer(c_int32_t) function adc2g_dims(h, ndim, n1, counts) bind(c, name='adc2g_dims')
       import :: c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int64_t), intent(out) :: ndim, n1, counts(5)
     end function adc2g_dims

     integer(c_int32_t) function adc2g_coupling_form(h, form) bind(c, name='adc2g_coupling_form')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), intent(out) :: form
     end function adc2g_coupling_form

     integer(c_int32_t) function adc2g_diagonal(h, hii) bind(c, name='adc2g_diagonal')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: h
       real(c_double), intent(out) :: hii(*)
     end function adc2g_diagonal

     integer(c_int32_t) function adc2g_sigma(h, nvec, X, W, precision) bind(c, name='adc2g_sigma')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nvec, precision
       real(c_double), intent(in) :: X(*)
       real(c_double), intent(out) :: W(*)
     end function adc2g_sigma

     integer(c_int32_t) function adc2g_matxvec_treal(h, v1, v2, precision) bind(c, name='adc2g_matxvec_treal')
       import :: c_int32_t, c_double_complex, c_ptr
       type(c_ptr), value :: h
       complex(c_double_complex), intent(in) :: v1(*)
       complex(c_double_complex), intent(out) :: v2(*)
       integer(c_int32_t), value :: precision
     end function adc2g_matxvec_treal

     integer(c_int32_t) function adc2g_davidson(h, p, eigval, eigvec, stats) bind(c, name='adc2g_davidson')
       import :: c_int32_t, c_double, c_ptr, adc2g_dav_params, adc2g_dav_stats
       type(c_ptr), value :: h
       type(adc2g_dav_params), intent(in) :: p
       real(c_double), intent(out) :: eigval(*), eigvec(*)
       type(adc2g_dav_stats), intent(out) :: stats
     end function adc2g_davidson

     integer(c_int32_t) function adc2g_sil_autocorr(h, p, fvec, auto0, nrec_cap, nrec_out, stats) &
          bind(c, name='adc2g_sil_autocorr')
       import :: c_int32_t, c_int64_t, c_double, c_double_complex, c_ptr, adc2g_sil_params, adc2g_sil_stats
       type(c_ptr), value :: h
       type(adc2g_sil_params), intent(in) :: p
       real(c_double), intent(in) :: fvec(*)
       complex(c_double_complex), intent(out) :: auto0(*)
       integer(c_int64_t), value :: nrec_cap
       integer(c_int64_t), intent(out) :: nrec_out
       type(adc2g_sil_stats), intent(out) :: stats
     end function adc2g_sil_autocorr

     ! auto1, auto2: c_loc of complex(c_double_complex) arrays (nrec+1) or c_null_ptr
     integer(c_int32_t) function adc2g_sil_autocorr_orders(h, p, fvec, autoord, auto0, auto1, auto2, &
          nrec_cap, nrec_out, stats) bind(c, name='adc2g_sil_autocorr_orders')
       import :: c_int32_t, c_int64_t, c_double, c_double_complex, c_ptr, adc2g_sil_params, adc2g_sil_stats
       type(c_ptr), value :: h
       type(adc2g_sil_params), intent(in) :: p
       real(c_double), intent(in) :: fvec(*)
       integer(c_int32_t), value :: autoord
       complex(c_double_complex), intent(out) :: auto0(*)
       type(c_ptr), value :: auto1, auto2
       integer(c_int64_t), value :: nrec_cap
       integer(c_int64_t), intent(out) :: nrec_out
       type(adc2g_sil_stats), intent(out) :: stats
     end function adc2g_sil_autocorr_orders

     integer(c_int32_t) function adc2g_sil_fdstates(h, p, fvec, nfbas, ebound, iwfunc, nsel, fbas2eig, vecs, &
          energies, stats) bind(c, name='adc2g_sil_fdstates')
       import :: c_int32_t, c_double, c_ptr, adc2g_sil_params, adc2g_sil_stats
       type(c_ptr), value :: h
       type(adc2g_sil_params), intent(in) :: p
       real(c_double), intent(in) :: fvec(*), ebound(2), fbas2eig(*)
       integer(c_int32_t), value :: nfbas, iwfunc, nsel
       real(c_double), intent(out) :: vecs(*), energies(*)
       type(adc2g_sil_stats), intent(out) :: stats
     end function adc2g_sil_fdstates
     integer(c_int32_t) function adc2g_cheby_fdstates(h, bounds, q0, order, nsel, coef, precision, vecs, &
          energies) bind(c, name='adc2g_cheby_fdstates')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: h
       real(c_double), intent(in) :: bounds(2), q0(*), coef(*)
       integer(c_int32_t), value :: order, nsel, precision
       real(c_double), intent(out) :: vecs(*), energies(*)
     end function adc2g_cheby_fdstates
     integer(c_int32_t) function adc2g_cheby_auto(h, bounds, q0, order, precision, auto_out) &
          bind(c, name='adc2g_cheby_auto')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: h
       real(c_double), intent(in) :: bounds(2)
       real(c_double), intent(in) :: q0(*)
       integer(c_int32_t), value :: order, precision
       real(c_double), intent(out) :: auto_out(*)
     end function adc2g_cheby_auto

     ! tmat, travec, tranmom, ritz: c_loc of the array or c_null_ptr
     integer(c_int32_t) function adc2g_block_lanczos(h, p, q0, eigval, tmat, travec, tranmom, &
          ritz, stats) bind(c, name='adc2g_block_lanczos')
       import :: c_int32_t, c_double, c_ptr, adc2g_lanczos_params, adc2g_lanczos_stats
       type(c_ptr), value :: h
       type(adc2g_lanczos_params), intent(in) :: p
       real(c_double), intent(in) :: q0(*)
       real(c_double), intent(out) :: eigval(*)
       type(c_ptr), value :: tmat, travec, tranmom, ritz
       type(adc2g_lanczos_stats), intent(out) :: stats
     end function adc2g_block_lanczos

     ! guess, eigvec, residual, niter: c_loc of the array or c_null_ptr
     integer(c_int32_t) function adc2g_relax(h, p, guess, eigval, eigvec, residual, niter, stats) &
          bind(c, name='adc2g_relax')
       import :: c_int32_t, c_double, c_ptr, adc2g_relax_params, adc2g_relax_stats
       type(c_ptr), value :: h
       type(adc2g_relax_params), intent(in) :: p
       type(c_ptr), value :: guess
       real(c_double), intent(out) :: eigval(*)
       type(c_ptr), value :: eigvec, residual, niter
       type(adc2g_relax_stats), intent(out) :: stats
     end function adc2g_relax

     ! dia_path, off_path: NUL-terminated (trim(name)//c_null_char)
     integer(c_int32_t) function adc2g_write_hmlt(h, dia_path, off_path, minc, maxbl, noffd, nbuf) &
          bind(c, name='adc2g_write_hmlt')
       import :: c_int32_t, c_int64_t, c_double, c_char, c_ptr
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: dia_path(*), off_path(*)
       real(c_double), value :: minc
       integer(c_int32_t), value :: maxbl
       integer(c_int64_t), intent(out) :: noffd
       integer(c_int32_t), intent(out) :: nbuf
     end function adc2g_write_hmlt

     ! initial vectors given (initvec_adc1): guess(ndim, blocksize)
     integer(c_int32_t) function adc2g_davidson_vectors(h, p, guess, eigval, eigvec, stats) &
          bind(c, name='adc2g_davidson_vectors')
       import :: c_int32_t, c_double, c_ptr, adc2g_dav_params, adc2g_dav_stats
       type(c_ptr), value :: h
       type(adc2g_dav_params), intent(in) :: p
       real(c_double), intent(in) :: guess(*)
       real(c_double), intent(out) :: eigval(*), eigvec(*)
       type(adc2g_dav_stats), intent(out) :: stats
     end function adc2g_davidson_vectors

     ! CVS view on the core orbitals core(1:ncore) (1-based occupied indices)
     integer(c_int32_t) function adc2g_create_cvs(full, ncore, core, out) bind(c, name='adc2g_create_cvs')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: full
       integer(c_int32_t), value :: ncore
       integer(c_int32_t), intent(in) :: core(*)
       type(c_ptr), intent(out) :: out
     end function adc2g_create_cvs
     integer(c_int32_t) function adc2g_create_frozen_core(full, nfrozen, frozen, out) bind(c, name='adc2g_create_frozen_core')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: full
       integer(c_int32_t), value :: nfrozen
       integer(c_int32_t), intent(in) :: frozen(*)
       type(c_ptr), intent(out) :: out
     end function adc2g_create_frozen_core
     integer(c_int32_t) function adc2g_set_adc2x(h, enable) bind(c, name='adc2g_set_adc2x')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: enable
     end function adc2g_set_adc2x

     ! laser / CAP propagation: IS operators in the layout of SCRATCH/dipole_<c>
     ! and SCRATCH/cap (0-based row/col, each off-diagonal element once);
     ! wavefunctions of ndim+1 complex components, ground state last
     integer(c_int32_t) function adc2g_set_dipole_is(h, component, diag, nnz, row, col, val, d00, d0j) &
          bind(c, name='adc2g_set_dipole_is')
       import :: c_int32_t, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: component
       real(c_double), intent(in) :: diag(*)
       integer(c_int64_t), value :: nnz
       integer(c_int64_t), intent(in) :: row(*), col(*)
       real(c_double), intent(in) :: val(*)
       real(c_double), value :: d00
       real(c_double), intent(in) :: d0j(*)
     end function adc2g_set_dipole_is
     integer(c_int32_t) function adc2g_set_cap_is(h, diag, nnz, row, col, val, w00, w0j) &
          bind(c, name='adc2g_set_cap_is')
       import :: c_int32_t, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: h
       real(c_double), intent(in) :: diag(*)
       integer(c_int64_t), value :: nnz
       integer(c_int64_t), intent(in) :: row(*), col(*)
       real(c_double), intent(in) :: val(*)
       real(c_double), value :: w00
       real(c_double), intent(in) :: w0j(*)
     end function adc2g_set_cap_is
     integer(c_int32_t) function adc2g_set_cap_projector(h, nvec, vecs, ground) &
          bind(c, name='adc2g_set_cap_projector')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nvec
       real(c_double), intent(in) :: vecs(*)
       integer(c_int32_t), value :: ground
     end function adc2g_set_cap_projector
     integer(c_int32_t) function adc2g_matxvec_treal_laser(h, time, npulse, pulses, v1, v2) &
          bind(c, name='adc2g_matxvec_treal_laser')
       import :: c_int32_t, c_double, c_double_complex, c_ptr, adc2g_pulse
       type(c_ptr), value :: h
       real(c_double), value :: time
       integer(c_int32_t), value :: npulse
       type(adc2g_pulse), intent(in) :: pulses(*)
       complex(c_double_complex), intent(in) :: v1(*)
       complex(c_double_complex), intent(out) :: v2(*)
     end function adc2g_matxvec_treal_laser
     ! flux: c_loc of nrec_cap doubles, or c_null_ptr; trdens: c_loc of
     ! nrec_cap * nvirt * nocc complex values (the TD-NTO input), or c_null_ptr
     integer(c_int32_t) function adc2g_sil_laser(h, p, npulse, pulses, psi, norms, flux, trdens, nrec_cap, &
          nrec_out, stats) bind(c, name='adc2g_sil_laser')
       import :: c_int32_t, c_int64_t, c_double, c_double_complex, c_ptr, adc2g_pulse, adc2g_laser_params, &
            adc2g_sil_stats
       type(c_ptr), value :: h
       type(adc2g_laser_params), intent(in) :: p
       integer(c_int32_t), value :: npulse
       type(adc2g_pulse), intent(in) :: pulses(*)
       complex(c_double_complex), intent(inout) :: psi(*)
       real(c_double), intent(out) :: norms(*)
       type(c_ptr), value :: flux, trdens
       integer(c_int64_t), value :: nrec_cap
       integer(c_int64_t), intent(out) :: nrec_out
       type(adc2g_sil_stats), intent(out) :: stats
     end function adc2g_sil_laser
     integer(c_int32_t) function adc2g_set_flux_theta(h, diag, nnz, row, col, val, theta00, theta0j) &
          bind(c, name='adc2g_set_flux_theta')
       import :: c_int32_t, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: h
       real(c_double), intent(in) :: diag(*)
       integer(c_int64_t), value :: nnz
       integer(c_int64_t), intent(in) :: row(*), col(*)
       real(c_double), intent(in) :: val(*)
       real(c_double), value :: theta00
       real(c_double), intent(in) :: theta0j(*)
     end function adc2g_set_flux_theta
     ! trdens(nvirt, nocc, nvec): adc2_trdens_gs_virt_occ of x(ldx, nvec)
     integer(c_int32_t) function adc2g_trdens_vo(h, nvec, x, ldx, trdens) bind(c, name='adc2g_trdens_vo')
       import :: c_int32_t, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nvec
       real(c_double), intent(in) :: x(*)
       integer(c_int64_t), value :: ldx
       real(c_double), intent(out) :: trdens(*)
     end function adc2g_trdens_vo

     ! MP2 energy, D2 diagnostic and density (rho: c_loc of nbas x nbas or c_null_ptr)
     integer(c_int32_t) function adc2g_mp2(h, e_mp2, d2, rho) bind(c, name='adc2g_mp2')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: h
       real(c_double), intent(out) :: e_mp2, d2
       type(c_ptr), value :: rho
     end function adc2g_mp2

     ! F-vector of get_modifiedtm_adc2 for the nbas x nbas operator matrix dpl
     integer(c_int32_t) function adc2g_fvector(h, dpl, f) bind(c, name='adc2g_fvector')
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: dpl(*)
       real(c_double), intent(out) :: f(*)
     end function adc2g_fvector

     ! block Lanczos start block: kind 8 init_vec_doublesubspace, 9 init_vec_singlesubspace
     integer(c_int32_t) function adc2g_lanczos_guess(h, kind, subdim, blocksize, q0) &
          bind(c, name='adc2g_lanczos_guess')
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: kind, subdim, blocksize
       real(c_double), intent(out) :: q0(*)
     end function adc2g_lanczos_guess

     ! spectral_bounds: method 1 'davlanc' (ncycles 0 = 100), 2 'davpower', 3 'lanczos' (0 = 50);
     ! widen by 0.9 / 1.1 as chebyspec.f90:82-83 when both are positive
     integer(c_int32_t) function adc2g_spectral_bounds(h, method, ncycles, bounds) &
          bind(c, name='adc2g_spectral_bounds')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: method, ncycles
       real(c_double), intent(out) :: bounds(2)
     end function adc2g_spectral_bounds

     ! power method (power_method): vec = c_loc of ndim doubles or c_null_ptr
     integer(c_int32_t) function adc2g_power(h, subdim, maxit, thrsh, eval, vec, niter) &
          bind(c, name='adc2g_power')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: h
       integer(c_int32_t), value :: subdim, maxit
       real(c_double), value :: thrsh
       real(c_double), intent(out) :: eval
       type(c_ptr), value :: vec
       integer(c_int32_t), intent(out) :: niter
     end function adc2g_power

     integer(c_int32_t) function adc2g_precon_nparams() bind(c, name='adc2g_precon_nparams')
       import :: c_int32_t
     end function adc2g_precon_nparams

     integer(c_int32_t) function adc2g_precon_set(h, params) bind(c, name='adc2g_precon_set')
       import :: c_int32_t, c_float, c_ptr
       type(c_ptr), value :: h
       real(c_float), intent(in) :: params(*)
     end function adc2g_precon_set

     integer(c_int32_t) function adc2g_last_error(buf, length) bind(c, name='adc2g_last_error')
       import :: c_int32_t, c_char
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int32_t), value :: length
     end function adc2g_last_error
  end interface

contains

  ! Last error as a Fortran string (for errmsg / error_control).
  function adc2g_errmsg() result(msg)
    character(len=256) :: msg
    character(kind=c_char) :: buf(256)
    integer :: i, rc
    rc = adc2g_last_error(buf, 256)
    msg = ' '
    do i = 1, 256
       if (buf(i) == c_null_char) exit
       msg(i:i) = buf(i)
    end do
  end function adc2g_errmsg

  ! Eigenpair file in the record layout of wreigenpairs: one unformatted
  ! sequential record (i, E_i, vec_i(1:ndim)) per state.  The Lanczos-state
  ! file of ritzvecs_incore (source/eigen/block_lanczos.f90:2297-2299, lancname)
  ! has the same records: call it with the Ritz values and vectors.
  subroutine adc2g_write_davstates(fname, ndim, nstates, eigval, eigvec)
    character(len=*), intent(in) :: fname
    integer(c_int64_t), intent(in) :: ndim
    integer, intent(in) :: nstates
    real(c_double), intent(in) :: eigval(nstates), eigvec(ndim, nstates)
    integer :: unit, i
    open(newunit=unit, file=fname, status='unknown', access='sequential', form='unformatted')
    do i = 1, nstates
       write(unit) i, eigval(i), eigvec(:, i)
    end do
    close(unit)
  end subroutine adc2g_write_davstates

  ! Autocorrelation file in the format of open_autofiles/wrauto
  ! (source/propagation/fvecprop.f90:281-304, 585-598).
  subroutine adc2g_write_auto(fname, nrec, tout, auto0)
    character(len=*), intent(in) :: fname
    integer(c_int64_t), intent(in) :: nrec
    real(c_double), intent(in) :: tout
    complex(c_double_complex), intent(in) :: auto0(nrec)
    integer :: unit
    integer(c_int64_t) :: i
    open(newunit=unit, file=fname, form='formatted', status='unknown')
    write(unit, '(a)') '#    time[fs]         Re(autocorrel)         Im(autocorrel)         Abs(autocorrel)'
    do i = 1, nrec
       write(unit, '(F15.8,4x,3(2x,E21.14))') real(i - 1, c_double) * tout * adc2g_au2fs, &
            real(auto0(i)), aimag(auto0(i)), abs(auto0(i))
    end do
    close(unit)
  end subroutine adc2g_write_auto

  ! the 'chebyauto' file of write_chebyshev_auto (chebyspec.f90:323-355):
  ! C_0 .. C_2K with the spectral bounds in the header
  subroutine adc2g_write_chebyauto(fname, bounds, order, auto_out)
    character(len=*), intent(in) :: fname
    real(c_double), intent(in) :: bounds(2)
    integer(c_int32_t), intent(in) :: order
    real(c_double), intent(in) :: auto_out(0:2*order)
    integer :: unit, k
    open(newunit=unit, file=fname, form='formatted', status='unknown')
    write(unit, '(a,2(2x,E21.14),/)') '#    Spectral bounds:', bounds(1), bounds(2)
    write(unit, '(a)') '#    Order [k]    C_k'
    do k = 0, 2 * order
       write(unit, '(i6,11x,E21.14)') k, auto_out(k)
    end do
    close(unit)
  end subroutine adc2g_write_chebyauto

end module adc2g_mod
