!=======================================================================
! adc2g_driver: Fortran host program for the GPU ADC(2)-s path.
!
! Mirrors the call sequence of adc2_ener (source/adc/adc2_ener.f90:7-156:
! configuration space -> matrix -> master_eig/davdiag_block -> eigenpair
! file) and of adc2_autospec (source/adc/adc2_autospec.f90:21-93:
! initial state -> propagate_fvec -> 'auto' file), with the matrix build,
! Davidson and SIL replaced by the library through adc2g_mod.  The orbital
! energies and density-fitting factors come from raw float64 files (stream
! access) instead of the GAMESS reader.
!
! Input: namelist /adc2g/ on standard input, e.g.
!   &adc2g nocc=5 nvirt=36 naux=164 eps_file='eps.bin' b_file='B.bin'
!          nstates=5 blocksize=5 maxvec=40 tol=1d-7 precon=1 davname='davstates'
!          sil=.true. fvec_file='f.bin' tout=0.1d0 tfinal=2d0 krydim=20 autotol=1d-10 /
!=======================================================================
program adc2g_driver

  use iso_c_binding
  use adc2g_mod
  implicit none

  integer :: nocc = 0, nvirt = 0, naux = 0, nstates = 5, blocksize = 0, maxvec = -1, maxiter = 300
  integer :: precon = 2, precision = 0, krydim = 10
  logical :: deflate = .true., sil = .false.
  real(c_double) :: tol = 1.0d-7, tout = 0.1d0, tfinal = 1.0d0, autotol = 1.0d-5
  character(len=256) :: eps_file = 'eps.bin', b_file = 'B.bin', davname = 'davstates'
  character(len=256) :: fvec_file = 'fvec.bin', auto_file = 'auto'
  namelist /adc2g/ nocc, nvirt, naux, eps_file, b_file, nstates, blocksize, maxvec, maxiter, tol, &
       precon, precision, deflate, davname, sil, fvec_file, auto_file, tout, tfinal, krydim, autotol

  type(c_ptr) :: h
  type(adc2g_opts) :: opts
  type(adc2g_dav_params) :: dp
  type(adc2g_dav_stats) :: ds
  type(adc2g_sil_params) :: sp
  type(adc2g_sil_stats) :: ss
  real(c_double), allocatable :: eps(:), B(:), eigval(:), eigvec(:,:), fvec(:)
  complex(c_double_complex), allocatable :: auto0(:)
  integer(c_int64_t) :: ndim, n1, counts(5), nrec, nout
  integer :: nbas, rc, i, u

  read(*, nml=adc2g)
  nbas = nocc + nvirt
  allocate(eps(nbas), B(int(naux, c_int64_t) * nbas * nbas))
  open(newunit=u, file=trim(eps_file), access='stream', form='unformatted', status='old')
  read(u) eps
  close(u)
  open(newunit=u, file=trim(b_file), access='stream', form='unformatted', status='old')
  read(u) B
  close(u)

  rc = adc2g_create(int(nocc, c_int32_t), int(nvirt, c_int32_t), int(naux, c_int32_t), eps, B, opts, h)
  call check(rc)
  deallocate(B)
  rc = adc2g_dims(h, ndim, n1, counts)
  call check(rc)
  write(*, '(a,i12,a,i8,a,5i10)') ' ADC(2)-s space: ndim =', ndim, '  1h1p =', n1, '  2h2p types =', counts

  dp%nstates = nstates
  dp%blocksize = blocksize
  dp%maxvec = maxvec
  dp%maxiter = maxiter
  dp%tol = tol
  dp%precon = precon
  dp%deflate = merge(1, 0, deflate)
  dp%precision = precision
  allocate(eigval(nstates), eigvec(ndim, nstates))
  rc = adc2g_davidson(h, dp, eigval, eigvec, ds)
  call check(rc)
  write(*, '(a,i6,a,i8,a,f10.3,a)') ' Davidson: iterations', ds%niter, ', sigma-vectors', ds%nsigma, &
       ', time', ds%time_total, ' s'
  do i = 1, nstates
     write(*, '(i4,2x,f18.12,2x,f14.7)') i, eigval(i), eigval(i) * adc2g_eh2ev
  end do
  call adc2g_write_davstates(trim(davname), ndim, nstates, eigval, eigvec)

  if (sil) then
     allocate(fvec(ndim))
     open(newunit=u, file=trim(fvec_file), access='stream', form='unformatted', status='old')
     read(u) fvec
     close(u)
     sp%tout = tout
     sp%tfinal = tfinal
     sp%kdim = krydim
     sp%autotol = autotol
     sp%precision = precision
     nrec = int(tfinal / (0.5d0 * tout), c_int64_t) + 1
     allocate(auto0(nrec))
     rc = adc2g_sil_autocorr(h, sp, fvec, auto0, nrec, nout, ss)
     call check(rc)
     call adc2g_write_auto(trim(auto_file), nout, tout, auto0)
     write(*, '(a,i10,a,i10)') ' SIL: steps', ss%nsteps, ', complex sigma-vectors', ss%nsigma
  end if

  rc = adc2g_destroy(h)

contains

  subroutine check(code)
    integer(c_int32_t), intent(in) :: code
    if (code /= ADC2G_OK) then
       write(*, '(a,i4,2a)') ' adc2g error ', code, ': ', trim(adc2g_errmsg())
       stop 1
    end if
  end subroutine check

end program adc2g_driver
