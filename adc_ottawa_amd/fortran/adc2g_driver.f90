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
! Further keys:
!   solver = 1 Davidson (davdiag_block) | 2 relaxation (relaxation.f90:36-133)
!            with stepsize, rlx_krydim and the same nstates / maxiter / tol;
!            integrator = 1 XSIL (default) | 2 Bulirsch-Stoer | 3 RKF45, with siltol;
!   guess = 0 unit vectors | 1 sub-diagonalisation (subdim ISs) for Davidson;
!   autoord = 0..2: SIL autocorrelation orders, files auto_file, 'auto1', 'auto2';
!   lanczos = .true.: block Lanczos (lancdiag_block) with lmain start vectors and
!            ncycles blocks, Lanczos states written to lancname, transition
!            moments fvec.r_n printed; lancguess 1 (default) unit vectors on
!            the 1h1p entries of largest |fvec|, 8 the double-subspace and 9
!            the single-subspace guess (800 ISs); orthotype 0, 1 (pro), 2 (mpro);
!   fd_nsel > 0 with fd_coef_file (stream, (chebyord + 1) x fd_nsel doubles,
!            the k2eig columns of the selected states): the Chebyshev
!            filter-diagonalisation states (calc_fdstates_cfd) from fvec,
!            written to 'fdstates' as (i, E_i, v_i) records;
!   fdtd_nsel > 0 (with sil): the time-domain filter-diagonalisation states
!            (calc_fdstates_tdfd) of the SIL periods of tout from fvec,
!            fdtd_nfbas filter energies on fdtd_ebound, window exponent
!            fdtd_iwfunc, fbas2eig rows in fdtd_coef_file (stream,
!            fdtd_nsel x fdtd_nfbas doubles), written to 'fdstates_td';
!   chebyord > 0: Chebyshev autocorrelation of fvec with the bounds
!            cheby_emin / cheby_emax, file 'chebyauto';
!   hmlt = .true.: the matrix files 'hmlt.diai' / 'hmlt.offi';
!   ehf: the HF energy added to the state energies of 'davstates.dat' (the
!            table2 file, written for the full space);
!   ncore = n, core = i1, ..., in: CVS on these 1-based core orbitals (lcvs);
!            every later step (solvers, F-vector, propagation) uses the CVS
!            configurations;
!   nfrozen = n: frozen core (lifrzcore) on the n lowest occupied orbitals:
!            every later step uses the configurations with no hole in them;
!   adc2x = .true.: ADC(2)-x (method = adc2-x): the first-order 2h2p-2h2p
!            block in the matrix (set before the CVS / frozen-core view);
!   dia_file = 'hmlt.diai', off_file = 'hmlt.offi': the matrix from the
!            reference's files instead of eps/B (rdham_on / rdham_off and
!            hxvec_incore); Davidson, SIL, Lanczos, Chebyshev and hmlt only;
!   nto_file = 'trd.bin': the occupied-virtual transition density of every
!            state (nvirt x nocc doubles per state, adc2_trdens_gs_virt_occ),
!            the input of the NTOs of adc2_nto_gs (full space only);
!   dpl_file = 'd.bin': the F-vector of this nbas x nbas column-major operator
!            matrix (get_modifiedtm_adc2), computed by the library, gives the
!            transition moments and oscillator strengths of the states
!            (adc2_ener.f90:138-146) and is the initial state instead of the
!            vector in fvec_file (contract_dipole_initial_state with
!            statenumber 0, adc2_autospec.f90:131-199).
!   laser = .true.: laser-driven propagation from the ground state
!            (propagate_laser_adc2, proplib_adc2.f90:37-830) over tout /
!            tfinal with krydim and proptol: the IS dipole operators from
!            the reference's files dipis_file(1:3) (SCRATCH/dipole_x/y/z; ''
!            leaves a component unset) with d0j_file(1:3) (ndim float64) and
!            d00(1:3); an optional CAP from cap_file with w0j_file and w00
!            (cap_ground = .true. projects the ground state out of it); npulse
!            pulses from carrier(:), envelope(:), pulse_vec(3,:), strength(:),
!            freq(:), t0(:), phase(:), envpar(2,:).  statenumber = 0 starts
!            from the ground state, statenumber = n from the n-th Davidson
!            state (initwf_exci_initstate); cap_nproj = n projects the n
!            lowest Davidson states out of the CAP (iprojcap); theta_file
!            (SCRATCH/theta) with theta0j_file and theta00 gives the CAP flux
!            in 'flux.dat' (lflux); trdens_file receives the TD-NTO input
!            (lnto: the transition density at every tout, nvirt x nocc
!            complex per record).  Prints Time / Norm at every tout
!            (wrstepinfo) and writes the final wavefunction (ndim + 1 complex,
!            ground state last) to psi_file.
!=======================================================================
program adc2g_driver

  use iso_c_binding
  use adc2g_mod
  implicit none

  integer :: nocc = 0, nvirt = 0, naux = 0, nstates = 5, blocksize = 0, maxvec = -1, maxiter = 300
  integer :: precon = 2, precision = 0, krydim = 10
  integer :: solver = 1, guess = 0, subdim = 800, rlx_krydim = 10, autoord = 0
  integer :: lmain = 0, ncycles = 0, chebyord = 0, lancguess = 1, orthotype = 0, integrator = 1
  integer :: ncore = 0, core(16) = 0, nfrozen = 0, ifr
  logical :: deflate = .true., sil = .false., lanczos = .false., hmlt = .false., adc2x = .false.
  real(c_double) :: tol = 1.0d-7, tout = 0.1d0, tfinal = 1.0d0, autotol = 1.0d-5
  real(c_double) :: stepsize = 10.0d0, cheby_emin = 0.0d0, cheby_emax = 1.0d0, siltol = 1.0d-5, ehf = 0.0d0
  character(len=256) :: eps_file = 'eps.bin', b_file = 'B.bin', davname = 'davstates'
  character(len=256) :: fvec_file = 'fvec.bin', auto_file = 'auto', lancname = 'lancstates'
  character(len=256) :: nto_file = '', dia_file = '', off_file = '', fd_coef_file = ''
  integer :: fd_nsel = 0, fdtd_nsel = 0, fdtd_nfbas = 0, fdtd_iwfunc = 1
  character(len=256) :: fdtd_coef_file = ''
  real(c_double) :: fdtd_ebound(2) = 0.0d0
  character(len=256) :: dpl_file = ''
  ! laser / CAP propagation
  integer, parameter :: maxpulse = 8
  logical :: laser = .false., cap_ground = .false.
  integer :: npulse = 0, carrier(maxpulse) = 1, envelope(maxpulse) = 1, statenumber = 0, cap_nproj = 0
  real(c_double) :: pulse_vec(3, maxpulse) = 0.0d0, strength(maxpulse) = 0.0d0, freq(maxpulse) = 0.0d0
  real(c_double) :: t0(maxpulse) = 0.0d0, phase(maxpulse) = 0.0d0, envpar(2, maxpulse) = 0.0d0
  real(c_double) :: d00(3) = 0.0d0, w00 = 0.0d0, proptol = 1.0d-8
  character(len=256) :: dipis_file(3) = '', d0j_file(3) = '', cap_file = '', w0j_file = '', psi_file = 'psi.bin'
  character(len=256) :: theta_file = '', theta0j_file = '', trdens_file = ''
  real(c_double) :: theta00 = 0.0d0
  namelist /adc2g/ nocc, nvirt, naux, eps_file, b_file, nstates, blocksize, maxvec, maxiter, tol, &
       precon, precision, deflate, davname, sil, fvec_file, auto_file, tout, tfinal, krydim, autotol, &
       solver, guess, subdim, stepsize, rlx_krydim, autoord, lanczos, lmain, ncycles, lancname, &
       chebyord, cheby_emin, cheby_emax, hmlt, dpl_file, nto_file, dia_file, off_file, fd_coef_file, fd_nsel, &
       fdtd_nsel, fdtd_nfbas, fdtd_iwfunc, fdtd_coef_file, fdtd_ebound, lancguess, orthotype, integrator, siltol, &
       ncore, core, ehf, nfrozen, adc2x, laser, cap_ground, npulse, carrier, envelope, pulse_vec, &
       strength, freq, t0, phase, envpar, d00, w00, proptol, dipis_file, d0j_file, cap_file, w0j_file, psi_file, &
       statenumber, cap_nproj, theta_file, theta0j_file, theta00, trdens_file

  type(c_ptr) :: h, hfull
  type(adc2g_opts) :: opts
  type(adc2g_dav_params) :: dp
  type(adc2g_dav_stats) :: ds
  type(adc2g_sil_params) :: sp
  type(adc2g_sil_stats) :: ss
  type(adc2g_relax_params) :: rp
  type(adc2g_relax_stats) :: rs
  type(adc2g_lanczos_params) :: lp
  type(adc2g_lanczos_stats) :: ls
  real(c_double), allocatable :: eps(:), B(:), eigval(:), fvec(:), q0(:,:), cauto(:), dpl(:,:), tmv(:), osc(:), trdvo(:,:,:)
  real(c_double), allocatable :: fdcoef(:,:), fdvec(:,:), fdener(:)
  real(c_double) :: emp2 = 0.0d0, d2mp = 0.0d0
  real(c_double), allocatable, target :: eigvec(:,:), lener(:), ritz(:,:), tmom(:), fvec_t(:)
  complex(c_double_complex), allocatable :: auto0(:)
  complex(c_double_complex), allocatable, target :: auto1(:), auto2(:)
  integer(c_int64_t) :: ndim, n1, counts(5), nrec, nout, noffd
  integer(c_int32_t) :: nbuf
  integer :: nbas, rc, i, u, nvec, k
  logical, allocatable :: used(:)
  real(c_double) :: bounds(2)

  read(*, nml=adc2g)
  if (ncore < 0 .or. ncore > size(core)) then
     write(*, '(a,i0,a,i0)') ' adc2g input error: ncore must be between 0 and ', size(core), ', got ', ncore
     stop 1
  end if
  nbas = nocc + nvirt
  if (len_trim(dia_file) > 0) then
     ! the matrix from the reference's hmlt files (rdham_on / rdham_off and
     ! hxvec_incore): no orbitals, so no view, ADC(2)-x, MP2 or F-vector
     if (adc2x .or. nfrozen > 0 .or. ncore > 0 .or. len_trim(dpl_file) > 0 .or. laser .or. &
         len_trim(nto_file) > 0) then
        write(*, '(a)') ' adc2g input error: dia_file allows Davidson, SIL, Lanczos, Chebyshev and hmlt only'
        stop 1
     end if
     rc = adc2g_create_from_hmlt(trim(dia_file)//c_null_char, trim(off_file)//c_null_char, 0_c_int32_t, h)
     call check(rc)
     write(*, '(a,a,a,a)') ' stored matrix from ', trim(dia_file), ' and ', trim(off_file)
  else
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
  end if
  hfull = c_null_ptr
  if (adc2x) then
     call check(adc2g_set_adc2x(h, 1_c_int32_t))
     write(*, '(a)') ' ADC(2)-x: first-order 2h2p-2h2p block included'
  end if
  if (nfrozen > 0 .and. ncore == 0) then
     ! frozen core (lifrzcore): the configurations with no hole in the frozen orbitals
     hfull = h
     call check(adc2g_create_frozen_core(hfull, int(nfrozen, c_int32_t), &
          int((/ (ifr, ifr = 1, nfrozen) /), c_int32_t), h))
     write(*, '(a,i4)') ' frozen core orbitals: 1 ..', nfrozen
  end if
  if (ncore > 0) then
     ! CVS (lcvs): every later step works on the core-hole configurations
     hfull = h
     call check(adc2g_create_cvs(hfull, int(ncore, c_int32_t), int(core(1:ncore), c_int32_t), h))
     write(*, '(a,16i4)') ' CVS core orbitals:', core(1:ncore)
  end if
  rc = adc2g_dims(h, ndim, n1, counts)
  call check(rc)
  write(*, '(a,i12,a,i8,a,5i10)') ' ADC(2)-s space: ndim =', ndim, '  1h1p =', n1, '  2h2p types =', counts

  allocate(eigval(nstates), eigvec(ndim, nstates))
  if (solver == 2) then
     ! master_eig with solver = 2: relaxation (adc_common.f90:952-993)
     rp%nstates = nstates
     rp%krydim = rlx_krydim
     rp%maxdim = -1
     rp%maxiter = maxiter
     rp%step = stepsize
     rp%tol = tol
     rp%precision = precision
     rp%integrator = integrator
     rp%toler = siltol
     rc = adc2g_relax(h, rp, c_null_ptr, eigval, c_loc(eigvec), c_null_ptr, c_null_ptr, rs)
     call check(rc)
     write(*, '(a,i8,a,f10.3,a)') ' Relaxation: sigma-vectors', rs%nsigma, ', time', rs%time_total, ' s'
  else
     dp%nstates = nstates
     dp%blocksize = blocksize
     dp%maxvec = maxvec
     dp%maxiter = maxiter
     dp%tol = tol
     dp%precon = precon
     dp%deflate = merge(1, 0, deflate)
     dp%precision = precision
     dp%guess = guess
     dp%subdim = subdim
     rc = adc2g_davidson(h, dp, eigval, eigvec, ds)
     call check(rc)
     write(*, '(a,i6,a,i8,a,f10.3,a)') ' Davidson: iterations', ds%niter, ', sigma-vectors', ds%nsigma, &
          ', time', ds%time_total, ' s'
  end if
  if (len_trim(dpl_file) > 0) then
     ! F-vector of the operator matrix (get_modifiedtm_adc2); with it the
     ! transition moments tm = x_n . f and oscillator strengths 2/3 E tm^2
     ! of adc2_ener (adc2_ener.f90:138-146, tm of get_moment.f90:16-28)
     allocate(fvec(ndim), dpl(nbas, nbas))
     open(newunit=u, file=trim(dpl_file), access='stream', form='unformatted', status='old')
     read(u) dpl
     close(u)
     call check(adc2g_fvector(h, dpl, fvec))
     write(*, '(a,es22.14)') ' F-vector: |f| =', sqrt(sum(fvec**2))
     allocate(tmv(nstates), osc(nstates))
     do i = 1, nstates
        tmv(i) = dot_product(eigvec(:, i), fvec)
        osc(i) = 2.0d0 / 3.0d0 * eigval(i) * tmv(i)**2
     end do
     write(*, '(a)') ' state        E [Eh]              E [eV]        transition moment    osc. strength'
     do i = 1, nstates
        write(*, '(i4,2x,f18.12,2x,f14.7,2x,es20.12,2x,es20.12)') i, eigval(i), eigval(i) * adc2g_eh2ev, &
             dot_product(eigvec(:, i), fvec), 2.0d0 / 3.0d0 * eigval(i) * dot_product(eigvec(:, i), fvec)**2
     end do
  else
     do i = 1, nstates
        write(*, '(i4,2x,f18.12,2x,f14.7)') i, eigval(i), eigval(i) * adc2g_eh2ev
     end do
  end if
  call adc2g_write_davstates(trim(davname), ndim, nstates, eigval, eigvec)
  if (len_trim(nto_file) > 0) then
     ! lnto (adc2_nto_gs, nto.f90:76-260): the occupied-virtual transition
     ! density of every state (nvirt x nocc per state), for its SVD
     allocate(trdvo(nvirt, nocc, nstates))
     call check(adc2g_trdens_vo(h, int(nstates, c_int32_t), eigvec, int(ndim, c_int64_t), trdvo))
     open(newunit=u, file=trim(nto_file), access='stream', form='unformatted', status='replace')
     write(u) trdvo
     close(u)
     write(*, '(a,a)') ' NTO transition densities written to ', trim(nto_file)
  end if
  if (ncore == 0 .and. nfrozen == 0 .and. len_trim(dia_file) == 0) then
     ! table2 / wrstateinfo (misc.f90:648-852): 'davstates.dat' (full space)
     call check(adc2g_mp2(h, emp2, d2mp, c_null_ptr))
     call write_state_table('davstates.dat')
  end if

  if (hmlt) then
     rc = adc2g_write_hmlt(h, 'hmlt.diai'//c_null_char, 'hmlt.offi'//c_null_char, 1.0d-12, 0_c_int32_t, &
          noffd, nbuf)
     call check(rc)
     write(*, '(a,i12,a,i6)') ' hmlt: off-diagonal elements', noffd, ', records', nbuf
  end if

  if ((sil .or. lanczos .or. chebyord > 0) .and. len_trim(dpl_file) == 0) then
     allocate(fvec(ndim))
     open(newunit=u, file=trim(fvec_file), access='stream', form='unformatted', status='old')
     read(u) fvec
     close(u)
  end if

  if (lanczos) then
     ! init_vec_unit: unit vectors on the lmain 1h1p entries of largest |fvec|
     nvec = lmain * ncycles
     allocate(q0(ndim, lmain), lener(nvec), ritz(ndim, nvec), tmom(nvec), fvec_t(ndim), used(n1))
     if (lancguess == 8 .or. lancguess == 9) then
        ! init_vec_doublesubspace / init_vec_singlesubspace (block_lanczos.f90:443-923)
        call check(adc2g_lanczos_guess(h, int(lancguess, c_int32_t), 800_c_int32_t, int(lmain, c_int32_t), q0))
     else
        q0 = 0.0d0
        used = .false.
        do i = 1, lmain
           k = maxloc(abs(fvec(1:n1)), dim=1, mask=.not. used)
           used(k) = .true.
           q0(k, i) = 1.0d0
        end do
     end if
     fvec_t = fvec
     lp%blocksize = lmain
     lp%ncycles = ncycles
     lp%precision = precision
     lp%orthotype = orthotype
     rc = adc2g_block_lanczos(h, lp, q0, lener, c_null_ptr, c_loc(fvec_t), c_loc(tmom), c_loc(ritz), ls)
     call check(rc)
     call adc2g_write_davstates(trim(lancname), ndim, nvec, lener, ritz)
     write(*, '(a,i6,a,f10.3,a)') ' Block Lanczos: states', nvec, ', time', ls%time_total, ' s'
     do i = 1, min(nvec, 5)
        write(*, '(i4,2x,f18.12,2x,es16.8)') i, lener(i), tmom(i)**2
     end do
  end if

  if (chebyord > 0) then
     bounds = (/ cheby_emin, cheby_emax /)
     allocate(cauto(0:2 * chebyord))
     rc = adc2g_cheby_auto(h, bounds, fvec, int(chebyord, c_int32_t), int(precision, c_int32_t), cauto)
     call check(rc)
     call adc2g_write_chebyauto('chebyauto', bounds, int(chebyord, c_int32_t), cauto)
     write(*, '(a,i6)') ' Chebyshev: orders', chebyord
     if (fd_nsel > 0) then
        ! calc_fdstates_cfd: the filter-diagonalisation states of the
        ! coefficients k2eig(0:chebyord, isel(:)) in fd_coef_file, written
        ! as SCRATCH/fdstates records (i, E_i, v_i)
        allocate(fdcoef(0:chebyord, fd_nsel), fdvec(ndim, fd_nsel), fdener(fd_nsel))
        open(newunit=u, file=trim(fd_coef_file), access='stream', form='unformatted', status='old')
        read(u) fdcoef
        close(u)
        call check(adc2g_cheby_fdstates(h, bounds, fvec, int(chebyord, c_int32_t), int(fd_nsel, c_int32_t), &
             fdcoef, int(precision, c_int32_t), fdvec, fdener))
        call adc2g_write_davstates('fdstates', ndim, int(fd_nsel, c_int32_t), fdener, fdvec)
        do i = 1, fd_nsel
           write(*, '(a,i4,f18.12,f14.7)') ' FD state', i, fdener(i), fdener(i) * adc2g_eh2ev
        end do
        deallocate(fdcoef, fdvec, fdener)
     end if
  end if

  if (sil) then
     sp%tout = tout
     sp%tfinal = tfinal
     sp%kdim = krydim
     sp%autotol = autotol
     sp%precision = precision
     nrec = int(tfinal / (0.5d0 * tout), c_int64_t) + 1
     allocate(auto0(nrec), auto1(nrec), auto2(nrec))
     if (autoord == 0) then
        rc = adc2g_sil_autocorr(h, sp, fvec, auto0, nrec, nout, ss)
     else
        rc = adc2g_sil_autocorr_orders(h, sp, fvec, int(autoord, c_int32_t), auto0, c_loc(auto1), &
             c_loc(auto2), nrec, nout, ss)
     end if
     call check(rc)
     call adc2g_write_auto(trim(auto_file), nout, tout, auto0)
     if (fdtd_nsel > 0) then
        ! calc_fdstates_tdfd: the time-domain filter-diagonalisation states
        ! of the fbas2eig rows in fdtd_coef_file (fdtd_nsel x fdtd_nfbas),
        ! written to 'fdstates_td' as (i, E_i, v_i) records
        allocate(fdcoef(fdtd_nsel, fdtd_nfbas), fdvec(ndim, fdtd_nsel), fdener(fdtd_nsel))
        open(newunit=u, file=trim(fdtd_coef_file), access='stream', form='unformatted', status='old')
        read(u) fdcoef
        close(u)
        call check(adc2g_sil_fdstates(h, sp, fvec, int(fdtd_nfbas, c_int32_t), fdtd_ebound, &
             int(fdtd_iwfunc, c_int32_t), int(fdtd_nsel, c_int32_t), fdcoef, fdvec, fdener, ss))
        call adc2g_write_davstates('fdstates_td', ndim, int(fdtd_nsel, c_int32_t), fdener, fdvec)
        do i = 1, fdtd_nsel
           write(*, '(a,i4,f18.12,f14.7)') ' TD-FD state', i, fdener(i), fdener(i) * adc2g_eh2ev
        end do
        deallocate(fdcoef, fdvec, fdener)
     end if
     if (autoord >= 1) call adc2g_write_auto('auto1', nout, tout, auto1)
     if (autoord >= 2) call adc2g_write_auto('auto2', nout, tout, auto2)
     write(*, '(a,i10,a,i10)') ' SIL: steps', ss%nsteps, ', complex sigma-vectors', ss%nsigma
  end if

  if (laser) call run_laser()

  rc = adc2g_destroy(h)
  if (c_associated(hfull)) rc = adc2g_destroy(hfull)

contains

  ! propagate_laser_adc2 from the ground state (statenumber = 0)
  subroutine run_laser()
    type(adc2g_laser_params) :: lpar
    type(adc2g_pulse) :: pulses(maxpulse)
    real(c_double), allocatable :: odiag(:), oval(:), vec0(:), norms(:)
    real(c_double), allocatable, target :: fluxv(:)
    complex(c_double_complex), allocatable, target :: trd(:,:,:)
    type(c_ptr) :: fptr, tptr
    integer(c_int64_t), allocatable :: orow(:), ocol(:)
    complex(c_double_complex), allocatable :: psi(:)
    real(c_double) :: dummy(1)
    integer(c_int64_t) :: onnz
    integer :: c, ip
    if (npulse < 0 .or. npulse > maxpulse) then
       write(*, '(a,i0)') ' adc2g input error: npulse must be between 0 and ', maxpulse
       stop 1
    end if
    allocate(vec0(ndim))
    do c = 1, 3
       if (len_trim(dipis_file(c)) == 0) cycle
       call read_is_operator(trim(dipis_file(c)), ndim, odiag, orow, ocol, oval, onnz)
       open(newunit=u, file=trim(d0j_file(c)), access='stream', form='unformatted', status='old')
       read(u) vec0
       close(u)
       call check(adc2g_set_dipole_is(h, int(c - 1, c_int32_t), odiag, onnz, orow, ocol, oval, d00(c), vec0))
       write(*, '(a,i2,a,i12)') ' laser: dipole component', c, ', off-diagonal elements', onnz
    end do
    if (len_trim(cap_file) > 0) then
       call read_is_operator(trim(cap_file), ndim, odiag, orow, ocol, oval, onnz)
       open(newunit=u, file=trim(w0j_file), access='stream', form='unformatted', status='old')
       read(u) vec0
       close(u)
       call check(adc2g_set_cap_is(h, odiag, onnz, orow, ocol, oval, w00, vec0))
       if (cap_nproj > 0) then
          ! iprojcap: the cap_nproj lowest Davidson states (and, with
          ! cap_ground, the ground state) projected out of the CAP
          if (cap_nproj > nstates) then
             write(*, '(a)') ' adc2g input error: cap_nproj exceeds nstates'
             stop 1
          end if
          call check(adc2g_set_cap_projector(h, int(cap_nproj, c_int32_t), eigvec(:, 1:cap_nproj), &
               merge(1_c_int32_t, 0_c_int32_t, cap_ground)))
       else if (cap_ground) then
          call check(adc2g_set_cap_projector(h, 0_c_int32_t, dummy, 1_c_int32_t))
       end if
       write(*, '(a,i12)') ' laser: CAP off-diagonal elements', onnz
    end if
    do ip = 1, npulse
       pulses(ip)%carrier = int(carrier(ip), c_int32_t)
       pulses(ip)%envelope = int(envelope(ip), c_int32_t)
       pulses(ip)%pulse_vec = pulse_vec(:, ip)
       pulses(ip)%strength = strength(ip)
       pulses(ip)%freq = freq(ip)
       pulses(ip)%t0 = t0(ip)
       pulses(ip)%phase = phase(ip)
       pulses(ip)%envpar = envpar(:, ip)
    end do
    lpar%tout = tout
    lpar%tfinal = tfinal
    lpar%kdim = int(krydim, c_int32_t)
    lpar%precision = int(precision, c_int32_t)
    lpar%proptol = proptol
    allocate(psi(ndim + 1), norms(int(tfinal / tout) + 1))
    psi = (0.0d0, 0.0d0)
    if (statenumber == 0) then
       psi(ndim + 1) = (1.0d0, 0.0d0)       ! the MP2 ground state
    else if (statenumber >= 1 .and. statenumber <= nstates) then
       psi(1:ndim) = cmplx(eigvec(:, statenumber), 0.0d0, c_double_complex)  ! initwf_exci_initstate
    else
       write(*, '(a)') ' adc2g input error: statenumber must be between 0 and nstates'
       stop 1
    end if
    allocate(fluxv(size(norms)))
    fptr = c_null_ptr
    tptr = c_null_ptr
    if (len_trim(theta_file) > 0) then
       ! lflux: the projector onto the absorbing region (SCRATCH/theta)
       call read_is_operator(trim(theta_file), ndim, odiag, orow, ocol, oval, onnz)
       open(newunit=u, file=trim(theta0j_file), access='stream', form='unformatted', status='old')
       read(u) vec0
       close(u)
       call check(adc2g_set_flux_theta(h, odiag, onnz, orow, ocol, oval, theta00, vec0))
       fptr = c_loc(fluxv)
    end if
    if (len_trim(trdens_file) > 0) then
       allocate(trd(nvirt, nocc, size(norms)))
       tptr = c_loc(trd)
    end if
    call check(adc2g_sil_laser(h, lpar, int(npulse, c_int32_t), pulses, psi, norms, fptr, tptr, &
         int(size(norms), c_int64_t), nout, ss))
    if (len_trim(trdens_file) > 0) then
       ! lnto: the transition densities of tdadc2_nto at every tout (records
       ! 0 .. nout - 1, each nvirt x nocc complex), for the NTO analysis
       open(newunit=u, file=trim(trdens_file), access='stream', form='unformatted', status='replace')
       write(u) trd(:, :, 1:nout)
       close(u)
    end if
    do i = 1, int(nout)
       write(*, '(a,f14.6,a,f12.9)') ' Time:', (i - 1) * tout, '   Norm:', norms(i)
    end do
    if (len_trim(theta_file) > 0) then
       ! flux.dat as wrstepinfo writes it (proplib_adc2.f90:133-138, 858)
       open(newunit=u, file='flux.dat', form='formatted', status='replace')
       write(u, '(72a)') ('#', i = 1, 72)
       write(u, '(a)') '# Time (au)    Flux'
       write(u, '(72a)') ('#', i = 1, 72)
       do i = 1, int(nout)
          write(u, '(F10.4,5x,ES15.8)') (i - 1) * tout, fluxv(i)
       end do
       close(u)
    end if
    write(*, '(a,i10,a,i10)') ' laser: steps', ss%nsteps, ', operator products', ss%nsigma
    open(newunit=u, file=trim(psi_file), access='stream', form='unformatted', status='replace')
    write(u) psi
    close(u)
  end subroutine run_laser

  ! An IS operator file as the reference writes SCRATCH/dipole_<c> and
  ! SCRATCH/cap and opxvec_ext reads them (tdselib.f90:1031-1153): buf_size,
  ! the diagonal, then records (oij, indxi, indxj, nlim), 1-based indices.
  subroutine read_is_operator(path, nd, diag, row, col, val, nnz)
    character(len=*), intent(in) :: path
    integer(c_int64_t), intent(in) :: nd
    real(c_double), allocatable, intent(out) :: diag(:), val(:)
    integer(c_int64_t), allocatable, intent(out) :: row(:), col(:)
    integer(c_int64_t), intent(out) :: nnz
    real(c_double), allocatable :: oij(:), tv(:)
    integer, allocatable :: ii(:), jj(:)
    integer(c_int64_t), allocatable :: tr(:), tc(:)
    integer :: uu, bs, nlim, ios
    open(newunit=uu, file=path, form='unformatted', access='sequential', status='old')
    read(uu) bs
    allocate(diag(nd), oij(bs), ii(bs), jj(bs), row(bs), col(bs), val(bs))
    read(uu) diag
    nnz = 0
    do
       read(uu, iostat=ios) oij, ii, jj, nlim
       if (ios /= 0) exit
       if (nnz + nlim > size(row)) then
          allocate(tr(2 * (nnz + nlim)), tc(2 * (nnz + nlim)), tv(2 * (nnz + nlim)))
          tr(1:nnz) = row(1:nnz)
          tc(1:nnz) = col(1:nnz)
          tv(1:nnz) = val(1:nnz)
          call move_alloc(tr, row)
          call move_alloc(tc, col)
          call move_alloc(tv, val)
       end if
       row(nnz + 1:nnz + nlim) = ii(1:nlim) - 1
       col(nnz + 1:nnz + nlim) = jj(1:nlim) - 1
       val(nnz + 1:nnz + nlim) = oij(1:nlim)
       nnz = nnz + nlim
    end do
    close(uu)
  end subroutine read_is_operator

  ! (lo, hi), 1-based, of pair p (0-based) of the ordered pairs lo < hi <= n
  subroutine pair_of(p, n, lo, hi)
    integer(c_int64_t), intent(in) :: p
    integer, intent(in) :: n
    integer, intent(out) :: lo, hi
    integer(c_int64_t) :: base
    lo = 1
    base = 0
    do while (base + (n - lo) <= p)
       base = base + (n - lo)
       lo = lo + 1
    end do
    hi = lo + int(p - base) + 1
  end subroutine pair_of

  ! kpq columns 3..6 (k, l, a, b as MO indices; l = -1 for 1h1p) of
  ! configuration idx (1-based) in the order of select_atom_is/select_atom_d
  subroutine config_of(idx, k, l, a, b)
    integer(c_int64_t), intent(in) :: idx
    integer, intent(out) :: k, l, a, b
    integer(c_int64_t) :: u, npv, off(6)
    integer :: t
    npv = int(nvirt, c_int64_t) * (nvirt - 1) / 2
    off(1) = n1
    do t = 1, 5
       off(t + 1) = off(t) + counts(t)
    end do
    if (idx <= n1) then
       u = idx - 1
       a = nocc + int(u / nocc) + 1
       k = int(mod(u, int(nocc, c_int64_t))) + 1
       l = -1
       b = -1
    else if (idx <= off(2)) then          ! t1: i outer, a inner
       u = idx - off(1) - 1
       k = int(u / nvirt) + 1
       l = k
       a = nocc + int(mod(u, int(nvirt, c_int64_t))) + 1
       b = a
    else if (idx <= off(3)) then          ! t2: (a<b) outer, i inner
       u = idx - off(2) - 1
       k = int(mod(u, int(nocc, c_int64_t))) + 1
       l = k
       call pair_of(u / nocc, nvirt, a, b)
       a = nocc + a
       b = nocc + b
    else if (idx <= off(4)) then          ! t3: (i<j) outer, a inner
       u = idx - off(3) - 1
       a = nocc + int(mod(u, int(nvirt, c_int64_t))) + 1
       b = a
       call pair_of(u / nvirt, nocc, k, l)
    else                                  ! t4, t5: (i<j) outer, (a<b) inner
       u = idx - merge(off(5), off(4), idx > off(5)) - 1
       call pair_of(u / npv, nocc, k, l)
       call pair_of(mod(u, npv), nvirt, a, b)
       a = nocc + a
       b = nocc + b
    end if
  end subroutine config_of

  ! table2 with wrstateinfo (source/utilities/misc.f90:648-852): per state the
  ! energy, excitation energy, transition moment and oscillator strength (when
  ! an operator matrix was given) and the configurations among the 50 largest
  ! |C| with |C| >= 0.05, in the reference's formats
  subroutine write_state_table(fname)
    character(len=*), intent(in) :: fname
    integer :: st, ks, io, kk, ll, aa, bb
    integer(c_int64_t) :: j, lim
    integer(c_int64_t), allocatable :: order(:)
    real(c_double), allocatable :: c2(:)
    character(len=120) :: fmat
    character(len=2) :: spin
    lim = n1 + counts(1) + counts(2) + counts(3) + counts(4)
    open(newunit=io, file=fname, form='formatted', status='replace')
    write(io, '(/,2x,a,2x,F14.8)') 'Ground state MP2 energy:', ehf + emp2
    allocate(order(ndim), c2(ndim))
    do st = 1, nstates
       c2 = eigvec(:, st)**2
       ! the indices of the 50 largest C^2, descending
       do ks = 1, int(min(50_c_int64_t, ndim))
          order(ks) = maxloc(c2, dim=1, kind=c_int64_t)
          c2(order(ks)) = -1.0d0
       end do
       if (st < 10) then
          fmat = '(2/,2x,a,3x,i1,a1,i1,2x,a,5x,F14.8)'
       else
          fmat = '(2/,2x,a,2x,i2,a1,i1,2x,a,5x,F14.8)'
       end if
       write(io, fmat) 'State', st, '.', 1, 'Energy:', eigval(st) + ehf + emp2
       if (eigval(st) * adc2g_eh2ev < 10.0d0) then
          write(io, '(2x,a,2x,F10.5)') 'Excitation Energy (eV):', eigval(st) * adc2g_eh2ev
       else
          write(io, '(2x,a,3x,F10.5)') 'Excitation Energy (eV):', eigval(st) * adc2g_eh2ev
       end if
       if (allocated(tmv)) then
          write(io, '(2x,a,F10.5)') 'Transition Dipole Moment:', tmv(st)
          write(io, '(2x,a,5x,F10.5)') 'Oscillator Strength:', osc(st)
       end if
       write(io, '(/,2x,a,/)') 'Dominant Configurations:'
       write(io, '(2x,29a)') ('*', ks = 1, 29)
       write(io, '(3x,a)') 'j   k -> a  b        C_jkab'
       write(io, '(2x,29a)') ('*', ks = 1, 29)
       do ks = 1, int(min(50_c_int64_t, ndim))
          j = order(ks)
          if (abs(eigvec(j, st)) < 0.05d0) cycle
          call config_of(j, kk, ll, aa, bb)
          if (ll == -1) then
             write(io, '(3x,i2,4x,a2,1x,i2,9x,F8.5)') kk, '->', aa, eigvec(j, st)
          else if (kk /= ll .and. aa /= bb) then
             spin = merge('I ', 'II', j <= lim)
             write(io, '(3x,2(i2,1x),a2,2(1x,i2),2x,a2,2x,F8.5)') kk, ll, '->', aa, bb, spin, eigvec(j, st)
          else
             write(io, '(3x,2(i2,1x),a2,2(1x,i2),6x,F8.5)') kk, ll, '->', aa, bb, eigvec(j, st)
          end if
       end do
       write(io, '(2x,29a)') ('*', ks = 1, 29)
    end do
    close(io)
  end subroutine write_state_table

  subroutine check(code)
    integer(c_int32_t), intent(in) :: code
    if (code /= ADC2G_OK) then
       write(*, '(a,i4,2a)') ' adc2g error ', code, ': ', trim(adc2g_errmsg())
       stop 1
    end if
  end subroutine check

end program adc2g_driver
