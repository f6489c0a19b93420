!------------------------------------------------------------------------------
! mattohf_main.f90 -- Fortran driver of the gfx950 TDHF path.
!
! Mirrors the reference's Program Main (src/mATTOHF.f90:2-132): the input file
! name is read from stdin, the keyword file is parsed as Read_input does
! (src/mATTOHF.f90:198-272), the initial orbitals come from Init_wf's LCG
! (src/mATTOHF.f90:159-196), and the ground state and the real-time run are
! computed by libtdhf.so through the iso_c_binding module tdhf_capi in place
! of GS_calculation and RT_calculation (src/mATTOHF.f90:104,114).  Output
! files follow the reference's names and formats:
!   <sys_name>_GSinfo.out     per-iteration columns and the SCF summary
!                             (src/GS_calculation.f90:22-24,48-51,61-91)
!   <sys_name>_DoS.out        band table and Make_DoS curves (:93-110, 239-265)
!   <sys_name>_band.out       ik, kx, residual, eps(:,ik)   (:112-117)
!   <sys_name>_EtotNJEAt.out  8 columns, e20.12          (src/RT_calculation.f90:22-24,52-53)
!   <sys_name>_JEAF.out, _JEAF_window.out, _sigmaeps.out (LR)
!                             FourierTr / WindowFourierTr (:246-351)
! TD_option 'TE', 'TEPC' or 'TEPC2' selects the reference's Taylor propagators
! (src/RT_calculation.f90:71-151, commented out in its time loop :39-41)
! instead of RK4.
!------------------------------------------------------------------------------
module drv_state
  implicit none
  ! defaults of src/Global_variables.f90:5-26
  character(256) :: input_data = '', sys_name = ''
  character(256) :: dir_name = './'
  character(32) :: DFT_option = 'hand', TD_option = 'RK4'
  logical :: Band_option = .false.
  character(256) :: functional = 'HF'
  integer :: Nscf = 100, NCG = 5, NT = 2000, Nomega = 4000
  real(8) :: dt = 0.5, omegaM = 4.0
  character(2) :: field_option = 'LP'
  real(8) :: dAc = 0.0001   ! single-precision literal, as src/Global_variables.f90:20
  real(8) :: omegac = 1.55d0, Tpulse = 16.0d0, phiCEP = 0.5d0, E0 = 1.0d0
  integer :: nenvelope = 4
  integer :: NG = 0, NK = 0, NB = 0, NBocc = 0
  real(8) :: acell = 0.d0
  integer :: Nspecies = 0, Nion = 0
  real(8), allocatable :: Zeffatom(:), alphaatom(:), Rion(:)
  integer, allocatable :: Kion(:)
  real(8) :: gs_tol = 0.d0
end module drv_state

program mattohf_drv
  use iso_c_binding
  use tdhf_capi
  use drv_state
  implicit none
  real(8), parameter :: pi = 4.d0*atan(1.d0)
  type(tdhf_params) :: p
  type(c_ptr) :: ctx
  complex(8), allocatable :: ubk(:,:,:)
  real(8), allocatable :: vion(:), zion(:), aion(:), kx(:), rx(:)
  real(8), allocatable :: etot_hist(:), epsbk(:,:), epserr(:,:)
  real(8), allocatable :: time(:), A(:), Aave(:), E(:), Etot(:), Nele(:), Jtot(:)
  real(8), allocatable :: diag(:,:), epsdos(:), dos(:), dosocc(:), spec(:,:)
  real(8) :: init3(3), bx, ecomps(4), etmp, ntmp, jtmp, Tau
  integer, parameter :: NDoS = 10000
  real(8), parameter :: Hartree = 27.2116d0, Atomtime = 0.02419d0
  integer(c_int) :: rc, niter
  integer :: ik, ib, it, ig, iseed, ii
  real(8) :: rnd, rxc
  ! wall times of the phases the reference reports (src/mATTOHF.f90:118-125)
  real(8) :: ts, tpreps, tprepe, tGSs, tGSe, tRTs, tRTe, te
  real(8), allocatable :: tchunk(:)
  integer :: done, nstep

  ts = wtime()
  tpreps = ts
  write(*,*) 'You can specify a file for input data just below(If you do not have, just input [""]'
  read(*,*) input_data
  if (trim(input_data) == '') then
    write(*,*) 'an input file is required (NG, NK, NB, NBocc, acell and atoms have no defaults)'
    stop 1
  end if
  call read_input()

  ! grids and occupations as src/Initialization.f90:93-112,176-186
  allocate(rx(NG), kx(NK))
  bx = 2.d0*pi/acell
  do ig = 1, NG
    rx(ig) = dble(ig-1)/dble(NG)*acell
  end do
  do ik = 1, NK
    kx(ik) = (-0.5d0 + (2.d0*ik - 1.0d0)/dble(2*NK))*bx
  end do

  p%ng = NG; p%nk = NK; p%nb = NB; p%nbocc = NBocc; p%acell = acell; p%dt = dt

  ! Prep_periodic_ion_potential on the device
  allocate(vion(NG), zion(Nion), aion(Nion))
  do ii = 1, Nion
    zion(ii) = Zeffatom(Kion(ii))
    aion(ii) = alphaatom(Kion(ii))
  end do
  rc = tdhf_ion_potential(p, int(Nion, c_int), zion, aion, Rion, vion)
  call tdhf_check(rc, c_null_ptr, 'tdhf_ion_potential')

  rc = tdhf_create(p, vion, c_null_ptr, c_null_ptr, ctx)
  call tdhf_check(rc, c_null_ptr, 'tdhf_create')

  ! Init_wf (src/mATTOHF.f90:159-184) and Gram_Schmidt
  allocate(ubk(NG,NB,NK))
  iseed = 123
  do ik = 1, NK
    do ib = 1, NB
      iseed = mod(iseed*106 + 1283, 6075)
      rnd = dble(iseed)/dble(6075)
      rxc = rnd*acell
      ubk(:,ib,ik) = exp(-0.5d0*(rx(:) - rxc)**2)
    end do
  end do
  rc = tdhf_gram_schmidt(ctx, ubk)
  call tdhf_check(rc, ctx, 'tdhf_gram_schmidt')

  ! ground state
  tprepe = wtime()
  tGSs = tprepe
  allocate(etot_hist(0:Nscf), epsbk(NB,NK), epserr(NB,NK), diag(5,max(Nscf,1)))
  if (gs_tol > 0.d0) then
    rc = tdhf_gs_run(ctx, ubk, int(Nscf, c_int), int(NCG, c_int), gs_tol, etot_hist, epsbk, &
         epserr, niter)
    call tdhf_check(rc, ctx, 'tdhf_gs_run')
  else
    rc = tdhf_gs_run_diag(ctx, ubk, int(Nscf, c_int), int(NCG, c_int), etot_hist, epsbk, &
         epserr, diag)
    call tdhf_check(rc, ctx, 'tdhf_gs_run_diag')
    niter = Nscf
  end if
  tGSe = wtime()
  write(*,'("########SCF procedure is finished.########")')
  rc = tdhf_observables(ctx, ubk, 0.d0, etmp, ecomps, ntmp, jtmp)
  call tdhf_check(rc, ctx, 'tdhf_observables')
  open(10, file=trim(sys_name)//'_GSinfo.out', action='write')
  write(10,'("#iter, Etotscf(iter), Etoscf(iter)-Etotscf(iter-1), ,sum(abs(dnsscf(:,iter+1)-dnsscf(:,iter)))/dble(NG) &
         sum(epserrorbk)/dble(NB*NK),maxval(epserrorbk),Fundamental gap, Optical Gap")')
  do it = 1, niter
    if (gs_tol > 0.d0) then
      write(10,'(2x,i5,100e20.12)') it, etot_hist(it), etot_hist(it) - etot_hist(it-1)
    else
      write(10,'(2x,i5,100e20.12)') it, etot_hist(it), etot_hist(it) - etot_hist(it-1), diag(:,it)
    end if
  end do
  write(10,'("#Ekin  at SCF:",f12.8)') ecomps(1)
  write(10,'("#Eion  at SCF:",f12.8)') ecomps(2)
  write(10,'("#EH    at SCF:",f12.8)') ecomps(3)
  write(10,'("#Exc   at SCF:",f12.8)') ecomps(4)
  write(10,'("#Min. value of the spectra: minval(epsbk)    =",f12.8," =",f12.8," [eV]")') &
       minval(epsbk), minval(epsbk)*Hartree
  write(10,'("#Max. value of the spectra: maxval(epsbk)    =",f12.8," =",f12.8," [eV]")') &
       maxval(epsbk), maxval(epsbk)*Hartree
  write(10,'("#Valence top: maxval(epsbk(:NBocc,:))        =",f12.8," =",f12.8," [eV]")') &
       maxval(epsbk(:NBocc,:)), maxval(epsbk(:NBocc,:))*Hartree
  ! the reference prints maxval here (src/GS_calculation.f90:84-85)
  write(10,'("#Conduction bottom: minval(epsbk(NBocc+1:,:))=",f12.8," =",f12.8," [eV]")') &
       maxval(epsbk(NBocc+1:,:)), maxval(epsbk(NBocc+1:,:))*Hartree
  write(10,'("#Fundamental gap: ",f12.8," =",f12.8," [eV]")') &
       minval(epsbk(NBocc+1:,:)) - maxval(epsbk(:NBocc,:)), &
       (minval(epsbk(NBocc+1:,:)) - maxval(epsbk(:NBocc,:)))*Hartree
  if (NB > NBocc) then
    write(10,'("#Min. gap among same k-point: ",f12.8," =",f12.8," [eV]")') &
         minval(epsbk(NBocc+1,:)-epsbk(NBocc,:)), (minval(epsbk(NBocc+1,:)-epsbk(NBocc,:)))*Hartree
  else
    write(10,'("#Min. gap among same k-point: ",f12.8," =",f12.8," [eV]")') &
         huge(0.d0), huge(0.d0)
  end if
  close(10)
  ! Make_DoS on the device, written as src/GS_calculation.f90:99-110
  allocate(epsdos(NDoS), dos(NDoS), dosocc(NDoS))
  rc = tdhf_make_dos(int(NB, c_int), int(NK, c_int), int(NBocc, c_int), epsbk, &
       int(NDoS, c_int), epsdos, dos, dosocc)
  call tdhf_check(rc, c_null_ptr, 'tdhf_make_dos')
  open(10, file=trim(sys_name)//'_DoS.out', action='write')
  write(10,*) '# ik,kx(ik),  sum(epserrorbk(:,ik))/dble(NB),  epsbk(:,ik)'
  do ik = 1, NK
    write(10,'("# ",i5,3e20.12,4x,1e20.12,4x,1000e20.12)') ik, kx(ik), sum(epserr(:,ik))/dble(NB), &
         epsbk(:,ik)
  end do
  do ii = 1, NDoS
    write(10,'(3e20.12e3)') epsdos(ii), dos(ii), dosocc(ii)
  end do
  close(10)
  open(10, file=trim(sys_name)//'_band.out', action='write')
  write(10,*) '# ik,kx(ik),  sum(epserrorbk(:,ik))/dble(NB),  epsbk(:,ik)'
  do ik = 1, NK
    write(10,'(i5,3e20.12,4x,1e20.12,4x,1000e20.12)') ik, kx(ik), sum(epserr(:,ik))/dble(NB), &
         epsbk(:,ik)
  end do
  close(10)
  write(*,'(" GS: ",i5," iterations, Etot =",f20.12)') niter, etot_hist(niter)

  ! real-time propagation
  allocate(time(NT), A(NT), Aave(NT), E(NT), Etot(NT), Nele(NT), Jtot(NT))
  allocate(tchunk(NT/1000 + 1))
  call define_afield(time, A, Aave, E)
  tRTs = wtime()
  select case (trim(TD_option))
  case ('TE', 'TEPC', 'TEPC2')
    ! the reference's Taylor propagators, one device step at a time
    rc = tdhf_observables(ctx, ubk, 0.d0, init3(1), ecomps, init3(2), init3(3))
    call tdhf_check(rc, ctx, 'tdhf_observables')
    do it = 1, NT
      if (trim(TD_option) == 'TE') then
        rc = tdhf_te_step(ctx, ubk, Aave(it), 0_c_int)
      else if (trim(TD_option) == 'TEPC') then
        rc = tdhf_te_step(ctx, ubk, Aave(it), 1_c_int)
      else
        rc = tdhf_te_step(ctx, ubk, Aave(it), 2_c_int)
      end if
      call tdhf_check(rc, ctx, 'tdhf_te_step')
      rc = tdhf_observables(ctx, ubk, A(it), Etot(it), ecomps, Nele(it), Jtot(it))
      call tdhf_check(rc, ctx, 'tdhf_observables')
      if (mod(it, 1000) == 0) tchunk(it/1000) = wtime() - tRTs
    end do
  case default
    ! tdhf_rt_run in pieces of 1000 steps (each call returns when the device
    ! has finished them), so that the reference's per-1000-step wall times can
    ! be reported
    call tdhf_check(tdhf_state_upload(ctx, ubk), ctx, 'tdhf_state_upload')
    call tdhf_check(tdhf_rt_begin(ctx, int(NT, c_int), A, Aave), ctx, 'tdhf_rt_begin')
    done = 0
    do while (done < NT)
      nstep = min(1000, NT - done)
      call tdhf_check(tdhf_rt_advance(ctx, int(nstep, c_int)), ctx, 'tdhf_rt_advance')
      done = done + nstep
      if (mod(done, 1000) == 0) tchunk(done/1000) = wtime() - tRTs
    end do
    call tdhf_check(tdhf_rt_end(ctx, Etot, Nele, Jtot, init3), ctx, 'tdhf_rt_end')
    call tdhf_check(tdhf_state_download(ctx, ubk), ctx, 'tdhf_state_download')
  end select
  tRTe = wtime()
  ! the reference's console lines of the time loop (src/RT_calculation.f90:37,54-60),
  ! written after the run
  write(*,'("#time(it), Etot(it)-Etot0, N(it)-N0, J(it), E(it), A(it)")')
  do it = 1, NT
    if (mod(it,100) == 0) write(*,'(i6,20e14.6)') it, time(it), Etot(it), Etot(it)-init3(1), &
         Nele(it), Nele(it)-init3(2), Jtot(it), E(it), A(it)
    if (mod(it,1000) == 0) &
         write(*,'("## calc-time for ",i4," iterations :",e12.4," [sec] =",f12.8," [hour]##")') &
         it, tchunk(it/1000), tchunk(it/1000)/3600.d0
  end do
  write(*,'("########Real-time propagation is finished.########")')
  open(10, file=trim(sys_name)//'_EtotNJEAt.out', action='write')
  write(10,'("#time(it), Etot(it)-Etot0, N(it)-N0, J(it), E(it), A(it)")')
  do it = 1, NT
    write(10,'(20e20.12)') time(it), Etot(it), Etot(it)-init3(1), Nele(it), Nele(it)-init3(2), &
         Jtot(it), E(it), A(it)
  end do
  close(10)
  write(*,'(" RT: ",i8," steps, Etot(NT) - Etot0 =",e14.6)') NT, Etot(NT) - init3(1)

  ! FourierTr and WindowFourierTr on the device (src/RT_calculation.f90:64-65)
  allocate(spec(7, Nomega))
  rc = tdhf_fourier(int(NT, c_int), dt, time, Jtot, init3(3), E, A, int(Nomega, c_int), omegaM, &
       0_c_int, 0.d0, int(nenvelope, c_int), spec)
  call tdhf_check(rc, c_null_ptr, 'tdhf_fourier')
  call write_jeaf(trim(sys_name)//'_JEAF.out')
  if (field_option == 'LR') then
    open(10, file=trim(sys_name)//'_sigmaeps.out', action='write')
    write(10,'("#omega(iomega), sigma(iomega), eps(iomega)")')
    do ii = 1, Nomega
      block
        complex(8) :: sig, eps
        sig = -cmplx(spec(2,ii), spec(3,ii), kind=8)/dAc
        eps = 4.d0*pi*(0.d0,1.d0)*sig/spec(1,ii)
        write(10,'(20e20.12)') spec(1,ii), real(sig), aimag(sig), real(eps), aimag(eps)
      end block
    end do
    close(10)
  else
    Tau = Tpulse/Atomtime
    rc = tdhf_fourier(int(NT, c_int), dt, time, Jtot, init3(3), E, A, int(Nomega, c_int), &
         omegaM, 1_c_int, Tau, int(nenvelope, c_int), spec)
    call tdhf_check(rc, c_null_ptr, 'tdhf_fourier')
    call write_jeaf(trim(sys_name)//'_JEAF_window.out')
  end if
  rc = tdhf_destroy(ctx)
  te = wtime()
  write(*,'("The execution is finished successfully!")')
  write(*,'("Time for preparation =",f12.8," [sec]")') tprepe-tpreps
  write(*,'("Time for GS calculatoin =",e12.4," [sec] =",f12.8," [min]")') tGSe-tGSs, (tGSe-tGSs)/60.d0
  write(*,'("Time for RT calculatoin =",e12.4," [sec] =",f12.8," [hour]")') tRTe-tRTs, (tRTe-tRTs)/3600.d0
  write(*,'("Time for an interaction in RT =",e12.4," [sec]")') (tRTe-tRTs)/dble(NT)
  write(*,'("Time for total  =",e12.4," [sec] =", f12.8," [hour]")') te-ts, (te-ts)/3600.d0

contains

  ! wall-clock seconds (the reference uses MPI_WTIME)
  real(8) function wtime()
    integer(8) :: c, r
    call system_clock(c, r)
    wtime = dble(c)/dble(r)
  end function wtime

  subroutine write_jeaf(fname)
    character(*), intent(in) :: fname
    integer :: iw
    open(10, file=fname, action='write')
    write(10,'("#omega(iomega), JF(iomega), EF(iomega), AF(iomega)")')
    do iw = 1, Nomega
      write(10,'(20e20.12)') spec(:, iw)
    end do
    close(10)
  end subroutine write_jeaf

  ! Read_input, src/mATTOHF.f90:198-272 (one keyword per line, values on the
  ! next; unknown lines are skipped).  'gs_tol' is an addition of this driver.
  subroutine read_input()
    character(256) :: str
    integer :: ios
    open(11, file=trim(input_data), status='old', action='read')
    do
      read(11, *, iostat=ios) str
      if (ios /= 0) exit
      select case (trim(str))
      case ('sys_name'); read(11,*) sys_name
      case ('dir_name'); read(11,*) dir_name
      case ('DFT_option'); read(11,*) DFT_option
      case ('Band_option'); read(11,*) Band_option
      case ('TD_option'); read(11,*) TD_option
      case ('acell'); read(11,*) acell
      case ('NB'); read(11,*) NB
      case ('NBocc'); read(11,*) NBocc
      case ('NG'); read(11,*) NG
      case ('NK'); read(11,*) NK
      case ('atoms')
        read(11,*) Nspecies
        allocate(Zeffatom(Nspecies), alphaatom(Nspecies))
        read(11,*) Zeffatom(1:Nspecies)
        read(11,*) alphaatom(1:Nspecies)
        read(11,*) Nion
        allocate(Rion(Nion), Kion(Nion))
        do ii = 1, Nion
          read(11,*) Rion(ii), Kion(ii)
        end do
      case ('Nscf'); read(11,*) Nscf
      case ('NCG'); read(11,*) NCG
      case ('NT'); read(11,*) NT
      case ('dt'); read(11,*) dt
      case ('Nomega'); read(11,*) Nomega
      case ('omegaM'); read(11,*) omegaM
      case ('field_option'); read(11,*) field_option
      case ('dAc'); read(11,*) dAc
      case ('omegac'); read(11,*) omegac
      case ('Tpulse'); read(11,*) Tpulse
      case ('phiCEP'); read(11,*) phiCEP
      case ('nenvelope'); read(11,*) nenvelope
      case ('E0'); read(11,*) E0
      case ('gs_tol'); read(11,*) gs_tol
      end select
    end do
    close(11)
  end subroutine read_input

  ! Define_Afield, src/RT_calculation.f90:175-244
  subroutine define_afield(time, A, Aave, E)
    real(8), intent(out) :: time(:), A(:), Aave(:), E(:)
    real(8), parameter :: Hartree = 27.2116d0, Atomtime = 0.02419d0, Atomfield = 514.2d0
    real(8) :: Am1, Aamp, w, T, Ef
    integer :: it, NTpulse
    do it = 1, NT
      time(it) = dt*(it-1)
    end do
    Am1 = 0.d0
    A(:) = 0.d0
    if (field_option == 'LR') then
      A(:) = dAc
    else if (field_option == 'LP') then
      w = omegac/Hartree
      T = Tpulse/Atomtime
      Ef = E0/Atomfield
      Aamp = Ef/w
      NTpulse = int(min(T/dt, dble(NT)))
      do it = 1, NTpulse
        A(it) = Aamp*sin(w*(time(it)-0.5*T)+2.0*pi*phiCEP)*(sin(pi*time(it)/T))**nenvelope
      end do
    else
      stop ('# Error stop: substituted field_option  is not acceptable.')
    end if
    Aave(1) = 0.5d0*(A(1)+Am1)
    do it = 2, NT
      Aave(it) = 0.5d0*(A(it)+A(it-1))
    end do
    E(1) = -(A(1)-Am1)/dt
    do it = 2, NT
      E(it) = -(A(it)-A(it-1))/dt
    end do
  end subroutine define_afield

end program mattohf_drv
