!------------------------------------------------------------------------------
! batch_check.f90 -- the batched-trajectory extension called from Fortran.
!
! Three trajectories of one cell (a = 8, NG = 32, NB = NBocc = 2, Z = 4) with
! different k-meshes (NK = 2, 4, 2) and fields run once through
! tdhf_batch_create / upload / propagate / download and once each through
! tdhf_rt_run (the drop-in for RT_calculation, src/RT_calculation.f90:2-69)
! on its own context.  Prints the largest differences of E(t), J(t) and the
! final orbitals and stops with status 1 if any exceeds 1e-12.
!------------------------------------------------------------------------------
program batch_check
  use iso_c_binding
  use tdhf_capi
  implicit none
  integer, parameter :: ng = 32, nb = 2, ntraj = 3, nt = 40
  integer(c_int) :: nks(ntraj) = [2, 4, 2]
  real(8), parameter :: acell = 8.d0, dt = 0.025d0, pi = 3.141592653589793d0
  type(tdhf_params) :: p
  type(c_ptr) :: ctx, b
  real(8) :: vion(ng, ntraj), A(nt, ntraj), Aave(nt, ntraj)
  real(8) :: e1(nt, ntraj), n1(nt, ntraj), j1(nt, ntraj), i31(3, ntraj)
  real(8) :: eb(nt, ntraj), nbt(nt, ntraj), jb(nt, ntraj), i3b(3, ntraj)
  complex(8), allocatable :: u0(:), u1(:), ub(:)
  integer :: off(ntraj + 1), t, it, ik, ib, ig, rc, m
  real(8) :: x, x0, de, dj, du, a0

  off(1) = 0
  do t = 1, ntraj
    off(t + 1) = off(t) + ng * nb * nks(t)
  end do
  allocate(u0(off(ntraj + 1)), u1(off(ntraj + 1)), ub(off(ntraj + 1)))

  p%ng = ng
  p%nb = nb
  p%nbocc = nb
  p%acell = acell
  p%dt = dt
  do t = 1, ntraj
    p%nk = nks(t)
    rc = tdhf_ion_potential(p, 1_c_int, [4.d0], [1.d0], [0.d0], vion(:, t))
    if (rc /= TDHF_OK) stop 1
    ! Gaussian orbitals at band- and k-dependent centres, then Gram-Schmidt
    m = off(t)
    do ik = 1, nks(t)
      do ib = 1, nb
        x0 = acell * (0.3d0 * ib + 0.1d0 * ik)
        do ig = 1, ng
          x = acell * (ig - 1) / ng
          m = m + 1
          u0(m) = exp(-0.5d0 * (x - x0)**2) * exp(cmplx(0.d0, 0.2d0 * ik * x, 8))
        end do
      end do
    end do
    call tdhf_check(tdhf_create(p, vion(:, t), c_null_ptr, c_null_ptr, ctx), c_null_ptr, 'tdhf_create')
    call tdhf_check(tdhf_gram_schmidt(ctx, u0(off(t) + 1:off(t + 1))), ctx, 'tdhf_gram_schmidt')
    ! a short sin^2 pulse per trajectory; Aave(it) = (A(it) + A(it-1))/2
    a0 = 0.1d0 * t
    do it = 1, nt
      A(it, t) = a0 * sin(pi * it / nt)**2 * sin(0.057d0 * dt * it)
      Aave(it, t) = 0.5d0 * A(it, t)
      if (it > 1) Aave(it, t) = Aave(it, t) + 0.5d0 * A(it - 1, t)
    end do
    u1(off(t) + 1:off(t + 1)) = u0(off(t) + 1:off(t + 1))
    call tdhf_check(tdhf_rt_run(ctx, u1(off(t) + 1:off(t + 1)), nt, A(:, t), Aave(:, t), &
                                e1(:, t), n1(:, t), j1(:, t), i31(:, t)), ctx, 'tdhf_rt_run')
    rc = tdhf_destroy(ctx)
  end do

  p%nk = 0
  rc = tdhf_batch_create(p, ntraj, nks, vion, b)
  if (rc /= TDHF_OK) then
    write(*,*) 'tdhf_batch_create failed', rc
    stop 1
  end if
  ! NG = 32: the pair-tile exchange is refused, the row-wise one is used
  if (tdhf_batch_set_exchange(b, 1_c_int) == TDHF_OK) stop 1
  if (tdhf_batch_upload(b, -1_c_int, u0) /= TDHF_OK) stop 1
  if (tdhf_batch_propagate(b, nt, A, Aave, eb, nbt, jb, i3b) /= TDHF_OK) stop 1
  if (tdhf_batch_download(b, -1_c_int, ub) /= TDHF_OK) stop 1
  rc = tdhf_batch_destroy(b)

  de = maxval(abs(eb - e1))
  dj = maxval(abs(jb - j1)) / maxval(abs(j1))
  du = maxval(abs(ub - u1))
  write(*,'(a,3es12.3)') 'batch_check: max|dE|, max|dJ|/max|J|, max|du| =', de, dj, du
  if (de > 1d-12 .or. dj > 1d-12 .or. du > 1d-12) stop 1
  write(*,'(a)') 'batch_check: OK'
end program batch_check
