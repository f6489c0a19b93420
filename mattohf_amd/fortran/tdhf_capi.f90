!------------------------------------------------------------------------------
! tdhf_capi.f90 -- iso_c_binding interface to libtdhf.so (include/tdhf.h).
!
! Lets a Fortran caller that holds ubk(NG,NB,NK) complex(8) -- the layout of
! the reference's GS_calculation / RT_calculation (src/GS_calculation.f90:8,
! src/RT_calculation.f90:10) -- call the gfx950 library directly.
!------------------------------------------------------------------------------
module tdhf_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: TDHF_OK = 0
  integer(c_int), parameter :: TDHF_OBJ_ENERGY = 0, TDHF_OBJ_JSQ = 1

  type, bind(C) :: tdhf_params
    integer(c_int) :: ng, nk, nb, nbocc
    real(c_double) :: acell, dt
  end type tdhf_params

  interface
    integer(c_int) function tdhf_create(p, vion, vcg, vcgkk, ctx) bind(C, name="tdhf_create")
      import :: c_int, c_double, c_ptr, tdhf_params
      type(tdhf_params), intent(in) :: p
      real(c_double), intent(in) :: vion(*)
      type(c_ptr), value :: vcg, vcgkk
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function tdhf_destroy(ctx) bind(C, name="tdhf_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function tdhf_last_error_c(ctx) bind(C, name="tdhf_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function tdhf_ion_potential(p, nion, z, alpha, rion, vion) &
         bind(C, name="tdhf_ion_potential")
      import :: c_int, c_double, tdhf_params
      type(tdhf_params), intent(in) :: p
      integer(c_int), value :: nion
      real(c_double), intent(in) :: z(*), alpha(*), rion(*)
      real(c_double), intent(out) :: vion(*)
    end function

    integer(c_int) function tdhf_get_tables(ctx, vcg, vq) bind(C, name="tdhf_get_tables")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: vcg(*), vq(*)
    end function

    integer(c_int) function tdhf_hpsi(ctx, u, w, A, hw) bind(C, name="tdhf_hpsi")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: u(*), w(*)
      real(c_double), value :: A
      complex(c_double_complex), intent(out) :: hw(*)
    end function

    integer(c_int) function tdhf_observables(ctx, u, A, etot, ecomps, nele, je) &
         bind(C, name="tdhf_observables")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: u(*)
      real(c_double), value :: A
      real(c_double), intent(out) :: etot, ecomps(4), nele, je
    end function

    integer(c_int) function tdhf_rk4_step(ctx, u, Aave, A) bind(C, name="tdhf_rk4_step")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(inout) :: u(*)
      real(c_double), value :: Aave, A
    end function

    integer(c_int) function tdhf_rt_run(ctx, u, nt, A, Aave, etot, nele, jtot, init3) &
         bind(C, name="tdhf_rt_run")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(inout) :: u(*)
      integer(c_int), value :: nt
      real(c_double), intent(in) :: A(*), Aave(*)
      real(c_double), intent(out) :: etot(*), nele(*), jtot(*), init3(3)
    end function

    ! the RT loop in pieces on the device-resident state and field
    integer(c_int) function tdhf_state_upload(ctx, u) bind(C, name="tdhf_state_upload")
      import :: c_int, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: u(*)
    end function

    integer(c_int) function tdhf_state_download(ctx, u) bind(C, name="tdhf_state_download")
      import :: c_int, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: u(*)
    end function

    integer(c_int) function tdhf_rt_begin(ctx, nt, A, Aave) bind(C, name="tdhf_rt_begin")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nt
      real(c_double), intent(in) :: A(*), Aave(*)
    end function

    integer(c_int) function tdhf_rt_advance(ctx, n) bind(C, name="tdhf_rt_advance")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
    end function

    integer(c_int) function tdhf_rt_end(ctx, etot, nele, jtot, init3) bind(C, name="tdhf_rt_end")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: etot(*), nele(*), jtot(*), init3(3)
    end function

    integer(c_int) function tdhf_gs_run(ctx, u, nscf, ncg, tol, etot_hist, epsbk, epserr, &
         niter) bind(C, name="tdhf_gs_run")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(inout) :: u(*)
      integer(c_int), value :: nscf, ncg
      real(c_double), value :: tol
      real(c_double), intent(out) :: etot_hist(*), epsbk(*), epserr(*)
      integer(c_int), intent(out) :: niter
    end function

    integer(c_int) function tdhf_gram_schmidt(ctx, u) bind(C, name="tdhf_gram_schmidt")
      import :: c_int, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(inout) :: u(*)
    end function

    integer(c_int) function tdhf_te_step(ctx, u, Aave, variant) bind(C, name="tdhf_te_step")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(inout) :: u(*)
      real(c_double), value :: Aave
      integer(c_int), value :: variant
    end function

    integer(c_int) function tdhf_gs_run_diag(ctx, u, nscf, ncg, etot_hist, epsbk, epserr, diag) &
         bind(C, name="tdhf_gs_run_diag")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(inout) :: u(*)
      integer(c_int), value :: nscf, ncg
      real(c_double), intent(out) :: etot_hist(*), epsbk(*), epserr(*), diag(*)
    end function

    integer(c_int) function tdhf_make_dos(nb, nk, nbocc, epsbk, ndos, epsdos, dos, dosocc) &
         bind(C, name="tdhf_make_dos")
      import :: c_int, c_double
      integer(c_int), value :: nb, nk, nbocc, ndos
      real(c_double), intent(in) :: epsbk(*)
      real(c_double), intent(out) :: epsdos(*), dos(*), dosocc(*)
    end function

    integer(c_int) function tdhf_fourier(nt, dt, time, J, j0, E, A, nomega, omegam, window, &
         tpulse, nenvelope, out) bind(C, name="tdhf_fourier")
      import :: c_int, c_double
      integer(c_int), value :: nt, nomega, window, nenvelope
      real(c_double), value :: dt, j0, omegam, tpulse
      real(c_double), intent(in) :: time(*), J(*), E(*), A(*)
      real(c_double), intent(out) :: out(*)
    end function

    ! adjoint pair: gradient of Etot(u_N) - eref with respect to A(1:nt) and u0
    integer(c_int) function tdhf_fwd(ctx, u, nt, A, eref, mem_budget, L, etot, nele, jtot) &
         bind(C, name="tdhf_fwd")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(inout) :: u(*)
      integer(c_int), value :: nt
      real(c_double), intent(in) :: A(*)
      real(c_double), value :: eref, mem_budget
      real(c_double), intent(out) :: L, etot(*), nele(*), jtot(*)
    end function

    integer(c_int) function tdhf_bwd(ctx, dLdA, dLdu0) bind(C, name="tdhf_bwd")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dLdA(*)
      complex(c_double_complex), intent(out) :: dLdu0(*)
    end function

    ! the same for ntraj pulses at once (u(NG,NB,NK,ntraj), A(nt,ntraj)) and
    ! either objective (TDHF_OBJ_ENERGY, TDHF_OBJ_JSQ = dt*sum J_n^2)
    integer(c_int) function tdhf_fwd_batch(ctx, ntraj, u, nt, A, objective, eref, mem_budget, &
         L, etot, nele, jtot) bind(C, name="tdhf_fwd_batch")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ntraj, nt, objective
      complex(c_double_complex), intent(inout) :: u(*)
      real(c_double), intent(in) :: A(*), eref(*)
      real(c_double), value :: mem_budget
      real(c_double), intent(out) :: L(*), etot(*), nele(*), jtot(*)
    end function

    integer(c_int) function tdhf_bwd_batch(ctx, dLdA, dLdu0) bind(C, name="tdhf_bwd_batch")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dLdA(*)
      complex(c_double_complex), intent(out) :: dLdu0(*)
    end function

    ! how the last forward stored its states (full tape, or checkpoints every
    ! segment steps)
    integer(c_int) function tdhf_adjoint_info(ctx, full_tape, segment) &
         bind(C, name="tdhf_adjoint_info")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: full_tape, segment
    end function

    ! progress callback of the checkpointed adjoint calls: fn is c_funloc of a
    ! bind(C) subroutine with the interface tdhf_progress below, or
    ! c_null_funptr to turn the reports off
    integer(c_int) function tdhf_set_progress(ctx, fn, user) bind(C, name="tdhf_set_progress")
      import :: c_int, c_ptr, c_funptr
      type(c_ptr), value :: ctx
      type(c_funptr), value :: fn
      type(c_ptr), value :: user
    end function
  end interface

  ! batched trajectories (include/tdhf.h, "Batched trajectories"): nk(ntraj),
  ! vion(NG,ntraj), u = the trajectories' ubk(NG,NB,nk(t)) back to back,
  ! A(nt,ntraj), Aave(nt,ntraj), etot/nele/jtot(nt,ntraj), init3(3,ntraj)
  interface
    integer(c_int) function tdhf_batch_create(p, ntraj, nk, vion, b) bind(C, name="tdhf_batch_create")
      import :: c_int, c_double, c_ptr, tdhf_params
      type(tdhf_params), intent(in) :: p
      integer(c_int), value :: ntraj
      integer(c_int), intent(in) :: nk(*)
      real(c_double), intent(in) :: vion(*)
      type(c_ptr), intent(out) :: b
    end function

    integer(c_int) function tdhf_batch_destroy(b) bind(C, name="tdhf_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function

    type(c_ptr) function tdhf_batch_last_error_c(b) bind(C, name="tdhf_batch_last_error")
      import :: c_ptr
      type(c_ptr), value :: b
    end function

    ! traj >= 0: one trajectory's orbitals; traj < 0: all of them
    integer(c_int) function tdhf_batch_upload(b, traj, u) bind(C, name="tdhf_batch_upload")
      import :: c_int, c_double_complex, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: traj
      complex(c_double_complex), intent(in) :: u(*)
    end function

    integer(c_int) function tdhf_batch_download(b, traj, u) bind(C, name="tdhf_batch_download")
      import :: c_int, c_double_complex, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: traj
      complex(c_double_complex), intent(out) :: u(*)
    end function

    integer(c_int) function tdhf_batch_propagate(b, nt, A, Aave, etot, nele, jtot, init3) &
         bind(C, name="tdhf_batch_propagate")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: nt
      real(c_double), intent(in) :: A(*), Aave(*)
      real(c_double), intent(out) :: etot(*), nele(*), jtot(*), init3(*)
    end function

    integer(c_int) function tdhf_batch_autotune(b, nt_trial, A, Aave, best_nw) &
         bind(C, name="tdhf_batch_autotune")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: nt_trial
      real(c_double), intent(in) :: A(*), Aave(*)
      integer(c_int), intent(out) :: best_nw
    end function

    integer(c_int) function tdhf_batch_set_waves(b, nw) bind(C, name="tdhf_batch_set_waves")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: nw
    end function

    integer(c_int) function tdhf_batch_set_schedule(b, sched) bind(C, name="tdhf_batch_set_schedule")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: sched
    end function

    integer(c_int) function tdhf_batch_set_exchange(b, mode) bind(C, name="tdhf_batch_set_exchange")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: mode
    end function
  end interface

  abstract interface
    subroutine tdhf_progress(user, phase, done, total) bind(C)
      import :: c_ptr, c_int, c_long
      type(c_ptr), value :: user
      integer(c_int), value :: phase
      integer(c_long), value :: done, total
    end subroutine
  end interface

contains

  ! Message of the last failed call (tdhf_last_error) as a Fortran string.
  function tdhf_last_error(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=512) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: i
    msg = ''
    p = tdhf_last_error_c(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [512])
    do i = 1, 512
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function tdhf_last_error

  ! Stops with the library's message when rc /= 0.
  subroutine tdhf_check(rc, ctx, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: what
    if (rc /= TDHF_OK) then
      write(*,'(a,a,i4,a,a)') trim(what), ' failed (', rc, '): ', trim(tdhf_last_error(ctx))
      stop 1
    end if
  end subroutine tdhf_check

end module tdhf_capi
