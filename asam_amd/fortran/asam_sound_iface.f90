!> ISO_C_BINDING interface to libasam_sound (include/asam_sound.h).
!>
!> Derived types mirror the C structs field by field (BIND(C)); every entry point of
!> the C-ABI that a Fortran host needs is declared here with VALUE/reference
!> arguments matching the C prototypes.  Host arrays are passed as assumed-size
!> REAL(c_double) dummies, i.e. by address in their native column-major order, which
!> is exactly the layout the C-ABI expects (component fastest, then x, y, z).
module asam_sound_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: ASAM_METHOD_BICGSTAB = 0
  integer(c_int), parameter :: ASAM_METHOD_FGMRES = 1
  integer(c_int), parameter :: ASAM_METHOD_GMRES = 2
  integer(c_int), parameter :: ASAM_METHOD_FGMRESV = 3
  integer(c_int), parameter :: ASAM_PREC_FP64 = 0
  integer(c_int), parameter :: ASAM_PREC_MIXED = 1
  integer(c_int), parameter :: ASAM_E_INVALID = -1

  type, bind(c) :: asam_sound_grid
    integer(c_int) :: nx, ny, nz
    integer(c_int) :: np1
    integer(c_int) :: grad_full
    real(c_double) :: beta0_dt
    real(c_double) :: grav
    real(c_double) :: fac_anela
  end type asam_sound_grid

  type, bind(c) :: asam_sound_control
    integer(c_int) :: method
    integer(c_int) :: step_gmres
    integer(c_int) :: precision
    real(c_double) :: inner_tol
    integer(c_int) :: max_inner
    integer(c_int) :: max_outer
  end type asam_sound_control

  type, bind(c) :: asam_sound_info
    integer(c_int) :: iterations
    integer(c_int) :: restarts
    integer(c_int) :: status
    integer(c_int) :: reserved
    real(c_double) :: rel_residual
    real(c_double) :: abs_residual
    real(c_double) :: ms
  end type asam_sound_info

  interface
    integer(c_int) function asam_sound_create(ctx, grid) bind(c, name="asam_sound_create")
      import :: c_int, c_ptr, asam_sound_grid
      type(c_ptr), intent(out) :: ctx
      type(asam_sound_grid), intent(in) :: grid
    end function

    integer(c_int) function asam_sound_destroy(ctx) bind(c, name="asam_sound_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function asam_sound_default_control(ctl) &
        bind(c, name="asam_sound_default_control")
      import :: c_int, asam_sound_control
      type(asam_sound_control), intent(out) :: ctl
    end function

    integer(c_int) function asam_sound_set_control(ctx, ctl) &
        bind(c, name="asam_sound_set_control")
      import :: c_int, c_ptr, asam_sound_control
      type(c_ptr), value :: ctx
      type(asam_sound_control), intent(in) :: ctl
    end function

    type(c_ptr) function asam_sound_last_error(ctx) bind(c, name="asam_sound_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function asam_sound_setup(ctx, FU, FV, FW, VolC, dx, dy, dz, rho, &
                                             rhotheta, T) bind(c, name="asam_sound_setup")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: FU(*), FV(*), FW(*), VolC(*), dx(*), dy(*), dz(*)
      real(c_double), intent(in) :: rho(*), rhotheta(*), T(*)
    end function

    integer(c_int) function asam_sound_setup_moist(ctx, FU, FV, FW, VolC, dx, dy, dz, rho, &
                                                   rhotheta, T, rhov, rhol, rhor) &
        bind(c, name="asam_sound_setup_moist")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: FU(*), FV(*), FW(*), VolC(*), dx(*), dy(*), dz(*)
      real(c_double), intent(in) :: rho(*), rhotheta(*), T(*)
      type(c_ptr), value :: rhov, rhol, rhor
    end function

    integer(c_int) function asam_sound_setup_weights(ctx, FU, FV, FW, VolC, dx, dy, dz, &
                                                     DTU, DTV, DTW, DUT) &
        bind(c, name="asam_sound_setup_weights")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: FU(*), FV(*), FW(*), VolC(*), dx(*), dy(*), dz(*)
      real(c_double), intent(in) :: DTU(*), DTV(*), DTW(*), DUT(*)
    end function

    integer(c_int) function asam_sound_solve(ctx, x, b, maxiter, tol, info) &
        bind(c, name="asam_sound_solve")
      import :: c_int, c_ptr, c_double, asam_sound_info
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: x(*)
      real(c_double), intent(in) :: b(*)
      integer(c_int), intent(inout) :: maxiter
      real(c_double), intent(inout) :: tol
      type(asam_sound_info), intent(out) :: info
    end function

    integer(c_int) function asam_sound_solve_T(ctx, x, b, maxiter, tol, info) &
        bind(c, name="asam_sound_solve_T")
      import :: c_int, c_ptr, c_double, asam_sound_info
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: x(*)
      real(c_double), intent(in) :: b(*)
      integer(c_int), intent(inout) :: maxiter
      real(c_double), intent(inout) :: tol
      type(asam_sound_info), intent(out) :: info
    end function

    integer(c_int) function asam_sound_apply(ctx, y, x) bind(c, name="asam_sound_apply")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: y(*)
      real(c_double), intent(in) :: x(*)
    end function

    integer(c_int) function asam_sound_apply_T(ctx, y, x) bind(c, name="asam_sound_apply_T")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: y(*)
      real(c_double), intent(in) :: x(*)
    end function

    integer(c_int) function asam_sound_rhs(ctx, b, U, V, W, incr_rt, incr_r) &
        bind(c, name="asam_sound_rhs")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: b(*)
      real(c_double), intent(in) :: U(*), V(*), W(*), incr_rt(*)
      type(c_ptr), value :: incr_r
    end function

    integer(c_int) function asam_sound_update(ctx, x, U, V, W, incr_rt, incr_r) &
        bind(c, name="asam_sound_update")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: x(*), U(*), V(*), W(*), incr_rt(*)
      type(c_ptr), value :: incr_r
    end function

    integer(c_int) function asam_sound_psolve(ctx, y, x) bind(c, name="asam_sound_psolve")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: y(*)
      real(c_double), intent(in) :: x(*)
    end function
  end interface

contains

  !> Last error message of a context as a Fortran string.
  function asam_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: i
    msg = ' '
    p = asam_sound_last_error(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [256])
    do i = 1, 256
      if (cs(i) == c_null_char) exit
      msg(i:i) = cs(i)
    end do
  end function asam_error_message

end module asam_sound_iface
