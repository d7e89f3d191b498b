!> ISO_C_BINDING interface to the transport stage solve of libasam_sound
!> (include/asam_transport.h).  BIND(C) types mirror the C structs field by field;
!> host arrays pass by address in their native column-major order, which is the layout
!> the C-ABI expects (SpDiag Val(m, NumDiag); vectors c(0:nx+1, 0:ny+1, 0:nz+1, ncomp)).
module asam_transport_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: ASAM_TR_MOM = 0
  integer(c_int), parameter :: ASAM_TR_POT = 1
  integer(c_int), parameter :: ASAM_TR_FALL = 2
  integer(c_int), parameter :: ASAM_TR_JACTMOM = 0
  integer(c_int), parameter :: ASAM_TR_JACTPOT = 1
  integer(c_int), parameter :: ASAM_TR_JACFALL = 2

  type, bind(c) :: asam_tr_grid
    integer(c_int) :: nx, ny, nz
    integer(c_int) :: ncomp
    integer(c_int) :: iterg
    integer(c_int) :: reserved(3)
  end type asam_tr_grid

  type, bind(c) :: asam_tr_info
    integer(c_int) :: iterations
    integer(c_int) :: status
    real(c_double) :: rnorm0
    real(c_double) :: rnorm
    integer(c_int) :: psolves
    integer(c_int) :: matvecs
  end type asam_tr_info

  interface
    integer(c_int) function asam_tr_create(ctx, grid) bind(c, name="asam_tr_create")
      import :: c_int, c_ptr, asam_tr_grid
      type(c_ptr), intent(out) :: ctx
      type(asam_tr_grid), intent(in) :: grid
    end function

    integer(c_int) function asam_tr_destroy(ctx) bind(c, name="asam_tr_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function asam_tr_last_error(ctx) bind(c, name="asam_tr_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function asam_tr_set_matrix(ctx, which, numdiag, diag, diagptr, val) &
        bind(c, name="asam_tr_set_matrix")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: which, numdiag, diag
      integer(c_int), intent(in) :: diagptr(*)
      real(c_double), intent(in) :: val(*)
    end function

    integer(c_int) function asam_tr_set_lu(ctx, lu) bind(c, name="asam_tr_set_lu")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: lu
    end function

    integer(c_int) function asam_tr_set_components(ctx, op_mv, op_ps, alpha) &
        bind(c, name="asam_tr_set_components")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: op_mv(*), op_ps(*)
      real(c_double), intent(in) :: alpha(*)
    end function

    integer(c_int) function asam_tr_matvec(ctx, x, y, init) bind(c, name="asam_tr_matvec")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
      integer(c_int), value :: init
    end function

    integer(c_int) function asam_tr_psolve(ctx, x, y) bind(c, name="asam_tr_psolve")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
    end function

    integer(c_int) function asam_tr_solve(ctx, x, b, maxiter, tol, info) &
        bind(c, name="asam_tr_solve")
      import :: c_int, c_ptr, c_double, asam_tr_info
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: x(*)
      real(c_double), intent(in) :: b(*)
      integer(c_int), intent(inout) :: maxiter
      real(c_double), intent(inout) :: tol
      type(asam_tr_info), intent(out) :: info
    end function
  end interface
end module asam_transport_iface
