!> Drop-in replacement of ASAM's transport stage solve for a Fortran host.
!>
!> Keeps the reference's call interface of the second implicit system of a stage:
!>   BiCGStabChemSp(A,x,b,MaxIter,Tol), BiCGStabMetSp(A,x,b,MaxIter,Tol)
!>     (reference SRC/MODULE/SpMatrix4CellPar_Mod.f90:140-348), operator MatVecChem
!>     (:56-69) and preconditioner PSolve (:86-102);
!> with the Jacobian of one block handed over once per step by AsamJacTransport (the
!> SpDiag matrices JacTMom / JacTPot / JacFall as JacMeteo and JacScaleCompute fill them,
!> reference SRC/ADVECT/Function_Mod.f90:1508-1939, and the per-cell LU after GefaMatrix).
!> The block's Jacobian lives in the library context, so A is not an argument here.
!> Vectors carry the components of the reference's Vector4Cell_T of one block as one
!> array c(0:nx+1, 0:ny+1, 0:nz+1, ncomp).  The arithmetic runs on the GPU.
module AsamTransport_Mod
  use iso_c_binding
  use asam_transport_iface
  implicit none
  private
  public :: Vector4Cell, AsamTransportInit, AsamJacTransport, BiCGStabChemSp, &
            BiCGStabMetSp, MatVecChem, PSolve, AsamTransportFinalize, AsamTrLastInfo, &
            AsamTrStatus

  !> The components of one block's Vector4Cell_T (Vector4Cell_Mod.f90), haloed.
  type Vector4Cell
    real(c_double), pointer, contiguous :: c(:,:,:,:) => null()
  end type Vector4Cell

  type(c_ptr), save :: ctx = c_null_ptr
  type(asam_tr_info), save :: AsamTrLastInfo
  integer, save :: AsamTrStatus = 0
  integer, save :: ncomp_ = 0, gx_ = 0, gy_ = 0
  integer(c_int), allocatable, save :: op_ps_(:), op_pot_(:)
  real(c_double), allocatable, save :: alpha_(:)

contains

  !> Create the solver context of one block (nx,ny,nz interior cells, ncomp components,
  !> IterG Gauss-Seidel sweeps after the first; reference MatSpDiag_Mod IterG = 5).
  subroutine AsamTransportInit(nx, ny, nz, ncomp, IterG)
    integer, intent(in) :: nx, ny, nz, ncomp, IterG
    type(asam_tr_grid) :: g
    if (c_associated(ctx)) call AsamTransportFinalize()
    g%nx = nx; g%ny = ny; g%nz = nz; g%ncomp = ncomp; g%iterg = IterG; g%reserved = 0
    call check(asam_tr_create(ctx, g), 'asam_tr_create')
    ncomp_ = ncomp; gx_ = nx; gy_ = ny
    allocate(op_ps_(ncomp), op_pot_(ncomp), alpha_(ncomp))
    op_ps_ = ASAM_TR_POT; op_pot_ = ASAM_TR_POT; alpha_ = 0.0d0
  end subroutine

  !> Hand over the block's transport Jacobian: Val(m,7) of JacTMom and JacTPot, optional
  !> Val(m,2) of JacFall and the per-cell LU factors (ncomp*ncomp interior-cell arrays),
  !> and per component the PSolveSp branch (ASAM_TR_MOM for uPosL..wPosR, ASAM_TR_FALL
  !> where ScaleMat2 < 0, else ASAM_TR_POT) with its ScaleMat2.
  subroutine AsamJacTransport(JacTMom, JacTPot, OpPSolve, ScaleMat2, JacFall, LU)
    real(c_double), intent(in) :: JacTMom(:,:), JacTPot(:,:)
    integer, intent(in) :: OpPSolve(:)
    real(c_double), intent(in) :: ScaleMat2(:)
    real(c_double), intent(in), optional :: JacFall(:,:)
    real(c_double), intent(in), optional, target :: LU(:,:)
    integer(c_int) :: p7(7), p2(2), nxy, nx2
    nx2 = int(gx_ + 2, c_int)
    nxy = int((gx_ + 2) * (gy_ + 2), c_int)
    p7 = [-nxy, -nx2, -1, 0, 1, nx2, nxy]
    p2 = [0, nxy]
    call check(asam_tr_set_matrix(ctx, ASAM_TR_JACTMOM, 7, 4, p7, JacTMom), 'JacTMom')
    call check(asam_tr_set_matrix(ctx, ASAM_TR_JACTPOT, 7, 4, p7, JacTPot), 'JacTPot')
    if (present(JacFall)) call check(asam_tr_set_matrix(ctx, ASAM_TR_JACFALL, 2, 1, p2, &
                                                        JacFall), 'JacFall')
    if (present(LU)) then
      call check(asam_tr_set_lu(ctx, c_loc(LU)), 'LU')
    else
      call check(asam_tr_set_lu(ctx, c_null_ptr), 'LU')
    end if
    op_ps_ = int(OpPSolve, c_int)
    alpha_ = ScaleMat2
  end subroutine

  !> BiCGStabChem_Vec4CellSp: MatVecChem (JacTPot for every component) in the Krylov loop.
  subroutine BiCGStabChemSp(x, b, MaxIter, Tol)
    type(Vector4Cell), intent(inout) :: x
    type(Vector4Cell), intent(in) :: b
    integer, intent(inout) :: MaxIter
    real(c_double), intent(inout) :: Tol
    call check(asam_tr_set_components(ctx, op_pot_, op_ps_, alpha_), 'components')
    call solve(x, b, MaxIter, Tol)
  end subroutine

  !> BiCGStabMet_Vec4CellSp: MatVecMet (each component's own operator) in the loop.
  subroutine BiCGStabMetSp(x, b, MaxIter, Tol)
    type(Vector4Cell), intent(inout) :: x
    type(Vector4Cell), intent(in) :: b
    integer, intent(inout) :: MaxIter
    real(c_double), intent(inout) :: Tol
    call check(asam_tr_set_components(ctx, op_ps_, op_ps_, alpha_), 'components')
    call solve(x, b, MaxIter, Tol)
  end subroutine

  subroutine solve(x, b, MaxIter, Tol)
    type(Vector4Cell), intent(inout) :: x
    type(Vector4Cell), intent(in) :: b
    integer, intent(inout) :: MaxIter
    real(c_double), intent(inout) :: Tol
    integer(c_int) :: mi
    mi = int(MaxIter, c_int)
    call check(asam_tr_solve(ctx, x%c, b%c, mi, Tol, AsamTrLastInfo), 'asam_tr_solve')
    MaxIter = mi
  end subroutine

  !> Ax = MatVecChem(A, x) (SpMatrix4CellPar_Mod.f90:56-69).
  subroutine MatVecChem(Ax, x)
    type(Vector4Cell), intent(inout) :: Ax
    type(Vector4Cell), intent(in) :: x
    call check(asam_tr_set_components(ctx, op_pot_, op_ps_, alpha_), 'components')
    call check(asam_tr_matvec(ctx, x%c, Ax%c, 0_c_int), 'asam_tr_matvec')
  end subroutine

  !> Px = PSolve(A, x) (PSolve_Vec4CellSp, SpMatrix4CellPar_Mod.f90:86-102).
  subroutine PSolve(Px, x)
    type(Vector4Cell), intent(inout) :: Px
    type(Vector4Cell), intent(in) :: x
    call check(asam_tr_psolve(ctx, x%c, Px%c), 'asam_tr_psolve')
  end subroutine

  subroutine AsamTransportFinalize()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = asam_tr_destroy(ctx)
    ctx = c_null_ptr
    if (allocated(op_ps_)) deallocate(op_ps_, op_pot_, alpha_)
  end subroutine

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=256) :: text
    type(c_ptr) :: p
    integer :: k
    AsamTrStatus = rc
    if (rc /= 0) then
      write(*, '(a,a,i0)') what, ' failed: ', rc
      p = asam_tr_last_error(ctx)
      if (c_associated(p)) then
        call c_f_pointer(p, msg, [256])
        text = ' '
        do k = 1, 256
          if (msg(k) == c_null_char) exit
          text(k:k) = msg(k)
        end do
        write(*, '(a)') trim(text)
      end if
    end if
  end subroutine
end module AsamTransport_Mod
