!> Drop-in replacement of ASAM's sound-solve entry points for a Fortran host.
!>
!> Keeps the reference's call interface for the stage solve:
!>   SolveSound(x,b,MaxIter,Tol)   (reference SRC/SOLU/Iter_Mod.f90:9-30)
!>   MatVec1(b,x), PSolve1(b,x)    (reference SRC/MATRIX/MatVec_Mod.f90:8, PreCond_Mod.f90:1555)
!> and the once-per-step operator set-up JacAccGrav (SRC/MULTI/JacAccGrav_Mod.f90:304) as
!> AsamJacAccGrav, with x and b carrying the reference's PressureVelocity%p component,
!> p(1:np1,nx,ny,nz).  The arithmetic runs on the GPU through libasam_sound.
module AsamSound_Mod
  use iso_c_binding
  use asam_sound_iface
  implicit none
  private
  public :: PressureVelocity, AsamSoundInit, AsamJacAccGrav, SolveSound, SolveSoundT, &
            MatVec1, PSolve1, Rhs, UpdateVelocity, AsamSoundFinalize, AsamLastInfo, &
            AsamStatus

  !> The p component of the reference's TYPE(PressureVelocity) (VecPrVel_Mod.f90:11-18);
  !> one closed block, so the interface velocities are absent.
  type PressureVelocity
    real(c_double), pointer, contiguous :: p(:,:,:,:) => null()
  end type PressureVelocity

  type(c_ptr), save :: ctx = c_null_ptr
  type(asam_sound_info), save :: AsamLastInfo
  integer, save :: AsamStatus = 0

contains

  !> Create the solver context for one block (nx,ny,nz interior cells).
  subroutine AsamSoundInit(nx, ny, nz, np1, GradFull, beta0dtP, GravComp, FacAnela, &
                           MethSound, StepGMRES, Mixed)
    integer, intent(in) :: nx, ny, nz, np1
    logical, intent(in) :: GradFull
    real(c_double), intent(in) :: beta0dtP, GravComp, FacAnela
    character(len=*), intent(in) :: MethSound
    integer, intent(in) :: StepGMRES
    logical, intent(in) :: Mixed
    type(asam_sound_grid) :: g
    type(asam_sound_control) :: ctl
    integer(c_int) :: rc
    logical :: missing
    if (c_associated(ctx)) call AsamSoundFinalize()
    g%nx = nx; g%ny = ny; g%nz = nz; g%np1 = np1
    g%grad_full = merge(1, 0, GradFull)
    g%beta0_dt = beta0dtP; g%grav = GravComp; g%fac_anela = FacAnela
    rc = asam_sound_create(ctx, g)
    call check(rc, 'asam_sound_create')
    rc = asam_sound_default_control(ctl)
    ! Iter_Mod.f90:14-29; BICGSTAB is this library's addition.  QMR, MINRES and CGR are
    ! not provided: AsamStatus reports ASAM_E_INVALID and the default method stays.
    missing = .false.
    select case (trim(MethSound))
    case ('FGMRES')
      ctl%method = ASAM_METHOD_FGMRES
    case ('FGMRESV')
      ctl%method = ASAM_METHOD_FGMRESV
    case ('BICGSTAB')
      ctl%method = ASAM_METHOD_BICGSTAB
    case ('QMR', 'MINRES', 'CGR')
      write(0, '(a,a)') 'AsamSoundInit: MethSound not provided: ', trim(MethSound)
      missing = .true.
    case default
      ctl%method = ASAM_METHOD_GMRES
    end select
    ctl%step_gmres = StepGMRES
    if (Mixed) ctl%precision = ASAM_PREC_MIXED
    rc = asam_sound_set_control(ctx, ctl)
    call check(rc, 'asam_sound_set_control')
    if (missing) AsamStatus = ASAM_E_INVALID
  end subroutine AsamSoundInit

  !> Operator set-up from geometry and thermodynamic state (JacAccGrav).  RhoV, RhoL and
  !> RhoR (vapour, cloud water, rain densities of the interior cells) make the weights those
  !> of moist air, as WeightDiagCompute 'Density' forms them; absent ones count as zero.
  subroutine AsamJacAccGrav(FU, FV, FW, VolC, dx, dy, dz, Rho, RhoTheta, T, RhoV, RhoL, RhoR)
    real(c_double), intent(in) :: FU(:,:,:), FV(:,:,:), FW(:,:,:), VolC(:,:,:)
    real(c_double), intent(in) :: dx(:), dy(:), dz(:)
    real(c_double), intent(in) :: Rho(:,:,:), RhoTheta(:,:,:), T(:,:,:)
    real(c_double), intent(in), optional, target, contiguous :: RhoV(:,:,:), RhoL(:,:,:), &
                                                                 RhoR(:,:,:)
    integer(c_int) :: rc
    type(c_ptr) :: pv, pl, pr
    if (.not. (present(RhoV) .or. present(RhoL) .or. present(RhoR))) then
      rc = asam_sound_setup(ctx, FU, FV, FW, VolC, dx, dy, dz, Rho, RhoTheta, T)
      call check(rc, 'asam_sound_setup')
      return
    end if
    pv = c_null_ptr
    pl = c_null_ptr
    pr = c_null_ptr
    if (present(RhoV)) pv = c_loc(RhoV)
    if (present(RhoL)) pl = c_loc(RhoL)
    if (present(RhoR)) pr = c_loc(RhoR)
    rc = asam_sound_setup_moist(ctx, FU, FV, FW, VolC, dx, dy, dz, Rho, RhoTheta, T, pv, pl, pr)
    call check(rc, 'asam_sound_setup_moist')
  end subroutine AsamJacAccGrav

  !> SolveSound(x,b,MaxIter,Tol): solves MatVec1(x) = -A x = b (Iter_Mod.f90:9-30).
  subroutine SolveSound(x, b, MaxIter, Tol)
    type(PressureVelocity), dimension(:) :: b, x
    integer :: MaxIter
    real(c_double) :: Tol
    integer(c_int) :: rc, mi
    mi = MaxIter
    rc = asam_sound_solve(ctx, x(1)%p, b(1)%p, mi, Tol, AsamLastInfo)
    call check(rc, 'asam_sound_solve')
    MaxIter = mi
  end subroutine SolveSound

  !> Adjoint of SolveSound: solves (-A)^T x = b with the transposed preconditioner.
  subroutine SolveSoundT(x, b, MaxIter, Tol)
    type(PressureVelocity), dimension(:) :: b, x
    integer :: MaxIter
    real(c_double) :: Tol
    integer(c_int) :: rc, mi
    mi = MaxIter
    rc = asam_sound_solve_T(ctx, x(1)%p, b(1)%p, mi, Tol, AsamLastInfo)
    call check(rc, 'asam_sound_solve_T')
    MaxIter = mi
  end subroutine SolveSoundT

  !> b = MatVec1(x) = -A x (MatVec_Mod.f90:8-18).
  subroutine MatVec1(b, x)
    type(PressureVelocity), dimension(:) :: b, x
    integer(c_int) :: rc
    rc = asam_sound_apply(ctx, b(1)%p, x(1)%p)
    call check(rc, 'asam_sound_apply')
  end subroutine MatVec1

  !> b = PSolve1(x) = -M^-1 x, M the vertical-line block of A (PreCond_Mod.f90:1555).
  subroutine PSolve1(b, x)
    type(PressureVelocity), dimension(:) :: b, x
    integer(c_int) :: rc
    rc = asam_sound_psolve(ctx, b(1)%p, x(1)%p)
    call check(rc, 'asam_sound_psolve')
  end subroutine PSolve1

  !> b = Rhs of the stage system (Rhs_Mod.f90:110-157): U, V, W are the predicted face
  !> velocities Vel%uF(0:nx,ny,nz), %vF(nx,0:ny,nz), %wF(nx,ny,0:nz); IncrRhoTheta and
  !> IncrRho the stage increments of rho*theta and rho (interior cells).
  subroutine Rhs(b, U, V, W, IncrRhoTheta, IncrRho)
    type(PressureVelocity), dimension(:) :: b
    real(c_double), intent(in) :: U(:,:,:), V(:,:,:), W(:,:,:), IncrRhoTheta(:,:,:)
    real(c_double), intent(in), optional, target, contiguous :: IncrRho(:,:,:)
    type(c_ptr) :: ir
    integer(c_int) :: rc
    ir = c_null_ptr
    if (present(IncrRho)) ir = c_loc(IncrRho)
    rc = asam_sound_rhs(ctx, b(1)%p, U, V, W, IncrRhoTheta, ir)
    call check(rc, 'asam_sound_rhs')
  end subroutine Rhs

  !> Post-solve update (UpdateVelocity, Rhs_Mod.f90:589-952): x(1)%p(1,...) becomes the
  !> pressure increment DUT*x1, the face velocities get the pressure-gradient (and, with
  !> two unknowns, buoyancy) correction, and the increments the corrected flux divergence.
  subroutine UpdateVelocity(x, U, V, W, IncrRhoTheta, IncrRho)
    type(PressureVelocity), dimension(:) :: x
    real(c_double), intent(inout) :: U(:,:,:), V(:,:,:), W(:,:,:), IncrRhoTheta(:,:,:)
    real(c_double), intent(inout), optional, target, contiguous :: IncrRho(:,:,:)
    type(c_ptr) :: ir
    integer(c_int) :: rc
    ir = c_null_ptr
    if (present(IncrRho)) ir = c_loc(IncrRho)
    rc = asam_sound_update(ctx, x(1)%p, U, V, W, IncrRhoTheta, ir)
    call check(rc, 'asam_sound_update')
  end subroutine UpdateVelocity

  subroutine AsamSoundFinalize()
    integer(c_int) :: rc
    if (c_associated(ctx)) then
      rc = asam_sound_destroy(ctx)
      ctx = c_null_ptr
    end if
  end subroutine AsamSoundFinalize

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    AsamStatus = rc
    if (rc /= 0) then
      write(0, '(a,a,i0,a,a)') what, ' failed: ', rc, ' ', trim(asam_error_message(ctx))
    end if
  end subroutine check

end module AsamSound_Mod
