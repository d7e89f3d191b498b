!> Fortran host driver: reads one stage system (geometry, state, b) from a stream
!> file, sets the operator up with AsamJacAccGrav, calls SolveSound exactly as
!> ProjectVelFace does (x = 0, MaxIter, Tol; reference SRC/ADVECT/IntSub_Mod.f90:46-56),
!> checks the residual with MatVec1 and writes x, MaxIter and Tol back.
!> Usage: drv_solve_sound <input file> <output file>
program drv_solve_sound
  use iso_c_binding
  use AsamSound_Mod
  implicit none
  integer :: nx, ny, nz, np1, gradfull, method, mixed, maxiter, u, ios
  real(c_double) :: beta0dt, grav, tol
  real(c_double), allocatable :: FU(:,:,:), FV(:,:,:), FW(:,:,:), VolC(:,:,:)
  real(c_double), allocatable :: dx(:), dy(:), dz(:), rho(:,:,:), rth(:,:,:), T(:,:,:)
  type(PressureVelocity) :: x(1), b(1), r(1)
  character(len=512) :: fin, fout
  character(len=8) :: meth
  real(c_double) :: rn, bn

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', &
       iostat=ios)
  if (ios /= 0) stop 'cannot open input'
  read(u) nx, ny, nz, np1, gradfull, method, mixed, maxiter
  read(u) beta0dt, grav, tol
  allocate(FU(0:nx, ny, nz), FV(nx, 0:ny, nz), FW(nx, ny, 0:nz), VolC(nx, ny, nz))
  allocate(dx(nx), dy(ny), dz(nz), rho(nx, ny, nz), rth(nx, ny, nz), T(nx, ny, nz))
  allocate(x(1)%p(np1, nx, ny, nz), b(1)%p(np1, nx, ny, nz), r(1)%p(np1, nx, ny, nz))
  read(u) FU, FV, FW, VolC, dx, dy, dz, rho, rth, T, b(1)%p
  close(u)

  meth = 'BICGSTAB'
  if (method == 1) meth = 'FGMRES'
  if (method == 2) meth = 'GMRES'
  if (method == 3) meth = 'FGMRESV'
  call AsamSoundInit(nx, ny, nz, np1, gradfull == 1, beta0dt, grav, 1.0d0, meth, 10, &
                     mixed == 1)
  call AsamJacAccGrav(FU, FV, FW, VolC, dx, dy, dz, rho, rth, T)
  x(1)%p = 0.0d0
  call SolveSound(x, b, maxiter, tol)
  if (AsamStatus /= 0) stop 1
  call MatVec1(r, x)
  r(1)%p = b(1)%p - r(1)%p
  rn = sqrt(sum(r(1)%p**2))
  bn = sqrt(sum(b(1)%p**2))
  write(*, '(a,i0,a,es12.4,a,i0,a,es12.4,a,f9.3,a)') 'SolveSound: MaxIter=', maxiter, &
      ' Tol=', tol, ' iterations=', AsamLastInfo%iterations, ' |b-MatVec1(x)|/|b|=', &
      rn / bn, ' (', AsamLastInfo%ms, ' ms)'
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) maxiter, tol, rn / bn, AsamLastInfo%iterations
  write(u) x(1)%p
  close(u)
  call AsamSoundFinalize()
end program drv_solve_sound
