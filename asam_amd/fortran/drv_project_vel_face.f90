!> Fortran host driver of one stage projection as ProjectVelFace performs it (reference
!> SRC/ADVECT/IntSub_Mod.f90:25-71): Rhs builds b from the predicted face velocities and
!> the stage increments, SolveSound solves with x = 0, UpdateVelocity corrects the face
!> velocities and the increments.  Input: the drv_solve_sound stream file followed by
!> U(0:nx,ny,nz), V(nx,0:ny,nz), W(nx,ny,0:nz), IncrRhoTheta(nx,ny,nz), IncrRho(nx,ny,nz)
!> (b in that file is ignored).  Output: maxiter, tol, then x, U, V, W, IncrRhoTheta,
!> IncrRho after the update.
!> Usage: drv_project_vel_face <input file> <output file>
program drv_project_vel_face
  use iso_c_binding
  use AsamSound_Mod
  implicit none
  integer :: nx, ny, nz, np1, gradfull, method, mixed, maxiter, u, ios
  real(c_double) :: beta0dt, grav, tol
  real(c_double), allocatable :: FU(:,:,:), FV(:,:,:), FW(:,:,:), VolC(:,:,:)
  real(c_double), allocatable :: dx(:), dy(:), dz(:), rho(:,:,:), rth(:,:,:), T(:,:,:)
  real(c_double), allocatable :: Uf(:,:,:), Vf(:,:,:), Wf(:,:,:), IncrTh(:,:,:), IncrRho(:,:,:)
  type(PressureVelocity) :: x(1), b(1)
  character(len=512) :: fin, fout
  character(len=8) :: meth

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', &
       iostat=ios)
  if (ios /= 0) stop 'cannot open input'
  read(u) nx, ny, nz, np1, gradfull, method, mixed, maxiter
  read(u) beta0dt, grav, tol
  allocate(FU(0:nx, ny, nz), FV(nx, 0:ny, nz), FW(nx, ny, 0:nz), VolC(nx, ny, nz))
  allocate(dx(nx), dy(ny), dz(nz), rho(nx, ny, nz), rth(nx, ny, nz), T(nx, ny, nz))
  allocate(x(1)%p(np1, nx, ny, nz), b(1)%p(np1, nx, ny, nz))
  allocate(Uf(0:nx, ny, nz), Vf(nx, 0:ny, nz), Wf(nx, ny, 0:nz))
  allocate(IncrTh(nx, ny, nz), IncrRho(nx, ny, nz))
  read(u) FU, FV, FW, VolC, dx, dy, dz, rho, rth, T, b(1)%p
  read(u) Uf, Vf, Wf, IncrTh, IncrRho
  close(u)

  meth = 'BICGSTAB'
  if (method == 1) meth = 'FGMRES'
  if (method == 2) meth = 'GMRES'
  if (method == 3) meth = 'FGMRESV'
  call AsamSoundInit(nx, ny, nz, np1, gradfull == 1, beta0dt, grav, 1.0d0, meth, 10, &
                     mixed == 1)
  call AsamJacAccGrav(FU, FV, FW, VolC, dx, dy, dz, rho, rth, T)
  b(1)%p = 0.0d0
  x(1)%p = 0.0d0
  call Rhs(b, Uf, Vf, Wf, IncrTh, IncrRho)
  call SolveSound(x, b, maxiter, tol)
  if (AsamStatus /= 0) stop 1
  call UpdateVelocity(x, Uf, Vf, Wf, IncrTh, IncrRho)
  write(*, '(a,i0,a,es12.4,a,i0)') 'ProjectVelFace: MaxIter=', maxiter, ' Tol=', tol, &
      ' iterations=', AsamLastInfo%iterations
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) maxiter, tol
  write(u) x(1)%p, Uf, Vf, Wf, IncrTh, IncrRho
  close(u)
  call AsamSoundFinalize()
end program drv_project_vel_face
