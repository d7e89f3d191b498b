!> Fortran host driver of the transport stage solve: reads one block's transport system
!> (SpDiag JacTMom / JacTPot / JacFall values, per-cell LU, component operators, b) from a
!> stream file, hands the Jacobian over with AsamJacTransport and calls BiCGStabChemSp or
!> BiCGStabMetSp from x = 0 as the stage does (reference SRC/ADVECT/Int_Mod.f90:314-327),
!> then writes MaxIter, Tol and x.
!> Usage: drv_transport <input file> <output file>
program drv_transport
  use iso_c_binding
  use AsamTransport_Mod
  implicit none
  integer :: nx, ny, nz, nc, iterg, met, hasfall, haslu, maxiter, u, ios, m, ncell
  real(c_double) :: tol
  real(c_double), allocatable :: tmom(:,:), tpot(:,:), fall(:,:), lu(:,:), alpha(:)
  integer, allocatable :: opps(:)
  type(Vector4Cell) :: x, b
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', &
       iostat=ios)
  if (ios /= 0) stop 'cannot open input'
  read(u) nx, ny, nz, nc, iterg, met, hasfall, haslu, maxiter
  read(u) tol
  m = (nx + 2) * (ny + 2) * (nz + 2)
  ncell = nx * ny * nz
  allocate(tmom(m, 7), tpot(m, 7), opps(nc), alpha(nc))
  allocate(x%c(0:nx+1, 0:ny+1, 0:nz+1, nc), b%c(0:nx+1, 0:ny+1, 0:nz+1, nc))
  read(u) tmom, tpot
  if (hasfall == 1) then
    allocate(fall(m, 2))
    read(u) fall
  end if
  if (haslu == 1) then
    allocate(lu(ncell, nc * nc))
    read(u) lu
  end if
  read(u) opps, alpha, b%c
  close(u)

  call AsamTransportInit(nx, ny, nz, nc, iterg)
  if (hasfall == 1 .and. haslu == 1) then
    call AsamJacTransport(tmom, tpot, opps, alpha, JacFall=fall, LU=lu)
  else if (hasfall == 1) then
    call AsamJacTransport(tmom, tpot, opps, alpha, JacFall=fall)
  else if (haslu == 1) then
    call AsamJacTransport(tmom, tpot, opps, alpha, LU=lu)
  else
    call AsamJacTransport(tmom, tpot, opps, alpha)
  end if
  if (AsamTrStatus /= 0) stop 1
  x%c = 0.0d0
  if (met == 1) then
    call BiCGStabMetSp(x, b, maxiter, tol)
  else
    call BiCGStabChemSp(x, b, maxiter, tol)
  end if
  if (AsamTrStatus /= 0) stop 1
  write(*, '(a,i0,a,es12.4,a,i0,a,i0)') 'BiCGStab: MaxIter=', maxiter, ' Tol=', tol, &
      ' iterations=', AsamTrLastInfo%iterations, ' psolves=', AsamTrLastInfo%psolves
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) maxiter, tol, AsamTrLastInfo%iterations
  write(u) x%c
  close(u)
  call AsamTransportFinalize()
end program drv_transport
