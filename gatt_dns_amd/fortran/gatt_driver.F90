! Fortran time loop over libgatt_dns: the main.F90 step sequence with the
! compute routines replaced by the C ABI (module gatt_hip).
!
! The case arrives as include files in the reference's generated format
! (parameters.F90, matrices.F90, boundaryInfo.F90) plus disturbInfo.F90; the
! initial state is read from flow_in.bin (U,V,W,R,E, raw little-endian FP64,
! column-major).  Usage: gatt_driver <nsteps> <input.bin> <output.bin>
program gatt_driver
  use iso_c_binding
  use gatt_hip
  implicit none

  include 'parameters.F90'

  integer :: nDerivBlocksX = 0, nDerivBlocksY = 0, nDerivBlocksZ = 0
  integer, dimension(:,:), allocatable :: derivBlocksX, derivBlocksY, derivBlocksZ
  integer :: derivnRHSx = 0, derivnRHSy = 0, derivnRHSz = 0, filternRHSx = 0, filternRHSy = 0, filternRHSz = 0
  integer :: ntypesX = 1, ntypesY = 1, ntypesZ = 1
  integer :: filterX = 0, filterY = 0, filterZ = 0, periodicX = 0, periodicY = 0, periodicZ = 0
  real(c_double), dimension(:,:), allocatable, target :: derivsAX, derivsBX, derivsCX, derivsDX
  real(c_double), dimension(:,:), allocatable, target :: derivsAY, derivsBY, derivsCY, derivsDY
  real(c_double), dimension(:,:), allocatable, target :: derivsAZ, derivsBZ, derivsCZ, derivsDZ
  real(c_double), dimension(:,:), allocatable, target :: filterAX, filterBX, filterCX, filterDX
  real(c_double), dimension(:,:), allocatable, target :: filterAY, filterBY, filterCY, filterDY
  real(c_double), dimension(:,:), allocatable, target :: filterAZ, filterBZ, filterCZ, filterDZ
  real(c_double), dimension(:,:,:), allocatable, target :: derivsRX, derivsRY, derivsRZ, filterRX, filterRY, filterRZ
  integer :: neumannLength = 0, neumann2Length = 0
  real(c_double), dimension(:), allocatable, target :: neumannCoeffs, neumann2Coeffs
  integer :: nUd, nVd, nWd, nPd, nEd, nUn, nVn, nWn, nPn, nEn, nUs, nVs, nWs, nPs, nEs, cN
  integer(c_int32_t), dimension(:,:), allocatable, target :: iUd, iVd, iWd, iPd, iEd, iUn, iVn, iWn, iPn, iEn
  integer(c_int32_t), dimension(:,:), allocatable, target :: iUs, iVs, iWs, iPs, iEs, cL, cD
  real(c_double), dimension(:), allocatable, target :: vUd, vVd, vWd, vPd, vEd
  integer(c_int32_t), dimension(:), allocatable, target :: dUn, dVn, dWn, dPn, dEn, dUs, dVs, dWs, dPs, dEs, cAdiabatic
  integer :: nDist = 0, nForc = 0
  integer :: distOff(64)
  type(gatt_disturbance), dimension(:), allocatable, target :: dist
  type(gatt_forcing), dimension(:), allocatable, target :: forc
  real(c_double), dimension(:), allocatable, target :: distX
  real(c_double), dimension(:), allocatable :: gbuf
  integer, dimension(:), allocatable :: ibuf

  integer(c_int32_t), dimension(:,:), allocatable, target :: typeX, typeY, typeZ
  real(c_double), dimension(:,:,:), allocatable, target :: U, V, W, R, E
  type(gatt_problem_desc), target :: desc
  type(c_ptr) :: ctx, q(5)
  real(c_double) :: t, dt, umax(3), maxChange(5), cfl, noTracked(5)
  character(kind=c_char, len=1024) :: logLine
  integer :: u_log, nLog
  integer(c_int32_t) :: diverged
  integer :: nsteps, tstep, i, rc, u_in, u_out
  character(len=512) :: arg, fin, fout

  include 'matrices.F90'
  include 'boundaryInfo.F90'
  include 'disturbInfo.F90'

  call get_command_argument(1, arg)
  read(arg, *) nsteps
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)

  ! line type maps from the derivative blocks (every line is in one block)
  allocate(typeX(ny, nz), typeY(nx, nz), typeZ(nx, ny))
  typeX = 1; typeY = 1; typeZ = 1
  do i = 1, nDerivBlocksX
    typeX(derivBlocksX(i,2):derivBlocksX(i,3), derivBlocksX(i,4):derivBlocksX(i,5)) = derivBlocksX(i,1)
  end do
  do i = 1, nDerivBlocksY
    typeY(derivBlocksY(i,2):derivBlocksY(i,3), derivBlocksY(i,4):derivBlocksY(i,5)) = derivBlocksY(i,1)
  end do
  do i = 1, nDerivBlocksZ
    typeZ(derivBlocksZ(i,2):derivBlocksZ(i,3), derivBlocksZ(i,4):derivBlocksZ(i,5)) = derivBlocksZ(i,1)
  end do

  desc%nx = nx; desc%ny = ny; desc%nz = nz
  desc%Re = Re; desc%Ma = Ma; desc%Pr = Pr; desc%gamma = gamma; desc%T0 = T0
  call set_op(desc%deriv(1), nx, ntypesX, derivnRHSx, periodicX, derivsAX, derivsBX, derivsCX, derivsDX, derivsRX)
  call set_op(desc%filter(1), nx, ntypesX, filternRHSx, periodicX, filterAX, filterBX, filterCX, filterDX, filterRX)
  call set_op(desc%deriv(2), ny, ntypesY, derivnRHSy, periodicY, derivsAY, derivsBY, derivsCY, derivsDY, derivsRY)
  call set_op(desc%filter(2), ny, ntypesY, filternRHSy, periodicY, filterAY, filterBY, filterCY, filterDY, filterRY)
  if (nz > 1) then
    call set_op(desc%deriv(3), nz, ntypesZ, derivnRHSz, periodicZ, derivsAZ, derivsBZ, derivsCZ, derivsDZ, derivsRZ)
    call set_op(desc%filter(3), nz, ntypesZ, filternRHSz, periodicZ, filterAZ, filterBZ, filterCZ, filterDZ, filterRZ)
  end if
  desc%types(1) = c_loc(typeX); desc%types(2) = c_loc(typeY); desc%types(3) = c_loc(typeZ)
  desc%filterDir = (/ filterX, filterY, filterZ /)
  desc%filterCharTime = FilterCharTime
  ! boundary lists, slots U,V,W,P,E
  call set_list(desc%bc%dir(1), nUd, iUd, vUd); call set_list(desc%bc%dir(2), nVd, iVd, vVd)
  call set_list(desc%bc%dir(3), nWd, iWd, vWd); call set_list(desc%bc%dir(4), nPd, iPd, vPd)
  call set_list(desc%bc%dir(5), nEd, iEd, vEd)
  call set_dlist(desc%bc%neu(1), nUn, iUn, dUn); call set_dlist(desc%bc%neu(2), nVn, iVn, dVn)
  call set_dlist(desc%bc%neu(3), nWn, iWn, dWn); call set_dlist(desc%bc%neu(4), nPn, iPn, dPn)
  call set_dlist(desc%bc%neu(5), nEn, iEn, dEn)
  call set_dlist(desc%bc%sec(1), nUs, iUs, dUs); call set_dlist(desc%bc%sec(2), nVs, iVs, dVs)
  call set_dlist(desc%bc%sec(3), nWs, iWs, dWs); call set_dlist(desc%bc%sec(4), nPs, iPs, dPs)
  call set_dlist(desc%bc%sec(5), nEs, iEs, dEs)
  desc%bc%nCorners = cN
  desc%bc%cL = c_loc(cL); desc%bc%cD = c_loc(cD); desc%bc%cAdiabatic = c_loc(cAdiabatic)
  desc%bc%neumannLength = neumannLength; desc%bc%neumann2Length = neumann2Length
  desc%bc%neumannCoeffs = c_loc(neumannCoeffs); desc%bc%neumann2Coeffs = c_loc(neumann2Coeffs)
  do i = 1, nDist
    dist(i)%x = c_loc(distX(distOff(i)))
  end do
  desc%bc%nDisturb = nDist
  desc%bc%disturb = c_loc(dist)
  desc%nForcing = nForc
  desc%forcing = c_loc(forc)
  desc%sfdType = SFD; desc%sfdDelta = SFD_Delta; desc%sfdXval = SFD_X_val
  desc%dxmin = dxmin
  desc%timeStepping = timeStepping

  rc = gatt_create(desc, 1_c_int32_t, ctx)
  call check(rc, 'gatt_create')

  allocate(U(nx,ny,nz), V(nx,ny,nz), W(nx,ny,nz), R(nx,ny,nz), E(nx,ny,nz))
  open(newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u_in) U, V, W, R, E
  close(u_in)
  q = (/ c_loc(U), c_loc(V), c_loc(W), c_loc(R), c_loc(E) /)
  rc = gatt_set_state(ctx, 0_c_int32_t, q)
  call check(rc, 'gatt_set_state')

  ! main.F90:112 -- boundary conditions on the initial flow
  t = 0
  rc = gatt_apply_bound_cond(ctx, t)
  call check(rc, 'gatt_apply_bound_cond')

  ! log.txt with the header of runDNS.m:58-66 and a line per step (logAll = 1, main.F90:467-506)
  noTracked = 0
  nLog = gatt_log_header(0_c_int32_t, logLine, int(len(logLine), c_int64_t))
  call check(min(nLog, 0), 'gatt_log_header')
  open(newunit=u_log, file='log.txt', access='stream', form='unformatted', status='replace')
  write(u_log) logLine(1:nLog)

  ! main.F90:192-533 with fixed dt (timeControl = 1)
  dt = dtmax
  do tstep = 1, nsteps
    rc = gatt_step(ctx, t, dt, umax, maxChange, diverged)
    call check(rc, 'gatt_step')
    t = t + dt
    ! CFL of the state the step started from (main.F90:207-214, 252)
    cfl = (1/Ma + umax(1))*dxmin(1) + (1/Ma + umax(2))*dxmin(2)
    if (nz > 1) cfl = cfl + (1/Ma + umax(3))*dxmin(3)
    cfl = dt*cfl
    nLog = gatt_log_line(0_c_int32_t, int(tstep, c_int32_t), t, dt, cfl, maxChange, 0_c_int32_t, noTracked, 0d0, &
                       & logLine, int(len(logLine), c_int64_t))
    call check(min(nLog, 0), 'gatt_log_line')
    write(u_log) logLine(1:nLog)
    if (mod(tstep, 100) == 0 .or. tstep == nsteps) then
      write(*,'(A9,I10,A13,ES14.6,A7,ES10.3)') 'Step = ', tstep, 'Time = ', t, 'dt = ', dt
      write(*,'(3(A13,ES10.4))') 'U change = ', maxChange(1), 'V change = ', maxChange(2), 'W change = ', maxChange(3)
      write(*,'(2(A13,ES10.4))') 'R change = ', maxChange(4), 'E change = ', maxChange(5)
    end if
    if (diverged /= 0) then
      write(*,*) 'Flow has diverged, stopping'
      exit
    end if
  end do

  close(u_log)
  rc = gatt_get_state(ctx, 0_c_int32_t, q)
  call check(rc, 'gatt_get_state')
  open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u_out) U, V, W, R, E
  close(u_out)
  rc = gatt_destroy(ctx)

contains

  subroutine check(code, what)
    integer, intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= 0) then
      write(*,*) what, ' failed: ', gatt_error_message()
      error stop 1
    end if
  end subroutine

  subroutine set_op(op, n, nt, nR, per, A, B, C, D, RR)
    type(gatt_lineop), intent(inout) :: op
    integer, intent(in) :: n, nt, nR, per
    real(c_double), dimension(:,:), allocatable, target, intent(in) :: A, B, C, D
    real(c_double), dimension(:,:,:), allocatable, target, intent(in) :: RR
    op%n = n; op%ntypes = nt; op%nR = nR; op%periodic = per
    op%A = c_loc(A); op%B = c_loc(B); op%C = c_loc(C); op%R = c_loc(RR)
    if (per == 1) then
      op%D = c_loc(D)
    else
      op%D = c_null_ptr
    end if
  end subroutine

  subroutine set_list(bl, n, ind, val)
    type(gatt_box_list), intent(inout) :: bl
    integer, intent(in) :: n
    integer(c_int32_t), dimension(:,:), allocatable, target, intent(in) :: ind
    real(c_double), dimension(:), allocatable, target, intent(in) :: val
    bl%count = n
    bl%box = c_loc(ind)
    bl%value = c_loc(val)
    bl%dir = c_null_ptr
  end subroutine

  subroutine set_dlist(bl, n, ind, d)
    type(gatt_box_list), intent(inout) :: bl
    integer, intent(in) :: n
    integer(c_int32_t), dimension(:,:), allocatable, target, intent(in) :: ind
    integer(c_int32_t), dimension(:), allocatable, target, intent(in) :: d
    bl%count = n
    bl%box = c_loc(ind)
    bl%value = c_null_ptr
    bl%dir = c_loc(d)
  end subroutine

end program gatt_driver
