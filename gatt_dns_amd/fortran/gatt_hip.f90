! gatt_hip -- ISO_C_BINDING interface of libgatt_dns (include/gatt_dns.h).
!
! A Fortran time loop replaces its calls to RK4_3D/RK4_2D, applyFilters3D/2D,
! applyBoundCond and the SFD block of main.F90 with the functions below; the
! derived types mirror the C descriptor field by field (bind(C) layout).
module gatt_hip
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: GATT_DERIV = 0, GATT_FILTER = 1
  integer(c_int32_t), parameter :: GATT_DIST_LID_DRIVEN = 1, GATT_DIST_HOLD_INLET = 2
  integer(c_int32_t), parameter :: GATT_FORCING_TURBULENCE = 1

  type, bind(C) :: gatt_lineop
    integer(c_int32_t) :: n = 0, ntypes = 0, nR = 0, periodic = 0
    type(c_ptr) :: A = c_null_ptr, B = c_null_ptr, C = c_null_ptr, D = c_null_ptr, R = c_null_ptr
  end type

  type, bind(C) :: gatt_box_list
    integer(c_int32_t) :: count = 0
    type(c_ptr) :: box = c_null_ptr, value = c_null_ptr, dir = c_null_ptr
  end type

  type, bind(C) :: gatt_disturbance
    integer(c_int32_t) :: kind = 0
    integer(c_int32_t) :: box(6) = 0
    type(c_ptr) :: x = c_null_ptr
    real(c_double) :: par(4) = 0
    integer(c_int32_t) :: var = 0
  end type

  type, bind(C) :: gatt_forcing
    integer(c_int32_t) :: kind = 0
    integer(c_int32_t) :: box(6) = 0
    real(c_double) :: par(4) = 0
    integer(c_int32_t) :: var(2) = 0
  end type

  type, bind(C) :: gatt_bc_desc
    type(gatt_box_list) :: dir(5), neu(5), sec(5)
    integer(c_int32_t) :: nCorners = 0
    type(c_ptr) :: cL = c_null_ptr, cD = c_null_ptr, cAdiabatic = c_null_ptr
    integer(c_int32_t) :: neumannLength = 0, neumann2Length = 0
    type(c_ptr) :: neumannCoeffs = c_null_ptr, neumann2Coeffs = c_null_ptr
    integer(c_int32_t) :: nDisturb = 0
    type(c_ptr) :: disturb = c_null_ptr
  end type

  type, bind(C) :: gatt_problem_desc
    integer(c_int32_t) :: nx = 0, ny = 0, nz = 0
    real(c_double) :: Re = 0, Ma = 0, Pr = 0, gamma = 0, T0 = 0
    type(gatt_lineop) :: deriv(3), filter(3)
    type(c_ptr) :: types(3) = c_null_ptr
    integer(c_int32_t) :: filterDir(3) = 0
    real(c_double) :: filterCharTime = -1
    type(gatt_bc_desc) :: bc
    integer(c_int32_t) :: nForcing = 0
    type(c_ptr) :: forcing = c_null_ptr
    integer(c_int32_t) :: sfdType = 0
    real(c_double) :: sfdDelta = 0, sfdXval = 0
    type(c_ptr) :: sfdX = c_null_ptr
    real(c_double) :: dxmin(3) = 0
    integer(c_int32_t) :: timeStepping = 1
  end type

  interface
    function gatt_last_error() bind(C, name="gatt_last_error")
      import :: c_ptr
      type(c_ptr) :: gatt_last_error
    end function
    integer(c_int) function gatt_create(desc, batch, ctx) bind(C, name="gatt_create")
      import :: gatt_problem_desc, c_int32_t, c_ptr, c_int
      type(gatt_problem_desc), intent(in) :: desc
      integer(c_int32_t), value :: batch
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function gatt_destroy(ctx) bind(C, name="gatt_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function gatt_set_state(ctx, member, q) bind(C, name="gatt_set_state")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: member
      type(c_ptr), intent(in) :: q(5)
    end function
    integer(c_int) function gatt_get_state(ctx, member, q) bind(C, name="gatt_get_state")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: member
      type(c_ptr), intent(in) :: q(5)
    end function
    integer(c_int) function gatt_rk4(ctx, t, dt) bind(C, name="gatt_rk4")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t, dt
    end function
    integer(c_int) function gatt_apply_filters(ctx, theta) bind(C, name="gatt_apply_filters")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: theta
    end function
    integer(c_int) function gatt_apply_bound_cond(ctx, t) bind(C, name="gatt_apply_bound_cond")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t
    end function
    ! flow files flow_%010d.mat (include/gatt_flow_io.h, libgatt_flowio.so): q holds
    ! c_loc(U), ..., c_loc(E); inside_wall is c_null_ptr or an integer(c_int8_t) mask
    integer(c_int) function gatt_flow_write(path, nx, ny, nz, t, q, inside_wall) bind(C, name="gatt_flow_write")
      import :: c_ptr, c_int, c_double, c_int32_t, c_char
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int32_t), value :: nx, ny, nz
      real(c_double), value :: t
      type(c_ptr), dimension(5), intent(in) :: q
      type(c_ptr), value :: inside_wall
    end function
    integer(c_int) function gatt_flow_read(path, nx, ny, nz, t, q, inside_wall) bind(C, name="gatt_flow_read")
      import :: c_ptr, c_int, c_double, c_int32_t, c_char
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int32_t), value :: nx, ny, nz
      real(c_double), intent(out) :: t
      type(c_ptr), dimension(5), intent(in) :: q
      type(c_ptr), value :: inside_wall
    end function
    integer(c_int) function gatt_flow_write_deriv(path, nx, ny, nz, t, q, qdot, inside_wall) &
        & bind(C, name="gatt_flow_write_deriv")
      import :: c_ptr, c_int, c_double, c_int32_t, c_char
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int32_t), value :: nx, ny, nz
      real(c_double), value :: t
      type(c_ptr), dimension(5), intent(in) :: q, qdot
      type(c_ptr), value :: inside_wall
    end function
    integer(c_int) function gatt_meanflow_write(path, nx, ny, nz, q, inside_wall) bind(C, name="gatt_meanflow_write")
      import :: c_ptr, c_int, c_int32_t, c_char
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int32_t), value :: nx, ny, nz
      type(c_ptr), dimension(5), intent(in) :: q
      type(c_ptr), value :: inside_wall
    end function
    integer(c_int) function gatt_meanflow_read(path, nx, ny, nz, q, inside_wall) bind(C, name="gatt_meanflow_read")
      import :: c_ptr, c_int, c_int32_t, c_char
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int32_t), value :: nx, ny, nz
      type(c_ptr), dimension(5), intent(in) :: q
      type(c_ptr), value :: inside_wall
    end function
    integer(c_int) function gatt_sfd_read(path, nx, ny, nz, sfd_x) bind(C, name="gatt_sfd_read")
      import :: c_int, c_int32_t, c_char, c_double
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int32_t), value :: nx, ny, nz
      real(c_double), dimension(*), intent(out) :: sfd_x
    end function
    integer(c_int) function gatt_sfd_write(path, nx, ny, nz, sfd_x) bind(C, name="gatt_sfd_write")
      import :: c_int, c_int32_t, c_char, c_double
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int32_t), value :: nx, ny, nz
      real(c_double), dimension(*), intent(in) :: sfd_x
    end function
    integer(c_int) function gatt_log_header(ntracked, buf, size) bind(C, name="gatt_log_header")
      import :: c_int, c_int32_t, c_int64_t, c_char
      integer(c_int32_t), value :: ntracked
      character(kind=c_char), dimension(*), intent(out) :: buf
      integer(c_int64_t), value :: size
    end function
    integer(c_int) function gatt_log_line(nsave, tstep, t, dt, cfl, max_change, ntracked, tracked, tracked_norm, &
        & buf, size) bind(C, name="gatt_log_line")
      import :: c_int, c_int32_t, c_int64_t, c_char, c_double
      integer(c_int32_t), value :: nsave, tstep, ntracked
      real(c_double), value :: t, dt, cfl, tracked_norm
      real(c_double), dimension(5), intent(in) :: max_change
      real(c_double), dimension(*), intent(in) :: tracked
      character(kind=c_char), dimension(*), intent(out) :: buf
      integer(c_int64_t), value :: size
    end function
    integer(c_int) function gatt_sfd(ctx, dt) bind(C, name="gatt_sfd")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function
    integer(c_int) function gatt_step(ctx, t, dt, umax, maxchange, diverged) bind(C, name="gatt_step")
      import :: c_ptr, c_int, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), value :: t, dt
      real(c_double), intent(out) :: umax(3), maxchange(5)
      integer(c_int32_t), intent(out) :: diverged
    end function
    integer(c_int) function gatt_advance(ctx, nsteps, t0, dt, maxchange, diverged) bind(C, name="gatt_advance")
      import :: c_ptr, c_int, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
      real(c_double), value :: t0, dt
      real(c_double), intent(out) :: maxchange(5)
      integer(c_int32_t), intent(out) :: diverged
    end function
  end interface

contains

  ! Message of the last failed call, as a Fortran string.
  function gatt_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: n
    call c_f_pointer(gatt_last_error(), p, [1024])
    n = 0
    do while (n < 1024)
      if (p(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function

end module gatt_hip
