!! The x3div benchmark driver (src/xcompact3d.f90:5-145) on libx3dhip:
!! same command line (nx ny nz p_row p_col ndt_max test_mode), the same
!! step loop through the reference's module names, and the same
!! "Elapse time min max" lines.  Fields are device arrays owned by the
!! library and mapped with c_f_pointer.
program xcompact3d
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  use variables, only : nx, ny, nz, nxm, nym, nzm, p_row, p_col, test_mode
  use x3d_operator_1d, only : x3d_operator_1d_init, x3d_operator_1d_finalize
  use x3d_derive, only : x3d_derive_init, x3d_derive_finalize
  use decomp_2d_poisson, only : decomp_2d_poisson_init, decomp_2d_poisson_finalize
  use transeq, only : calculate_transeq_rhs
  use time_integrators, only : int_time, gdt1
  use navier, only : solve_poisson, cor_vel
  implicit none

  real(mytype), dimension(:, :, :), pointer, contiguous :: ux1, uy1, uz1, dux1, duy1, duz1
  real(mytype), dimension(:, :, :), pointer, contiguous :: pp3, px1, py1, pz1
  type(c_ptr) :: dp(10)
  double precision :: tstart, tend, telapsed, tmin, tmax
  integer :: ndt, ndt_max, i

  call init_xcompact3d(ndt_max)

  telapsed = 0
  tmin = telapsed
  tmax = telapsed
  ndt = 1
  ! init_flowfield -> case::init -> mom::vel (src/x3d_tools.f90:195-203)
  call x3d_check('init_fields', x3d_init_fields(x3d_ctx_h, 0_c_int, 0_c_int64_t, 0._c_double, &
                                                dp(1), dp(2), dp(3)))
  call x3d_check('sync', x3d_sync(x3d_ctx_h))
  do while (ndt <= ndt_max)
    tstart = x3d_wtime()
    call calculate_transeq_rhs(dux1, duy1, duz1, ux1, uy1, uz1)
    call int_time(ux1, uy1, uz1, dux1, duy1, duz1)
    call solve_poisson(pp3, px1, py1, pz1, ux1, uy1, uz1)
    call cor_vel(ux1, uy1, uz1, px1, py1, pz1)
    call x3d_check('sync', x3d_sync(x3d_ctx_h))
    tend = x3d_wtime()
    telapsed = telapsed + (tend - tstart)
    tmin = telapsed
    tmax = telapsed
    print *, "Elapse time min ", tmin, " max ", tmax
    ndt = ndt + 1
  end do

  print *, "Elapsed time (min-max) [s]: ", tmin, tmax
  print *, "Timesteps completed: ", ndt
  print *, "Compute rate (min-max)[dt/s]: ", ndt / tmin, ndt / tmax

  do i = 1, 10
    call x3d_check('field_free', x3d_field_free(x3d_ctx_h, dp(i)))
  end do
  call decomp_2d_poisson_finalize()
  call x3d_derive_finalize()
  call x3d_operator_1d_finalize()
  call x3d_check('destroy', x3d_destroy(x3d_ctx_h))

contains

  !> init_xcompact3d (src/x3d_tools.f90:58-186): command line, periodic
  !> parameters, operators, derivatives, Poisson, field allocation.
  subroutine init_xcompact3d(ndt_max)
    integer, intent(inout) :: ndt_max
    integer :: nargin, arg, status, decind, flen
    character(len=80) :: inputfn
    nargin = command_argument_count()
    nx = 32; ny = 32; nz = 32
    p_row = 0; p_col = 0
    ndt_max = 5
    test_mode = .false.
    do arg = 1, nargin
      call get_command_argument(arg, inputfn, flen, status)
      read (inputfn, *, iostat=status) decind
      select case (arg)
      case (1); nx = decind
      case (2); ny = decind
      case (3); nz = decind
      case (4); p_row = decind
      case (5); p_col = decind
      case (6); ndt_max = decind
      case (7)
        test_mode = decind /= 0
        write (*, *) 'Test mode ', test_mode
      case default
        print *, "Error: Too many arguments!"
        print *, "  x3div accepts"
        print *, "  1) nx (default=16)"
        print *, "  2) ny (default=16)"
        print *, "  3) nz (default=16)"
        print *, "  4) p_row (default=0)"
        print *, "  5) p_col (default=0)"
        print *, "  6) ndt_max (default=5)"
        print *, "  7) test_mode logical 0/1 (default=0)"
      end select
    end do
    ! parameter() with periodic defaults (src/parameters.f90:24-136): nxm = nx,
    ! Lx = Ly = Lz = 1, dt = 0 so gdt(1) = 0 (src/var.f90:183-186).
    nxm = nx; nym = ny; nzm = nz
    gdt1 = 0._mytype
    if (p_row * p_col > 1) print *, "p_row x p_col = ", p_row * p_col, " ranks requested; running on one GPU"
    call x3d_check('x3d_create', x3d_create(int(nx, c_int), int(ny, c_int), int(nz, c_int), &
                                            1._c_double, 1._c_double, 1._c_double, x3d_ctx_h))
    call x3d_operator_1d_init()
    call x3d_derive_init()
    call decomp_2d_poisson_init(nx, ny, nz)
    do i = 1, 10
      call x3d_check('field_alloc', x3d_field_alloc(x3d_ctx_h, int(8, c_size_t) * nx * ny * nz, dp(i)))
    end do
    call c_f_pointer(dp(1), ux1, [nx, ny, nz]);  call c_f_pointer(dp(2), uy1, [nx, ny, nz])
    call c_f_pointer(dp(3), uz1, [nx, ny, nz]);  call c_f_pointer(dp(4), dux1, [nx, ny, nz])
    call c_f_pointer(dp(5), duy1, [nx, ny, nz]); call c_f_pointer(dp(6), duz1, [nx, ny, nz])
    call c_f_pointer(dp(7), pp3, [nx, ny, nz]);  call c_f_pointer(dp(8), px1, [nx, ny, nz])
    call c_f_pointer(dp(9), py1, [nx, ny, nz]);  call c_f_pointer(dp(10), pz1, [nx, ny, nz])
  end subroutine init_xcompact3d

end program xcompact3d
