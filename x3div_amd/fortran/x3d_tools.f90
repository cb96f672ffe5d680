!! Start-up and shut-down of the x3div driver (src/x3d_tools.f90:45-260), as
!! external subroutines like the reference's: the command line, parameter()
!! and listing(), the one-rank decomposition, the library context on the
!! current GPU, the step's device arrays, the operators and Poisson.

!> boot_xcompact3d (src/x3d_tools.f90:45-63): MPI_INIT there; one rank here.
subroutine boot_xcompact3d()
  use decomp_2d_mpi, only : nrank, nproc
  implicit none
  nrank = 0
  nproc = 1
end subroutine boot_xcompact3d

!> init_xcompact3d (src/x3d_tools.f90:81-186).
subroutine init_xcompact3d(ndt_max)
  use, intrinsic :: iso_c_binding
  use decomp_2d, only : decomp_2d_init, decomp_info_init, ph1, ph2, ph3, phG
  use decomp_2d_poisson, only : decomp_2d_poisson_init
  use x3d_operator_1d, only : x3d_operator_1d_init
  use x3d_derive, only : x3d_derive_init
  use x3d_hip, only : x3d_ctx_h, x3d_check, x3d_create_bc, x3d_set_stretching
  use parameters, only : parameter, listing
  use param, only : xlx, yly, zlz, nclx1, nclxn, ncly1, nclyn, nclz1, nclzn, istret
  use var, only : var_init
  use variables, only : nx, ny, nz, nxm, nym, nzm, p_row, p_col, test_mode, ppy, ppyi
  implicit none
  integer, intent(inout) :: ndt_max
  integer :: nargin, arg, flen, status, decind
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

  call parameter()
  call listing()

  call decomp_2d_init(nx, ny, nz, p_row, p_col)
  call decomp_info_init(nxm, nym, nzm, ph1)
  call decomp_info_init(nxm, ny, nz, ph2)
  call decomp_info_init(nxm, nym, nz, ph3)

  ! the library context: operator tables, line plans, FFT plans on this GPU
  ! with parameter()'s boundary codes (all periodic in this driver)
  call x3d_check('x3d_create', x3d_create_bc(int(nx, c_int), int(ny, c_int), int(nz, c_int), &
                                             real(xlx, c_double), real(yly, c_double), real(zlz, c_double), &
                                             int([nclx1, ncly1, nclz1], c_int), int([nclxn, nclyn, nclzn], c_int), &
                                             x3d_ctx_h))
  call var_init()
  ! a stretched y grid: the step's y-derivatives take ppy / ppyi (istret = 0
  ! in this driver, src/parameters.f90:117)
  if (istret /= 0) call x3d_check('x3d_set_stretching', x3d_set_stretching(x3d_ctx_h, int(istret, c_int), &
                                                                            c_loc(ppy), c_loc(ppyi)))
  call x3d_operator_1d_init()
  call x3d_derive_init()

  call decomp_2d_poisson_init()
  call decomp_info_init(nxm, nym, nzm, phG)
end subroutine init_xcompact3d

!> init_flowfield (src/x3d_tools.f90:189-203).
subroutine init_flowfield()
  use case, only : init
  use var, only : ux1, uy1, uz1, dux1, duy1, duz1, pp3, px1, py1, pz1
  use param, only : itime
  implicit none
  call init(ux1, uy1, uz1, dux1, duy1, duz1, pp3, px1, py1, pz1)
  itime = 0
end subroutine init_flowfield

!> finalise_xcompact3d (src/x3d_tools.f90:206-246).
subroutine finalise_xcompact3d(flag)
  use decomp_2d, only : decomp_2d_finalize, decomp_info_finalize, ph1, ph2, ph3, phG
  use decomp_2d_poisson, only : decomp_2d_poisson_finalize
  use x3d_operator_1d, only : x3d_operator_1d_finalize
  use x3d_derive, only : x3d_derive_finalize
  use x3d_hip, only : x3d_ctx_h, x3d_check, x3d_destroy
  use var, only : var_finalize
  use, intrinsic :: iso_c_binding, only : c_null_ptr
  implicit none
  logical, intent(in) :: flag
  call decomp_info_finalize(ph1)
  call decomp_info_finalize(ph2)
  call decomp_info_finalize(ph3)
  call decomp_info_finalize(phG)
  call decomp_2d_poisson_finalize()
  call x3d_derive_finalize()
  call x3d_operator_1d_finalize()
  call var_finalize()
  call decomp_2d_finalize()
  call x3d_check('x3d_destroy', x3d_destroy(x3d_ctx_h))
  x3d_ctx_h = c_null_ptr
  if (flag) continue  ! MPI_FINALIZE at the reference; one rank here
end subroutine finalise_xcompact3d
