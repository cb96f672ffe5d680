!! One-rank forms of the 2decomp-fft modules and the x3div parameter modules
!! that the reference's callers use (decomp_2d_mpi, decomp_2d, variables,
!! param; src/module_param.f90).  With prow = pcol = 1 every pencil of a
!! grid is the whole (nx, ny, nz) array in the same order, so the pencil
!! index ranges all start at 1 and the transposes are device copies.

!! nrank / nproc and the abort and warning calls of the reference's callers
!! (src/x3d_tools.f90:63-70, src/navier.f90:220-223).
module decomp_2d_mpi
  implicit none
  private
  integer, save, public :: nrank = 0, nproc = 1
  public :: decomp_2d_abort, decomp_2d_warning

  interface decomp_2d_abort
    module procedure abort_basic
    module procedure abort_file_line
  end interface decomp_2d_abort

  interface decomp_2d_warning
    module procedure warning_basic
    module procedure warning_file_line
  end interface decomp_2d_warning

contains

  subroutine abort_basic(errorcode, msg)
    integer, intent(in) :: errorcode
    character(len=*), intent(in) :: msg
    write (*, '(a,i0,a,a)') 'decomp_2d error - errorcode: ', errorcode, ' ', msg
    error stop 1
  end subroutine abort_basic

  subroutine abort_file_line(file, line, errorcode, msg)
    character(len=*), intent(in) :: file, msg
    integer, intent(in) :: line, errorcode
    write (*, '(a,a,a,i0,a,i0,a,a)') 'decomp_2d error in ', file, ':', line, ' - errorcode: ', errorcode, ' ', msg
    error stop 1
  end subroutine abort_file_line

  subroutine warning_basic(errorcode, msg)
    integer, intent(in) :: errorcode
    character(len=*), intent(in) :: msg
    write (*, '(a,i0,a,a)') 'decomp_2d warning - errorcode: ', errorcode, ' ', msg
  end subroutine warning_basic

  subroutine warning_file_line(file, line, errorcode, msg)
    character(len=*), intent(in) :: file, msg
    integer, intent(in) :: line, errorcode
    write (*, '(a,a,a,i0,a,i0,a,a)') 'decomp_2d warning in ', file, ':', line, ' - errorcode: ', errorcode, ' ', msg
  end subroutine warning_file_line

end module decomp_2d_mpi

!! The pencil bookkeeping of 2decomp-fft at one rank: decomp_info with its
!! x/y/z start, end and size triples, the main decomposition (xsize, ...),
!! the pressure-grid decompositions ph1..ph4 and phG the reference keeps
!! there (src/x3d_tools.f90:170-176), and the transposes, which copy the whole
!! device array (a no-op when source and destination are the same array).
module decomp_2d
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  implicit none
  private

  type, public :: decomp_info
    integer, dimension(3) :: xst = 1, xen = 0, xsz = 0
    integer, dimension(3) :: yst = 1, yen = 0, ysz = 0
    integer, dimension(3) :: zst = 1, zen = 0, zsz = 0
  end type decomp_info

  integer, save, public :: nx_global = 0, ny_global = 0, nz_global = 0
  integer, save, dimension(3), public :: xstart = 1, xend = 0, xsize = 0
  integer, save, dimension(3), public :: ystart = 1, yend = 0, ysize = 0
  integer, save, dimension(3), public :: zstart = 1, zend = 0, zsize = 0
  type(decomp_info), save, target, public :: decomp_main, ph1, ph2, ph3, ph4, phG

  public :: decomp_2d_init, decomp_2d_finalize, decomp_info_init, decomp_info_finalize
  public :: transpose_x_to_y, transpose_y_to_z, transpose_z_to_y, transpose_y_to_x

  interface transpose_x_to_y
    module procedure x_to_y_real
    module procedure x_to_y_cplx
  end interface transpose_x_to_y
  interface transpose_y_to_z
    module procedure y_to_z_real
    module procedure y_to_z_cplx
  end interface transpose_y_to_z
  interface transpose_z_to_y
    module procedure z_to_y_real
    module procedure z_to_y_cplx
  end interface transpose_z_to_y
  interface transpose_y_to_x
    module procedure y_to_x_real
    module procedure y_to_x_cplx
  end interface transpose_y_to_x

contains

  !> decomp_2d_init(nx, ny, nz, p_row, p_col): one rank owns every pencil.
  subroutine decomp_2d_init(nx, ny, nz, p_row, p_col)
    integer, intent(in) :: nx, ny, nz, p_row, p_col
    if (p_row * p_col > 1) print *, "p_row x p_col = ", p_row * p_col, " ranks requested; running on one GPU"
    nx_global = nx; ny_global = ny; nz_global = nz
    call decomp_info_init(nx, ny, nz, decomp_main)
    xstart = decomp_main%xst; xend = decomp_main%xen; xsize = decomp_main%xsz
    ystart = decomp_main%yst; yend = decomp_main%yen; ysize = decomp_main%ysz
    zstart = decomp_main%zst; zend = decomp_main%zen; zsize = decomp_main%zsz
  end subroutine decomp_2d_init

  subroutine decomp_2d_finalize()
    call decomp_info_finalize(decomp_main)
    xsize = 0; ysize = 0; zsize = 0
    xend = 0; yend = 0; zend = 0
  end subroutine decomp_2d_finalize

  !> decomp_info_init(nx, ny, nz, decomp): every pencil is 1..n in each index.
  subroutine decomp_info_init(nx, ny, nz, decomp)
    integer, intent(in) :: nx, ny, nz
    type(decomp_info), intent(inout) :: decomp
    integer, dimension(3) :: n
    n = [nx, ny, nz]
    decomp%xst = 1; decomp%xen = n; decomp%xsz = n
    decomp%yst = 1; decomp%yen = n; decomp%ysz = n
    decomp%zst = 1; decomp%zen = n; decomp%zsz = n
  end subroutine decomp_info_init

  subroutine decomp_info_finalize(decomp)
    type(decomp_info), intent(inout) :: decomp
    decomp%xen = 0; decomp%xsz = 0
    decomp%yen = 0; decomp%ysz = 0
    decomp%zen = 0; decomp%zsz = 0
  end subroutine decomp_info_finalize

  ! ---- transposes: one device copy of the whole array ------------------------
  subroutine copy(src, dst, n, eb)
    type(c_ptr), intent(in) :: src, dst
    integer, intent(in) :: n(3)
    integer(c_size_t), intent(in) :: eb
    call x3d_check('transpose', x3d_copy_field(x3d_ctx_h, src, dst, eb, int(n(1), c_int), &
                                                     int(n(2), c_int), int(n(3), c_int)))
  end subroutine copy

  subroutine x_to_y_real(src, dst, opt_decomp)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: src
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: dst
    type(decomp_info), intent(in), optional :: opt_decomp
    call copy(c_loc(src), c_loc(dst), shape(src), int(storage_size(src) / 8, c_size_t))
  end subroutine x_to_y_real

  subroutine x_to_y_cplx(src, dst, opt_decomp)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: src
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: dst
    type(decomp_info), intent(in), optional :: opt_decomp
    call copy(c_loc(src), c_loc(dst), shape(src), int(storage_size(src) / 8, c_size_t))
  end subroutine x_to_y_cplx

  subroutine y_to_z_real(src, dst, opt_decomp)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: src
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: dst
    type(decomp_info), intent(in), optional :: opt_decomp
    call copy(c_loc(src), c_loc(dst), shape(src), int(storage_size(src) / 8, c_size_t))
  end subroutine y_to_z_real

  subroutine y_to_z_cplx(src, dst, opt_decomp)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: src
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: dst
    type(decomp_info), intent(in), optional :: opt_decomp
    call copy(c_loc(src), c_loc(dst), shape(src), int(storage_size(src) / 8, c_size_t))
  end subroutine y_to_z_cplx

  subroutine z_to_y_real(src, dst, opt_decomp)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: src
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: dst
    type(decomp_info), intent(in), optional :: opt_decomp
    call copy(c_loc(src), c_loc(dst), shape(src), int(storage_size(src) / 8, c_size_t))
  end subroutine z_to_y_real

  subroutine z_to_y_cplx(src, dst, opt_decomp)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: src
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: dst
    type(decomp_info), intent(in), optional :: opt_decomp
    call copy(c_loc(src), c_loc(dst), shape(src), int(storage_size(src) / 8, c_size_t))
  end subroutine z_to_y_cplx

  subroutine y_to_x_real(src, dst, opt_decomp)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: src
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: dst
    type(decomp_info), intent(in), optional :: opt_decomp
    call copy(c_loc(src), c_loc(dst), shape(src), int(storage_size(src) / 8, c_size_t))
  end subroutine y_to_x_real

  subroutine y_to_x_cplx(src, dst, opt_decomp)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: src
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: dst
    type(decomp_info), intent(in), optional :: opt_decomp
    call copy(c_loc(src), c_loc(dst), shape(src), int(storage_size(src) / 8, c_size_t))
  end subroutine y_to_x_cplx

end module decomp_2d

!! Grid sizes and run-time switches (module variables, src/module_param.f90:5-76):
!! the names the hot path's callers read.
module variables
  use decomp_2d_constants, only : mytype
  implicit none
  integer, save, public :: nx = 32, ny = 32, nz = 32, nxm = 32, nym = 32, nzm = 32
  integer, save, public :: p_row = 0, p_col = 0
  integer, save, public :: nstat = 1, nvisu = 1, nprobe = 1
  logical, save, public :: test_mode = .false.
  real(mytype), save, public :: nu0nu = 0._mytype, cnu = 0._mytype
  ! stretching vectors (src/var.f90:138-140), passed by dery/deryvp/derypv
  ! callers; with istret /= 0 the library multiplies y-derivatives by them
  real(mytype), save, allocatable, dimension(:), public, target :: ppy, pp2y, pp4y, ppyi, pp2yi, pp4yi
end module variables

!! Numerical parameters (module param, src/module_param.f90:79-146).
module param
  use decomp_2d_constants, only : mytype
  implicit none
  integer, save, public :: nclx1 = 0, nclxn = 0, ncly1 = 0, nclyn = 0, nclz1 = 0, nclzn = 0
  integer, save, public :: nclxS1 = 0, nclxSn = 0, nclyS1 = 0, nclySn = 0, nclzS1 = 0, nclzSn = 0
  logical, save, public :: nclx = .true., ncly = .true., nclz = .true.
  integer, save, public :: itype = 2, itr = 1, itime = 0, istret = 0
  integer, save, public :: itimescheme = 1, iimplicit = 0, ifirst = 0, ilast = 0, iscalar = 0
  integer, save, public :: ifirstder = 4, isecondder = 4, ipinter = 3
  integer, save, public :: ntime = 1, npress = 1
  integer, parameter, public :: itype_tgv = 2, itype_tgv2d = 10
  real(mytype), save, public :: xlx = 1._mytype, yly = 1._mytype, zlz = 1._mytype
  real(mytype), save, public :: dx = 0._mytype, dy = 0._mytype, dz = 0._mytype
  real(mytype), save, public :: dx2 = 0._mytype, dy2 = 0._mytype, dz2 = 0._mytype
  real(mytype), save, public :: dt = 0._mytype, re = 0._mytype, xnu = 0._mytype, t = 0._mytype
  real(mytype), save, public :: u1 = 2._mytype, u2 = 1._mytype
  real(mytype), save, dimension(5), public :: adt = 0._mytype, bdt = 0._mytype, cdt = 0._mytype, &
                                              ddt = 0._mytype, gdt = 0._mytype
  real(mytype), parameter, public :: zero = 0._mytype, half = 0.5_mytype, one = 1._mytype
  real(mytype), parameter, public :: onepfive = 1.5_mytype, two = 2._mytype, twopfive = 2.5_mytype
  real(mytype), parameter, public :: three = 3._mytype, threepfive = 3.5_mytype, four = 4._mytype
  real(mytype), parameter, public :: five = 5._mytype, twothird = 2._mytype / 3._mytype
  real(mytype), parameter, public :: pi = 3.14159265358979323846_mytype, twopi = 2._mytype * pi
end module param
