!! Reference-named modules for the staggered operators, transposes, Poisson
!! and the timestep callers, bound to libx3dhip.  Signatures follow
!! src/x3d_staggered.f90, src/x3d_transpose.f90, src/poisson.f90,
!! src/transeq.f90, src/time_integrators.f90 and src/navier.f90.

!! Staggered derivatives and mid-point interpolations (src/x3d_staggered.f90).
module x3d_staggered
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  use x3d_operator_1d, only : x3doperator1d
  implicit none
  private
  public :: derxvp, interxvp, derxpv, interxpv, interyvp, deryvp, interypv, derypv, &
            derzvp, interzvp, derzpv, interzpv

  interface
    integer(c_int) function c_derxvp(ctx, t, u, op, n1, n2, n3, n4) bind(C, name="x3d_derxvp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u
      integer(c_int), value :: op, n1, n2, n3, n4
    end function
    integer(c_int) function c_interxvp(ctx, t, u, op, n1, n2, n3, n4) bind(C, name="x3d_interxvp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u
      integer(c_int), value :: op, n1, n2, n3, n4
    end function
    integer(c_int) function c_derxpv(ctx, t, u, op, n1, n2, n3, n4) bind(C, name="x3d_derxpv")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u
      integer(c_int), value :: op, n1, n2, n3, n4
    end function
    integer(c_int) function c_interxpv(ctx, t, u, op, n1, n2, n3, n4) bind(C, name="x3d_interxpv")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u
      integer(c_int), value :: op, n1, n2, n3, n4
    end function
    integer(c_int) function c_interyvp(ctx, t, u, op, n1, n2, n3, n4) bind(C, name="x3d_interyvp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u
      integer(c_int), value :: op, n1, n2, n3, n4
    end function
    integer(c_int) function c_interypv(ctx, t, u, op, n1, n2, n3, n4) bind(C, name="x3d_interypv")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u
      integer(c_int), value :: op, n1, n2, n3, n4
    end function
    integer(c_int) function c_derzvp(ctx, t, u, op, n1, n2, n3, n4) bind(C, name="x3d_derzvp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u
      integer(c_int), value :: op, n1, n2, n3, n4
    end function
    integer(c_int) function c_interzvp(ctx, t, u, op, n1, n2, n3, n4) bind(C, name="x3d_interzvp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u
      integer(c_int), value :: op, n1, n2, n3, n4
    end function
    integer(c_int) function c_derzpv(ctx, t, u, op, n1, n2, n3, n4) bind(C, name="x3d_derzpv")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u
      integer(c_int), value :: op, n1, n2, n3, n4
    end function
    integer(c_int) function c_interzpv(ctx, t, u, op, n1, n2, n3, n4) bind(C, name="x3d_interzpv")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u
      integer(c_int), value :: op, n1, n2, n3, n4
    end function
  end interface

contains

  subroutine derxvp(tx, ux, x3dop, nx, nxm, ny, nz)
    integer, intent(in) :: nx, nxm, ny, nz
    real(mytype), intent(out), dimension(nxm, ny, nz), target :: tx
    real(mytype), intent(in), dimension(nx, ny, nz), target :: ux
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('derxvp', c_derxvp(x3d_ctx_h, c_loc(tx), c_loc(ux), int(x3dop%id, c_int), nx, nxm, ny, nz))
  end subroutine derxvp

  subroutine interxvp(tx, ux, x3dop, nx, nxm, ny, nz)
    integer, intent(in) :: nx, nxm, ny, nz
    real(mytype), intent(out), dimension(nxm, ny, nz), target :: tx
    real(mytype), intent(in), dimension(nx, ny, nz), target :: ux
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('interxvp', c_interxvp(x3d_ctx_h, c_loc(tx), c_loc(ux), int(x3dop%id, c_int), nx, nxm, ny, nz))
  end subroutine interxvp

  subroutine derxpv(tx, ux, x3dop, nxm, nx, ny, nz)
    integer, intent(in) :: nx, nxm, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: tx
    real(mytype), intent(in), dimension(nxm, ny, nz), target :: ux
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('derxpv', c_derxpv(x3d_ctx_h, c_loc(tx), c_loc(ux), int(x3dop%id, c_int), nxm, nx, ny, nz))
  end subroutine derxpv

  subroutine interxpv(tx, ux, x3dop, nxm, nx, ny, nz)
    integer, intent(in) :: nx, nxm, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: tx
    real(mytype), intent(in), dimension(nxm, ny, nz), target :: ux
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('interxpv', c_interxpv(x3d_ctx_h, c_loc(tx), c_loc(ux), int(x3dop%id, c_int), nxm, nx, ny, nz))
  end subroutine interxpv

  subroutine interyvp(ty, uy, x3dop, nx, ny, nym, nz)
    integer, intent(in) :: nx, ny, nym, nz
    real(mytype), intent(out), dimension(nx, nym, nz), target :: ty
    real(mytype), intent(in), dimension(nx, ny, nz), target :: uy
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('interyvp', c_interyvp(x3d_ctx_h, c_loc(ty), c_loc(uy), int(x3dop%id, c_int), nx, ny, nym, nz))
  end subroutine interyvp

  subroutine deryvp(ty, uy, x3dop, ppyi, nx, ny, nym, nz)
    integer, intent(in) :: nx, ny, nym, nz
    real(mytype), intent(out), dimension(nx, nym, nz), target :: ty
    real(mytype), intent(in), dimension(nx, ny, nz), target :: uy
    real(mytype), intent(in), dimension(nym) :: ppyi
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('deryvp', x3d_deryvp(x3d_ctx_h, c_loc(ty), c_loc(uy), int(x3dop%id, c_int), c_null_ptr, &
                                        nx, ny, nym, nz))
  end subroutine deryvp

  subroutine interypv(ty, uy, x3dop, nx, nym, ny, nz)
    integer, intent(in) :: nx, nym, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: ty
    real(mytype), intent(in), dimension(nx, nym, nz), target :: uy
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('interypv', c_interypv(x3d_ctx_h, c_loc(ty), c_loc(uy), int(x3dop%id, c_int), nx, nym, ny, nz))
  end subroutine interypv

  subroutine derypv(ty, uy, x3dop, ppy, nx, nym, ny, nz)
    integer, intent(in) :: nx, nym, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: ty
    real(mytype), intent(in), dimension(nx, nym, nz), target :: uy
    real(mytype), intent(in), dimension(ny) :: ppy
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('derypv', x3d_derypv(x3d_ctx_h, c_loc(ty), c_loc(uy), int(x3dop%id, c_int), c_null_ptr, &
                                        nx, nym, ny, nz))
  end subroutine derypv

  subroutine derzvp(tz, uz, x3dop, nx, ny, nz, nzm)
    integer, intent(in) :: nx, ny, nz, nzm
    real(mytype), intent(out), dimension(nx, ny, nzm), target :: tz
    real(mytype), intent(in), dimension(nx, ny, nz), target :: uz
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('derzvp', c_derzvp(x3d_ctx_h, c_loc(tz), c_loc(uz), int(x3dop%id, c_int), nx, ny, nz, nzm))
  end subroutine derzvp

  subroutine interzvp(tz, uz, x3dop, nx, ny, nz, nzm)
    integer, intent(in) :: nx, ny, nz, nzm
    real(mytype), intent(out), dimension(nx, ny, nzm), target :: tz
    real(mytype), intent(in), dimension(nx, ny, nz), target :: uz
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('interzvp', c_interzvp(x3d_ctx_h, c_loc(tz), c_loc(uz), int(x3dop%id, c_int), nx, ny, nz, nzm))
  end subroutine interzvp

  subroutine derzpv(tz, uz, x3dop, nx, ny, nzm, nz)
    integer, intent(in) :: nx, ny, nzm, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: tz
    real(mytype), intent(in), dimension(nx, ny, nzm), target :: uz
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('derzpv', c_derzpv(x3d_ctx_h, c_loc(tz), c_loc(uz), int(x3dop%id, c_int), nx, ny, nzm, nz))
  end subroutine derzpv

  subroutine interzpv(tz, uz, x3dop, nx, ny, nzm, nz)
    integer, intent(in) :: nx, ny, nzm, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: tz
    real(mytype), intent(in), dimension(nx, ny, nzm), target :: uz
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('interzpv', c_interzpv(x3d_ctx_h, c_loc(tz), c_loc(uz), int(x3dop%id, c_int), nx, ny, nzm, nz))
  end subroutine interzpv

end module x3d_staggered

!! Pencil transposes (src/x3d_transpose.f90:16-47), short forms (no decomp
!! argument).  At one rank each is a device copy of the whole array.
module x3d_transpose
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  implicit none
  private
  public :: x3d_transpose_x_to_y, x3d_transpose_y_to_z, x3d_transpose_z_to_y, x3d_transpose_y_to_x

  interface
    integer(c_int) function c_x_to_y(ctx, a, b, eb, n1, n2, n3) bind(C, name="x3d_transpose_x_to_y")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, a, b
      integer(c_size_t), value :: eb
      integer(c_int), value :: n1, n2, n3
    end function
    integer(c_int) function c_y_to_z(ctx, a, b, eb, n1, n2, n3) bind(C, name="x3d_transpose_y_to_z")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, a, b
      integer(c_size_t), value :: eb
      integer(c_int), value :: n1, n2, n3
    end function
    integer(c_int) function c_z_to_y(ctx, a, b, eb, n1, n2, n3) bind(C, name="x3d_transpose_z_to_y")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, a, b
      integer(c_size_t), value :: eb
      integer(c_int), value :: n1, n2, n3
    end function
    integer(c_int) function c_y_to_x(ctx, a, b, eb, n1, n2, n3) bind(C, name="x3d_transpose_y_to_x")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, a, b
      integer(c_size_t), value :: eb
      integer(c_int), value :: n1, n2, n3
    end function
  end interface

  interface x3d_transpose_x_to_y
    module procedure x3d_transpose_x_to_y_real_short
    module procedure x3d_transpose_x_to_y_cplx_short
  end interface x3d_transpose_x_to_y
  interface x3d_transpose_y_to_z
    module procedure x3d_transpose_y_to_z_real_short
    module procedure x3d_transpose_y_to_z_cplx_short
  end interface x3d_transpose_y_to_z
  interface x3d_transpose_z_to_y
    module procedure x3d_transpose_z_to_y_real_short
    module procedure x3d_transpose_z_to_y_cplx_short
  end interface x3d_transpose_z_to_y
  interface x3d_transpose_y_to_x
    module procedure x3d_transpose_y_to_x_real_short
    module procedure x3d_transpose_y_to_x_cplx_short
  end interface x3d_transpose_y_to_x

contains

  subroutine x3d_transpose_x_to_y_real_short(data_in, data_out)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call x3d_check('x3d_transpose_x_to_y', c_x_to_y(x3d_ctx_h, c_loc(data_in), c_loc(data_out), &
                   int(storage_size(data_in) / 8, c_size_t), size(data_in, 1), size(data_in, 2), size(data_in, 3)))
  end subroutine x3d_transpose_x_to_y_real_short

  subroutine x3d_transpose_x_to_y_cplx_short(data_in, data_out)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call x3d_check('x3d_transpose_x_to_y', c_x_to_y(x3d_ctx_h, c_loc(data_in), c_loc(data_out), &
                   int(storage_size(data_in) / 8, c_size_t), size(data_in, 1), size(data_in, 2), size(data_in, 3)))
  end subroutine x3d_transpose_x_to_y_cplx_short

  subroutine x3d_transpose_y_to_z_real_short(data_in, data_out)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call x3d_check('x3d_transpose_y_to_z', c_y_to_z(x3d_ctx_h, c_loc(data_in), c_loc(data_out), &
                   int(storage_size(data_in) / 8, c_size_t), size(data_in, 1), size(data_in, 2), size(data_in, 3)))
  end subroutine x3d_transpose_y_to_z_real_short

  subroutine x3d_transpose_y_to_z_cplx_short(data_in, data_out)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call x3d_check('x3d_transpose_y_to_z', c_y_to_z(x3d_ctx_h, c_loc(data_in), c_loc(data_out), &
                   int(storage_size(data_in) / 8, c_size_t), size(data_in, 1), size(data_in, 2), size(data_in, 3)))
  end subroutine x3d_transpose_y_to_z_cplx_short

  subroutine x3d_transpose_z_to_y_real_short(data_in, data_out)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call x3d_check('x3d_transpose_z_to_y', c_z_to_y(x3d_ctx_h, c_loc(data_in), c_loc(data_out), &
                   int(storage_size(data_in) / 8, c_size_t), size(data_in, 1), size(data_in, 2), size(data_in, 3)))
  end subroutine x3d_transpose_z_to_y_real_short

  subroutine x3d_transpose_z_to_y_cplx_short(data_in, data_out)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call x3d_check('x3d_transpose_z_to_y', c_z_to_y(x3d_ctx_h, c_loc(data_in), c_loc(data_out), &
                   int(storage_size(data_in) / 8, c_size_t), size(data_in, 1), size(data_in, 2), size(data_in, 3)))
  end subroutine x3d_transpose_z_to_y_cplx_short

  subroutine x3d_transpose_y_to_x_real_short(data_in, data_out)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call x3d_check('x3d_transpose_y_to_x', c_y_to_x(x3d_ctx_h, c_loc(data_in), c_loc(data_out), &
                   int(storage_size(data_in) / 8, c_size_t), size(data_in, 1), size(data_in, 2), size(data_in, 3)))
  end subroutine x3d_transpose_y_to_x_real_short

  subroutine x3d_transpose_y_to_x_cplx_short(data_in, data_out)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call x3d_check('x3d_transpose_y_to_x', c_y_to_x(x3d_ctx_h, c_loc(data_in), c_loc(data_out), &
                   int(storage_size(data_in) / 8, c_size_t), size(data_in, 1), size(data_in, 2), size(data_in, 3)))
  end subroutine x3d_transpose_y_to_x_cplx_short

end module x3d_transpose

!! Spectral Poisson (src/poisson.f90:5-769): procedure pointer poisson and
!! the public module data ax..bz, kxyz (real part, reference spectral grid).
module decomp_2d_poisson
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  implicit none
  private

  real(mytype), save, allocatable, dimension(:), public, target :: az, bz, ay, by, ax, bx
  real(mytype), save, allocatable, dimension(:, :, :), public, target :: kxyz

  abstract interface
    subroutine poisson_xxx(rhs)
      import :: mytype
      real(mytype), dimension(:, :, :), intent(inout), target, contiguous :: rhs
    end subroutine poisson_xxx
  end interface
  procedure(poisson_xxx), pointer, public :: poisson => null()

  public :: decomp_2d_poisson_init, decomp_2d_poisson_finalize

contains

  subroutine decomp_2d_poisson_init(nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    allocate (ax(nx), bx(nx), ay(ny), by(ny), az(nz), bz(nz), kxyz(nx, ny, nz / 2 + 1))
    call x3d_check('x3d_poisson_tables', x3d_poisson_tables(x3d_ctx_h, c_loc(ax), c_loc(bx), c_loc(ay), &
                                                            c_loc(by), c_loc(az), c_loc(bz), c_loc(kxyz)))
    poisson => poisson_000
  end subroutine decomp_2d_poisson_init

  subroutine decomp_2d_poisson_finalize()
    nullify (poisson)
    if (allocated(ax)) deallocate (ax, bx, ay, by, az, bz, kxyz)
  end subroutine decomp_2d_poisson_finalize

  subroutine poisson_000(rhs)
    real(mytype), dimension(:, :, :), intent(inout), target, contiguous :: rhs
    call x3d_check('poisson', x3d_poisson(x3d_ctx_h, c_loc(rhs)))
  end subroutine poisson_000

end module decomp_2d_poisson

!! Run-time settings read by the callers (module variables, src/module_param.f90).
module variables
  implicit none
  integer, save, public :: nx = 32, ny = 32, nz = 32, nxm = 32, nym = 32, nzm = 32
  integer, save, public :: p_row = 0, p_col = 0
  logical, save, public :: test_mode = .false.
end module variables

!! test_du / test_dv / test_dw (src/mom.f90:59-243): RMS error of a
!! derivative against the analytic one, printed in test mode.
module mom
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  use variables, only : test_mode
  implicit none
  private
  public :: test_du, test_dv, test_dw
contains
  subroutine test_one(d, which, label)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: d
    integer, intent(in) :: which
    character(len=*), intent(in) :: label
    real(c_double) :: err
    if (.not. test_mode) return
    call x3d_check('test_err', x3d_test_err(x3d_ctx_h, c_loc(d), int(which, c_int), err))
    print *, "RMS error in "//label//": ", err
  end subroutine test_one
  subroutine test_du(du)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: du
    call test_one(du, 0, 'dudx')
  end subroutine test_du
  subroutine test_dv(dv)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: dv
    call test_one(dv, 1, 'dvdy')
  end subroutine test_dv
  subroutine test_dw(dw)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: dw
    call test_one(dw, 2, 'dwdz')
  end subroutine test_dw
end module mom

!! calculate_transeq_rhs (src/transeq.f90:18-34).  In test mode the three
!! collocated derivatives that momentum_rhs_eq checks (:109, :214, :319) are
!! formed in a scratch field and reported before the fused right-hand side.
module transeq
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  use variables, only : test_mode
  use x3d_operator_1d, only : x3d_op_derx, x3d_op_dery, x3d_op_derz
  use mom, only : test_du, test_dv, test_dw
  implicit none
  private
  public :: calculate_transeq_rhs
  type(c_ptr), save :: scratch = c_null_ptr
contains
  subroutine calculate_transeq_rhs(dux1, duy1, duz1, ux1, uy1, uz1)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: ux1, uy1, uz1
    real(mytype), dimension(:, :, :), target, contiguous :: dux1, duy1, duz1
    real(mytype), dimension(:, :, :), pointer, contiguous :: t
    integer(c_int) :: n1, n2, n3
    if (test_mode) then
      n1 = size(ux1, 1); n2 = size(ux1, 2); n3 = size(ux1, 3)
      if (.not. c_associated(scratch)) &
        call x3d_check('x3d_field_alloc', x3d_field_alloc(x3d_ctx_h, int(8, c_size_t) * n1 * n2 * n3, scratch))
      call c_f_pointer(scratch, t, [n1, n2, n3])
      call x3d_check('derx', x3d_derx(x3d_ctx_h, scratch, c_loc(ux1), int(x3d_op_derx%id, c_int), n1, n2, n3))
      call test_du(t)
      call x3d_check('dery', x3d_dery(x3d_ctx_h, scratch, c_loc(uy1), int(x3d_op_dery%id, c_int), c_null_ptr, &
                                      n1, n2, n3))
      call test_dv(t)
      call x3d_check('derz', x3d_derz(x3d_ctx_h, scratch, c_loc(uz1), int(x3d_op_derz%id, c_int), n1, n2, n3))
      call test_dw(t)
    end if
    call x3d_check('calculate_transeq_rhs', x3d_momentum_rhs(x3d_ctx_h, c_loc(dux1), c_loc(duy1), c_loc(duz1), &
                                                             c_loc(ux1), c_loc(uy1), c_loc(uz1)))
  end subroutine calculate_transeq_rhs
end module transeq

!! int_time, Euler (src/time_integrators.f90:20-106).
module time_integrators
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  implicit none
  private
  public :: int_time
  real(mytype), save, public :: gdt1 = 0._mytype   ! gdt(1) = adt(1) + bdt(1) = dt (src/var.f90:183-186)
contains
  subroutine int_time(ux1, uy1, uz1, dux1, duy1, duz1)
    real(mytype), dimension(:, :, :), target, contiguous :: ux1, uy1, uz1
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: dux1, duy1, duz1
    call x3d_check('int_time', x3d_int_time(x3d_ctx_h, real(gdt1, c_double), c_loc(ux1), c_loc(uy1), &
                                            c_loc(uz1), c_loc(dux1), c_loc(duy1), c_loc(duz1)))
  end subroutine int_time
end module time_integrators

!! solve_poisson, divergence, gradp, cor_vel (src/navier.f90:21-359).
!! In test mode solve_poisson runs its three parts separately so that the
!! "DIV U* max mean" line of divergence (:213-231) can be printed.
module navier
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  use variables, only : test_mode
  implicit none
  private
  public :: solve_poisson, divergence, gradp, cor_vel
contains
  subroutine solve_poisson(pp3, px1, py1, pz1, ux1, uy1, uz1)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: ux1, uy1, uz1
    real(mytype), dimension(:, :, :), target, contiguous :: pp3, px1, py1, pz1
    if (test_mode) then
      call divergence(pp3, ux1, uy1, uz1, 1)
      call x3d_check('poisson', x3d_poisson(x3d_ctx_h, c_loc(pp3)))
      call gradp(px1, py1, pz1, pp3)
    else
      call x3d_check('solve_poisson', x3d_solve_poisson(x3d_ctx_h, c_loc(pp3), c_loc(px1), c_loc(py1), &
                                                        c_loc(pz1), c_loc(ux1), c_loc(uy1), c_loc(uz1)))
    end if
  end subroutine solve_poisson

  subroutine divergence(pp3, ux1, uy1, uz1, nlock)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: ux1, uy1, uz1
    real(mytype), dimension(:, :, :), target, contiguous :: pp3
    integer, intent(in) :: nlock
    real(c_double) :: tmax, tmoy
    call x3d_check('divergence', x3d_divergence(x3d_ctx_h, c_loc(pp3), c_loc(ux1), c_loc(uy1), c_loc(uz1)))
    if (test_mode .and. nlock > 0) then
      call x3d_check('div_stats', x3d_div_stats(x3d_ctx_h, c_loc(pp3), tmax, tmoy))
      if (nlock == 2) then
        print *, 'DIV U  max mean=', real(tmax, 4), real(tmoy, 4)
      else
        print *, 'DIV U* max mean=', real(tmax, 4), real(tmoy, 4)
      end if
    end if
  end subroutine divergence

  subroutine gradp(px1, py1, pz1, pp3)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: pp3
    real(mytype), dimension(:, :, :), target, contiguous :: px1, py1, pz1
    call x3d_check('gradp', x3d_gradp(x3d_ctx_h, c_loc(px1), c_loc(py1), c_loc(pz1), c_loc(pp3)))
  end subroutine gradp

  subroutine cor_vel(ux, uy, uz, px, py, pz)
    real(mytype), dimension(:, :, :), target, contiguous :: ux, uy, uz
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: px, py, pz
    call x3d_check('cor_vel', x3d_cor_vel(x3d_ctx_h, c_loc(ux), c_loc(uy), c_loc(uz), c_loc(px), c_loc(py), &
                                          c_loc(pz)))
  end subroutine cor_vel
end module navier
