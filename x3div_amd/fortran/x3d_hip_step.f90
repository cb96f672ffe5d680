!! Reference-named modules for the staggered operators, transposes, Poisson
!! and the timestep callers, bound to libx3dhip.  Signatures follow
!! src/x3d_staggered.f90, src/x3d_transpose.f90, src/poisson.f90, src/var.f90,
!! src/mom.f90, src/case.f90, src/transeq.f90, src/time_integrators.f90,
!! src/navier.f90 and src/parameters.f90, with the reference's array shapes.

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
    real(mytype), intent(in), dimension(nym), target :: ppyi
    type(x3doperator1d), intent(in) :: x3dop
    ! ppyi scales the result when istret /= 0 (src/x3d_staggered.f90:628-634)
    call x3d_check('deryvp', x3d_deryvp(x3d_ctx_h, c_loc(ty), c_loc(uy), int(x3dop%id, c_int), c_loc(ppyi), &
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
    real(mytype), intent(in), dimension(ny), target :: ppy
    type(x3doperator1d), intent(in) :: x3dop
    ! ppy scales the result when istret /= 0 (src/x3d_staggered.f90:874-880)
    call x3d_check('derypv', x3d_derypv(x3d_ctx_h, c_loc(ty), c_loc(uy), int(x3dop%id, c_int), c_loc(ppy), &
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

!! Pencil transposes (src/x3d_transpose.f90:16-378): the short forms and the
!! forms with a decomp_info argument, real and complex.  At one rank each is
!! a device copy of the whole array (nothing when in and out are the same).
module x3d_transpose
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use decomp_2d, only : decomp_info
  use x3d_hip, only : x3d_ctx_h, x3d_check
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
    module procedure x_to_y_real_short, x_to_y_real, x_to_y_cplx_short, x_to_y_cplx
  end interface x3d_transpose_x_to_y
  interface x3d_transpose_y_to_z
    module procedure y_to_z_real_short, y_to_z_real, y_to_z_cplx_short, y_to_z_cplx
  end interface x3d_transpose_y_to_z
  interface x3d_transpose_z_to_y
    module procedure z_to_y_real_short, z_to_y_real, z_to_y_cplx_short, z_to_y_cplx
  end interface x3d_transpose_z_to_y
  interface x3d_transpose_y_to_x
    module procedure y_to_x_real_short, y_to_x_real, y_to_x_cplx_short, y_to_x_cplx
  end interface x3d_transpose_y_to_x

contains

  !> One device copy: dir 0 x->y, 1 y->z, 2 z->y, 3 y->x.
  subroutine move(dir, a, b, eb, n)
    integer, intent(in) :: dir
    type(c_ptr), intent(in) :: a, b
    integer, intent(in) :: eb, n(3)
    integer(c_int) :: rc
    integer(c_size_t) :: e
    e = int(eb, c_size_t)
    select case (dir)
    case (0); rc = c_x_to_y(x3d_ctx_h, a, b, e, int(n(1), c_int), int(n(2), c_int), int(n(3), c_int))
    case (1); rc = c_y_to_z(x3d_ctx_h, a, b, e, int(n(1), c_int), int(n(2), c_int), int(n(3), c_int))
    case (2); rc = c_z_to_y(x3d_ctx_h, a, b, e, int(n(1), c_int), int(n(2), c_int), int(n(3), c_int))
    case default; rc = c_y_to_x(x3d_ctx_h, a, b, e, int(n(1), c_int), int(n(2), c_int), int(n(3), c_int))
    end select
    call x3d_check('x3d_transpose', rc)
  end subroutine move

  subroutine x_to_y_real_short(data_in, data_out)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call move(0, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine x_to_y_real_short
  subroutine x_to_y_real(data_in, data_out, decomp)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    type(decomp_info), intent(in) :: decomp
    call move(0, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine x_to_y_real
  subroutine x_to_y_cplx_short(data_in, data_out)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call move(0, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine x_to_y_cplx_short
  subroutine x_to_y_cplx(data_in, data_out, decomp)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    type(decomp_info), intent(in) :: decomp
    call move(0, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine x_to_y_cplx

  subroutine y_to_z_real_short(data_in, data_out)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call move(1, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine y_to_z_real_short
  subroutine y_to_z_real(data_in, data_out, decomp)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    type(decomp_info), intent(in) :: decomp
    call move(1, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine y_to_z_real
  subroutine y_to_z_cplx_short(data_in, data_out)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call move(1, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine y_to_z_cplx_short
  subroutine y_to_z_cplx(data_in, data_out, decomp)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    type(decomp_info), intent(in) :: decomp
    call move(1, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine y_to_z_cplx

  subroutine z_to_y_real_short(data_in, data_out)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call move(2, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine z_to_y_real_short
  subroutine z_to_y_real(data_in, data_out, decomp)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    type(decomp_info), intent(in) :: decomp
    call move(2, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine z_to_y_real
  subroutine z_to_y_cplx_short(data_in, data_out)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call move(2, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine z_to_y_cplx_short
  subroutine z_to_y_cplx(data_in, data_out, decomp)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    type(decomp_info), intent(in) :: decomp
    call move(2, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine z_to_y_cplx

  subroutine y_to_x_real_short(data_in, data_out)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call move(3, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine y_to_x_real_short
  subroutine y_to_x_real(data_in, data_out, decomp)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    type(decomp_info), intent(in) :: decomp
    call move(3, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine y_to_x_real
  subroutine y_to_x_cplx_short(data_in, data_out)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    call move(3, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine y_to_x_cplx_short
  subroutine y_to_x_cplx(data_in, data_out, decomp)
    complex(mytype), dimension(:, :, :), intent(in), target, contiguous :: data_in
    complex(mytype), dimension(:, :, :), intent(out), target, contiguous :: data_out
    type(decomp_info), intent(in) :: decomp
    call move(3, c_loc(data_in), c_loc(data_out), storage_size(data_in) / 8, shape(data_in))
  end subroutine y_to_x_cplx

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

  !> decomp_2d_poisson_init() (src/poisson.f90:73-231) on the global grid of
  !> decomp_2d_init; the spectral tables come from the library.
  subroutine decomp_2d_poisson_init()
    use decomp_2d, only : nx_global, ny_global, nz_global
    integer :: nx, ny, nz
    nx = nx_global; ny = ny_global; nz = nz_global
    if (allocated(ax)) deallocate (ax, bx, ay, by, az, bz, kxyz)
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

!! The step's arrays (module var, src/var.f90:12-31, 39-201): device arrays
!! owned by the library and mapped with c_f_pointer, in the reference's
!! shapes: ux1..pz1 (xsize), dux1..duz1 (xsize, ntime), pp3 (the ph1
!! z-pencil, nzm, npress).
module var
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  implicit none
  private
  real(mytype), save, pointer, contiguous, dimension(:, :, :), public :: ux1 => null(), uy1 => null(), &
    uz1 => null(), px1 => null(), py1 => null(), pz1 => null()
  real(mytype), save, pointer, contiguous, dimension(:, :, :, :), public :: dux1 => null(), duy1 => null(), &
    duz1 => null(), pp3 => null()
  type(c_ptr), save :: mem(10) = c_null_ptr
  public :: var_init, var_finalize
contains
  !> var_init (src/var.f90:39-201): Euler time scheme (gdt(1) = dt, ntime = 1)
  !> and the step's arrays.
  subroutine var_init()
    use decomp_2d, only : xsize, ph1
    use param, only : adt, bdt, gdt, dt, ntime, npress, zero, one
    use variables, only : ny, nzm, ppy, pp2y, pp4y, ppyi, pp2yi, pp4yi
    integer :: i
    integer(c_size_t) :: nv, np
    ! stretching vectors, zero as at the reference (src/var.f90:138-140)
    if (.not. allocated(ppy)) then
      allocate (ppy(ny))
      ppy = zero
      allocate (pp2y, pp4y, ppyi, pp2yi, pp4yi, source=ppy)
    end if
    adt = zero; bdt = zero; gdt = zero
    adt(1) = one * dt
    gdt(1) = adt(1) + bdt(1)
    gdt(3) = gdt(1)
    ntime = 1
    nv = int(8, c_size_t) * xsize(1) * xsize(2) * xsize(3)
    np = int(8, c_size_t) * (ph1%zen(1) - ph1%zst(1) + 1) * (ph1%zen(2) - ph1%zst(2) + 1) * nzm
    do i = 1, 10
      if (i >= 4 .and. i <= 6) then
        call x3d_check('x3d_field_alloc', x3d_field_alloc(x3d_ctx_h, nv * ntime, mem(i)))
      else if (i == 7) then
        call x3d_check('x3d_field_alloc', x3d_field_alloc(x3d_ctx_h, np * npress, mem(i)))
      else
        call x3d_check('x3d_field_alloc', x3d_field_alloc(x3d_ctx_h, nv, mem(i)))
      end if
    end do
    call c_f_pointer(mem(1), ux1, xsize); call c_f_pointer(mem(2), uy1, xsize); call c_f_pointer(mem(3), uz1, xsize)
    call c_f_pointer(mem(4), dux1, [xsize, ntime]); call c_f_pointer(mem(5), duy1, [xsize, ntime])
    call c_f_pointer(mem(6), duz1, [xsize, ntime])
    call c_f_pointer(mem(7), pp3, [ph1%zen(1) - ph1%zst(1) + 1, ph1%zen(2) - ph1%zst(2) + 1, nzm, npress])
    call c_f_pointer(mem(8), px1, xsize); call c_f_pointer(mem(9), py1, xsize); call c_f_pointer(mem(10), pz1, xsize)
  end subroutine var_init

  subroutine var_finalize()
    use variables, only : ppy, pp2y, pp4y, ppyi, pp2yi, pp4yi
    integer :: i
    if (allocated(ppy)) deallocate (ppy, pp2y, pp4y, ppyi, pp2yi, pp4yi)
    do i = 1, 10
      if (c_associated(mem(i))) call x3d_check('x3d_field_free', x3d_field_free(x3d_ctx_h, mem(i)))
      mem(i) = c_null_ptr
    end do
    nullify (ux1, uy1, uz1, px1, py1, pz1, dux1, duy1, duz1, pp3)
  end subroutine var_finalize
end module var

!! vel and test_du / test_dv / test_dw (src/mom.f90:26-243): the reference's
!! initial field on the device, and in test mode the RMS error of a
!! derivative against the analytic one.
module mom
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use decomp_2d, only : xsize
  use x3d_hip
  use variables, only : test_mode
  implicit none
  private
  public :: vel, test_du, test_dv, test_dw
contains
  subroutine vel(u, v, w)
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), intent(out), target :: u, v, w
    call x3d_check('init_fields', x3d_init_fields(x3d_ctx_h, 0_c_int, 0_c_int64_t, 0._c_double, &
                                                  c_loc(u), c_loc(v), c_loc(w)))
  end subroutine vel
  subroutine test_one(d, which, label)
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), intent(in), target :: d
    integer, intent(in) :: which
    character(len=*), intent(in) :: label
    real(c_double) :: err
    if (.not. test_mode) return
    call x3d_check('test_err', x3d_test_err(x3d_ctx_h, c_loc(d), int(which, c_int), err))
    print *, "RMS error in "//label//": ", err
  end subroutine test_one
  subroutine test_du(du)
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), intent(in), target :: du
    call test_one(du, 0, 'dudx')
  end subroutine test_du
  subroutine test_dv(dv)
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), intent(in), target :: dv
    call test_one(dv, 1, 'dvdy')
  end subroutine test_dv
  subroutine test_dw(dw)
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), intent(in), target :: dw
    call test_one(dw, 2, 'dwdz')
  end subroutine test_dw
end module mom

!! init (src/case.f90:20-58): zero pressure and gradients, the reference's
!! velocity field, and the old right-hand sides set to it.
module case
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use decomp_2d, only : xsize, ph1
  use x3d_hip
  use param, only : ntime, npress
  use variables, only : nzm
  implicit none
  private
  public :: init
contains
  subroutine init(ux1, uy1, uz1, dux1, duy1, duz1, pp3, px1, py1, pz1)
    use mom, only : vel
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), target :: ux1, uy1, uz1
    real(mytype), dimension(xsize(1), xsize(2), xsize(3), ntime), target :: dux1, duy1, duz1
    real(mytype), dimension(ph1%zst(1):ph1%zen(1), ph1%zst(2):ph1%zen(2), nzm, npress), target :: pp3
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), target :: px1, py1, pz1
    integer(c_size_t) :: nv
    integer :: it
    nv = int(xsize(1), c_size_t) * xsize(2) * xsize(3)
    call x3d_check('x3d_fill', x3d_fill(x3d_ctx_h, c_loc(pp3), &
                   int(ph1%zen(1) - ph1%zst(1) + 1, c_size_t) * (ph1%zen(2) - ph1%zst(2) + 1) * nzm, 0._c_double))
    call x3d_check('x3d_fill', x3d_fill(x3d_ctx_h, c_loc(px1), nv, 0._c_double))
    call x3d_check('x3d_fill', x3d_fill(x3d_ctx_h, c_loc(py1), nv, 0._c_double))
    call x3d_check('x3d_fill', x3d_fill(x3d_ctx_h, c_loc(pz1), nv, 0._c_double))
    call vel(ux1, uy1, uz1)
    do it = 1, ntime
      call copy(ux1, dux1(:, :, :, it)); call copy(uy1, duy1(:, :, :, it)); call copy(uz1, duz1(:, :, :, it))
    end do
  end subroutine init
  subroutine copy(a, b)
    real(mytype), dimension(:, :, :), intent(in), target, contiguous :: a
    real(mytype), dimension(:, :, :), intent(out), target, contiguous :: b
    call x3d_check('copy', x3d_copy_field(x3d_ctx_h, c_loc(a), c_loc(b), 8_c_size_t, &
                   int(size(a, 1), c_int), int(size(a, 2), c_int), int(size(a, 3), c_int)))
  end subroutine copy
end module case

!! calculate_transeq_rhs (src/transeq.f90:18-34).  In test mode the three
!! collocated derivatives that momentum_rhs_eq checks (:107-109, :212-214,
!! :317-319) are formed in a scratch field and reported first.
module transeq
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use decomp_2d, only : xsize
  use x3d_hip
  use param, only : ntime
  use variables, only : test_mode
  use x3d_operator_1d, only : x3d_op_derx, x3d_op_dery, x3d_op_derz
  use mom, only : test_du, test_dv, test_dw
  implicit none
  private
  public :: calculate_transeq_rhs
  type(c_ptr), save :: scratch = c_null_ptr
contains
  subroutine calculate_transeq_rhs(dux1, duy1, duz1, ux1, uy1, uz1)
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), intent(in), target :: ux1, uy1, uz1
    real(mytype), dimension(xsize(1), xsize(2), xsize(3), ntime), target :: dux1, duy1, duz1
    real(mytype), dimension(:, :, :), pointer, contiguous :: t
    integer(c_int) :: n1, n2, n3
    if (test_mode) then
      n1 = xsize(1); n2 = xsize(2); n3 = xsize(3)
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

!! int_time, Euler: var1 = gdt(itr) dvar1(:,:,:,1) + var1
!! (src/time_integrators.f90:20-106).
module time_integrators
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use decomp_2d, only : xsize
  use x3d_hip
  use param, only : itr, ntime, gdt
  implicit none
  private
  public :: int_time
contains
  subroutine int_time(ux1, uy1, uz1, dux1, duy1, duz1)
    real(mytype), dimension(xsize(1), xsize(2), xsize(3), ntime), target :: dux1, duy1, duz1
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), target :: ux1, uy1, uz1
    call x3d_check('int_time', x3d_int_time(x3d_ctx_h, real(gdt(itr), c_double), c_loc(ux1), c_loc(uy1), &
                                            c_loc(uz1), c_loc(dux1), c_loc(duy1), c_loc(duz1)))
  end subroutine int_time
end module time_integrators

!! solve_poisson, divergence, gradp, cor_vel (src/navier.f90:21-359).
!! In test mode solve_poisson runs its three parts separately so that the
!! "DIV U* max mean" line of divergence (:213-231) can be printed.
module navier
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use decomp_2d, only : xsize, ph1
  use x3d_hip
  use param, only : npress
  use variables, only : test_mode, nzm
  use decomp_2d_poisson, only : poisson
  implicit none
  private
  public :: solve_poisson, divergence, gradp, cor_vel
contains
  subroutine solve_poisson(pp3, px1, py1, pz1, ux1, uy1, uz1)
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), intent(in), target :: ux1, uy1, uz1
    real(mytype), dimension(ph1%zst(1):ph1%zen(1), ph1%zst(2):ph1%zen(2), nzm, npress), target :: pp3
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), target :: px1, py1, pz1
    integer :: nlock
    nlock = 1
    if (test_mode) then
      call divergence(pp3(:, :, :, 1), ux1, uy1, uz1, nlock)
      call poisson(pp3(:, :, :, 1))
      call gradp(px1, py1, pz1, pp3(:, :, :, 1))
    else
      call x3d_check('solve_poisson', x3d_solve_poisson(x3d_ctx_h, c_loc(pp3), c_loc(px1), c_loc(py1), &
                                                        c_loc(pz1), c_loc(ux1), c_loc(uy1), c_loc(uz1)))
    end if
  end subroutine solve_poisson

  subroutine divergence(pp3, ux1, uy1, uz1, nlock)
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), intent(in), target :: ux1, uy1, uz1
    real(mytype), dimension(ph1%zst(1):ph1%zen(1), ph1%zst(2):ph1%zen(2), nzm), target :: pp3
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
    real(mytype), dimension(ph1%zst(1):ph1%zen(1), ph1%zst(2):ph1%zen(2), nzm), intent(in), target :: pp3
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), target :: px1, py1, pz1
    call x3d_check('gradp', x3d_gradp(x3d_ctx_h, c_loc(px1), c_loc(py1), c_loc(pz1), c_loc(pp3)))
  end subroutine gradp

  subroutine cor_vel(ux, uy, uz, px, py, pz)
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), target :: ux, uy, uz
    real(mytype), dimension(xsize(1), xsize(2), xsize(3)), intent(in), target :: px, py, pz
    call x3d_check('cor_vel', x3d_cor_vel(x3d_ctx_h, c_loc(ux), c_loc(uy), c_loc(uz), c_loc(px), c_loc(py), &
                                          c_loc(pz)))
  end subroutine cor_vel
end module navier

!! parameter() and listing() (src/parameters.f90:24-228): the periodic
!! defaults (Lx = Ly = Lz = 1, dt = 0, Euler), the derived sizes, and the
!! banner the reference prints at start-up.
module parameters
  use decomp_2d_constants, only : mytype
  use decomp_2d_mpi, only : nrank, nproc, decomp_2d_abort
  use param
  use variables
  implicit none
  private
  public :: parameter, listing
contains
  subroutine parameter()
#ifdef DEBUG
    if (nrank == 0) write (*, *) '# parameter start'
#endif
    ifirstder = 4; isecondder = 4; ipinter = 3
    itimescheme = 1; iimplicit = 0; istret = 0; iscalar = 0
    u1 = 2; u2 = 1
    dt = zero
    xlx = one; yly = one; zlz = one
    nclx1 = 0; nclxn = 0; ncly1 = 0; nclyn = 0; nclz1 = 0; nclzn = 0
    npress = 1
    if (nz == 1) then
      nclz1 = 0; nclzn = 0
      p_row = nproc; p_col = 1
    end if
    nclxS1 = nclx1; nclxSn = nclxn; nclyS1 = ncly1; nclySn = nclyn; nclzS1 = nclz1; nclzSn = nclzn
    nu0nu = four
    cnu = 0.44_mytype
    ! periodic in every direction: the pressure grid equals the velocity grid
    nclx = nclx1 == 0 .and. nclxn == 0
    ncly = ncly1 == 0 .and. nclyn == 0
    nclz = nclz1 == 0 .and. nclzn == 0
    nxm = merge(nx, nx - 1, nclx)
    nym = merge(ny, ny - 1, ncly)
    nzm = merge(nz, nz - 1, nclz)
    dx = xlx / real(nxm, mytype); dy = yly / real(nym, mytype); dz = zlz / real(nzm, mytype)
    dx2 = dx * dx; dy2 = dy * dy; dz2 = dz * dz
    xnu = zero
    if (abs(re) > epsilon(re)) xnu = one / re
    if (itimescheme > 1) call decomp_2d_abort(itimescheme, "itimescheme must be specified as 1-6")
  end subroutine parameter

  subroutine listing()
    use, intrinsic :: iso_fortran_env, only : compiler_version, compiler_options
    if (nrank /= 0) return
    print *, '==========================================================='
    print *, '========================x3div=============================='
    print *, '==========================================================='
    print *, '==========================================================='
    print *, '===============mini app of Xcompact3D======================'
    print *, '==========================================================='
    print *, '==========================================================='
    write (*, *) 'Git version        : unknown'
    write (*, "(' xnu                    : ',F17.8)") xnu
    print *, '==========================================================='
    write (*, "(' p_row, p_col           : ',I9, I8)") p_row, p_col
    print *, '==========================================================='
    write (*, "(' Time step dt           : ',F17.8)") dt
    if (itimescheme == 1) then
      write (*, "(' Temporal scheme        : ',A20)") "Forwards Euler"
    else
      print *, 'Error: itimescheme must be specified as 1-6'
      stop
    end if
    write (*, *) '==========================================================='
    write (*, "(' ifirst                 : ',I17)") ifirst
    write (*, "(' ilast                  : ',I17)") ilast
    write (*, *) '==========================================================='
    write (*, "(' Lx                     : ',F17.8)") xlx
    write (*, "(' Ly                     : ',F17.8)") yly
    write (*, "(' Lz                     : ',F17.8)") zlz
    write (*, "(' nx                     : ',I17)") nx
    write (*, "(' ny                     : ',I17)") ny
    write (*, "(' nz                     : ',I17)") nz
    write (*, *) '==========================================================='
    write (*, "(' nu0nu                  : ',F17.8)") nu0nu
    write (*, "(' cnu                    : ',F17.8)") cnu
    write (*, *) '==========================================================='
    write (*, "(' High and low speed : u1=',F6.2,' and u2=',F6.2)") u1, u2
    write (*, *) '==========================================================='
    write (*, *) 'Numerical precision: Double'
    write (*, *) 'Compiled with ', compiler_version()
    write (*, *) 'Compiler options : ', compiler_options()
    write (*, '(" Version of the MPI library : ",A)') 'none (one rank, libx3dhip on one GPU)'
#ifdef DEBUG
    write (*, *) 'Compile flag DEBUG detected'
#endif
    write (*, *) '==========================================================='
  end subroutine listing
end module parameters
