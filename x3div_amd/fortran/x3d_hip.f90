!! iso_c_binding interface to libx3dhip (include/x3dhip.h) and the modules
!! that keep x3div's public names and call signatures for the hot path:
!!   decomp_2d_constants, x3d_hip, x3d_operator_1d, thomas, x3d_derive,
!!   x3d_staggered, x3d_transpose, decomp_2d_poisson, transeq,
!!   time_integrators, navier.
!! Every field argument is a device array (allocated through x3d_field_alloc
!! and mapped with c_f_pointer); the host never dereferences it.  A nonzero
!! status from the library stops the program the way decomp_2d_abort does.

module decomp_2d_constants
  use, intrinsic :: iso_fortran_env, only : real64
  implicit none
  integer, parameter, public :: mytype = real64
  integer, parameter, public :: real_type = 0
end module decomp_2d_constants

module x3d_hip
  use, intrinsic :: iso_c_binding
  implicit none
  private :: c_abort
  type(c_ptr), save, public :: x3d_ctx_h = c_null_ptr

  interface
    integer(c_int) function x3d_create(nx, ny, nz, xlx, yly, zlz, ctx) bind(C, name="x3d_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nx, ny, nz
      real(c_double), value :: xlx, yly, zlz
      type(c_ptr) :: ctx
    end function
    integer(c_int) function x3d_destroy(ctx) bind(C, name="x3d_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function x3d_sync(ctx) bind(C, name="x3d_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    real(c_double) function x3d_wtime() bind(C, name="x3d_wtime")
      import :: c_double
    end function
    integer(c_int) function x3d_field_alloc(ctx, nbytes, p) bind(C, name="x3d_field_alloc")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_size_t), value :: nbytes
      type(c_ptr) :: p
    end function
    integer(c_int) function x3d_field_free(ctx, p) bind(C, name="x3d_field_free")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, p
    end function
    integer(c_int) function x3d_operator_get(ctx, op, f, s, w, p, alfa, n, npaire) bind(C, name="x3d_operator_get")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: op
      type(c_ptr), value :: f, s, w, p
      real(c_double) :: alfa
      integer(c_int) :: n, npaire
    end function
    integer(c_int) function x3d_xthomas(ctx, tt, op, nx, ny, nz) bind(C, name="x3d_xthomas")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, tt
      integer(c_int), value :: op, nx, ny, nz
    end function
    integer(c_int) function x3d_ythomas(ctx, tt, op, nx, ny, nz) bind(C, name="x3d_ythomas")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, tt
      integer(c_int), value :: op, nx, ny, nz
    end function
    integer(c_int) function x3d_zthomas(ctx, tt, op, nx, ny, nz) bind(C, name="x3d_zthomas")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, tt
      integer(c_int), value :: op, nx, ny, nz
    end function
    integer(c_int) function x3d_derx(ctx, t, u, op, nx, ny, nz) bind(C, name="x3d_derx")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u
      integer(c_int), value :: op, nx, ny, nz
    end function
    integer(c_int) function x3d_dery(ctx, t, u, op, ppy, nx, ny, nz) bind(C, name="x3d_dery")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u, ppy
      integer(c_int), value :: op, nx, ny, nz
    end function
    integer(c_int) function x3d_derz(ctx, t, u, op, nx, ny, nz) bind(C, name="x3d_derz")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u
      integer(c_int), value :: op, nx, ny, nz
    end function
    integer(c_int) function x3d_deryvp(ctx, t, u, op, pp, nx, ny, nym, nz) bind(C, name="x3d_deryvp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u, pp
      integer(c_int), value :: op, nx, ny, nym, nz
    end function
    integer(c_int) function x3d_derypv(ctx, t, u, op, pp, nx, nym, ny, nz) bind(C, name="x3d_derypv")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u, pp
      integer(c_int), value :: op, nx, nym, ny, nz
    end function
    integer(c_int) function x3d_poisson(ctx, rhs) bind(C, name="x3d_poisson")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, rhs
    end function
    integer(c_int) function x3d_poisson_tables(ctx, ax, bx, ay, by, az, bz, kre) bind(C, name="x3d_poisson_tables")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ax, bx, ay, by, az, bz, kre
    end function
    integer(c_int) function x3d_momentum_rhs(ctx, dux, duy, duz, ux, uy, uz) bind(C, name="x3d_momentum_rhs")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dux, duy, duz, ux, uy, uz
    end function
    integer(c_int) function x3d_int_time(ctx, gdt, ux, uy, uz, dux, duy, duz) bind(C, name="x3d_int_time")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, ux, uy, uz, dux, duy, duz
      real(c_double), value :: gdt
    end function
    integer(c_int) function x3d_solve_poisson(ctx, pp3, px, py, pz, ux, uy, uz) bind(C, name="x3d_solve_poisson")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, pp3, px, py, pz, ux, uy, uz
    end function
    integer(c_int) function x3d_divergence(ctx, pp3, ux, uy, uz) bind(C, name="x3d_divergence")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, pp3, ux, uy, uz
    end function
    integer(c_int) function x3d_gradp(ctx, px, py, pz, pp3) bind(C, name="x3d_gradp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, px, py, pz, pp3
    end function
    integer(c_int) function x3d_cor_vel(ctx, ux, uy, uz, px, py, pz) bind(C, name="x3d_cor_vel")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ux, uy, uz, px, py, pz
    end function
    integer(c_int) function x3d_test_err(ctx, d, which, err) bind(C, name="x3d_test_err")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, d
      integer(c_int), value :: which
      real(c_double) :: err
    end function
    integer(c_int) function x3d_div_stats(ctx, pp3, tmax, tmean) bind(C, name="x3d_div_stats")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, pp3
      real(c_double) :: tmax, tmean
    end function
    integer(c_int) function x3d_init_fields(ctx, kind, seed, amp, ux, uy, uz) bind(C, name="x3d_init_fields")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx, ux, uy, uz
      integer(c_int), value :: kind
      integer(c_int64_t), value :: seed
      real(c_double), value :: amp
    end function
    function x3d_strerror(code) bind(C, name="x3d_strerror") result(s)
      import :: c_int, c_ptr
      integer(c_int), value :: code
      type(c_ptr) :: s
    end function
  end interface

contains

  subroutine c_abort(where, code)
    character(len=*), intent(in) :: where
    integer(c_int), intent(in) :: code
    write (*, '(a,a,a,i0)') 'x3d_hip: ', where, ' failed with status ', code
    error stop 1
  end subroutine c_abort

  !> Abort on a nonzero library status (the decomp_2d_abort path).
  subroutine x3d_check(where, code)
    character(len=*), intent(in) :: where
    integer(c_int), intent(in) :: code
    if (code /= 0) call c_abort(where, code)
  end subroutine x3d_check

end module x3d_hip

!! The x3doperator1d type and its 18 instances (src/x3d_operator_1d.f90:12-47).
!! Components keep the reference's names; `id` is the library's operator id.
module x3d_operator_1d
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  implicit none
  private

  type, public :: x3doperator1d
    integer :: n = 0
    integer :: npaire = 0
    real(mytype) :: alfa = 0._mytype
    real(mytype), dimension(:), pointer :: f => null(), s => null(), w => null()
    real(mytype), dimension(:), allocatable :: periodic
    integer :: id = -1
  end type x3doperator1d

  type(x3doperator1d), save, public, target :: x3d_op_derx, x3d_op_derxp
  type(x3doperator1d), save, public, target :: x3d_op_dery, x3d_op_deryp
  type(x3doperator1d), save, public, target :: x3d_op_derz, x3d_op_derzp
  type(x3doperator1d), save, public, target :: x3d_op_derxvp, x3d_op_deryvp, x3d_op_derzvp
  type(x3doperator1d), save, public, target :: x3d_op_derxpv, x3d_op_derypv, x3d_op_derzpv
  type(x3doperator1d), save, public, target :: x3d_op_intxvp, x3d_op_intyvp, x3d_op_intzvp
  type(x3doperator1d), save, public, target :: x3d_op_intxpv, x3d_op_intypv, x3d_op_intzpv

  public :: x3d_operator_1d_init, x3d_operator_1d_finalize, x3d_op_by_id

contains

  !> x3d_operator_1d_init (src/x3d_operator_1d.f90:59-195): the coefficient
  !> vectors come from the library, which builds them on context creation.
  subroutine x3d_operator_1d_init()
    call fill(x3d_op_derx, 0);   call fill(x3d_op_derxp, 1)
    call fill(x3d_op_dery, 2);   call fill(x3d_op_deryp, 3)
    call fill(x3d_op_derz, 4);   call fill(x3d_op_derzp, 5)
    call fill(x3d_op_derxvp, 6); call fill(x3d_op_deryvp, 7);  call fill(x3d_op_derzvp, 8)
    call fill(x3d_op_derxpv, 9); call fill(x3d_op_derypv, 10); call fill(x3d_op_derzpv, 11)
    call fill(x3d_op_intxvp, 12); call fill(x3d_op_intyvp, 13); call fill(x3d_op_intzvp, 14)
    call fill(x3d_op_intxpv, 15); call fill(x3d_op_intypv, 16); call fill(x3d_op_intzpv, 17)
  end subroutine x3d_operator_1d_init

  subroutine fill(op, id)
    type(x3doperator1d), intent(inout) :: op
    integer, intent(in) :: id
    real(c_double), allocatable, target :: f(:), s(:), w(:), p(:)
    real(c_double) :: alfa
    integer(c_int) :: n, npaire
    ! the length first (null vectors are not written), then the vectors
    call x3d_check('x3d_operator_get', x3d_operator_get(x3d_ctx_h, int(id, c_int), c_null_ptr, c_null_ptr, &
                                                         c_null_ptr, c_null_ptr, alfa, n, npaire))
    allocate (f(n), s(n), w(n), p(n))
    call x3d_check('x3d_operator_get', x3d_operator_get(x3d_ctx_h, int(id, c_int), c_loc(f), c_loc(s), &
                                                         c_loc(w), c_loc(p), alfa, n, npaire))
    op%n = n; op%npaire = npaire; op%alfa = alfa; op%id = id
    allocate (op%f(n), op%s(n), op%w(n))
    op%f = f; op%s = s; op%w = w
    call move_alloc(p, op%periodic)
  end subroutine fill

  subroutine x3d_operator_1d_finalize()
    call drop(x3d_op_derx); call drop(x3d_op_derxp); call drop(x3d_op_dery); call drop(x3d_op_deryp)
    call drop(x3d_op_derz); call drop(x3d_op_derzp)
    call drop(x3d_op_derxvp); call drop(x3d_op_deryvp); call drop(x3d_op_derzvp)
    call drop(x3d_op_derxpv); call drop(x3d_op_derypv); call drop(x3d_op_derzpv)
    call drop(x3d_op_intxvp); call drop(x3d_op_intyvp); call drop(x3d_op_intzvp)
    call drop(x3d_op_intxpv); call drop(x3d_op_intypv); call drop(x3d_op_intzpv)
  end subroutine x3d_operator_1d_finalize

  subroutine drop(op)
    type(x3doperator1d), intent(inout) :: op
    if (associated(op%f)) deallocate (op%f, op%s, op%w)
    if (allocated(op%periodic)) deallocate (op%periodic)
    op%n = 0; op%npaire = 0; op%id = -1
  end subroutine drop

  !> The operator instance whose coefficient vector f is the given array
  !> (lets the Thomas generics keep the reference's array arguments).
  integer function x3d_op_by_id(ff) result(id)
    real(mytype), dimension(:), target, intent(in) :: ff
    type(c_ptr) :: a
    a = c_loc(ff(1))
    id = -1
    if (c_associated(a, c_loc(x3d_op_derx%f(1)))) id = 0
    if (c_associated(a, c_loc(x3d_op_derxp%f(1)))) id = 1
    if (c_associated(a, c_loc(x3d_op_dery%f(1)))) id = 2
    if (c_associated(a, c_loc(x3d_op_deryp%f(1)))) id = 3
    if (c_associated(a, c_loc(x3d_op_derz%f(1)))) id = 4
    if (c_associated(a, c_loc(x3d_op_derzp%f(1)))) id = 5
    if (c_associated(a, c_loc(x3d_op_derxvp%f(1)))) id = 6
    if (c_associated(a, c_loc(x3d_op_deryvp%f(1)))) id = 7
    if (c_associated(a, c_loc(x3d_op_derzvp%f(1)))) id = 8
    if (c_associated(a, c_loc(x3d_op_derxpv%f(1)))) id = 9
    if (c_associated(a, c_loc(x3d_op_derypv%f(1)))) id = 10
    if (c_associated(a, c_loc(x3d_op_derzpv%f(1)))) id = 11
    if (c_associated(a, c_loc(x3d_op_intxvp%f(1)))) id = 12
    if (c_associated(a, c_loc(x3d_op_intyvp%f(1)))) id = 13
    if (c_associated(a, c_loc(x3d_op_intzvp%f(1)))) id = 14
    if (c_associated(a, c_loc(x3d_op_intxpv%f(1)))) id = 15
    if (c_associated(a, c_loc(x3d_op_intypv%f(1)))) id = 16
    if (c_associated(a, c_loc(x3d_op_intzpv%f(1)))) id = 17
  end function x3d_op_by_id

end module x3d_operator_1d

!! Thomas generics (src/thomas.f90:15-28) with the periodic variants
!! xthomas_0/ythomas_0/zthomas_0 (:34-58, 87-111, 139-163).
module thomas
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  use x3d_operator_1d, only : x3d_op_by_id
  implicit none
  private
  public :: xthomas, ythomas, zthomas

  interface xthomas
    module procedure xthomas_0
  end interface xthomas
  interface ythomas
    module procedure ythomas_0
  end interface ythomas
  interface zthomas
    module procedure zthomas_0
  end interface zthomas

contains

  subroutine xthomas_0(tt, ff, fs, fw, perio, alfa, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(inout), dimension(nx, ny, nz), target :: tt
    real(mytype), intent(in), dimension(nx), target :: ff, fs, fw, perio
    real(mytype), intent(in) :: alfa
    call x3d_check('xthomas', x3d_xthomas(x3d_ctx_h, c_loc(tt), int(op_of(ff), c_int), nx, ny, nz))
  end subroutine xthomas_0

  subroutine ythomas_0(tt, ff, fs, fw, perio, alfa, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(inout), dimension(nx, ny, nz), target :: tt
    real(mytype), intent(in), dimension(ny), target :: ff, fs, fw, perio
    real(mytype), intent(in) :: alfa
    call x3d_check('ythomas', x3d_ythomas(x3d_ctx_h, c_loc(tt), int(op_of(ff), c_int), nx, ny, nz))
  end subroutine ythomas_0

  subroutine zthomas_0(tt, ff, fs, fw, perio, alfa, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(inout), dimension(nx, ny, nz), target :: tt
    real(mytype), intent(in), dimension(nz), target :: ff, fs, fw, perio
    real(mytype), intent(in) :: alfa
    call x3d_check('zthomas', x3d_zthomas(x3d_ctx_h, c_loc(tt), int(op_of(ff), c_int), nx, ny, nz))
  end subroutine zthomas_0

  integer function op_of(ff)
    real(mytype), dimension(:), target, intent(in) :: ff
    op_of = x3d_op_by_id(ff)
    if (op_of < 0) call x3d_check('thomas: coefficient vectors of no known operator', -1_c_int)
  end function op_of

end module thomas

!! Derivative procedure pointers (src/x3d_derive.f90:18-109), periodic derx_00,
!! dery_00, derz_00 (:111-152, :279-339, :503-558).
module x3d_derive
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  use x3d_operator_1d, only : x3doperator1d
  implicit none
  public

  abstract interface
    subroutine DERIVATIVE_X(t, u, x3dop, nx, ny, nz)
      import :: mytype, x3doperator1d
      integer, intent(in) :: nx, ny, nz
      real(mytype), intent(out), dimension(nx, ny, nz), target :: t
      real(mytype), intent(in), dimension(nx, ny, nz), target :: u
      type(x3doperator1d), intent(in) :: x3dop
    end subroutine DERIVATIVE_X
    subroutine DERIVATIVE_Y(t, u, x3dop, pp, nx, ny, nz)
      import :: mytype, x3doperator1d
      integer, intent(in) :: nx, ny, nz
      real(mytype), intent(out), dimension(nx, ny, nz), target :: t
      real(mytype), intent(in), dimension(nx, ny, nz), target :: u
      real(mytype), intent(in), dimension(ny) :: pp
      type(x3doperator1d), intent(in) :: x3dop
    end subroutine DERIVATIVE_Y
    subroutine DERIVATIVE_Z(t, u, x3dop, nx, ny, nz)
      import :: mytype, x3doperator1d
      integer, intent(in) :: nx, ny, nz
      real(mytype), intent(out), dimension(nx, ny, nz), target :: t
      real(mytype), intent(in), dimension(nx, ny, nz), target :: u
      type(x3doperator1d), intent(in) :: x3dop
    end subroutine DERIVATIVE_Z
  end interface

  procedure(DERIVATIVE_X), pointer :: derx => null()
  procedure(DERIVATIVE_Y), pointer :: dery => null()
  procedure(DERIVATIVE_Z), pointer :: derz => null()

contains

  !> x3d_derive_init (src/x3d_derive.f90:66-93), periodic BCs.
  subroutine x3d_derive_init()
    derx => derx_00
    dery => dery_00
    derz => derz_00
  end subroutine x3d_derive_init

  subroutine x3d_derive_finalize()
    nullify (derx, dery, derz)
  end subroutine x3d_derive_finalize

  subroutine derx_00(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('derx', x3d_derx(x3d_ctx_h, c_loc(t), c_loc(u), int(x3dop%id, c_int), nx, ny, nz))
  end subroutine derx_00

  subroutine dery_00(t, u, x3dop, pp, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    real(mytype), intent(in), dimension(ny) :: pp
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('dery', x3d_dery(x3d_ctx_h, c_loc(t), c_loc(u), int(x3dop%id, c_int), c_null_ptr, nx, ny, nz))
  end subroutine dery_00

  subroutine derz_00(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('derz', x3d_derz(x3d_ctx_h, c_loc(t), c_loc(u), int(x3dop%id, c_int), nx, ny, nz))
  end subroutine derz_00

end module x3d_derive
