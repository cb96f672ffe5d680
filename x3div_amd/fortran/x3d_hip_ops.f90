!! The operator modules of the hot path with the reference's public names and
!! call signatures, on libx3dhip: x3d_operator_1d, thomas, x3d_derive,
!! x3d_staggered, x3d_transpose, decomp_2d_poisson.  Field arguments are device
!! arrays; coefficient vectors (x3doperator1d components, Thomas tables) are
!! host arrays, as in the reference.

!! The x3doperator1d type and its 18 instances (src/x3d_operator_1d.f90:12-47).
!! Components keep the reference's names; `id` is the library's operator id
!! of a module instance (-1 for a caller's own object).
module x3d_operator_1d
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use decomp_2d_mpi, only : nrank
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
    logical :: owned = .false.   ! f, s, w allocated here (x3d_operator_1d_init)
  contains
    procedure, public :: init
    procedure, public :: finalize
  end type x3doperator1d

  type(x3doperator1d), save, public, target :: x3d_op_derx, x3d_op_derxp
  type(x3doperator1d), save, public, target :: x3d_op_dery, x3d_op_deryp
  type(x3doperator1d), save, public, target :: x3d_op_derz, x3d_op_derzp
  type(x3doperator1d), save, public, target :: x3d_op_derxvp, x3d_op_deryvp, x3d_op_derzvp
  type(x3doperator1d), save, public, target :: x3d_op_derxpv, x3d_op_derypv, x3d_op_derzpv
  type(x3doperator1d), save, public, target :: x3d_op_intxvp, x3d_op_intyvp, x3d_op_intzvp
  type(x3doperator1d), save, public, target :: x3d_op_intxpv, x3d_op_intypv, x3d_op_intzpv
  ! second derivatives for derxx / deryy / derzz: this build's instances (the
  ! reference declares the procedure pointers, src/x3d_derive.f90:54-59, and
  ! no operator for them)
  type(x3doperator1d), save, public, target :: x3d_op_derxx, x3d_op_deryy, x3d_op_derzz

  public :: x3d_operator_1d_init, x3d_operator_1d_finalize, x3d_op_by_id

contains

  !> x3d_operator_1d_init (src/x3d_operator_1d.f90:59-195): the coefficient
  !> vectors come from the library, which builds them on context creation.
  subroutine x3d_operator_1d_init()
#ifdef DEBUG
    if (nrank == 0) write (*, *) '# x3d_operator_1d_init start'
#endif
    call fill(x3d_op_derx, 0);   call fill(x3d_op_derxp, 1)
    call fill(x3d_op_dery, 2);   call fill(x3d_op_deryp, 3)
    call fill(x3d_op_derz, 4);   call fill(x3d_op_derzp, 5)
    call fill(x3d_op_derxvp, 6); call fill(x3d_op_deryvp, 7);  call fill(x3d_op_derzvp, 8)
    call fill(x3d_op_derxpv, 9); call fill(x3d_op_derypv, 10); call fill(x3d_op_derzpv, 11)
    call fill(x3d_op_intxvp, 12); call fill(x3d_op_intyvp, 13); call fill(x3d_op_intzvp, 14)
    call fill(x3d_op_intxpv, 15); call fill(x3d_op_intypv, 16); call fill(x3d_op_intzpv, 17)
    call fill(x3d_op_derxx, 18); call fill(x3d_op_deryy, 19); call fill(x3d_op_derzz, 20)
#ifdef DEBUG
    if (nrank == 0) write (*, *) '# x3d_operator_1d_init done'
#endif
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
    call op%finalize()
    op%n = n; op%npaire = npaire; op%alfa = alfa; op%id = id
    allocate (op%f(n), op%s(n), op%w(n))
    op%owned = .true.
    op%f = f; op%s = s; op%w = w
    call move_alloc(p, op%periodic)
  end subroutine fill

  subroutine x3d_operator_1d_finalize()
    call x3d_op_derx%finalize(); call x3d_op_derxp%finalize()
    call x3d_op_dery%finalize(); call x3d_op_deryp%finalize()
    call x3d_op_derz%finalize(); call x3d_op_derzp%finalize()
    call x3d_op_derxvp%finalize(); call x3d_op_deryvp%finalize(); call x3d_op_derzvp%finalize()
    call x3d_op_derxpv%finalize(); call x3d_op_derypv%finalize(); call x3d_op_derzpv%finalize()
    call x3d_op_intxvp%finalize(); call x3d_op_intyvp%finalize(); call x3d_op_intzvp%finalize()
    call x3d_op_intxpv%finalize(); call x3d_op_intypv%finalize(); call x3d_op_intzpv%finalize()
    call x3d_op_derxx%finalize(); call x3d_op_deryy%finalize(); call x3d_op_derzz%finalize()
  end subroutine x3d_operator_1d_finalize

  !> x3doperator1d%init (src/x3d_operator_1d.f90:234-266): point f, s, w at
  !> the caller's vectors, and for a periodic direction (ncl) form the
  !> periodic vector thomas1d((-1, 0, ..., 0, alfa)).  On one of the 18 module
  !> instances the tables also go to the library (x3d_operator_set), which
  !> keeps its line solver when they are its own and otherwise solves that
  !> operator with the caller's tables.
  subroutine init(x3dop, f, s, w, n, ncl, alfa, paire)
    class(x3doperator1d), target :: x3dop
    real(mytype), dimension(:), target, intent(in) :: f, s, w
    integer, intent(in) :: n, paire
    logical, intent(in) :: ncl
    real(mytype), intent(in) :: alfa
    real(c_double), allocatable, target :: per(:)
    if (n == 1) return
    call x3dop%finalize()
    x3dop%id = instance_id(x3dop)
    x3dop%n = n
    x3dop%npaire = paire
    x3dop%f => f
    x3dop%s => s
    x3dop%w => w
    x3dop%alfa = alfa
    if (ncl) then
      allocate (per(n))
      per = 0._mytype
      per(1) = -1._mytype
      per(n) = alfa
      call x3d_check('thomas1d', x3d_thomas1d(c_loc(per), c_loc(x3dop%f), c_loc(x3dop%s), c_loc(x3dop%w), &
                                              int(n, c_int)))
      call move_alloc(per, x3dop%periodic)
      if (x3dop%id >= 0 .and. c_associated(x3d_ctx_h)) &
        call x3d_check('x3d_operator_set', x3d_operator_set(x3d_ctx_h, int(x3dop%id, c_int), c_loc(x3dop%f), &
                                                             c_loc(x3dop%s), c_loc(x3dop%w), &
                                                             c_loc(x3dop%periodic), real(alfa, c_double)))
    end if
  end subroutine init

  !> x3doperator1d%finalize (src/x3d_operator_1d.f90:271-283).
  subroutine finalize(x3dop)
    class(x3doperator1d) :: x3dop
    if (x3dop%owned .and. associated(x3dop%f)) deallocate (x3dop%f, x3dop%s, x3dop%w)
    nullify (x3dop%f, x3dop%s, x3dop%w)
    x3dop%owned = .false.
    x3dop%n = 0
    x3dop%npaire = 0
    if (allocated(x3dop%periodic)) deallocate (x3dop%periodic)
  end subroutine finalize

  !> The library id of a module instance, by identity; -1 for other objects.
  integer function instance_id(x3dop) result(id)
    class(x3doperator1d), target, intent(in) :: x3dop
    type(c_ptr) :: a
    integer :: k
    a = c_loc(x3dop%n)
    id = -1
    do k = 0, 20
      if (c_associated(a, c_loc(instance_n(k)))) id = k
    end do
  end function instance_id

  !> The operator instance whose coefficient vector f (and periodic vector)
  !> are the given arrays, or -1: lets the Thomas generics take the library's
  !> device tables when called with an instance's own components.
  integer function x3d_op_by_id(ff, perio) result(id)
    real(mytype), dimension(:), target, intent(in) :: ff
    real(mytype), dimension(:), target, intent(in), optional :: perio
    type(x3doperator1d), pointer :: o
    integer :: k
    id = -1
    do k = 0, 20
      o => instance(k)
      if (.not. associated(o%f) .or. .not. allocated(o%periodic)) cycle
      if (size(o%f) /= size(ff)) cycle
      if (.not. c_associated(c_loc(ff(1)), c_loc(o%f(1)))) cycle
      if (present(perio)) then
        if (.not. c_associated(c_loc(perio(1)), c_loc(o%periodic(1)))) cycle
      end if
      id = k
      return
    end do
  end function x3d_op_by_id

  function instance_n(k) result(p)
    integer, intent(in) :: k
    integer, pointer :: p
    type(x3doperator1d), pointer :: o
    o => instance(k)
    p => o%n
  end function instance_n

  function instance(k) result(o)
    integer, intent(in) :: k
    type(x3doperator1d), pointer :: o
    select case (k)
    case (0); o => x3d_op_derx
    case (1); o => x3d_op_derxp
    case (2); o => x3d_op_dery
    case (3); o => x3d_op_deryp
    case (4); o => x3d_op_derz
    case (5); o => x3d_op_derzp
    case (6); o => x3d_op_derxvp
    case (7); o => x3d_op_deryvp
    case (8); o => x3d_op_derzvp
    case (9); o => x3d_op_derxpv
    case (10); o => x3d_op_derypv
    case (11); o => x3d_op_derzpv
    case (12); o => x3d_op_intxvp
    case (13); o => x3d_op_intyvp
    case (14); o => x3d_op_intzvp
    case (15); o => x3d_op_intxpv
    case (16); o => x3d_op_intypv
    case (17); o => x3d_op_intzpv
    case (18); o => x3d_op_derxx
    case (19); o => x3d_op_deryy
    case default; o => x3d_op_derzz
    end select
  end function instance

end module x3d_operator_1d

!! Thomas generics (src/thomas.f90:15-28): the periodic members
!! xthomas_0/ythomas_0/zthomas_0 (:34-58, 87-111, 139-163), the plain sweeps
!! xthomas_12/ythomas_12/zthomas_12 (:61-84, 114-136, 166-188) and thomas1d
!! (:200-218).  Called with an operator instance's own components they use
!! the library's device tables; with any other vectors they send those.
module thomas
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  use x3d_operator_1d, only : x3d_op_by_id
  implicit none
  private
  public :: xthomas, ythomas, zthomas, thomas1d

  interface xthomas
    module procedure xthomas_0
    module procedure xthomas_12
  end interface xthomas
  interface ythomas
    module procedure ythomas_0
    module procedure ythomas_12
  end interface ythomas
  interface zthomas
    module procedure zthomas_0
    module procedure zthomas_12
  end interface zthomas

contains

  subroutine xthomas_0(tt, ff, fs, fw, perio, alfa, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(inout), dimension(nx, ny, nz), target :: tt
    real(mytype), intent(in), dimension(nx), target :: ff, fs, fw, perio
    real(mytype), intent(in) :: alfa
    integer :: id
    id = x3d_op_by_id(ff, perio)
    if (id >= 0) then
      call x3d_check('xthomas', x3d_xthomas(x3d_ctx_h, c_loc(tt), int(id, c_int), nx, ny, nz))
    else
      call x3d_check('xthomas', x3d_xthomas_tab(x3d_ctx_h, c_loc(tt), c_loc(ff), c_loc(fs), c_loc(fw), &
                                                c_loc(perio), real(alfa, c_double), nx, ny, nz))
    end if
  end subroutine xthomas_0

  subroutine xthomas_12(tt, ff, fs, fw, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(inout), dimension(nx, ny, nz), target :: tt
    real(mytype), intent(in), dimension(nx), target :: ff, fs, fw
    call x3d_check('xthomas', x3d_xthomas_tab(x3d_ctx_h, c_loc(tt), c_loc(ff), c_loc(fs), c_loc(fw), &
                                              c_null_ptr, 0._c_double, nx, ny, nz))
  end subroutine xthomas_12

  subroutine ythomas_0(tt, ff, fs, fw, perio, alfa, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(inout), dimension(nx, ny, nz), target :: tt
    real(mytype), intent(in), dimension(ny), target :: ff, fs, fw, perio
    real(mytype), intent(in) :: alfa
    integer :: id
    id = x3d_op_by_id(ff, perio)
    if (id >= 0) then
      call x3d_check('ythomas', x3d_ythomas(x3d_ctx_h, c_loc(tt), int(id, c_int), nx, ny, nz))
    else
      call x3d_check('ythomas', x3d_ythomas_tab(x3d_ctx_h, c_loc(tt), c_loc(ff), c_loc(fs), c_loc(fw), &
                                                c_loc(perio), real(alfa, c_double), nx, ny, nz))
    end if
  end subroutine ythomas_0

  subroutine ythomas_12(tt, ff, fs, fw, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(inout), dimension(nx, ny, nz), target :: tt
    real(mytype), intent(in), dimension(ny), target :: ff, fs, fw
    call x3d_check('ythomas', x3d_ythomas_tab(x3d_ctx_h, c_loc(tt), c_loc(ff), c_loc(fs), c_loc(fw), &
                                              c_null_ptr, 0._c_double, nx, ny, nz))
  end subroutine ythomas_12

  subroutine zthomas_0(tt, ff, fs, fw, perio, alfa, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(inout), dimension(nx, ny, nz), target :: tt
    real(mytype), intent(in), dimension(nz), target :: ff, fs, fw, perio
    real(mytype), intent(in) :: alfa
    integer :: id
    id = x3d_op_by_id(ff, perio)
    if (id >= 0) then
      call x3d_check('zthomas', x3d_zthomas(x3d_ctx_h, c_loc(tt), int(id, c_int), nx, ny, nz))
    else
      call x3d_check('zthomas', x3d_zthomas_tab(x3d_ctx_h, c_loc(tt), c_loc(ff), c_loc(fs), c_loc(fw), &
                                                c_loc(perio), real(alfa, c_double), nx, ny, nz))
    end if
  end subroutine zthomas_0

  subroutine zthomas_12(tt, ff, fs, fw, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(inout), dimension(nx, ny, nz), target :: tt
    real(mytype), intent(in), dimension(nz), target :: ff, fs, fw
    call x3d_check('zthomas', x3d_zthomas_tab(x3d_ctx_h, c_loc(tt), c_loc(ff), c_loc(fs), c_loc(fw), &
                                              c_null_ptr, 0._c_double, nx, ny, nz))
  end subroutine zthomas_12

  !> thomas1d on a host vector (the library's host routine).
  pure subroutine thomas1d(tt, ff, fs, fw, nn)
    integer, intent(in) :: nn
    real(mytype), intent(inout), dimension(nn) :: tt
    real(mytype), intent(in), dimension(nn) :: ff, fs, fw
    integer :: k
    do k = 2, nn
      tt(k) = tt(k) - tt(k - 1) * fs(k)
    end do
    tt(nn) = tt(nn) * fw(nn)
    do k = nn - 1, 1, -1
      tt(k) = (tt(k) - ff(k) * tt(k + 1)) * fw(k)
    end do
  end subroutine thomas1d

end module thomas

!! Derivative procedure pointers (src/x3d_derive.f90:18-109), periodic derx_00,
!! dery_00, derz_00 (:111-152, :279-339, :503-558) and the non-periodic
!! der?_11/12/21/22 (:219-275, :386-499, :608-725).
module x3d_derive
  use, intrinsic :: iso_c_binding
  use decomp_2d_constants, only : mytype
  use x3d_hip
  use x3d_operator_1d, only : x3doperator1d
  implicit none
  private
  public :: DERIVATIVE_X, DERIVATIVE_Y, DERIVATIVE_YY, DERIVATIVE_Z, derx, dery, derz
  public :: derxx, deryy, derzz
  public :: x3d_derive_init, x3d_derive_finalize, derx_00, dery_00, derz_00
  public :: derxx_00, deryy_00, derzz_00
  public :: derx_11, derx_12, derx_21, derx_22, dery_11, dery_12, dery_21, dery_22
  public :: derz_11, derz_12, derz_21, derz_22

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
    subroutine DERIVATIVE_YY(t, u, x3dop, nx, ny, nz)
      import :: mytype, x3doperator1d
      integer, intent(in) :: nx, ny, nz
      real(mytype), intent(out), dimension(nx, ny, nz), target :: t
      real(mytype), intent(in), dimension(nx, ny, nz), target :: u
      type(x3doperator1d), intent(in) :: x3dop
    end subroutine DERIVATIVE_YY
    subroutine DERIVATIVE_Z(t, u, x3dop, nx, ny, nz)
      import :: mytype, x3doperator1d
      integer, intent(in) :: nx, ny, nz
      real(mytype), intent(out), dimension(nx, ny, nz), target :: t
      real(mytype), intent(in), dimension(nx, ny, nz), target :: u
      type(x3doperator1d), intent(in) :: x3dop
    end subroutine DERIVATIVE_Z
  end interface

  procedure(DERIVATIVE_X), pointer :: derx => null(), derxx => null()
  procedure(DERIVATIVE_Y), pointer :: dery => null()
  procedure(DERIVATIVE_YY), pointer :: deryy => null()
  procedure(DERIVATIVE_Z), pointer :: derz => null(), derzz => null()

contains

  !> x3d_derive_init (src/x3d_derive.f90:66-93): the pointers chosen by the
  !> boundary codes of module param.
  subroutine x3d_derive_init()
    use param, only : nclx1, nclxn, ncly1, nclyn, nclz1, nclzn
    if (nclx1 == 0 .and. nclxn == 0) derx => derx_00
    if (nclx1 == 1 .and. nclxn == 1) derx => derx_11
    if (nclx1 == 1 .and. nclxn == 2) derx => derx_12
    if (nclx1 == 2 .and. nclxn == 1) derx => derx_21
    if (nclx1 == 2 .and. nclxn == 2) derx => derx_22
    if (ncly1 == 0 .and. nclyn == 0) dery => dery_00
    if (ncly1 == 1 .and. nclyn == 1) dery => dery_11
    if (ncly1 == 1 .and. nclyn == 2) dery => dery_12
    if (ncly1 == 2 .and. nclyn == 1) dery => dery_21
    if (ncly1 == 2 .and. nclyn == 2) dery => dery_22
    if (nclz1 == 0 .and. nclzn == 0) derz => derz_00
    if (nclz1 == 1 .and. nclzn == 1) derz => derz_11
    if (nclz1 == 1 .and. nclzn == 2) derz => derz_12
    if (nclz1 == 2 .and. nclzn == 1) derz => derz_21
    if (nclz1 == 2 .and. nclzn == 2) derz => derz_22
    ! second derivatives: the reference leaves these pointers unassociated
    derxx => derxx_00
    deryy => deryy_00
    derzz => derzz_00
  end subroutine x3d_derive_init

  subroutine x3d_derive_finalize()
    nullify (derx, dery, derz, derxx, deryy, derzz)
  end subroutine x3d_derive_finalize

  !> derxx / deryy / derzz, periodic: the sixth-order compact second derivative
  !> (x3d_op_derxx / deryy / derzz) on the library's line kernels.
  subroutine derxx_00(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('derxx', x3d_derxx(x3d_ctx_h, c_loc(t), c_loc(u), int(x3dop%id, c_int), nx, ny, nz))
  end subroutine derxx_00

  subroutine deryy_00(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('deryy', x3d_deryy(x3d_ctx_h, c_loc(t), c_loc(u), int(x3dop%id, c_int), nx, ny, nz))
  end subroutine deryy_00

  subroutine derzz_00(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('derzz', x3d_derzz(x3d_ctx_h, c_loc(t), c_loc(u), int(x3dop%id, c_int), nx, ny, nz))
  end subroutine derzz_00

  subroutine derx_00(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('derx', x3d_derx(x3d_ctx_h, c_loc(t), c_loc(u), int(x3dop%id, c_int), nx, ny, nz))
  end subroutine derx_00

  subroutine dery_00(t, u, x3dop, pp, nx, ny, nz)
    use param, only : istret
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    real(mytype), intent(in), dimension(ny) :: pp
    type(x3doperator1d), intent(in) :: x3dop
    real(mytype), allocatable, target :: v(:)
    type(c_ptr) :: pv
    ! pp scales the result when istret /= 0 (src/x3d_derive.f90:331-337); the
    ! DERIVATIVE_Y interface gives it no TARGET, so its values go through a copy
    pv = c_null_ptr
    if (istret /= 0) then
      v = pp
      pv = c_loc(v)
    end if
    call x3d_check('dery', x3d_dery(x3d_ctx_h, c_loc(t), c_loc(u), int(x3dop%id, c_int), pv, nx, ny, nz))
  end subroutine dery_00

  subroutine derz_00(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call x3d_check('derz', x3d_derz(x3d_ctx_h, c_loc(t), c_loc(u), int(x3dop%id, c_int), nx, ny, nz))
  end subroutine derz_00

  !> der?_11/12/21/22: the context was made with these codes
  !> (x3d_create_bc), so the library's derivative entry points run the
  !> non-periodic rows; the routines keep the reference's names.
  subroutine derx_11(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call derx_00(t, u, x3dop, nx, ny, nz)
  end subroutine derx_11
  subroutine derx_12(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call derx_00(t, u, x3dop, nx, ny, nz)
  end subroutine derx_12
  subroutine derx_21(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call derx_00(t, u, x3dop, nx, ny, nz)
  end subroutine derx_21
  subroutine derx_22(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call derx_00(t, u, x3dop, nx, ny, nz)
  end subroutine derx_22

  subroutine dery_11(t, u, x3dop, pp, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    real(mytype), intent(in), dimension(ny) :: pp
    type(x3doperator1d), intent(in) :: x3dop
    call dery_00(t, u, x3dop, pp, nx, ny, nz)
  end subroutine dery_11
  subroutine dery_12(t, u, x3dop, pp, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    real(mytype), intent(in), dimension(ny) :: pp
    type(x3doperator1d), intent(in) :: x3dop
    call dery_00(t, u, x3dop, pp, nx, ny, nz)
  end subroutine dery_12
  subroutine dery_21(t, u, x3dop, pp, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    real(mytype), intent(in), dimension(ny) :: pp
    type(x3doperator1d), intent(in) :: x3dop
    call dery_00(t, u, x3dop, pp, nx, ny, nz)
  end subroutine dery_21
  subroutine dery_22(t, u, x3dop, pp, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    real(mytype), intent(in), dimension(ny) :: pp
    type(x3doperator1d), intent(in) :: x3dop
    call dery_00(t, u, x3dop, pp, nx, ny, nz)
  end subroutine dery_22

  subroutine derz_11(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call derz_00(t, u, x3dop, nx, ny, nz)
  end subroutine derz_11
  subroutine derz_12(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call derz_00(t, u, x3dop, nx, ny, nz)
  end subroutine derz_12
  subroutine derz_21(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call derz_00(t, u, x3dop, nx, ny, nz)
  end subroutine derz_21
  subroutine derz_22(t, u, x3dop, nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    real(mytype), intent(out), dimension(nx, ny, nz), target :: t
    real(mytype), intent(in), dimension(nx, ny, nz), target :: u
    type(x3doperator1d), intent(in) :: x3dop
    call derz_00(t, u, x3dop, nx, ny, nz)
  end subroutine derz_22

end module x3d_derive
