!! iso_c_binding interface to libx3dhip (include/x3dhip.h): modules
!! decomp_2d_constants and x3d_hip.  The modules that keep x3div's public
!! names and call signatures are built on it in decomp_2d.f90,
!! x3d_hip_ops.f90 and x3d_hip_step.f90.
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
    integer(c_int) function x3d_create_bc(nx, ny, nz, xlx, yly, zlz, ncl1, ncln, ctx) bind(C, name="x3d_create_bc")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nx, ny, nz
      real(c_double), value :: xlx, yly, zlz
      integer(c_int), intent(in) :: ncl1(3), ncln(3)
      type(c_ptr) :: ctx
    end function
    integer(c_int) function x3d_set_stretching(ctx, istret, ppy, ppyi) bind(C, name="x3d_set_stretching")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, ppy, ppyi
      integer(c_int), value :: istret
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
    integer(c_int) function x3d_memcpy_d2h(ctx, dst, src, nbytes) bind(C, name="x3d_memcpy_d2h")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, dst, src
      integer(c_size_t), value :: nbytes
    end function
    integer(c_int) function x3d_memcpy_h2d(ctx, dst, src, nbytes) bind(C, name="x3d_memcpy_h2d")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, dst, src
      integer(c_size_t), value :: nbytes
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
    integer(c_int) function x3d_derxx(ctx, t, u, op, nx, ny, nz) bind(C, name="x3d_derxx")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u
      integer(c_int), value :: op, nx, ny, nz
    end function
    integer(c_int) function x3d_deryy(ctx, t, u, op, nx, ny, nz) bind(C, name="x3d_deryy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t, u
      integer(c_int), value :: op, nx, ny, nz
    end function
    integer(c_int) function x3d_derzz(ctx, t, u, op, nx, ny, nz) bind(C, name="x3d_derzz")
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
    integer(c_int) function x3d_xthomas_tab(ctx, tt, ff, fs, fw, perio, alfa, nx, ny, nz) &
        bind(C, name="x3d_xthomas_tab")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, tt, ff, fs, fw, perio
      real(c_double), value :: alfa
      integer(c_int), value :: nx, ny, nz
    end function
    integer(c_int) function x3d_ythomas_tab(ctx, tt, ff, fs, fw, perio, alfa, nx, ny, nz) &
        bind(C, name="x3d_ythomas_tab")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, tt, ff, fs, fw, perio
      real(c_double), value :: alfa
      integer(c_int), value :: nx, ny, nz
    end function
    integer(c_int) function x3d_zthomas_tab(ctx, tt, ff, fs, fw, perio, alfa, nx, ny, nz) &
        bind(C, name="x3d_zthomas_tab")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, tt, ff, fs, fw, perio
      real(c_double), value :: alfa
      integer(c_int), value :: nx, ny, nz
    end function
    integer(c_int) function x3d_thomas1d(tt, ff, fs, fw, n) bind(C, name="x3d_thomas1d")
      import :: c_int, c_ptr
      type(c_ptr), value :: tt, ff, fs, fw
      integer(c_int), value :: n
    end function
    integer(c_int) function x3d_operator_set(ctx, op, f, s, w, p, alfa) bind(C, name="x3d_operator_set")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, f, s, w, p
      integer(c_int), value :: op
      real(c_double), value :: alfa
    end function
    integer(c_int) function x3d_copy_field(ctx, a, b, eb, n1, n2, n3) bind(C, name="x3d_transpose_x_to_y")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, a, b
      integer(c_size_t), value :: eb
      integer(c_int), value :: n1, n2, n3
    end function
    integer(c_int) function x3d_fill(ctx, p, n, v) bind(C, name="x3d_fill")
      import :: c_int, c_ptr, c_size_t, c_double
      type(c_ptr), value :: ctx, p
      integer(c_size_t), value :: n
      real(c_double), value :: v
    end function
    integer(c_int) function x3d_graph_begin(ctx) bind(C, name="x3d_graph_begin")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function x3d_graph_end(ctx) bind(C, name="x3d_graph_end")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function x3d_graph_launch(ctx) bind(C, name="x3d_graph_launch")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
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

