!! The x3div benchmark driver (src/xcompact3d.f90:5-145) on libx3dhip: the
!! same command line (nx ny nz p_row p_col ndt_max test_mode), start-up
!! banner, step loop through the reference's module names and "Elapse time"
!! lines.  The fields are device arrays of module var.  Outside test mode the
!! step's four calls are recorded once into the library's graph and each step
!! replays it; in test mode they run one by one, since they print host values.
program xcompact3d
  use, intrinsic :: iso_c_binding
  use var
  use decomp_2d_mpi, only : nrank
  use param, only : itr
  use transeq, only : calculate_transeq_rhs
  use navier, only : solve_poisson, cor_vel
  use time_integrators, only : int_time
  use variables, only : test_mode
  use x3d_hip, only : x3d_ctx_h, x3d_check, x3d_wtime, x3d_sync, x3d_graph_begin, x3d_graph_end, &
                      x3d_graph_launch
  implicit none

  double precision :: tstart, tend, telapsed, tmin, tmax
  integer :: ndt, ndt_max
  logical :: replay

  call boot_xcompact3d()
  call init_xcompact3d(ndt_max)

  telapsed = 0
  tmin = telapsed
  ndt = 1

  call init_flowfield()
  replay = .not. test_mode
  if (replay) then
    itr = 1
    call x3d_check('x3d_graph_begin', x3d_graph_begin(x3d_ctx_h))
    call calculate_transeq_rhs(dux1, duy1, duz1, ux1, uy1, uz1)
    call int_time(ux1, uy1, uz1, dux1, duy1, duz1)
    call solve_poisson(pp3, px1, py1, pz1, ux1, uy1, uz1)
    call cor_vel(ux1, uy1, uz1, px1, py1, pz1)
    call x3d_check('x3d_graph_end', x3d_graph_end(x3d_ctx_h))
  end if
  call x3d_check('x3d_sync', x3d_sync(x3d_ctx_h))

  do while (ndt <= ndt_max)
    itr = 1 ! no inner iterations

    tstart = x3d_wtime()
    if (replay) then
      call x3d_check('x3d_graph_launch', x3d_graph_launch(x3d_ctx_h))
    else
      call calculate_transeq_rhs(dux1, duy1, duz1, ux1, uy1, uz1)
      call int_time(ux1, uy1, uz1, dux1, duy1, duz1)
      call solve_poisson(pp3, px1, py1, pz1, ux1, uy1, uz1)
      call cor_vel(ux1, uy1, uz1, px1, py1, pz1)
    end if
    call x3d_check('x3d_sync', x3d_sync(x3d_ctx_h))
    tend = x3d_wtime()
    telapsed = telapsed + (tend - tstart)
    tmin = telapsed
    tmax = telapsed
    if (nrank == 0) print *, "Elapse time min ", tmin, " max ", tmax

    ndt = ndt + 1
  end do

  if (nrank == 0) then
    print *, "Elapsed time (min-max) [s]: ", tmin, tmax
    print *, "Timesteps completed: ", ndt
    print *, "Compute rate (min-max)[dt/s]: ", ndt / tmin, ndt / tmax
  end if

  call finalise_xcompact3d(.true.)

end program xcompact3d
