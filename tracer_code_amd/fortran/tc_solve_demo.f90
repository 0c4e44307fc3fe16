! A Fortran caller of the C-ABI shaped like the reference's start-up and
! first solve (main.f90:121-183, restart=1): it loads T and the tracers,
! bins them, filters T and C, solves for SF and computes the velocities,
! then advances one time step on the device.
!
!   tc_solve_demo <input.bin> <output.bin>
!
! input.bin (stream): tc_params bytes, ntr (int64), T(-2:nx+2,-2:nz+2),
! rtr(ntr), str(ntr), Ttype(ntr) (int32).
! output.bin (stream): q (int32), resid, resref, SF, u, w after the solve,
! then SF and rtr, str after one tc_time_step.
program tc_solve_demo
  use iso_c_binding
  use tracer_hip
  implicit none
  type(tc_params) :: p
  type(c_ptr) :: ctx
  integer(c_int64_t) :: ntr
  real(c_double), allocatable, target :: T(:,:), SF(:,:), u(:,:), w(:,:)
  real(c_double), allocatable :: rtr(:), str(:)
  integer(c_int32_t), allocatable :: ty(:)
  integer(c_int8_t), allocatable :: raw(:)
  integer(c_int32_t) :: q, n_temp, n_comp(TC_MAX_TYPES)
  real(c_double) :: resid, resref
  character(len=512) :: fin, fout
  integer :: ui, uo, nx, nz

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=ui, file=trim(fin), access='stream', form='unformatted', status='old')
  ! the C struct bytes, padding included (a derived-type read would skip it)
  allocate(raw(c_sizeof(p)))
  read(ui) raw, ntr
  p = transfer(raw, p)
  nx = p%nx; nz = p%nz
  allocate(T(-2:nx+2,-2:nz+2), SF(-2:nx+2,-2:nz+2), u(-2:nx+2,-2:nz+2), w(-2:nx+2,-2:nz+2))
  allocate(rtr(ntr), str(ntr), ty(ntr))
  read(ui) T, rtr, str, ty
  close(ui)

  call check(tc_init(p, ntr, ctx), 'tc_init')
  SF = 0.0d0
  call check(tc_put_field(ctx, TC_F_T, c_loc(T)), 'put T')
  call check(tc_put_field(ctx, TC_F_SF, c_loc(SF)), 'put SF')
  call check(tc_put_tracers(ctx, rtr, str, ty), 'put tracers')
  call check(tc_tracers_to_corners(ctx, TC_TR_X, 0, n_comp), 'tracers_to_corners')
  call check(tc_smoother_time_T(ctx, TC_F_T, n_temp), 'smoother_time_T')
  call check(tc_smoother_time(ctx, n_comp), 'smoother_time')
  call check(tc_load_conductivity(ctx), 'load_conductivity_array')
  call check(tc_set_time(ctx, 0.0d0, 0.0d0), 'set_time')
  call check(tc_multigrid(ctx, q, resid, resref), 'multigrid')
  call check(tc_compute_velocities(ctx), 'compute_velocities')
  call check(tc_get_field(ctx, TC_F_SF, c_loc(SF)), 'get SF')
  call check(tc_get_field(ctx, TC_F_U, c_loc(u)), 'get u')
  call check(tc_get_field(ctx, TC_F_W, c_loc(w)), 'get w')
  write(*, '(a,i0,a,es12.4,a,es12.4)') 'tc_solve_demo: q=', q, ' resid=', resid, ' res_ref=', resref
  open(newunit=uo, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(uo) q, resid, resref, SF, u, w
  call check(tc_time_step(ctx, 1), 'time_step')
  call check(tc_get_field(ctx, TC_F_SF, c_loc(SF)), 'get SF')
  call check(tc_get_tracers(ctx, TC_TR_X, rtr, str), 'get tracers')
  write(uo) SF, rtr, str
  close(uo)
  call check(tc_finalize(ctx), 'tc_finalize')

contains

  subroutine check(st, what)
    integer(c_int), intent(in) :: st
    character(len=*), intent(in) :: what
    if (st /= TC_OK) then
      write(*, '(a,a,a,i0,a,a)') 'tc_solve_demo: ', what, ' failed, status ', st, ': ', tc_message(ctx)
      stop 1
    end if
  end subroutine check

end program tc_solve_demo
