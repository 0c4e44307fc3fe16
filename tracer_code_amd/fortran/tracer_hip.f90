! ISO_C_BINDING interface to libtracer_hip.so (include/tracer_hip.h) for the
! reference's Fortran driver.  The derived types are interoperable with the
! C structs tc_params / tc_stats; every entry point returns the C status
! (0 = ok, >0 = the reference's stop conditions, <0 = errors).
module tracer_hip
  use iso_c_binding
  implicit none
  private :: c_char

  integer(c_int), parameter :: TC_MAX_TYPES = 4, TC_MAX_LEVELS = 6, TC_MAX_MULTI = 52
  integer(c_int), parameter :: TC_OK = 0, TC_ERR_MAX_CYCLES = 1, TC_ERR_RESID_BLOWUP = 2, &
                               TC_ERR_TRACER_DOMAIN = 3, TC_ERR_SINGULAR = 4, TC_ERR_ARG = -1, &
                               TC_ERR_HIP = -2, TC_ERR_UNSUPPORTED = -3, TC_ERR_NOMEM = -4
  integer(c_int), parameter :: TC_MODE_FP64 = 0, TC_MODE_MIXED = 1, TC_MODE_DC64 = 2
  integer(c_int), parameter :: TC_GUESS_PLAIN = 0, TC_GUESS_EXTRAP = 1, TC_GUESS_LEARNED = 2
  integer(c_int), parameter :: TC_DC_FAST_MAX_CYCLES = 200
  integer(c_int), parameter :: TC_F_T = 0, TC_F_SF = 1, TC_F_TVIS = 2, TC_F_TBUOY = 3, TC_F_CNEW = 4, &
                               TC_F_CVIS = 5, TC_F_CBUOY = 6, TC_F_U = 7, TC_F_W = 8, TC_F_STRAIN = 9, &
                               TC_F_VIS = 10, TC_F_CONDUCT = 11, TC_F_NEMPTY = 12, TC_F_COUNT_TYPE = 13, &
                               TC_F_T0 = 14, TC_F_T1 = 15, TC_F_T2 = 16, TC_F_SF_LO = 17, TC_F_RHS = 18
  integer(c_int), parameter :: TC_FIX_ZERO_COARSE = 1
  integer(c_int), parameter :: TC_TR_X = 0, TC_TR_X0 = 1, TC_TR_X1 = 2, TC_TR_X2 = 3
  integer(c_int), parameter :: TC_K_SWEEP_L1 = 0

  type, bind(C) :: tc_params
    integer(c_int32_t) :: nx, nz, grids, order, iterate_order, Vbc, Tbot
    integer(c_int32_t) :: Nmulti, vis_iter, max_cycles, nsteps
    integer(c_int32_t) :: multi(0:TC_MAX_MULTI-1)
    integer(c_int32_t) :: comp, ntype, ivis, torder, RK_update
    integer(c_int32_t) :: fixes
    real(c_double) :: aspect
    real(c_double) :: tolerance, eig_ratio, eig_ratio_min
    real(c_double) :: courant, courantT, courantC, courant_stream
    real(c_double) :: RaT, visT, visP, H
    real(c_double) :: RaC(TC_MAX_TYPES), visC(TC_MAX_TYPES)
    real(c_double) :: conduct_factor(TC_MAX_TYPES), Htype(TC_MAX_TYPES)
    real(c_double) :: delta_T
    real(c_double) :: delta_C(TC_MAX_TYPES)
    real(c_double) :: D1(-2:2), D2(-2:2), D3(-2:2), D4(-2:2)
    real(c_double) :: Dint(-1:1, 1:2)      ! Dint(j,m) = D(m,j) of the interpolation table
    real(c_double) :: Tr_base, Trr_base, Trrr_base, Trrrr_base, Trs_base, Trrs_base, Trrss_base
  end type tc_params

  type, bind(C) :: tc_stats
    integer(c_int32_t) :: q_last, n_eta, n_temp, n_comp(TC_MAX_TYPES), guess_used, pad
    real(c_double) :: residual_last, res_ref_last, eig_fine, dt_last, time
    integer(c_int64_t) :: sweeps_exec(TC_MAX_LEVELS), residuals_exec(TC_MAX_LEVELS)
    integer(c_int64_t) :: coarse_solves, vis_iters, c_iters, t_iters, mg_cycles, guesses_used(3)
    real(c_double) :: guess_res(3)
    integer(c_int64_t) :: dc_retries
    integer(c_int32_t) :: q_step(4)
  end type tc_stats

  ! one row of stats.dat (basics.f90:236-324)
  type, bind(C) :: tc_stats_rec
    real(c_double) :: time, Tavg, vrms, Nu, Nu_bot, RaT_avg, mobility, strain_max, q(4)
    integer(c_int32_t) :: empty_cells, tracer_min, tracer_max, kent
    real(c_double) :: Cavg(TC_MAX_TYPES), ent(TC_MAX_TYPES)
  end type tc_stats_rec

  ! equalize_tracers record: the columns of equalize.dat (composition.f90:566-574)
  type, bind(C) :: tc_equalize_info
    integer(c_int32_t) :: did_work, passes, tpc_type(0:TC_MAX_TYPES), minexcess(0:TC_MAX_TYPES)
    integer(c_int64_t) :: deficit, moved, nfree, count(0:TC_MAX_TYPES)
  end type tc_equalize_info

  interface
    integer(c_int) function tc_init(p, ntr, ctx) bind(C, name="tc_init")
      import :: c_int, c_int64_t, c_ptr, tc_params
      type(tc_params), intent(in) :: p
      integer(c_int64_t), value :: ntr
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function tc_finalize(ctx) bind(C, name="tc_finalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function tc_last_error(ctx) bind(C, name="tc_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_set_mode(ctx, mode) bind(C, name="tc_set_mode")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function
    integer(c_int) function tc_get_stats(ctx, st) bind(C, name="tc_get_stats")
      import :: c_int, c_ptr, tc_stats
      type(c_ptr), value :: ctx
      type(tc_stats), intent(out) :: st
    end function
    integer(c_int) function tc_sync(ctx) bind(C, name="tc_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_put_field(ctx, field, host) bind(C, name="tc_put_field")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      type(c_ptr), value :: host
    end function
    integer(c_int) function tc_get_field(ctx, field, host) bind(C, name="tc_get_field")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      type(c_ptr), value :: host
    end function
    integer(c_int) function tc_put_tracers(ctx, rtr, str, ttype) bind(C, name="tc_put_tracers")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: rtr(*), str(*)
      integer(c_int32_t), intent(in) :: ttype(*)
    end function
    integer(c_int) function tc_get_tracers(ctx, which, r, s) bind(C, name="tc_get_tracers")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
      real(c_double), intent(out) :: r(*), s(*)
    end function
    integer(c_int) function tc_get_tracer_types(ctx, ttype) bind(C, name="tc_get_tracer_types")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: ttype(*)
    end function
    integer(c_int) function tc_set_time(ctx, time, dt) bind(C, name="tc_set_time")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: time, dt
    end function
    integer(c_int) function tc_multigrid(ctx, q, resid, resref) bind(C, name="tc_multigrid")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: q
      real(c_double), intent(out) :: resid, resref
    end function
    integer(c_int) function tc_compute_velocities(ctx) bind(C, name="tc_compute_velocities")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_compute_strain_rate_invariant(ctx) bind(C, name="tc_compute_strain_rate_invariant")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_smoother_time_T(ctx, src_field, n_temp) bind(C, name="tc_smoother_time_T")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: src_field
      integer(c_int32_t), intent(out) :: n_temp
    end function
    integer(c_int) function tc_tracers_to_corners(ctx, which, smooth, n_comp) bind(C, name="tc_tracers_to_corners")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: which, smooth
      integer(c_int32_t), intent(out) :: n_comp(*)
    end function
    integer(c_int) function tc_smoother_time(ctx, n_comp) bind(C, name="tc_smoother_time")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: n_comp(*)
    end function
    integer(c_int) function tc_load_conductivity(ctx) bind(C, name="tc_load_conductivity")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_tracer_stage(ctx, stage, dt) bind(C, name="tc_tracer_stage")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: stage
      real(c_double), value :: dt
    end function
    integer(c_int) function tc_stability_RK4(ctx, dt) bind(C, name="tc_stability_RK4")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dt
    end function
    integer(c_int) function tc_energy_stage(ctx, stage, dt) bind(C, name="tc_energy_stage")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: stage
      real(c_double), value :: dt
    end function
    integer(c_int) function tc_time_step(ctx, nsteps) bind(C, name="tc_time_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function tc_derivative_eigenvalues(D1, D2, D3, D4, base) bind(C, name="tc_derivative_eigenvalues")
      import :: c_int, c_double
      real(c_double), intent(in) :: D1(5), D2(5), D3(5), D4(5)
      real(c_double), intent(out) :: base(7)
    end function
    integer(c_int) function tc_set_profiling(ctx, enable) bind(C, name="tc_set_profiling")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function
    integer(c_int) function tc_kernel_timing(ctx, which, avg_us, count) bind(C, name="tc_kernel_timing")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: which
      real(c_double), intent(out) :: avg_us
      integer(c_int64_t), intent(out) :: count
    end function
    integer(c_int) function tc_guess_enable(ctx, mode) bind(C, name="tc_guess_enable")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function
    ! binning on device buffers without allocation (no reference counterpart;
    ! for callers that cannot allocate); pointers are device addresses
    integer(c_int) function tc_bin_tracers_noalloc(nx, nz, aspect, ntype, comp, r, s, ttype, ntr, count_type, &
                                                   nempty, Cnew, stream) bind(C, name="tc_bin_tracers_noalloc")
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      integer(c_int32_t), value :: nx, nz, ntype, comp
      real(c_double), value :: aspect
      type(c_ptr), value :: r, s, ttype, count_type, nempty, Cnew, stream
      integer(c_int64_t), value :: ntr
    end function
    integer(c_int) function tc_ensemble_step(ctxs, n, nsteps, threads, failed) bind(C, name="tc_ensemble_step")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), intent(in) :: ctxs(*)
      integer(c_int32_t), value :: n, nsteps, threads
      integer(c_int32_t), intent(out) :: failed
    end function
    integer(c_int) function tc_restart_read(path, nx, nz, has_T, ntype_c, xg, zg, T, Cnew, SF) &
        bind(C, name="tc_restart_read")
      import :: c_int, c_int32_t, c_char, c_ptr
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: nx, nz, has_T, ntype_c
      type(c_ptr), value :: xg, zg, T, Cnew, SF
    end function
    integer(c_int) function tc_restart_write(path, nx, nz, has_T, ntype_c, xg, zg, T, Cnew, SF) &
        bind(C, name="tc_restart_write")
      import :: c_int, c_int32_t, c_char, c_ptr
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: nx, nz, has_T, ntype_c
      type(c_ptr), value :: xg, zg, T, Cnew, SF
    end function
    integer(c_int) function tc_restart_tracers_read(path, ntr, rtr, str, ttype) bind(C, name="tc_restart_tracers_read")
      import :: c_int, c_int64_t, c_char, c_ptr
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int64_t), value :: ntr
      type(c_ptr), value :: rtr, str, ttype
    end function
    integer(c_int) function tc_restart_tracers_write(path, ntr, rtr, str, ttype) &
        bind(C, name="tc_restart_tracers_write")
      import :: c_int, c_int64_t, c_char, c_ptr
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int64_t), value :: ntr
      type(c_ptr), value :: rtr, str, ttype
    end function
    integer(c_int) function tc_diagnostics(ctx, out) bind(C, name="tc_diagnostics")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(9)
    end function
    integer(c_int) function tc_stats_record(ctx, kent, mass, rec) bind(C, name="tc_stats_record")
      import :: c_int, c_int32_t, c_ptr, tc_stats_rec
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kent
      type(c_ptr), value :: mass
      type(tc_stats_rec), intent(out) :: rec
    end function
    integer(c_int) function tc_snapshot_stage(ctx) bind(C, name="tc_snapshot_stage")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_snapshot_get(ctx, field, host) bind(C, name="tc_snapshot_get")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      type(c_ptr), value :: host
    end function
    integer(c_int) function tc_set_equalize(ctx, tpc, Cinit, seed, in_time_step) bind(C, name="tc_set_equalize")
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tpc, in_time_step
      real(c_double), intent(in) :: Cinit(*)
      integer(c_int64_t), value :: seed
    end function
    integer(c_int) function tc_equalize_tracers(ctx, info) bind(C, name="tc_equalize_tracers")
      import :: c_int, c_ptr, tc_equalize_info
      type(c_ptr), value :: ctx
      type(tc_equalize_info), intent(out) :: info
    end function
    integer(c_int) function tc_set_graphs(ctx, enable) bind(C, name="tc_set_graphs")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: enable
    end function
    integer(c_int) function tc_graph_counts(ctx, captures, replays) bind(C, name="tc_graph_counts")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: captures, replays
    end function
    integer(c_int) function tc_set_gcr(ctx, directions) bind(C, name="tc_set_gcr")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: directions
    end function
    integer(c_int) function tc_set_coarse_refresh(ctx, solves) bind(C, name="tc_set_coarse_refresh")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: solves
    end function
    integer(c_int) function tc_set_filter_blocking(ctx, enable) bind(C, name="tc_set_filter_blocking")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: enable
    end function
    integer(c_int) function tc_gcr_counts(ctx, updates, restarts) bind(C, name="tc_gcr_counts")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: updates, restarts
    end function
    integer(c_int) function tc_set_dc_level_sweeps(ctx, level, sweeps) bind(C, name="tc_set_dc_level_sweeps")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: level, sweeps
    end function
    integer(c_int) function tc_set_tb_tile(ctx, level, tile) bind(C, name="tc_set_tb_tile")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: level, tile
    end function
    integer(c_int) function tc_autotune_tb_tiles(ctx, steps_each, tiles) bind(C, name="tc_autotune_tb_tiles")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: steps_each
      integer(c_int32_t), intent(out) :: tiles(*)
    end function
    integer(c_int) function tc_sort_tracers(ctx) bind(C, name="tc_sort_tracers")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function tc_set_tracer_sort(ctx, interval) bind(C, name="tc_set_tracer_sort")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: interval
    end function
    integer(c_int) function tc_set_gs_smoother(ctx, sweeps, per_launch, omega) bind(C, name="tc_set_gs_smoother")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sweeps, per_launch
      real(c_double), value :: omega
    end function
    integer(c_int) function tc_set_dc_cheb_interval(ctx, lb, la) bind(C, name="tc_set_dc_cheb_interval")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: lb, la
    end function
    integer(c_int) function tc_tracer_slots(ctx, caller_index) bind(C, name="tc_tracer_slots")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: caller_index(*)
    end function
    integer(c_int) function tc_set_dc_sweeps(ctx, sweeps) bind(C, name="tc_set_dc_sweeps")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: sweeps
    end function
    integer(c_int) function tc_autotune_dc_sweeps(ctx, candidates, n, steps_each, best, ms_per_step) &
        bind(C, name="tc_autotune_dc_sweeps")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: candidates(*)
      integer(c_int32_t), value :: n, steps_each
      integer(c_int32_t), intent(out) :: best
      real(c_double), intent(out) :: ms_per_step(*)
    end function
    integer(c_int) function tc_autotune_dc_level_sweeps(ctx, first_level, last_level, candidates, n, steps_each, &
        best, ms_per_step) bind(C, name="tc_autotune_dc_level_sweeps")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: first_level, last_level
      integer(c_int32_t), intent(in) :: candidates(*)
      integer(c_int32_t), value :: n, steps_each
      integer(c_int32_t), intent(out) :: best
      real(c_double), intent(out) :: ms_per_step(*)
    end function
    ! learned initial guess: the network handle (tc_nn*) is a c_ptr
    integer(c_int) function tc_nn_create(H, W, trainable, nn) bind(C, name="tc_nn_create")
      import :: c_int, c_int32_t, c_ptr
      integer(c_int32_t), value :: H, W, trainable
      type(c_ptr), intent(out) :: nn
    end function
    integer(c_int) function tc_nn_destroy(nn) bind(C, name="tc_nn_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: nn
    end function
    type(c_ptr) function tc_nn_last_error(nn) bind(C, name="tc_nn_last_error")
      import :: c_ptr
      type(c_ptr), value :: nn
    end function
    integer(c_int) function tc_nn_param_count(nn, count) bind(C, name="tc_nn_param_count")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: nn
      integer(c_int64_t), intent(out) :: count
    end function
    integer(c_int) function tc_nn_set_weights(nn, w, count) bind(C, name="tc_nn_set_weights")
      import :: c_int, c_int64_t, c_float, c_ptr
      type(c_ptr), value :: nn
      real(c_float), intent(in) :: w(*)
      integer(c_int64_t), value :: count
    end function
    integer(c_int) function tc_nn_get_weights(nn, w, count) bind(C, name="tc_nn_get_weights")
      import :: c_int, c_int64_t, c_float, c_ptr
      type(c_ptr), value :: nn
      real(c_float), intent(out) :: w(*)
      integer(c_int64_t), value :: count
    end function
    integer(c_int) function tc_nn_get_grads(nn, g, count) bind(C, name="tc_nn_get_grads")
      import :: c_int, c_int64_t, c_float, c_ptr
      type(c_ptr), value :: nn
      real(c_float), intent(out) :: g(*)
      integer(c_int64_t), value :: count
    end function
    integer(c_int) function tc_nn_set_sample(nn, x, target, mask) bind(C, name="tc_nn_set_sample")
      import :: c_int, c_ptr
      type(c_ptr), value :: nn, x, target, mask
    end function
    integer(c_int) function tc_nn_forward(nn, out) bind(C, name="tc_nn_forward")
      import :: c_int, c_ptr
      type(c_ptr), value :: nn, out
    end function
    integer(c_int) function tc_nn_backward(nn, loss) bind(C, name="tc_nn_backward")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: nn
      real(c_double), intent(out) :: loss
    end function
    integer(c_int) function tc_nn_adam(nn, lr, grad_scale) bind(C, name="tc_nn_adam")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: nn
      real(c_float), value :: lr, grad_scale
    end function
    integer(c_int) function tc_learn_attach(ctx, nn) bind(C, name="tc_learn_attach")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, nn
    end function
    integer(c_int) function tc_learn_record(ctx, capacity) bind(C, name="tc_learn_record")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: capacity
    end function
    integer(c_int) function tc_learn_samples(ctx, count) bind(C, name="tc_learn_samples")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: count
    end function
    integer(c_int) function tc_learn_train(ctx, epochs, lr, batch, seed, mean_loss) bind(C, name="tc_learn_train")
      import :: c_int, c_int32_t, c_int64_t, c_float, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: epochs, batch
      real(c_float), value :: lr
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: mean_loss
    end function
  end interface

contains

  ! The library's message for the last failed call as a Fortran string.
  function tc_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: n
    interface
      integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
        import :: c_size_t, c_ptr
        type(c_ptr), value :: s
      end function
    end interface
    cp = tc_last_error(ctx)
    n = int(c_strlen(cp))
    call c_f_pointer(cp, p, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function tc_message

end module tracer_hip
