"""ctypes mirror of include/tracer_hip.h (the C-ABI of the HIP library).

The struct layout and the enum values here must match the header exactly;
tests/test_abi.py checks the struct size and every exported symbol.
"""
import ctypes
import os

TC_MAX_TYPES = 4
TC_MAX_LEVELS = 6
TC_MAX_MULTI = 52

# status codes
TC_OK = 0
TC_ERR_MAX_CYCLES = 1
TC_ERR_RESID_BLOWUP = 2
TC_ERR_TRACER_DOMAIN = 3
TC_ERR_SINGULAR = 4
TC_ERR_ARG = -1
TC_ERR_HIP = -2
TC_ERR_UNSUPPORTED = -3
TC_ERR_NOMEM = -4

TC_BIN_NOALLOC_MAX_COUNTS = 8 << 20  # int32 words of tc_bin_tracers_noalloc's per-GPU work array

TC_MODE_FP64 = 0
TC_MODE_MIXED = 1
TC_MODE_DC64 = 2
TC_GUESS_PLAIN, TC_GUESS_EXTRAP, TC_GUESS_LEARNED = 0, 1, 2

# field ids
TC_F_T, TC_F_SF, TC_F_TVIS, TC_F_TBUOY, TC_F_CNEW, TC_F_CVIS, TC_F_CBUOY = range(7)
TC_F_U, TC_F_W, TC_F_STRAIN, TC_F_VIS, TC_F_CONDUCT, TC_F_NEMPTY, TC_F_COUNT_TYPE = range(7, 14)
TC_F_T0, TC_F_T1, TC_F_T2 = 14, 15, 16
TC_F_SF_LO, TC_F_RHS = 17, 18

TC_TR_X, TC_TR_X0, TC_TR_X1, TC_TR_X2 = 0, 1, 2, 3
TC_K_SWEEP_L1 = 0
TC_FIX_ZERO_COARSE = 1

c_double5 = ctypes.c_double * 5


class TcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("nz", ctypes.c_int32), ("grids", ctypes.c_int32),
        ("order", ctypes.c_int32), ("iterate_order", ctypes.c_int32),
        ("Vbc", ctypes.c_int32), ("Tbot", ctypes.c_int32),
        ("Nmulti", ctypes.c_int32), ("vis_iter", ctypes.c_int32), ("max_cycles", ctypes.c_int32),
        ("nsteps", ctypes.c_int32), ("multi", ctypes.c_int32 * TC_MAX_MULTI),
        ("comp", ctypes.c_int32), ("ntype", ctypes.c_int32), ("ivis", ctypes.c_int32),
        ("torder", ctypes.c_int32), ("RK_update", ctypes.c_int32), ("fixes", ctypes.c_int32),
        ("aspect", ctypes.c_double),
        ("tolerance", ctypes.c_double), ("eig_ratio", ctypes.c_double), ("eig_ratio_min", ctypes.c_double),
        ("courant", ctypes.c_double), ("courantT", ctypes.c_double), ("courantC", ctypes.c_double),
        ("courant_stream", ctypes.c_double),
        ("RaT", ctypes.c_double), ("visT", ctypes.c_double), ("visP", ctypes.c_double), ("H", ctypes.c_double),
        ("RaC", ctypes.c_double * TC_MAX_TYPES), ("visC", ctypes.c_double * TC_MAX_TYPES),
        ("conduct_factor", ctypes.c_double * TC_MAX_TYPES), ("Htype", ctypes.c_double * TC_MAX_TYPES),
        ("delta_T", ctypes.c_double), ("delta_C", ctypes.c_double * TC_MAX_TYPES),
        ("D1", c_double5), ("D2", c_double5), ("D3", c_double5), ("D4", c_double5),
        ("Dint", (ctypes.c_double * 3) * 2),
        ("Tr_base", ctypes.c_double), ("Trr_base", ctypes.c_double), ("Trrr_base", ctypes.c_double),
        ("Trrrr_base", ctypes.c_double), ("Trs_base", ctypes.c_double), ("Trrs_base", ctypes.c_double),
        ("Trrss_base", ctypes.c_double),
    ]


class TcStats(ctypes.Structure):
    _fields_ = [
        ("q_last", ctypes.c_int32), ("n_eta", ctypes.c_int32), ("n_temp", ctypes.c_int32),
        ("n_comp", ctypes.c_int32 * TC_MAX_TYPES), ("guess_used", ctypes.c_int32), ("pad", ctypes.c_int32),
        ("residual_last", ctypes.c_double), ("res_ref_last", ctypes.c_double), ("eig_fine", ctypes.c_double),
        ("dt_last", ctypes.c_double), ("time", ctypes.c_double),
        ("sweeps_exec", ctypes.c_int64 * TC_MAX_LEVELS), ("residuals_exec", ctypes.c_int64 * TC_MAX_LEVELS),
        ("coarse_solves", ctypes.c_int64), ("vis_iters", ctypes.c_int64), ("c_iters", ctypes.c_int64),
        ("t_iters", ctypes.c_int64), ("mg_cycles", ctypes.c_int64), ("guesses_used", ctypes.c_int64 * 3),
        ("guess_res", ctypes.c_double * 3), ("dc_retries", ctypes.c_int64), ("q_step", ctypes.c_int32 * 4),
    ]


class TcEqualizeInfo(ctypes.Structure):
    _fields_ = [
        ("did_work", ctypes.c_int32), ("passes", ctypes.c_int32),
        ("tpc_type", ctypes.c_int32 * (TC_MAX_TYPES + 1)), ("minexcess", ctypes.c_int32 * (TC_MAX_TYPES + 1)),
        ("deficit", ctypes.c_int64), ("moved", ctypes.c_int64), ("nfree", ctypes.c_int64),
        ("count", ctypes.c_int64 * (TC_MAX_TYPES + 1)),
    ]


class TcStatsRow(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_double), ("Tavg", ctypes.c_double), ("vrms", ctypes.c_double), ("Nu", ctypes.c_double),
        ("Nu_bot", ctypes.c_double), ("RaT_avg", ctypes.c_double), ("mobility", ctypes.c_double),
        ("strain_max", ctypes.c_double), ("q", ctypes.c_double * 4),
        ("empty_cells", ctypes.c_int32), ("tracer_min", ctypes.c_int32), ("tracer_max", ctypes.c_int32),
        ("kent", ctypes.c_int32),
        ("Cavg", ctypes.c_double * TC_MAX_TYPES), ("ent", ctypes.c_double * TC_MAX_TYPES),
    ]


# every C-ABI entry point declared in include/tracer_hip.h: name -> (restype, argtypes)
_VP = ctypes.c_void_p
_P = ctypes.POINTER
SIGNATURES = {
    "tc_init": (ctypes.c_int, [_P(TcParams), ctypes.c_int64, _P(_VP)]),
    "tc_finalize": (ctypes.c_int, [_VP]),
    "tc_last_error": (ctypes.c_char_p, [_VP]),
    "tc_set_mode": (ctypes.c_int, [_VP, ctypes.c_int]),
    "tc_set_dc_sweeps": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "tc_set_dc_level_sweeps": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32]),
    "tc_set_graphs": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "tc_graph_counts": (ctypes.c_int, [_VP, _P(ctypes.c_int64), _P(ctypes.c_int64)]),
    "tc_set_gcr": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "tc_set_coarse_refresh": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "tc_set_filter_blocking": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "tc_gcr_counts": (ctypes.c_int, [_VP, _P(ctypes.c_int64), _P(ctypes.c_int64)]),
    "tc_sort_tracers": (ctypes.c_int, [_VP]),
    "tc_diagnostics": (ctypes.c_int, [_VP, _VP]),
    "tc_restart_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, _VP, _VP, _VP, _VP, _VP]),
    "tc_restart_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_int32, _VP, _VP, _VP, _VP, _VP]),
    "tc_restart_tracers_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64, _VP, _VP, _VP]),
    "tc_restart_tracers_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64, _VP, _VP, _VP]),
    "tc_stats_record": (ctypes.c_int, [_VP, ctypes.c_int32, _VP, _P(TcStatsRow)]),
    "tc_snapshot_stage": (ctypes.c_int, [_VP]),
    "tc_snapshot_get": (ctypes.c_int, [_VP, ctypes.c_int, _VP]),
    "tc_set_equalize": (ctypes.c_int, [_VP, ctypes.c_int32, _VP, ctypes.c_uint64, ctypes.c_int32]),
    "tc_equalize_tracers": (ctypes.c_int, [_VP, _P(TcEqualizeInfo)]),
    "tc_ensemble_step": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _VP]),
    "tc_set_tracer_sort": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "tc_tracer_slots": (ctypes.c_int, [_VP, _VP]),
    "tc_set_gs_smoother": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_double]),
    "tc_set_dc_cheb_interval": (ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_double]),
    "tc_autotune_dc_sweeps": (ctypes.c_int, [_VP, _VP, ctypes.c_int32, ctypes.c_int32, _VP, _VP]),
    "tc_autotune_dc_level_sweeps": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32, _VP, ctypes.c_int32,
                                                   ctypes.c_int32, _VP, _VP]),
    "tc_set_tb_tile": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32]),
    "tc_autotune_tb_tiles": (ctypes.c_int, [_VP, ctypes.c_int32, _VP]),
    "tc_get_stats": (ctypes.c_int, [_VP, _P(TcStats)]),
    "tc_sync": (ctypes.c_int, [_VP]),
    "tc_put_field": (ctypes.c_int, [_VP, ctypes.c_int, _VP]),
    "tc_get_field": (ctypes.c_int, [_VP, ctypes.c_int, _VP]),
    "tc_put_tracers": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
    "tc_get_tracers": (ctypes.c_int, [_VP, ctypes.c_int, _VP, _VP]),
    "tc_get_tracer_types": (ctypes.c_int, [_VP, _VP]),
    "tc_set_time": (ctypes.c_int, [_VP, ctypes.c_double, ctypes.c_double]),
    "tc_multigrid": (ctypes.c_int, [_VP, _P(ctypes.c_int32), _P(ctypes.c_double), _P(ctypes.c_double)]),
    "tc_compute_velocities": (ctypes.c_int, [_VP]),
    "tc_compute_strain_rate_invariant": (ctypes.c_int, [_VP]),
    "tc_smoother_time_T": (ctypes.c_int, [_VP, ctypes.c_int, _P(ctypes.c_int32)]),
    "tc_tracers_to_corners": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _P(ctypes.c_int32)]),
    "tc_smoother_time": (ctypes.c_int, [_VP, _P(ctypes.c_int32)]),
    "tc_load_conductivity": (ctypes.c_int, [_VP]),
    "tc_tracer_stage": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_double]),
    "tc_stability_RK4": (ctypes.c_int, [_VP, _P(ctypes.c_double)]),
    "tc_energy_stage": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_double]),
    "tc_time_step": (ctypes.c_int, [_VP, ctypes.c_int]),
    "tc_guess_enable": (ctypes.c_int, [_VP, ctypes.c_int]),
    "tc_bin_tracers_noalloc": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                              ctypes.c_int32, _VP, _VP, _VP, ctypes.c_int64, _VP, _VP, _VP, _VP]),
    "tc_set_profiling": (ctypes.c_int, [_VP, ctypes.c_int]),
    "tc_kernel_timing": (ctypes.c_int, [_VP, ctypes.c_int, _P(ctypes.c_double), _P(ctypes.c_int64)]),
    "tc_derivative_eigenvalues": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
    "tc_nn_create": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _P(_VP)]),
    "tc_nn_destroy": (ctypes.c_int, [_VP]),
    "tc_nn_last_error": (ctypes.c_char_p, [_VP]),
    "tc_nn_param_count": (ctypes.c_int, [_VP, _P(ctypes.c_int64)]),
    "tc_nn_set_weights": (ctypes.c_int, [_VP, _VP, ctypes.c_int64]),
    "tc_nn_get_weights": (ctypes.c_int, [_VP, _VP, ctypes.c_int64]),
    "tc_nn_get_grads": (ctypes.c_int, [_VP, _VP, ctypes.c_int64]),
    "tc_nn_set_sample": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
    "tc_nn_forward": (ctypes.c_int, [_VP, _VP]),
    "tc_nn_backward": (ctypes.c_int, [_VP, _P(ctypes.c_double)]),
    "tc_nn_adam": (ctypes.c_int, [_VP, ctypes.c_float, ctypes.c_float]),
    "tc_learn_attach": (ctypes.c_int, [_VP, _VP]),
    "tc_learn_record": (ctypes.c_int, [_VP, ctypes.c_int32]),
    "tc_learn_samples": (ctypes.c_int, [_VP, _P(ctypes.c_int32)]),
    "tc_learn_train": (ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_float, ctypes.c_int32, ctypes.c_uint64,
                                      _P(ctypes.c_double)]),
}

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtracer_hip.so")


def load_library(path=None):
    """Load the HIP library and bind every declared entry point.

    Raises OSError when the library has not been built: there is no
    fallback implementation of the hot path.
    """
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise OSError(f"tracer_hip library not built: {path} (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
