"""CPU oracle of the tracer_code time-step hot path -- TEST INFRASTRUCTURE.

A plain-C restatement (tc_oracle.c) of the reference Fortran, loaded with
ctypes.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may use it, and only as the checker or the timed CPU baseline; the product
path (tracer_code_amd) never imports it.

Parity status of each function is documented in tc_oracle.c and in the tests
that pin it (tests/test_oracle_*.py).
"""
import ctypes
import os
import subprocess

import numpy as np

from tracer_code_amd._abi import TcEqualizeInfo, TcParams, TcStatsRow

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build of tc_oracle.c (the mutation test of the
# pins loads deliberately broken copies this way)
LIB = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

# ids (tc_oracle.h)
OR_T, OR_SF, OR_U, OR_W, OR_STRAIN, OR_TVIS, OR_TBUOY, OR_CONDUCT = range(8)
OR_CONDUCT_R, OR_CONDUCT_S, OR_DTS, OR_DERR_U, OR_DERR_W, OR_T0, OR_T1, OR_T2 = range(8, 16)
OR_RESIDUAL = 16
OR_CNEW, OR_CVIS, OR_CBUOY, OR_C0 = 32, 33, 34, 35
(OR_ERROR, OR_RHS, OR_VIS, OR_VISX, OR_VISZ, OR_VISXX, OR_VISZZ, OR_VISXZ,
 OR_DTSTREAM, OR_DTVIS, OR_RATIO) = range(64, 75)
OR_COUNT, OR_NEMPTY = 96, 97
OR_RTR, OR_STR, OR_RTR0, OR_STR0, OR_RTR1, OR_STR1, OR_RTR2, OR_STR2, OR_TTYPE = range(128, 137)
OR_MCOARSE, OR_ROWB, OR_IPIV, OR_BCOARSE = 160, 161, 162, 163
ORS_DT, ORS_TIME, ORS_RESID1, ORS_RESREF1, ORS_EIG_FINE, ORS_DELTA_T = range(6)
(ORS_Q, ORS_N_ETA, ORS_N_TEMP, ORS_N_COMP1, ORS_STATUS, ORS_KL, ORS_KU, ORS_NGC, ORS_TSTEP,
 ORS_SMOOTH_VIS, ORS_NVIS_STORE, ORS_NN_STORE, ORS_CYCLES_TOTAL, ORS_VIS_ITERS, ORS_C_ITERS,
 ORS_T_ITERS) = range(16, 32)
ORS_Q_STEP = tuple(range(32, 36))
ORS_N_COMP_T = tuple(range(36, 40))  # n_comp of types 1..4
ORS_INTERP_CLAMPS = 40
ORS_MAX_CYCLES = 41

_lib = None


def build():
    """Compile liboracle.so from tc_oracle.c (gcc, no GPU)."""
    r = subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        L.or_create.restype = vp
        L.or_create.argtypes = [ctypes.POINTER(TcParams), ctypes.c_int64]
        L.or_destroy.argtypes = [vp]
        L.or_ptr.restype = vp
        L.or_ptr.argtypes = [vp, i, i]
        L.or_get_scalar.restype = d
        L.or_get_scalar.argtypes = [vp, i]
        L.or_set_scalar.argtypes = [vp, i, d]
        for n in ("or_ld",):
            getattr(L, n).restype = i
            getattr(L, n).argtypes = [vp]
        for n in ("or_nx_grid", "or_nz_grid"):
            getattr(L, n).restype = i
            getattr(L, n).argtypes = [vp, i]
        L.or_residual_mag.restype = d
        L.or_residual_mag.argtypes = [vp, i]
        L.or_res_ref.restype = d
        L.or_res_ref.argtypes = [vp, i]
        L.or_fd_coefficients.argtypes = [i, vp, vp, vp, vp, vp]
        L.or_optimize_grid.argtypes = [ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i), i]
        L.or_multigrid_setup.restype = i
        L.or_multigrid_setup.argtypes = [i, vp]
        L.or_derivative_eigenvalues.argtypes = [vp, vp, vp, vp, vp]
        L.or_dgbequ.restype = i
        L.or_dgbequ.argtypes = [i, i, i, i, vp, i, vp, vp, vp, vp, vp]
        L.or_dgbtrf.restype = i
        L.or_dgbtrf.argtypes = [i, i, i, i, vp, i, vp]
        L.or_dgbtrs.argtypes = [i, i, i, vp, i, vp, vp]
        L.or_stream_space.restype = d
        L.or_stream_space.argtypes = [vp, i, i, i, vp]
        L.or_interpolate.restype = d
        L.or_interpolate.argtypes = [vp, d, d, vp, vp]
        L.or_stability_RK4.restype = d
        L.or_stability_RK4.argtypes = [vp]
        for n, a in [
            ("or_stability_vis_smoother", [vp, i]), ("or_stability_RK4_smoother", [vp]),
            ("or_enforce_bcs", [vp, vp]), ("or_enforce_bcs_stream", [vp, i, vp]), ("or_enforce_bcs_comp", [vp, vp]),
            ("or_iterate_stream", [vp, i, i, i, i]), ("or_sweep", [vp, i, i, i, i]),
            ("or_compute_residual", [vp, i, i]), ("or_compute_RHS", [vp, i]),
            ("or_compute_reference_residual", [vp, i]), ("or_restrict", [vp, i, vp, vp]),
            ("or_prolong", [vp, i]), ("or_viscosity_gradients", [vp, i]), ("or_test_eigenvalues", [vp, i]),
            ("or_iterate_local_time_steps", [vp, i]), ("or_smoother_vis", [vp, i, i]),
            ("or_compute_viscosity", [vp, i]), ("or_viscosity_bcs", [vp, i]),
            ("or_viscosity_bcs_initialize", [vp, i]), ("or_compute_matrix_coarse", [vp]), ("or_coarse_solve", [vp]),
            ("or_compute_velocities", [vp]), ("or_compute_strain_rate_invariant", [vp]),
            ("or_smoother_time", [vp]), ("or_load_conductivity_array", [vp]),
            ("or_compute_derivatives", [vp, vp, vp]), ("or_energy_stage", [vp, i, d]), ("or_stats", [vp, vp]),
            ("or_visf_plane", [vp, vp]),
        ]:
            getattr(L, n).argtypes = a
            getattr(L, n).restype = None
        for n, a in [("or_multigrid", [vp]), ("or_smoother_time_T", [vp, vp]), ("or_tracers_to_corners", [vp, i]),
                     ("or_tracer_stage", [vp, i, d]), ("or_energy_time", [vp]), ("or_time_step", [vp]),
                     ("or_initialize_state", [vp])]:
            getattr(L, n).argtypes = a
            getattr(L, n).restype = i
        u64 = ctypes.c_uint64
        L.or_equalize_tracers.restype = i
        L.or_equalize_tracers.argtypes = [vp, i, vp, u64, u64, ctypes.POINTER(TcEqualizeInfo)]
        L.or_stats_record.restype = None
        L.or_stats_record.argtypes = [vp, i, vp, ctypes.POINTER(TcStatsRow)]
        L.or_set_equalize.restype = None
        L.or_set_equalize.argtypes = [vp, i, i, vp, u64]
        L.or_eq_uniform.restype = d
        L.or_eq_uniform.argtypes = [u64, u64, ctypes.c_uint, u64, u64]
        L.or_eq_perm40.restype = u64
        L.or_eq_perm40.argtypes = [u64, u64, u64]
        L.or_eq_shuffle.restype = None
        L.or_eq_shuffle.argtypes = [vp, i, u64, u64, ctypes.c_uint, u64]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def fd_coefficients(order=2):
    D1, D2, D3, D4 = (np.zeros(5) for _ in range(4))
    Dint = np.zeros((2, 3))
    lib().or_fd_coefficients(order, _p(D1), _p(D2), _p(D3), _p(D4), _p(Dint))
    return D1, D2, D3, D4, Dint


def optimize_grid(nx, nz, grids, mingrid=4):
    a, b, c = ctypes.c_int(nx), ctypes.c_int(nz), ctypes.c_int(grids)
    lib().or_optimize_grid(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), mingrid)
    return a.value, b.value, c.value


def multigrid_setup(grids):
    m = np.zeros(52, dtype=np.int32)
    n = lib().or_multigrid_setup(grids, _p(m))
    return n, m


def derivative_eigenvalues(D1, D2, D3, D4):
    base = np.zeros(7)
    lib().or_derivative_eigenvalues(_p(np.ascontiguousarray(D1)), _p(np.ascontiguousarray(D2)),
                                    _p(np.ascontiguousarray(D3)), _p(np.ascontiguousarray(D4)), _p(base))
    return base


def dgbtrf(ab, n, kl, ku):
    """LAPACK-semantics banded LU of ab (2kl+ku+1, n) (Fortran order copy)."""
    a = np.asfortranarray(ab.copy())
    ipiv = np.zeros(n, dtype=np.int32)
    info = lib().or_dgbtrf(n, n, kl, ku, _p(a), a.shape[0], _p(ipiv))
    return a, ipiv, info


def dgbtrs(a, n, kl, ku, ipiv, b):
    x = np.ascontiguousarray(b, dtype=np.float64).copy()
    lib().or_dgbtrs(n, kl, ku, _p(a), a.shape[0], _p(ipiv), _p(x))
    return x


class Oracle:
    """State of one oracle run: ghosted planes as numpy views shaped
    (nz+5, nx+5) (row k, column i; element [k+2, i+2] is f(i,k))."""

    def __init__(self, params, ntr=0):
        self.L = lib()
        self.params = params
        self.ntr = int(ntr)
        self.h = self.L.or_create(ctypes.byref(params), self.ntr)
        self.nx, self.nz = params.nx, params.nz
        self.ld = self.L.or_ld(self.h)

    def close(self):
        if self.h:
            self.L.or_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def plane(self, which, idx=0, dtype=np.float64):
        p = self.L.or_ptr(self.h, which, idx)
        n = (self.nz + 5) * self.ld
        buf = (ctypes.c_char * (n * np.dtype(dtype).itemsize)).from_address(p)
        return np.frombuffer(buf, dtype=dtype).reshape(self.nz + 5, self.ld)

    def tracers(self, which, dtype=np.float64):
        p = self.L.or_ptr(self.h, which, 0)
        buf = (ctypes.c_char * (self.ntr * np.dtype(dtype).itemsize)).from_address(p)
        return np.frombuffer(buf, dtype=dtype)

    def scalar(self, which):
        return self.L.or_get_scalar(self.h, which)

    def set_scalar(self, which, v):
        self.L.or_set_scalar(self.h, which, float(v))

    def equalize_tracers(self, tpc, Cinit, seed, call):
        """equalize_tracers in its counter-based form; returns (status, info)."""
        info = TcEqualizeInfo()
        c = np.ascontiguousarray(np.atleast_1d(np.asarray(Cinit, dtype=np.float64)))
        st = self.L.or_equalize_tracers(self.h, int(tpc), _p(c), int(seed), int(call), ctypes.byref(info))
        return st, info

    def stats_record(self, kent, mass=None):
        """The stats.dat row (basics.f90:236-324) as a TcStatsRow."""
        out = TcStatsRow()
        m = None if mass is None else np.ascontiguousarray(np.atleast_1d(np.asarray(mass, dtype=np.float64)))
        self.L.or_stats_record(self.h, int(kent), None if m is None else _p(m), ctypes.byref(out))
        return out

    def set_equalize(self, on, tpc, Cinit, seed):
        c = np.ascontiguousarray(np.atleast_1d(np.asarray(Cinit, dtype=np.float64)))
        self.L.or_set_equalize(self.h, int(on), int(tpc), _p(c), int(seed))

    def __getattr__(self, name):
        fn = getattr(self.L, "or_" + name)

        def call(*args):
            return fn(self.h, *args)
        return call
