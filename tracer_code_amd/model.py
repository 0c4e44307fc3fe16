"""Host-side mirror of the reference's hot-path routines over the C-ABI.

`Solver` owns one library context (device state) and exposes the reference's
routine names: multigrid, compute_velocities, compute_strain_rate_invariant,
smoother_time_T, tracers_to_corners, smoother_time, load_conductivity_array,
the RK stages of energy_time and whole time steps.  Arrays cross the boundary
in the reference's Fortran layout: a ghosted field f(-2:nx+2,-2:nz+2) is a
numpy array of shape (nz+5, nx+5) whose element [k+2, i+2] is f(i,k);
composition fields are (nz+5, nx+5, ntype) (type fastest) and count_type is
(nz+5, nx+5, ntype+1).

Errors are raised as TcError carrying the library status (the reference's
`stop` conditions are positive codes).  There is no CPU fallback: without the
built library the constructor raises.
"""
import ctypes

import numpy as np

from . import _abi


class TcError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"tracer_hip status {status}: {msg}")
        self.status = status


_FIELD = {
    "T": _abi.TC_F_T, "SF": _abi.TC_F_SF, "Tvis": _abi.TC_F_TVIS, "Tbuoy": _abi.TC_F_TBUOY,
    "Cnew": _abi.TC_F_CNEW, "Cvis": _abi.TC_F_CVIS, "Cbuoy": _abi.TC_F_CBUOY, "u": _abi.TC_F_U,
    "w": _abi.TC_F_W, "strain": _abi.TC_F_STRAIN, "vis": _abi.TC_F_VIS, "conduct": _abi.TC_F_CONDUCT,
    "nempty": _abi.TC_F_NEMPTY, "count_type": _abi.TC_F_COUNT_TYPE, "T0": _abi.TC_F_T0, "T1": _abi.TC_F_T1,
    "T2": _abi.TC_F_T2, "SF_lo": _abi.TC_F_SF_LO, "RHS": _abi.TC_F_RHS,
}
_TRACERS = {"x": _abi.TC_TR_X, "x0": _abi.TC_TR_X0, "x1": _abi.TC_TR_X1, "x2": _abi.TC_TR_X2}

_lib = None


def library():
    global _lib
    if _lib is None:
        _lib = _abi.load_library()
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Solver:
    def __init__(self, params, ntr=0, lib=None):
        self.lib = lib or library()
        self.params = params
        self.nx, self.nz, self.ntype = params.nx, params.nz, params.ntype
        self.ntr = int(ntr)
        h = ctypes.c_void_p()
        st = self.lib.tc_init(ctypes.byref(params), self.ntr, ctypes.byref(h))
        if st != 0:
            raise TcError(st, "tc_init failed")
        self.h = h

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.tc_finalize(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != 0:
            msg = self.lib.tc_last_error(self.h)
            raise TcError(st, f"{what}: {msg.decode() if msg else ''}")

    # -- data movement -------------------------------------------------------
    def _shape(self, name):
        base = (self.nz + 5, self.nx + 5)
        if name in ("Cnew", "Cvis", "Cbuoy"):
            return base + (self.ntype,), np.float64
        if name == "count_type":
            return base + (self.ntype + 1,), np.int32
        if name == "nempty":
            return base, np.int32
        return base, np.float64

    def put(self, name, arr):
        shape, dt = self._shape(name)
        a = np.ascontiguousarray(arr, dtype=dt)
        if a.shape != shape:
            raise ValueError(f"{name}: expected shape {shape}, got {a.shape}")
        self._check(self.lib.tc_put_field(self.h, _FIELD[name], _ptr(a)), f"put {name}")

    def get(self, name):
        shape, dt = self._shape(name)
        a = np.empty(shape, dtype=dt)
        self._check(self.lib.tc_get_field(self.h, _FIELD[name], _ptr(a)), f"get {name}")
        return a

    def put_tracers(self, rtr, str_, ttype):
        r = np.ascontiguousarray(rtr, dtype=np.float64)
        s = np.ascontiguousarray(str_, dtype=np.float64)
        t = np.ascontiguousarray(ttype, dtype=np.int32)
        if not (r.shape == s.shape == t.shape == (self.ntr,)):
            raise ValueError("tracer arrays must have length ntr")
        self._check(self.lib.tc_put_tracers(self.h, _ptr(r), _ptr(s), _ptr(t)), "put tracers")

    def get_tracers(self, which="x"):
        r = np.empty(self.ntr)
        s = np.empty(self.ntr)
        self._check(self.lib.tc_get_tracers(self.h, _TRACERS[which], _ptr(r), _ptr(s)), "get tracers")
        return r, s

    def get_tracer_types(self):
        t = np.empty(self.ntr, dtype=np.int32)
        self._check(self.lib.tc_get_tracer_types(self.h, _ptr(t)), "get tracer types")
        return t

    def set_equalize(self, tpc, Cinit=(), seed=0, in_time_step=False):
        """equalize_tracers settings (composition.f90:83-580): tpc for comp=0,
        Cinit = the initial mean composition of each type for comp=1; with
        in_time_step, time_step equalizes after stability_RK4."""
        c = np.zeros(max(1, self.params.ntype))
        c[:len(Cinit)] = Cinit
        self._check(self.lib.tc_set_equalize(self.h, int(tpc), _ptr(c), int(seed), int(bool(in_time_step))),
                    "set_equalize")

    def equalize_tracers(self):
        """One equalisation of the current tracers; returns the record whose
        fields are the columns of equalize.dat (tc_equalize_info)."""
        info = _abi.TcEqualizeInfo()
        self._check(self.lib.tc_equalize_tracers(self.h, ctypes.byref(info)), "equalize_tracers")
        return info

    def set_time(self, time, dt=0.0):
        self._check(self.lib.tc_set_time(self.h, float(time), float(dt)), "set_time")

    MODES = {"fp64": _abi.TC_MODE_FP64, "mixed": _abi.TC_MODE_MIXED, "dc64": _abi.TC_MODE_DC64}

    def set_mode(self, mode):
        """'fp64' (the reference's cycle), 'mixed' or 'dc64' (defect correction
        with FP32 or FP64 corrections), or the TC_MODE_* integer."""
        if isinstance(mode, str):
            mode = self.MODES[mode]
        self._check(self.lib.tc_set_mode(self.h, int(mode)), "set_mode")

    def stats(self):
        s = _abi.TcStats()
        self._check(self.lib.tc_get_stats(self.h, ctypes.byref(s)), "stats")
        return s

    def sync(self):
        self._check(self.lib.tc_sync(self.h), "sync")

    # -- the reference's routines ------------------------------------------
    def multigrid(self):
        """multigrid (iterative_solver.f90:79-173) -> (q, residual_mag(1), res_ref(1))."""
        q = ctypes.c_int32()
        r = ctypes.c_double()
        rr = ctypes.c_double()
        self._check(self.lib.tc_multigrid(self.h, ctypes.byref(q), ctypes.byref(r), ctypes.byref(rr)), "multigrid")
        return q.value, r.value, rr.value

    def compute_velocities(self):
        self._check(self.lib.tc_compute_velocities(self.h), "compute_velocities")

    def compute_strain_rate_invariant(self):
        self._check(self.lib.tc_compute_strain_rate_invariant(self.h), "compute_strain_rate_invariant")

    def smoother_time_T(self, src="T"):
        n = ctypes.c_int32()
        self._check(self.lib.tc_smoother_time_T(self.h, _FIELD[src], ctypes.byref(n)), "smoother_time_T")
        return n.value

    def tracers_to_corners(self, which="x", smooth=True):
        n = (ctypes.c_int32 * _abi.TC_MAX_TYPES)()
        self._check(self.lib.tc_tracers_to_corners(self.h, _TRACERS[which], int(bool(smooth)), n),
                    "tracers_to_corners")
        return list(n)[: self.ntype]

    def smoother_time(self):
        n = (ctypes.c_int32 * _abi.TC_MAX_TYPES)()
        self._check(self.lib.tc_smoother_time(self.h, n), "smoother_time")
        return list(n)[: self.ntype]

    def load_conductivity_array(self):
        self._check(self.lib.tc_load_conductivity(self.h), "load_conductivity")

    def tracer_stage(self, stage, dt):
        self._check(self.lib.tc_tracer_stage(self.h, int(stage), float(dt)), "tracer_stage")

    def energy_stage(self, stage, dt):
        self._check(self.lib.tc_energy_stage(self.h, int(stage), float(dt)), "energy_stage")

    def stability_RK4(self):
        dt = ctypes.c_double()
        self._check(self.lib.tc_stability_RK4(self.h, ctypes.byref(dt)), "stability_RK4")
        return dt.value

    # -- learned initial guess ----------------------------------------------
    def learn_attach(self, net):
        """Attach a tracer_code_amd.learn.Network (H = nz, W = nx) or None."""
        self._check(self.lib.tc_learn_attach(self.h, net.h if net is not None else None), "learn_attach")
        self._net = net

    def learn_record(self, capacity):
        self._check(self.lib.tc_learn_record(self.h, int(capacity)), "learn_record")

    def learn_samples(self):
        n = ctypes.c_int32()
        self._check(self.lib.tc_learn_samples(self.h, ctypes.byref(n)), "learn_samples")
        return n.value

    def learn_train(self, epochs, lr, batch=1, seed=0):
        loss = ctypes.c_double()
        self._check(self.lib.tc_learn_train(self.h, int(epochs), float(lr), int(batch), int(seed),
                                            ctypes.byref(loss)), "learn_train")
        return loss.value

    def diagnostics(self):
        """stats (basics.f90:236-324) on the device: dict of vrms, Nu, Tavg,
        Cavg, empty, tmin, tmax, Nu_bot, strain_max."""
        out = np.zeros(9)
        self._check(self.lib.tc_diagnostics(self.h, _ptr(out)), "diagnostics")
        keys = ("vrms", "Nu", "Tavg", "Cavg", "empty", "tmin", "tmax", "Nu_bot", "strain_max")
        return dict(zip(keys, out.tolist()))

    def stats_record(self, kent, mass=None):
        """One stats.dat row (basics.f90:236-324) computed on the device, as a
        TcStatsRow: kent from config.entrainment_row, mass = Cavg of each type
        at the first call (None: this call's Cavg, as the first call uses)."""
        out = _abi.TcStatsRow()
        m = None if mass is None else np.ascontiguousarray(np.atleast_1d(mass), dtype=np.float64)
        self._check(self.lib.tc_stats_record(self.h, int(kent), None if m is None else _ptr(m), ctypes.byref(out)),
                    "stats_record")
        return out

    def snapshot_stage(self):
        """Stage the snapshot fields (main.f90:234-268) for an asynchronous
        copy to the host; time steps may run before snapshot_get."""
        self._check(self.lib.tc_snapshot_stage(self.h), "snapshot_stage")

    def snapshot_get(self, name):
        """One staged field (as get()), waiting for the copy."""
        shape, dt = self._shape(name)
        a = np.empty(shape, dtype=dt)
        self._check(self.lib.tc_snapshot_get(self.h, _FIELD[name], _ptr(a)), f"snapshot_get {name}")
        return a

    def set_tb_tile(self, level, tile):
        self._check(self.lib.tc_set_tb_tile(self.h, int(level), int(tile)), "set_tb_tile")

    def autotune_tb_tiles(self, steps_each=1):
        """Level-by-level choice of the temporally blocked tile (16 or 32);
        returns the tile of each level (index 0 = level 1)."""
        t = np.zeros(_abi.TC_MAX_LEVELS, dtype=np.int32)
        self._check(self.lib.tc_autotune_tb_tiles(self.h, int(steps_each), t.ctypes.data), "autotune_tb_tiles")
        return t.tolist()

    def sort_tracers(self):
        """Reorder the device tracers by node (get_tracers keeps the caller's order)."""
        self._check(self.lib.tc_sort_tracers(self.h), "sort_tracers")

    def set_tracer_sort(self, interval):
        """Sort the tracers at the start of every interval-th time step (0: never)."""
        self._check(self.lib.tc_set_tracer_sort(self.h, int(interval)), "set_tracer_sort")

    def set_gs_smoother(self, sweeps, per_launch=2, omega=1.0):
        """Multicolour Gauss-Seidel sweeps per level-1 visit of the mixed
        mode's correction cycle (0: the Jacobi-Richardson sweeps)."""
        self._check(self.lib.tc_set_gs_smoother(self.h, int(sweeps), int(per_launch), float(omega)),
                    "set_gs_smoother")

    def set_dc_cheb_interval(self, lb=-2.0, la=-1.0):
        """Chebyshev interval of the mixed mode's level-1 visits (the
        reference's -2, -1 by default; la closer to 0 damps more of the
        spectrum per visit)."""
        self._check(self.lib.tc_set_dc_cheb_interval(self.h, float(lb), float(la)), "set_dc_cheb_interval")

    def tracer_slots(self):
        """The caller's index of the tracer in each device slot."""
        out = np.empty(self.ntr, dtype=np.int64)
        self._check(self.lib.tc_tracer_slots(self.h, out.ctypes.data), "tracer_slots")
        return out

    def set_dc_sweeps(self, sweeps):
        """Sweeps per level visit of the defect-correction cycle (0: Nmulti)."""
        self._check(self.lib.tc_set_dc_sweeps(self.h, int(sweeps)), "set_dc_sweeps")

    def set_graphs(self, enable):
        """Replay the coarse part of each FP32 defect-correction cycle as one
        captured HIP graph (bit-identical results; for many concurrent cases)."""
        self._check(self.lib.tc_set_graphs(self.h, int(bool(enable))), "set_graphs")

    def graph_counts(self):
        """(captures, replays) of the cycle-segment graphs so far."""
        c, r = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.tc_graph_counts(self.h, ctypes.byref(c), ctypes.byref(r)), "graph_counts")
        return c.value, r.value

    def set_gcr(self, directions):
        """GCR over the defect-correction cycles with up to `directions` kept
        directions (0..16; 0 runs the plain cycles).  The stop test and the
        tolerance are unchanged."""
        self._check(self.lib.tc_set_gcr(self.h, int(directions)), "set_gcr")

    def set_coarse_refresh(self, solves):
        """Refactorise the coarsest operator every `solves`-th solve of the
        defect-correction modes (1: every solve)."""
        self._check(self.lib.tc_set_coarse_refresh(self.h, int(solves)), "set_coarse_refresh")

    def set_filter_blocking(self, enable):
        """Several C-filter iterations per launch (True, the default) or one."""
        self._check(self.lib.tc_set_filter_blocking(self.h, int(bool(enable))), "set_filter_blocking")

    def gcr_counts(self):
        """(updates, restarts) of the GCR acceleration so far."""
        u, r = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.tc_gcr_counts(self.h, ctypes.byref(u), ctypes.byref(r)), "gcr_counts")
        return u.value, r.value

    def set_dc_level_sweeps(self, level, sweeps):
        """Sweeps per visit of one level of the tuned cycle (0: the common count)."""
        self._check(self.lib.tc_set_dc_level_sweeps(self.h, int(level), int(sweeps)), "set_dc_level_sweeps")

    def autotune_dc_sweeps(self, candidates, steps_each=1):
        """Runs steps_each time steps with each candidate sweep count and
        keeps the best; returns (best, {candidate: ms per step})."""
        c = np.ascontiguousarray(candidates, dtype=np.int32)
        ms = np.zeros(len(c))
        best = ctypes.c_int32()
        self._check(self.lib.tc_autotune_dc_sweeps(self.h, c.ctypes.data, len(c), int(steps_each), ctypes.byref(best),
                                                   ms.ctypes.data), "autotune_dc_sweeps")
        return best.value, {int(k): float(v) for k, v in zip(c, ms)}

    def autotune_dc_level_sweeps(self, first_level, last_level, candidates, steps_each=1):
        """After autotune_dc_sweeps: one sweep count for the levels
        first_level..last_level (0: the common count), tuned the same way;
        returns (best, {candidate: ms per step})."""
        c = np.ascontiguousarray(candidates, dtype=np.int32)
        ms = np.zeros(len(c))
        best = ctypes.c_int32()
        self._check(self.lib.tc_autotune_dc_level_sweeps(self.h, int(first_level), int(last_level), c.ctypes.data,
                                                         len(c), int(steps_each), ctypes.byref(best),
                                                         ms.ctypes.data), "autotune_dc_level_sweeps")
        return best.value, {int(k): float(v) for k, v in zip(c, ms)}

    GUESS_MODES = {"plain": _abi.TC_GUESS_PLAIN, "extrap": _abi.TC_GUESS_EXTRAP, "learned": _abi.TC_GUESS_LEARNED}

    def guess_enable(self, mode):
        """Initial-guess candidates of the defect-correction solves inside
        time_step: "plain" (the last solution only), "extrap" (also the RK
        stage's extrapolation in time), "learned" (also the extrapolation
        plus the attached network's correction); True means "learned"."""
        if mode is True:
            mode = "learned"
        elif mode is False:
            mode = "plain"
        self._check(self.lib.tc_guess_enable(self.h, self.GUESS_MODES[mode]), "guess_enable")

    def set_profiling(self, enable):
        self._check(self.lib.tc_set_profiling(self.h, int(bool(enable))), "set_profiling")

    def kernel_timing(self, which=_abi.TC_K_SWEEP_L1):
        avg = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self.lib.tc_kernel_timing(self.h, int(which), ctypes.byref(avg), ctypes.byref(n)), "kernel_timing")
        return avg.value, n.value

    def time_step(self, nsteps=1):
        self._check(self.lib.tc_time_step(self.h, int(nsteps)), "time_step")

    # -- start-up ------------------------------------------------------------
    def initialize(self, T, rtr, str_, ttype, SF=None):
        """The restart start-up of the reference (main.f90:121-163 with
        restart=1, grid.f90:56-95): load T (boundary conditions applied),
        SF and the tracers, bin the tracers without smoothing, then
        smoother_time_T(T), smoother_time and load_conductivity_array."""
        self.put("T", T)
        sf = np.zeros((self.nz + 5, self.nx + 5)) if SF is None else SF
        self.put("SF", sf)
        if self.ntr:
            self.put_tracers(rtr, str_, ttype)
            self.tracers_to_corners("x", smooth=False)
        self.smoother_time_T("T")
        if self.params.comp:
            self.smoother_time()
        self.load_conductivity_array()
        self.set_time(0.0, 0.0)
