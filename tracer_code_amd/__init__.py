"""MI355X-native (gfx950, HIP) implementation of the time-step hot path of
seantrim/tracer_code: the variable-viscosity streamfunction multigrid solve
and the tracer-ratio composition tracers, behind the C-ABI declared in
include/tracer_hip.h.

    from tracer_code_amd import config, ic, Solver
    params, cfg = config.preset_params("cfg0")
    with Solver(params, ntr) as s:
        s.initialize(T, rtr, str_, ttype)
        s.time_step(10)
"""
from . import _abi, config, ic  # noqa: F401
from .model import Solver, TcError, library  # noqa: F401

__all__ = ["Solver", "TcError", "config", "ic", "library"]
