import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_oracle_state(params, T, r, s, ty, SF=None):
    """Oracle run initialised exactly like tracer_code_amd.Solver.initialize."""
    import oracle
    o = oracle.Oracle(params, len(r))
    o.plane(oracle.OR_T)[:, :] = T
    if SF is not None:
        o.plane(oracle.OR_SF)[:, :] = SF
    if len(r):
        o.tracers(oracle.OR_RTR)[:] = r
        o.tracers(oracle.OR_STR)[:] = s
        o.tracers(oracle.OR_TTYPE, np.int32)[:] = ty
    st = o.initialize_state()
    assert st == 0
    return o


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    if den == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / den)
