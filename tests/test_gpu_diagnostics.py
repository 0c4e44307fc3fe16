"""Device stats (tc_diagnostics, basics.f90:236-324) against the oracle's
stats evaluated on the same device fields after two time steps: the
integrals (vrms, Nu, Tavg, Cavg, Nu at the bottom) and the surface strain
maximum to 1e-12 relative (the device sums in another order), the tracer
counts exactly."""
import numpy as np
import pytest

import oracle
from tracer_code_amd import Solver, config, ic

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("preset,ov", [("cfg0", dict(nx=64, nz=64, tpc=4)), ("cfg2", dict(nx=128, nz=64, tpc=4))])
def test_device_stats_match_oracle(preset, ov):
    params, cfg = config.preset_params(preset, **ov)
    T, r, s, ty = ic.workload(cfg, seed=1)
    print(f"[diagnostics {preset} {params.nx}x{params.nz}]", flush=True)
    with Solver(params, len(r)) as g:
        g.initialize(T, r, s, ty)
        g.time_step(2)
        g.compute_velocities()
        dev = g.diagnostics()
        f = {k: g.get(k) for k in ("T", "u", "w", "conduct", "Cnew", "nempty")}
    nx = params.nx
    o = oracle.Oracle(params, 0)
    for oid, key in ((oracle.OR_T, "T"), (oracle.OR_U, "u"), (oracle.OR_W, "w"), (oracle.OR_CONDUCT, "conduct")):
        o.plane(oid)[:, :nx + 5] = f[key]
    o.plane(oracle.OR_CNEW, 1)[:, :nx + 5] = f["Cnew"][..., 0]
    o.plane(oracle.OR_NEMPTY, 0, np.int32)[:, :nx + 5] = f["nempty"]
    st = np.zeros(9)
    o.stats(st.ctypes.data)
    o.close()
    ref = dict(zip(("vrms", "Nu", "Tavg", "Cavg", "empty", "tmin", "tmax", "Nu_bot", "strain_max"), st.tolist()))
    print({k: (dev[k], ref[k]) for k in ref}, flush=True)
    for k in ("vrms", "Nu", "Tavg", "Cavg", "Nu_bot", "strain_max"):
        assert abs(dev[k] - ref[k]) <= 1e-12 * max(abs(ref[k]), 1e-300), k
    for k in ("empty", "tmin", "tmax"):
        assert dev[k] == ref[k], k
