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


@pytest.mark.parametrize("preset,ov", [("cfg0", dict(nx=64, nz=64, tpc=4)), ("cfg2", dict(nx=128, nz=64, tpc=4))])
def test_stats_row_matches_oracle(preset, ov):
    """The whole stats.dat row (tc_stats_record): the integrals to 1e-12
    relative, the corner fluxes to 1e-13, the counts exactly, with the
    entrainment row of dent = 0.1 and with a given mass."""
    params, cfg = config.preset_params(preset, **ov)
    T, r, s, ty = ic.workload(cfg, seed=1)
    kent = config.entrainment_row(params.nz)
    print(f"[stats row {preset} {params.nx}x{params.nz} kent={kent}]", flush=True)
    with Solver(params, len(r)) as g:
        g.initialize(T, r, s, ty)
        g.time_step(2)
        g.compute_velocities()
        first = g.stats_record(kent)
        later = g.stats_record(kent, mass=[0.9 * first.Cavg[0]])
        f = {k: g.get(k) for k in ("T", "u", "w", "conduct", "Cnew", "nempty", "vis")}
    nx = params.nx
    o = oracle.Oracle(params, 0)
    for oid, key in ((oracle.OR_T, "T"), (oracle.OR_U, "u"), (oracle.OR_W, "w"), (oracle.OR_CONDUCT, "conduct")):
        o.plane(oid)[:, :nx + 5] = f[key]
    o.plane(oracle.OR_VIS, 1)[:, :nx + 5] = f["vis"]
    o.plane(oracle.OR_CNEW, 1)[:, :nx + 5] = f["Cnew"][..., 0]
    o.plane(oracle.OR_NEMPTY, 0, np.int32)[:, :nx + 5] = f["nempty"]
    ref1 = o.stats_record(kent)
    ref2 = o.stats_record(kent, mass=[0.9 * ref1.Cavg[0]])
    o.close()
    for dev, ref in ((first, ref1), (later, ref2)):
        for k in ("Tavg", "vrms", "Nu", "Nu_bot", "RaT_avg", "mobility", "strain_max"):
            a, b = getattr(dev, k), getattr(ref, k)
            assert abs(a - b) <= 1e-12 * max(abs(b), 1e-300), (k, a, b)
        for j in range(4):
            assert abs(dev.q[j] - ref.q[j]) <= 1e-13 * max(abs(ref.q[j]), 1e-300), ("q", j)
        for k in ("Cavg", "ent"):
            a, b = getattr(dev, k)[0], getattr(ref, k)[0]
            assert abs(a - b) <= 1e-12 * max(abs(b), 1e-300), (k, a, b)
        assert (dev.empty_cells, dev.tracer_min, dev.tracer_max, dev.kent) == (
            ref.empty_cells, ref.tracer_min, ref.tracer_max, ref.kent)
    print("RaT_avg", first.RaT_avg, "mobility", first.mobility, "q", list(first.q), "ent", first.ent[0], flush=True)
