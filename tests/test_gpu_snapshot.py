"""Snapshot staging (tc_snapshot_stage / tc_snapshot_get, main.f90:234-268).

The fields staged after a time step come back bit for bit equal to the ones
read synchronously at that moment, although another time step ran on the
compute stream before they were fetched; the staging leaves the time steps
themselves bit-identical to a run without it; and the snapshot writer turns
the staged fields into the reference's records (g14.5E3 columns, one blank
record per column of nodes).
"""
import time

import numpy as np
import pytest

from tracer_code_amd import Solver, config, ic
from tracer_code_amd.outputs import snapshot_lines, staged_snapshot_fields

pytestmark = pytest.mark.gpu

FIELDS = ("T", "SF", "u", "w", "vis", "Tvis", "nempty", "Cnew", "Cvis")


def test_staged_fields_are_the_step_they_were_staged_at():
    params, cfg = config.preset_params("cfg0", nx=64, nz=64, tpc=4)
    T, r, s, ty = ic.workload(cfg, seed=1)
    print(f"[snapshot cfg0 {params.nx}x{params.nz}]", flush=True)
    with Solver(params, len(r)) as g, Solver(params, len(r)) as ref:
        for x in (g, ref):
            x.initialize(T, r, s, ty)
            x.time_step(2)
        g.compute_strain_rate_invariant()
        now = {k: g.get(k) for k in FIELDS + ("strain",)}
        g.snapshot_stage()
        g.time_step(1)
        ref.time_step(1)
        for k in FIELDS + ("strain",):
            assert np.array_equal(g.snapshot_get(k), now[k]), k
        assert np.array_equal(g.get("SF"), ref.get("SF"))
        assert np.array_equal(g.get("T"), ref.get("T"))
        g.snapshot_stage()
        g.time_step(1)
        ref.time_step(1)
        f = staged_snapshot_fields(g)
    nx, nz = params.nx, params.nz
    xg = np.arange(nx + 1) * (params.aspect / nx)
    zg = np.arange(nz + 1) / nz
    lines = snapshot_lines(xg, zg, f, comp=1, ntype=1)
    assert len(lines) == (nx + 1) * (nz + 2)
    row = lines[5 * (nz + 2) + 7]          # i = 5, k = 7
    assert len(row) == 9 * 14 + 6 + 2 * 14
    vals = [float(row[14 * j:14 * (j + 1)]) for j in range(9)]
    want = [xg[5], zg[7], f["T"][7, 5], f["SF"][7, 5], f["u"][7, 5], f["w"][7, 5], f["strain"][7, 5],
            f["vis"][7, 5], f["Tvis"][7, 5]]
    for a, b in zip(vals, want):
        assert a == pytest.approx(b, rel=5e-5, abs=1e-300)   # 5 significant digits
    assert int(row[126:132]) == f["nempty"][7, 5]
    assert lines[nz + 1] == "  "


def test_stage_cost_at_full_size():
    params, cfg = config.preset_params("cfg2")
    T, r, s, ty = ic.workload(cfg, seed=1)
    print(f"[snapshot full size {params.nx}x{params.nz}]", flush=True)
    with Solver(params, len(r)) as g:
        g.initialize(T, r, s, ty)
        g.set_mode("mixed")
        g.time_step(1)
        g.snapshot_stage()
        g.snapshot_get("T")          # the first stage's host copy is done
        g.sync()
        t0 = time.perf_counter()
        g.snapshot_stage()
        t1 = time.perf_counter()
        g.sync()
        t2 = time.perf_counter()
        sf = g.snapshot_get("SF")
        t3 = time.perf_counter()
        assert np.array_equal(sf, g.get("SF"))
    print(f"stage call {1e3 * (t1 - t0):.2f} ms, compute stream done {1e3 * (t2 - t0):.2f} ms, "
          f"host copy ready {1e3 * (t3 - t0):.2f} ms", flush=True)
