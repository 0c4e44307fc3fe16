"""The residual+restriction fused into the last blocked launch of a visit.

In the FP32 defect-correction cycle, a visit of a level >= 2 that is followed
by a descent forms the next coarser level's right-hand side in its last
blocked-sweep launch (16x16 tiles, a halo 3 cells wider, on levels of at
most one tile per CU), instead of a separate k_residual_restrict32 launch
after it; for the coarsest level it also forms the coarse solve's
equilibrated right-hand side.  The values are the
separate launch's, so whole time steps must be bit-identical with the fusion
off (TC_FUSED_RR=0, read when a context is created): streamfunction,
temperature, tracers, cycle, sweep and residual counts.  Cases: the common
sweep count (4 sweeps in the last launch), the reference's 70 (2 in the last
launch), per-level counts with 8-sweep launches on the small levels, GCR,
graph replay, and an aspect-3 grid.
"""
import os

import numpy as np
import pytest

from tracer_code_amd import Solver, config, ic

pytestmark = pytest.mark.gpu


def run(preset, ov, fused, steps=3, sweeps=32, gcr=0, graphs=False, levels=None):
    params, cfg = config.preset_params(preset, **ov)
    T, r, s, ty = ic.workload(cfg, seed=1)
    old = os.environ.get("TC_FUSED_RR")
    os.environ["TC_FUSED_RR"] = "1" if fused else "0"
    try:
        g = Solver(params, len(r))
    finally:
        if old is None:
            del os.environ["TC_FUSED_RR"]
        else:
            os.environ["TC_FUSED_RR"] = old
    with g:
        g.initialize(T, r, s, ty)
        g.set_mode("mixed")
        g.set_dc_sweeps(sweeps)
        for lv, n in (levels or {}).items():
            g.set_dc_level_sweeps(lv, n)
        g.set_gcr(gcr)
        g.set_graphs(graphs)
        g.time_step(steps)
        st = g.stats()
        return {"SF": g.get("SF"), "T": g.get("T"), "r": g.get_tracers("x")[0], "cycles": st.mg_cycles,
                "sweeps": list(st.sweeps_exec), "residuals": list(st.residuals_exec), "grids": params.grids}


CASES = [
    ("cfg0", dict(nx=64, nz=64, tpc=4), dict()),
    ("cfg2", dict(nx=256, nz=128, tpc=4), dict()),
    ("cfg0", dict(nx=64, nz=64, tpc=4), dict(sweeps=70)),
    ("cfg2", dict(nx=256, nz=128, tpc=4), dict(sweeps=20, levels={2: 16, 3: 8, 4: 8, 5: 8})),
    ("cfg2", dict(nx=256, nz=128, tpc=4), dict(sweeps=20, gcr=16)),
    ("cfg0", dict(nx=64, nz=64, tpc=4), dict(graphs=True)),
    ("cfg1", dict(nx=192, nz=64, tpc=4), dict()),
]


@pytest.mark.parametrize("preset,ov,kw", CASES)
def test_fused_restriction_is_bit_identical(preset, ov, kw):
    print(f"[fused restriction {preset} {ov} {kw}]", flush=True)
    a = run(preset, ov, False, **kw)
    b = run(preset, ov, True, **kw)
    print(f"grids {b['grids']} cycles {b['cycles']} residuals {b['residuals']}", flush=True)
    assert b["grids"] >= 3
    assert a["cycles"] == b["cycles"]
    assert a["sweeps"] == b["sweeps"]
    assert a["residuals"] == b["residuals"]
    for k in ("SF", "T", "r"):
        assert np.array_equal(a[k], b[k]), k
