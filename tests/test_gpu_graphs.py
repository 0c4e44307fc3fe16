"""HIP-graph replay of the coarse segment of the FP32 defect-correction cycle.

The part of every cycle below level 1 is captured once per solve and replayed
in the solve's other cycles, with the eager launches of level 1, the stop
tests, the filters and the tracer stages in between.  Every replay must give
what the launches give: the runs are compared bit for bit (streamfunction,
temperature, tracers, cycle counts) with graphs off.  Cases: two presets at
the common sweep count (even launches per visit, one graph per solve), the
reference's 70 sweeps (odd launches per visit on the small levels, so the
buffer parity alternates between cycles and two graphs are used), and an
ensemble of contexts stepped concurrently.
"""
import numpy as np
import pytest

from tracer_code_amd import Solver, config, ic
from tracer_code_amd.model import library

pytestmark = pytest.mark.gpu


def run(preset, ov, graphs, steps=3, sweeps=32, gcr=0):
    params, cfg = config.preset_params(preset, **ov)
    T, r, s, ty = ic.workload(cfg, seed=1)
    with Solver(params, len(r)) as g:
        g.initialize(T, r, s, ty)
        g.set_mode("mixed")
        g.set_dc_sweeps(sweeps)
        g.set_graphs(graphs)
        g.set_gcr(gcr)
        g.time_step(steps)
        out = {"SF": g.get("SF"), "T": g.get("T"), "r": g.get_tracers("x")[0], "cycles": g.stats().mg_cycles,
               "counts": g.graph_counts(), "sweeps": list(g.stats().sweeps_exec)}
    return out


@pytest.mark.parametrize("preset,ov,sweeps", [("cfg0", dict(nx=64, nz=64, tpc=4), 32),
                                              ("cfg2", dict(nx=128, nz=64, tpc=4), 32),
                                              ("cfg0", dict(nx=64, nz=64, tpc=4), 70)])
def test_graph_replay_is_bit_identical(preset, ov, sweeps):
    print(f"[graphs {preset} {ov} sweeps={sweeps}]", flush=True)
    a = run(preset, ov, False, sweeps=sweeps)
    b = run(preset, ov, True, sweeps=sweeps)
    captures, replays = b["counts"]
    print(f"captures {captures} replays {replays} cycles {b['cycles']}", flush=True)
    assert a["counts"] == (0, 0)
    assert captures >= 1 and replays > captures
    assert a["cycles"] == b["cycles"]
    assert a["sweeps"] == b["sweeps"]
    for k in ("SF", "T", "r"):
        assert np.array_equal(a[k], b[k]), k


def test_graph_replay_with_gcr_is_bit_identical():
    """GCR's step runs after each replayed segment and changes psi and the
    residual in place; the replays stay bit-identical to the launches."""
    ov = dict(nx=64, nz=64, tpc=4)
    print(f"[graphs + gcr 16 cfg0 {ov}]", flush=True)
    a = run("cfg0", ov, False, gcr=16)
    b = run("cfg0", ov, True, gcr=16)
    print(f"captures/replays {b['counts']} cycles {b['cycles']}", flush=True)
    assert b["counts"][1] > b["counts"][0] >= 1
    assert a["cycles"] == b["cycles"]
    for k in ("SF", "T", "r"):
        assert np.array_equal(a[k], b[k]), k


def test_graph_replay_in_an_ensemble():
    import ctypes
    lib = library()
    cases = [("cfg0", dict(nx=64, nz=64, tpc=4)), ("cfg0", dict(nx=96, nz=48, aspect=2.0, tpc=4)),
             ("cfg2", dict(nx=128, nz=64, tpc=4)), ("cfg0", dict(nx=128, nz=64, aspect=2.0, tpc=4))]
    alone = [run(p, ov, False, steps=2) for p, ov in cases]
    solvers = []
    try:
        for p, ov in cases:
            params, cfg = config.preset_params(p, **ov)
            T, r, s, ty = ic.workload(cfg, seed=1)
            g = Solver(params, len(r))
            g.initialize(T, r, s, ty)
            g.set_mode("mixed")
            g.set_dc_sweeps(32)
            g.set_graphs(True)
            solvers.append(g)
        arr = (ctypes.c_void_p * len(solvers))(*[g.h for g in solvers])
        failed = ctypes.c_int32(-1)
        assert lib.tc_ensemble_step(arr, len(solvers), 2, len(solvers), ctypes.byref(failed)) == 0
        for g, ref in zip(solvers, alone):
            assert g.graph_counts()[1] > 0
            assert np.array_equal(g.get("SF"), ref["SF"])
            assert np.array_equal(g.get_tracers("x")[0], ref["r"])
    finally:
        for g in solvers:
            g.close()
