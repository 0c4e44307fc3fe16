"""The level-1 Chebyshev interval of the FP32 correction cycle
(tc_set_dc_cheb_interval; the reference's sweep factors,
iterative_solver.f90:873-899, take lb = -2, la = -1).

* Setting the reference's interval explicitly gives the same bits as the
  default.
* A wider interval (la = -0.25) with 10 sweeps per visit solves every
  solve to the reference's criterion, on the device residual and on the
  independent extended-precision host residual.
* Bounds: lb <= la < 0, else the call fails and the context is unchanged.
"""
import numpy as np
import pytest

from test_gpu_dc import check_dc_solution, solver_for
from tracer_code_amd.model import TcError

pytestmark = pytest.mark.gpu


def steps(preset, ov, interval, sweeps, n=2):
    params, g, _ = solver_for(preset, **ov)
    with g:
        g.set_mode("mixed")
        g.set_dc_sweeps(sweeps)
        g.set_gcr(16)
        if interval is not None:
            g.set_dc_cheb_interval(*interval)
        g.time_step(n)
        return g.get("SF"), g.stats().mg_cycles


def test_reference_interval_is_the_default():
    ov = dict(nx=128, nz=64, tpc=4)
    a, ca = steps("cfg2", ov, None, 20)
    b, cb = steps("cfg2", ov, (-2.0, -1.0), 20)
    print(f"cycles {ca} {cb}", flush=True)
    assert ca == cb and np.array_equal(a, b)


@pytest.mark.parametrize("preset,ov", [("cfg2", dict(nx=256, nz=128, tpc=4)), ("cfg0", dict(nx=64, nz=64)),
                                       ("cfg1", dict(nx=96, nz=32, tpc=8))])
def test_wide_interval_solves_to_tolerance(preset, ov):
    params, g, _ = solver_for(preset, **ov)
    with g:
        g.set_mode("mixed")
        g.set_gcr(16)
        g.set_dc_sweeps(10)
        g.set_dc_cheb_interval(-2.0, -0.25)
        for _ in range(2):
            g.put("SF", np.zeros_like(g.get("SF")))
            c0 = g.stats().mg_cycles
            g.multigrid()
            check_dc_solution(params, g, f"{preset} la=-0.25 10 sweeps ({g.stats().mg_cycles - c0} cycles)")
        assert g.stats().dc_retries == 0


def test_interval_bounds():
    params, g, _ = solver_for("cfg0", nx=32, nz=32)
    with g:
        for lb, la in ((-2.0, 0.0), (-2.0, 0.5), (-1.0, -2.0), (float("nan"), -1.0)):
            with pytest.raises(TcError):
                g.set_dc_cheb_interval(lb, la)
        g.set_dc_cheb_interval(-2.0, -0.25)
