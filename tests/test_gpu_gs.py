"""The multicolour Gauss-Seidel level-1 smoother of the mixed mode
(tc_set_gs_smoother; the reference smooths with Jacobi-Richardson sweeps).
It is an accelerated form, not held to bit parity: every solve meets the
reference's stop test on the device residual and on an independent
extended-precision host residual, the solution agrees with the
reference-form FP64 solve to the accuracy the tolerance allows (the same
1e-5 relative bar as the other accelerated forms), and a repeated solve
gives the same bits (the tiles read the previous launch's halo, so the
result does not depend on workgroup timing)."""
import numpy as np
import pytest

from conftest import rel_err
from test_gpu_dc import check_dc_solution, solver_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("preset,ov", [("cfg0", dict(nx=64, nz=64)), ("cfg2", dict(nx=256, nz=128, tpc=4))])
def test_gs_smoother_solves_to_tolerance(preset, ov):
    params, g, _ = solver_for(preset, **ov)
    with g:
        g.multigrid()
        sf64 = g.get("SF")
        g.set_mode("mixed")
        g.set_gs_smoother(12, 2, 0.5)
        sols = []
        for rep in range(2):
            g.put("SF", np.zeros_like(sf64))
            g.multigrid()
            sols.append(check_dc_solution(params, g, f"mixed+gs rep {rep}"))
        e = rel_err(sols[0], sf64)
        print(f"gs vs fp64 reference form: {e:.2e}", flush=True)
        assert e <= 1e-5
        assert np.array_equal(sols[0], sols[1])
        g.set_gs_smoother(0)


def test_gs_smoother_rejects_bad_arguments():
    params, g, _ = solver_for("cfg0", nx=32, nz=32)
    with g:
        for args in ((-1, 2, 0.5), (4, 0, 0.5), (4, 9, 0.5), (4, 2, 0.0), (4, 2, 2.0)):
            with pytest.raises(Exception):
                g.set_gs_smoother(*args)
