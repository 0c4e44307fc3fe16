"""Concurrent ensembles (tc_ensemble_step): several independent cases on
different grids, stepped concurrently from host threads on their own
streams, give bit-identical results to stepping each case alone."""
import numpy as np
import pytest

from tracer_code_amd import ensemble

pytestmark = pytest.mark.gpu


def small_cases():
    """Three cases on different grids with moderate physics (the drawn
    distribution's extremes need the larger grids)."""
    cases = ensemble.draw_cases(3, seed=7)
    for c, (nx, nz), ra, vt in zip(cases, ((64, 64), (128, 64), (96, 32)), (1e5, 3e5, 1e6), (1e2, 1e3, 1e2)):
        c.update(nx=nx, nz=nz, aspect=nx / nz, tpc=4, RaT=ra, RaC=(0.5 * ra,), visT=vt)
    return cases


@pytest.mark.parametrize("mode,guess", [("fp64", "plain"), ("mixed", "extrap")])
def test_concurrent_equals_alone(mode, guess):
    cases = small_cases()
    print(f"[ensemble {[(c['nx'], c['nz']) for c in cases]} mode={mode}]", flush=True)
    fields = []
    for concurrent in (False, True):
        solvers, _ = ensemble.make_solvers(cases, mode, guess)
        if concurrent:
            ensemble.ensemble_step(solvers, 3)
        else:
            for g in solvers:
                g.time_step(3)
        fields.append([{k: g.get(k) for k in ("SF", "T", "Cnew")} | dict(zip(("r", "s"), g.get_tracers("x")))
                       for g in solvers])
        for g in solvers:
            g.close()
    for a, b in zip(*fields):
        for k in a:
            assert np.array_equal(a[k], b[k]), k


def test_drawn_cases_at_full_size():
    """Six cases of the batched-ensemble distribution (full grids from
    {256x256, 512x256, 768x256, 1024x512}, drawn physics, 10-40 tracers per
    node), mixed mode with GCR(16) and the learned guess's candidates, two
    steps: the concurrent arm is bit-identical to running the cases one
    after another, and the last solve of every case met the reference's
    stop test (mean |r| <= tolerance * res_ref; a solve that cannot meet it
    returns an error)."""
    cases = ensemble.draw_cases(6, seed=0)
    print(f"[ensemble full size {[(c['nx'], c['nz']) for c in cases]}]", flush=True)
    res = ensemble.run(cases, 2, mode="mixed", guess="extrap", gcr=16, log=lambda m: print(m, flush=True))
    assert res["identical_to_sequential"] == len(cases)
    for arm in ("sequential", "concurrent"):
        assert res[arm]["max_last_residual_over_tolerance"] <= 1.0
        assert min(res[arm]["cycles_per_step"]) > 0
