"""Pins of the oracle's boundary routines for T and C.

* enforceBCs (energy.f90:455-475): mirrored side ghosts on rows 0..nz; the
  bottom ghost rows reflect T about 1 (isothermal bottom, Tbot=0) or copy it
  (insulating, Tbot=1); the top ghost rows reflect T about 0; the walls are
  then set to 1 (bottom, Tbot=0) and 0 (top) over the whole ghosted row.
* enforceBCs_comp (composition.f90:792-803): mirrors on all four sides, with
  the reference's interleaved loop: the first ghost rows are filled before the
  second ghost columns, so the four cells (-2 or nx+2, -1 or nz+1) keep the
  value the second ghost column held at row 1 (nz-1) before the call.  A
  restatement that did the sides first and the rows after would give the
  mirrored interior value there instead.
"""
import numpy as np
import pytest

import oracle
from conftest import load_oracle_state
from tracer_code_amd import config, ic


def state(nx=12, nz=10, **kw):
    params, cfg = config.preset_params("cfg0", nx=nx, nz=nz, tpc=2, grids=2, **kw)
    T, r, s, ty = ic.workload(cfg, seed=5)
    return params, load_oracle_state(params, T, r, s, ty)


def at(f, i, k):
    return f[k + 2, i + 2]


@pytest.mark.parametrize("tbot", [0, 1])
def test_temperature_bcs(tbot):
    params, o = state(Tbot=tbot)
    nx, nz = params.nx, params.nz
    rng = np.random.default_rng(3)
    T = o.plane(oracle.OR_T)
    T[:] = rng.uniform(-0.5, 1.5, T.shape)
    before = T.copy()
    o.enforce_bcs(T.ctypes.data)
    for k in range(0, nz + 1):
        for i in (1, 2):
            if 0 < k < nz:
                assert at(T, -i, k) == at(before, i, k)
                assert at(T, nx + i, k) == at(before, nx - i, k)
    for i in range(-2, nx + 3):
        for k in (1, 2):
            up = at(T, i, k)
            assert at(T, i, nz + k) == -at(T, i, nz - k)
            assert at(T, i, -k) == (-(up - 1.0) + 1.0 if tbot == 0 else up)
        assert at(T, i, nz) == 0.0
        assert at(T, i, 0) == (1.0 if tbot == 0 else at(T, i, 0))
    # the interior is untouched
    assert np.array_equal(T[3:nz + 2, 2:nx + 3], before[3:nz + 2, 2:nx + 3])
    if tbot == 1:
        assert np.array_equal(T[2, 2:nx + 3], before[2, 2:nx + 3])


def test_composition_bcs_interleaved_corners():
    params, o = state()
    nx, nz = params.nx, params.nz
    rng = np.random.default_rng(4)
    C = o.plane(oracle.OR_CNEW, 1)
    C[:] = rng.uniform(0.0, 1.0, C.shape)
    before = C.copy()
    o.enforce_bcs_comp(C.ctypes.data)
    for k in range(0, nz + 1):
        for i in (1, 2):
            assert at(C, -i, k) == at(before, i, k)
            assert at(C, nx + i, k) == at(before, nx - i, k)
    for i in range(-2, nx + 3):
        for k in (1, 2):
            if k == 1 and i in (-2, nx + 2):
                # the second ghost column before its refresh
                assert at(C, i, -1) == at(before, i, 1)
                assert at(C, i, nz + 1) == at(before, i, nz - 1)
                continue
            assert at(C, i, -k) == at(C, i, k)
            assert at(C, i, nz + k) == at(C, i, nz - k)
    assert at(C, -2, -2) == at(before, 2, 2) and at(C, nx + 2, nz + 2) == at(before, nx - 2, nz - 2)
    assert np.array_equal(C[2:nz + 3, 2:nx + 3], before[2:nz + 3, 2:nx + 3])
