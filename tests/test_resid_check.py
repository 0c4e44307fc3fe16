"""Pins of the extended-precision residual checker (tests/resid_ld.py) that
the defect-correction GPU tests use at full grid sizes.

* It equals the oracle's FP64 compute_residual on a random field with a
  variable viscosity (same operator, different rounding).
* It is exact on a quartic streamfunction with constant viscosity, where
  the 5-point stencils have no truncation error: L(x^4 + x^2 z^2 + z^4) =
  vis (24 + 24 + 2*4).
"""
import numpy as np

import oracle
from conftest import load_oracle_state
from resid_ld import residual_ld
from tracer_code_amd import config, ic


def test_matches_oracle_fp64_residual():
    params, cfg = config.preset_params("cfg0", nx=32, nz=32, tpc=2, grids=2)
    T, r, s, ty = ic.workload(cfg, seed=4)
    o = load_oracle_state(params, T, r, s, ty)
    o.compute_viscosity(1)
    o.compute_RHS(1)
    nx, nz = params.nx, params.nz
    E = o.plane(oracle.OR_ERROR, 1)
    E[:, :nx + 5] = np.random.default_rng(1).standard_normal((nz + 5, nx + 5))
    o.enforce_bcs_stream(1, E.ctypes.data)
    o.compute_residual(1, 1)
    R = o.plane(oracle.OR_RESIDUAL)[3:nz + 2, 3:nx + 2]
    vis = o.plane(oracle.OR_VIS, 1)[:, :nx + 5]
    assert np.ptp(vis) > 0.1 * np.max(vis)  # the gradient terms are exercised
    rl, mag = residual_ld(params, E[:, :nx + 5], None, o.plane(oracle.OR_RHS, 1)[:, :nx + 5], vis)
    scale = float(np.max(np.abs(R)))
    assert float(np.max(np.abs(rl.astype(np.float64) - R))) <= 1e-13 * scale
    assert abs(mag - float(np.mean(np.abs(R)))) <= 1e-13 * mag


def test_exact_on_quartic():
    params, _ = config.preset_params("cfg0", nx=16, nz=8, aspect=2.0, grids=2)
    nx, nz = params.nx, params.nz
    x = np.arange(-2, nx + 3) * (2.0 / nx)
    z = np.arange(-2, nz + 3) * (1.0 / nz)
    X, Z = np.meshgrid(x, z)
    S = X ** 4 + X ** 2 * Z ** 2 + Z ** 4
    vis = np.full_like(S, 2.0)
    rhs = np.full_like(S, 2.0 * (24.0 + 24.0 + 2.0 * 4.0))
    rl, mag = residual_ld(params, S, None, rhs, vis)
    assert mag < 1e-9
