"""Residual RHS - L(psi) of the streamfunction equation in extended precision
(numpy longdouble, 64-bit significand on x86), for checking solves on the
host at any grid size.

L is the variable-viscosity operator of stream_space (build=0, scaling=0,
iterative_solver.f90:483-560) on a uniform grid:
  L(S) = (vis_xx - vis_zz)(S_rr - S_ss) + vis (S_rrrr + S_ssss) + 2 vis S_rrss
         + 4 vis_xz S_rs + 2 vis_x (S_rrr + S_rss) + 2 vis_z (S_sss + S_rrs)
with the viscosity derivatives taken by the same finite differences from vis
(viscosity_gradients, iterative_solver.f90:750-790).  Fields use the
reference's ghosted layout f[k+2, i+2] = f(i,k).
"""
import numpy as np

LD = np.longdouble


def _shift(f, di, dk, nx, nz):
    return f[2 + 1 + dk:2 + nz + dk, 2 + 1 + di:2 + nx + di]


def _line(f, D, axis, nx, nz, width):
    acc = np.zeros((nz - 1, nx - 1), dtype=LD)
    for j in range(-width, width + 1):
        w = LD(D[j + 2])
        if w == 0:
            continue
        acc += w * (_shift(f, j, 0, nx, nz) if axis == 0 else _shift(f, 0, j, nx, nz))
    return acc


def _cross(f, Dk, Di, nx, nz):
    acc = np.zeros((nz - 1, nx - 1), dtype=LD)
    for dk in (-1, 0, 1):
        for di in (-1, 0, 1):
            w = LD(Dk[dk + 2]) * LD(Di[di + 2])
            if w != 0:
                acc += w * _shift(f, di, dk, nx, nz)
    return acc


def operator_ld(params, S, vis):
    """L(S) on the interior i=1..nx-1, k=1..nz-1 as a (nz-1, nx-1) longdouble array."""
    nx, nz = params.nx, params.nz
    D1, D2, D3, D4 = list(params.D1), list(params.D2), list(params.D3), list(params.D4)
    dr, ds = LD(params.aspect) / LD(nx), LD(1) / LD(nz)
    S = np.asarray(S, dtype=LD)
    V = np.asarray(vis, dtype=LD)
    Srr = _line(S, D2, 0, nx, nz, 1) / dr ** 2
    Sss = _line(S, D2, 1, nx, nz, 1) / ds ** 2
    Srrr = _line(S, D3, 0, nx, nz, 2) / dr ** 3
    Ssss = _line(S, D3, 1, nx, nz, 2) / ds ** 3
    Srrrr = _line(S, D4, 0, nx, nz, 2) / dr ** 4
    Sssss = _line(S, D4, 1, nx, nz, 2) / ds ** 4
    Srs = _cross(S, D1, D1, nx, nz) / (dr * ds)
    Srss = _cross(S, D2, D1, nx, nz) / (dr * ds ** 2)
    Srrs = _cross(S, D1, D2, nx, nz) / (dr ** 2 * ds)
    Srrss = _cross(S, D2, D2, nx, nz) / (dr ** 2 * ds ** 2)
    v = _shift(V, 0, 0, nx, nz)
    L = v * (Srrrr + Sssss) + 2 * v * Srrss
    if params.ivis != 0:
        vx = _line(V, D1, 0, nx, nz, 1) / dr
        vz = _line(V, D1, 1, nx, nz, 1) / ds
        vxx = _line(V, D2, 0, nx, nz, 1) / dr ** 2
        vzz = _line(V, D2, 1, nx, nz, 1) / ds ** 2
        vxz = _cross(V, D1, D1, nx, nz) / (dr * ds)
        L += (vxx - vzz) * (Srr - Sss) + 4 * vxz * Srs + 2 * vx * (Srrr + Srss) + 2 * vz * (Ssss + Srrs)
    return L


def residual_ld(params, psi_hi, psi_lo, rhs, vis):
    """Interior residual RHS - L(psi_hi + psi_lo) and its mean magnitude
    (the reference's residual_mag: sum |r| / ((nx-1)(nz-1)))."""
    S = np.asarray(psi_hi, dtype=LD)
    if psi_lo is not None:
        S = S + np.asarray(psi_lo, dtype=LD)
    r = _shift(np.asarray(rhs, dtype=LD), 0, 0, params.nx, params.nz) - operator_ld(params, S, vis)
    mag = np.sum(np.abs(r)) / LD(params.nx - 1) / LD(params.nz - 1)
    return r, float(mag)
