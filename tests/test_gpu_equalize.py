"""GPU parity of equalize_tracers (composition.f90:83-580) against the oracle.

Both run the counter-based form from identical states, so the bar is bit for
bit: tracer positions and types, the binned counts, nempty and Cnew, and the
equalize.dat record.  Covered: the hand-built lattices of the oracle's pins
(one rich and one poor node, wall donors, donors running short, two types
converted toward the initial totals), evolved states of three presets, the
device tracer order (node-sorted), two calls in a row, and equalisation
inside the time step.  At 2048x1024 with 42M tracers the size-independent
properties hold: tracers conserved, a node only gives or only gets, and
every moved tracer sits in the control volume of a node that gained.
"""
import time

import numpy as np
import pytest

import oracle
from conftest import load_oracle_state, rel_err
from test_oracle_equalize import lattice, params_for
from tracer_code_amd import Solver, config, ic

pytestmark = pytest.mark.gpu

INFO_FIELDS = ("did_work", "passes", "deficit", "moved", "nfree")


def info_tuple(info, nt):
    return (tuple(getattr(info, f) for f in INFO_FIELDS) + tuple(info.tpc_type[:nt + 1]) +
            tuple(info.minexcess[:nt + 1]) + tuple(info.count[:nt + 1]))


def pair(params, T, r, s, ty, SF=None):
    o = load_oracle_state(params, T, r, s, ty, SF=SF)
    g = Solver(params, len(r))
    g.initialize(T, r, s, ty, SF=SF)
    return o, g


def assert_same(o, g, params, info_o, info_g):
    nt = params.ntype if params.comp else 0
    assert info_tuple(info_g, nt) == info_tuple(info_o, nt)
    r, s = g.get_tracers("x")
    assert np.array_equal(r, o.tracers(oracle.OR_RTR))
    assert np.array_equal(s, o.tracers(oracle.OR_STR))
    assert np.array_equal(g.get_tracer_types(), o.tracers(oracle.OR_TTYPE, np.int32))
    assert np.array_equal(g.get("nempty"), o.plane(oracle.OR_NEMPTY, 0, np.int32))
    ct = g.get("count_type")
    for t in range(params.ntype + 1):
        assert np.array_equal(ct[..., t], o.plane(oracle.OR_COUNT, t, np.int32))
    cn = g.get("Cnew")
    for t in range(params.ntype):
        assert np.array_equal(cn[..., t], o.plane(oracle.OR_CNEW, t + 1))


def lattice_case(name):
    nx = nz = 12
    tpc = 6
    m = np.full((nz + 1, nx + 1), tpc)
    params = params_for(nx, nz, tpc)
    if name == "pair":
        m[3, 2] += 3
        m[5, 6] -= 3
    elif name == "short":
        m[2, 2] += 2
        m[6, 2] -= 2
        m[6, 6] -= 2
        m[9, 9] -= 2
    elif name == "wall":
        m[4, 0] += 2
        m[4, 4] -= 2
        m[0, 7] += 3
        m[8, 8] -= 1
    elif name == "types":
        m[2, 6] += 2
        m[9, 6] -= 2
        m[5, 3] -= 1
    r, s, node = lattice(nx, nz, m, seed=3)
    ty = np.zeros(len(r), dtype=np.int32)
    Cinit = [0.0]
    if name == "wall":
        wall = np.flatnonzero(node == 4 * (nx + 1))
        r[wall[[1, 4]]] = 0.3e-4 / nx
    if name == "types":
        ty = ((node // (nx + 1)) <= 4).astype(np.int32)
        Cinit = [(int(ty.sum()) - 2) / len(r)]
    T = ic.initial_temperature(nx, nz, 1.0, seed=1)
    return params, T, r, s, ty, Cinit, tpc


@pytest.mark.parametrize("name", ["pair", "short", "wall", "types"])
def test_lattice_cases_match_oracle(name):
    params, T, r, s, ty, Cinit, tpc = lattice_case(name)
    print(f"[equalize lattice {name}] ntr={len(r)}", flush=True)
    o, g = pair(params, T, r, s, ty)
    with g:
        g.set_equalize(tpc, Cinit, seed=1)
        for call in range(2):
            st, info_o = o.equalize_tracers(tpc, Cinit, 1, call)
            assert st == 0
            info_g = g.equalize_tracers()
            assert_same(o, g, params, info_o, info_g)


EVOLVED = {
    # name: (preset, overrides, oracle time steps before equalising)
    "cfg0_32": ("cfg0", dict(nx=32, nz=32, tpc=6), 8),
    "cfg0_48x24": ("cfg0", dict(nx=48, nz=24, aspect=2.0, tpc=5), 10),
    "cfg1_192x64": ("cfg1", dict(nx=192, nz=64, tpc=6), 1),
}


def evolved_state(case):
    preset, ov, steps = EVOLVED[case]
    params, cfg = config.preset_params(preset, **ov)
    T, r, s, ty = ic.workload(cfg, seed=1)
    o = load_oracle_state(params, T, r, s, ty)
    for _ in range(steps):
        assert o.time_step() == 0
    T1 = o.plane(oracle.OR_T).copy()
    SF1 = o.plane(oracle.OR_SF).copy()
    r1 = o.tracers(oracle.OR_RTR).copy()
    s1 = o.tracers(oracle.OR_STR).copy()
    ty1 = o.tracers(oracle.OR_TTYPE, np.int32).copy()
    o.close()
    return params, T1, SF1, r1, s1, ty1, [float(ty.mean())], ov["tpc"]


@pytest.mark.parametrize("sort", [False, True])
@pytest.mark.parametrize("case", list(EVOLVED))
def test_evolved_state_matches_oracle(case, sort):
    params, T, SF, r, s, ty, Cinit, tpc = evolved_state(case)
    print(f"[equalize evolved {case} sort={sort}] ntr={len(r)}", flush=True)
    o, g = pair(params, T, r, s, ty, SF=SF)
    with g:
        if sort:
            g.sort_tracers()
        g.set_equalize(tpc, Cinit, seed=42)
        for call in range(2):
            st, info_o = o.equalize_tracers(tpc, Cinit, 42, call)
            assert st == 0
            info_g = g.equalize_tracers()
            assert info_o.did_work == 1 or call == 1
            assert_same(o, g, params, info_o, info_g)


def test_equalize_inside_time_step():
    params, T, SF, r, s, ty, Cinit, tpc = evolved_state("cfg0_32")
    print(f"[equalize in time step cfg0_32] ntr={len(r)}", flush=True)
    o, g = pair(params, T, r, s, ty, SF=SF)
    with g:
        o.set_equalize(1, tpc, Cinit, 9)
        g.set_equalize(tpc, Cinit, seed=9, in_time_step=True)
        for _ in range(2):
            assert o.time_step() == 0
            g.time_step(1)
        assert rel_err(g.get("SF"), o.plane(oracle.OR_SF)) <= 1e-10
        assert rel_err(g.get("T"), o.plane(oracle.OR_T)) <= 1e-10
        rg, sg = g.get_tracers("x")
        assert np.max(np.abs(rg - o.tracers(oracle.OR_RTR))) <= 1e-12
        assert np.max(np.abs(sg - o.tracers(oracle.OR_STR))) <= 1e-12
        assert np.array_equal(g.get_tracer_types(), o.tracers(oracle.OR_TTYPE, np.int32))
        assert np.array_equal(g.get("nempty"), o.plane(oracle.OR_NEMPTY, 0, np.int32))


def test_full_size_properties():
    params, cfg = config.preset_params("cfg2")
    T, r, s, ty = ic.workload(cfg, seed=1)
    print(f"[equalize full size] {params.nx}x{params.nz} ntr={len(r)}", flush=True)
    with Solver(params, len(r)) as g:
        g.initialize(T, r, s, ty)
        g.set_mode("mixed")
        g.time_step(1)
        r0, s0 = g.get_tracers("x")
        n0 = g.get("nempty")[2:-2, 2:-2].copy()
        g.set_equalize(cfg["tpc"], [float(ty.mean())], seed=3)
        g.sync()
        t0 = time.perf_counter()
        info = g.equalize_tracers()
        ms = (time.perf_counter() - t0) * 1e3
        print(f"equalize: {ms:.1f} ms, passes {info.passes}, nfree {info.nfree}, moved {info.moved}", flush=True)
        r1, s1 = g.get_tracers("x")
        n1 = g.get("nempty")[2:-2, 2:-2]
    nx, nz = params.nx, params.nz
    assert n0.sum() == n1.sum() == len(r)
    d = (n1 - n0).ravel()
    moved = (r1 != r0) | (s1 != s0)
    assert moved.sum() == info.moved == d[d > 0].sum() == -d[d < 0].sum()
    assert info.moved <= info.nfree and info.count[0] + info.count[1] == len(r)
    dst = np.rint(s1[moved] * nz).astype(np.int64) * (nx + 1) + np.rint(r1[moved] / (params.aspect / nx)).astype(np.int64)
    src = np.rint(s0[moved] * nz).astype(np.int64) * (nx + 1) + np.rint(r0[moved] / (params.aspect / nx)).astype(np.int64)
    assert np.array_equal(np.bincount(dst, minlength=d.size)[d > 0], d[d > 0])
    assert np.array_equal(np.bincount(src, minlength=d.size)[d < 0], -d[d < 0])
