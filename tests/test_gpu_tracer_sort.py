"""Node-sorted tracer order (tc_sort_tracers): the order of the device
tracer arrays carries no meaning in the reference, so a run that sorts the
tracers by node before every step must give bit-identical fields, counts,
composition and tracer positions (returned in the caller's order) to a run
that never sorts.  The tracers are first shuffled on the host, so the sort
really moves them."""
import numpy as np
import pytest

from tracer_code_amd import Solver, config, ic

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_sorted_run_equals_unsorted(mode):
    params, cfg = config.preset_params("cfg2", nx=128, nz=64, tpc=4)
    T, r, s, ty = ic.workload(cfg, seed=3)
    perm = np.random.default_rng(5).permutation(len(r))
    r, s, ty = r[perm], s[perm], ty[perm]
    print(f"[tracer sort {params.nx}x{params.nz} ntr={len(r)} mode={mode}]", flush=True)
    out = []
    for sort in (False, True):
        with Solver(params, len(r)) as g:
            g.initialize(T, r, s, ty)
            g.set_mode(mode)
            if sort:
                g.sort_tracers()
                g.set_tracer_sort(1)
            g.time_step(3)
            out.append({k: g.get(k) for k in ("SF", "T", "Cnew", "count_type", "nempty")}
                       | dict(zip(("r", "s"), g.get_tracers("x"))))
    a, b = out
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def _node_keys(r, s, params):
    dr, ds = params.aspect / params.nx, 1.0 / params.nz
    i = np.floor(r / dr + 0.5).astype(np.int64)   # NINT of a positive value
    k = np.floor(s / ds + 0.5).astype(np.int64)
    return k * (params.nx + 1) + i


def test_sort_order_is_node_order_and_stable():
    """The device slot order after tc_sort_tracers (the hand-written radix
    sort, 8-bit digits over the key's 18 bits at 256x128) is the row-major
    node order, and stable: tracers of one node keep the order they had in
    the slots before the sort -- the caller's (shuffled) order for the first
    sort, the previous sorted order after a time step."""
    params, cfg = config.preset_params("cfg2", nx=256, nz=128, tpc=5)
    T, r, s, ty = ic.workload(cfg, seed=4)
    perm = np.random.default_rng(6).permutation(len(r))
    r, s, ty = r[perm], s[perm], ty[perm]
    print(f"[radix sort order {params.nx}x{params.nz} ntr={len(r)}]", flush=True)
    with Solver(params, len(r)) as g:
        g.initialize(T, r, s, ty)
        assert np.array_equal(g.tracer_slots(), np.arange(len(r)))
        prev = np.arange(len(r))
        for step in range(2):
            g.sort_tracers()
            slots = g.tracer_slots()
            assert np.array_equal(np.sort(slots), np.arange(len(r)))
            rc, sc = g.get_tracers("x")
            key = _node_keys(rc[slots], sc[slots], params)
            assert np.all(np.diff(key) >= 0)
            where_before = np.empty(len(r), dtype=np.int64)
            where_before[prev] = np.arange(len(r))
            same = np.diff(key) == 0
            assert np.all(np.diff(where_before[slots])[same] > 0)
            prev = slots
            g.time_step(1)
