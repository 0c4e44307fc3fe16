"""Batched ensembles of independent convection cases.

    python -m tracer_code_amd.ensemble --cases 16 --steps 5 [--mode mixed] [--guess plain|extrap|learned]

Cases are drawn from a seeded distribution: grids from {256x256, 512x256,
768x256, 1024x512}, Ra log-uniform in [1e5, 1e7], thermal viscosity
contrast log-uniform in [1e2, 1e5], buoyancy ratio uniform in [0.3, 0.8],
compositional viscosity contrast uniform in [10, 100], 10-40 tracers per
node.  Each case is its own context; tc_ensemble_step runs their time
steps concurrently (one host thread and one stream per case), so the GPU
overlaps the launches of different cases -- one case alone leaves most of
the chip idle on its coarse multigrid levels.  The printed JSON compares
the throughput with running the same cases one after another.
"""
import argparse
import ctypes
import hashlib
import json
import math
import os
import sys
import time

import numpy as np

from . import Solver, config, ic, learn

GRIDS = ((256, 256), (512, 256), (768, 256), (1024, 512))


def draw_cases(n, seed=0):
    """The case parameters (dicts for config.make_params plus IC keys)."""
    rng = np.random.default_rng(seed)
    cases = []
    for j in range(n):
        nx, nz = GRIDS[int(rng.integers(len(GRIDS)))]
        ra = float(math.exp(rng.uniform(math.log(1e5), math.log(1e7))))
        visT = float(math.exp(rng.uniform(math.log(1e2), math.log(1e5))))
        B = float(rng.uniform(0.3, 0.8))
        cc = float(rng.uniform(10.0, 100.0))
        tpc = int(rng.integers(10, 41))
        cases.append(dict(aspect=nx / nz, nx=nx, nz=nz, RaT=ra, RaC=(B * ra,), ivis=2, visT=visT, visC=(1.0 / cc,),
                          tpc=tpc, dlayer=0.15, initial_comp=0, t_variant="fourier4", fixes=1, delta_visC=(1.2,),
                          seed=1000 + j))
    return cases


def make_solvers(cases, mode="mixed", guess="plain", weights=None, level=4, tune=None, gcr=0, log=None, cheb=None):
    """(solvers, networks) initialised from each case's seeded inputs; `tune`
    (per case: dict of sweeps, tiles, level2, coarse) applies autotuned
    settings; gcr: GCR directions over the defect-correction cycles; cheb:
    the mixed mode's level-1 Chebyshev interval (lb, la)."""
    solvers, nets = [], []
    for j, c in enumerate(cases):
        if log and j % 16 == 0:
            log(f"[ensemble] initialising cases {j}..{min(j + 16, len(cases)) - 1} of {len(cases)}")
        params = config.make_params(**c)
        assert (params.nx, params.nz) == (c["nx"], c["nz"]), "grid changed by optimize_grid"
        T, r, s, ty = ic.workload(c, seed=c["seed"])
        g = Solver(params, len(r))
        g.initialize(T, r, s, ty)
        g.set_mode(mode)
        if mode != "fp64":
            if guess == "learned":
                net = learn.Network(*learn.level_shape(params.nx, params.nz, level), trainable=False)
                net.set_weights(weights)
                g.learn_attach(net)
                nets.append(net)
            g.guess_enable(guess)
            g.set_gcr(gcr)
            if cheb is not None and mode == "mixed":
                g.set_dc_cheb_interval(*cheb)
            if tune:
                t = tune[len(solvers)]
                g.set_dc_sweeps(t["sweeps"])
                for L in range(2, params.grids):
                    n = t["level2"] if L == 2 else t["coarse"]
                    if n:
                        g.set_dc_level_sweeps(L, n)
                for L, tile in enumerate(t["tiles"], start=1):
                    if 2 <= L < params.grids and tile:
                        g.set_tb_tile(L, tile)
        solvers.append(g)
    return solvers, nets


def autotune_cases(cases, mode="mixed", candidates=(28, 32, 40, 48, 70), gcr=0, log=None, cheb=None):
    """Per case: the sweep counts and the tiles the autotuners pick (on a
    scratch copy of the case, advanced by the tuning steps).  With GCR the
    levels from 3 down and level 2 get counts of their own, and the common
    count is tuned again after them.  cheb: the level-1 interval the cases
    run with (the tuning uses it too)."""
    out = []
    for j, c in enumerate(cases):
        if log and j % 8 == 0:
            log(f"[ensemble] autotuning cases {j}..{min(j + 8, len(cases)) - 1} of {len(cases)}")
        (g,), _ = make_solvers([c], mode, gcr=gcr, cheb=cheb)
        sweeps, _ = g.autotune_dc_sweeps(candidates, 1)
        level2 = coarse = 0
        if gcr and g.params.grids > 3:
            coarse, _ = g.autotune_dc_level_sweeps(3, g.params.grids - 1, (4, 8, 12, 16, 24, 0), 1)
            level2, _ = g.autotune_dc_level_sweeps(2, 2, (8, 12, 16, 24, 0), 1)
            sweeps, _ = g.autotune_dc_sweeps(candidates, 1)
        tiles = g.autotune_tb_tiles(1) if mode == "mixed" else [0] * 6
        out.append(dict(sweeps=sweeps, tiles=tiles, level2=level2, coarse=coarse))
        g.close()
    return out


def ensemble_step(solvers, steps, threads=0):
    """tc_ensemble_step over the solvers' contexts."""
    lib = solvers[0].lib
    arr = (ctypes.c_void_p * len(solvers))(*[g.h for g in solvers])
    failed = ctypes.c_int32(-1)
    st = lib.tc_ensemble_step(arr, len(solvers), int(steps), int(threads), ctypes.byref(failed))
    if st != 0:
        msg = lib.tc_last_error(solvers[failed.value].h) if failed.value >= 0 else b""
        raise RuntimeError(f"tc_ensemble_step: status {st} in case {failed.value}: {msg.decode() if msg else ''}")


def run(cases, steps, mode="mixed", guess="plain", weights=None, threads=0, tune=None, log=print, graphs=False,
        gcr=0, cheb=None):
    """Throughput of the cases run concurrently and one after another, from
    the same initial states (each after one untimed warm-up step).  graphs:
    replay the coarse part of each defect-correction cycle as a HIP graph;
    cheb: the mixed mode's level-1 Chebyshev interval."""
    out, hashes = {}, []
    for concurrent in (False, True):
        solvers, nets = make_solvers(cases, mode, guess, weights, tune=tune, gcr=gcr, log=log, cheb=cheb)
        for g in solvers:
            g.set_graphs(graphs)
        ensemble_step(solvers, 1, threads)  # warm-up, concurrently in both arms
        for g in solvers:
            g.sync()
        c0 = [g.stats().mg_cycles for g in solvers]
        t0 = time.perf_counter()
        if concurrent:
            ensemble_step(solvers, steps, threads)
        else:
            for j, g in enumerate(solvers):
                g.time_step(steps)
                if j % 16 == 15:
                    log(f"[ensemble] sequential: {j + 1} of {len(solvers)} cases done")
        for g in solvers:
            g.sync()
        el = time.perf_counter() - t0
        cells = sum(c["nx"] * c["nz"] for c in cases)
        arm = "concurrent" if concurrent else "sequential"
        # every solve stops on the reference's test (a solve that cannot
        # reach it returns an error); the last one's residual is recorded
        resid = [g.stats().residual_last / (g.params.tolerance * g.stats().res_ref_last) for g in solvers]
        out[arm] = dict(seconds=el, case_steps_per_s=len(cases) * steps / el, cell_steps_per_s=cells * steps / el,
                        cycles_per_step=[(g.stats().mg_cycles - c) / steps for g, c in zip(solvers, c0)],
                        max_last_residual_over_tolerance=max(resid))
        hashes.append([state_hash(g) for g in solvers])
        log(f"[ensemble] {arm}: {el:.2f}s for {len(cases)} cases x {steps} steps")
        for g in solvers:
            if guess == "learned":
                g.learn_attach(None)
            g.close()
        for net in nets:
            net.close()
    out["speedup"] = out["sequential"]["seconds"] / out["concurrent"]["seconds"]
    out["identical_to_sequential"] = sum(a == b for a, b in zip(*hashes))
    return out


def state_hash(g):
    """SHA-256 of a case's streamfunction, temperature, composition and
    tracer positions."""
    h = hashlib.sha256()
    for k in ("SF", "T", "Cnew"):
        h.update(np.ascontiguousarray(g.get(k)).tobytes())
    for a in g.get_tracers("x"):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", type=int, default=16)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--mode", default="mixed", choices=["fp64", "mixed", "dc64"])
    ap.add_argument("--guess", default="plain", choices=["plain", "extrap", "learned"])
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--autotune", action="store_true", help="tune sweeps and tiles per case first (untimed)")
    ap.add_argument("--graphs", action="store_true", help="HIP-graph replay of each cycle's coarse part")
    ap.add_argument("--gcr", type=int, default=0, help="GCR directions over the defect-correction cycles")
    ap.add_argument("--cheb-la", type=float, default=None,
                    help="la of the mixed mode's level-1 Chebyshev interval (-2, la); the autotuner then also "
                         "tries 10 to 20 sweeps per visit (default: the reference's interval)")
    a = ap.parse_args(argv)
    cheb = (-2.0, a.cheb_la) if a.cheb_la is not None else None
    cases = draw_cases(a.cases, a.seed)
    w = None
    if a.guess == "learned":
        w, _ = learn.load_weights(os.path.join(os.path.dirname(__file__), "weights", "unet_cfg1_L4.npz"))
    gcr = a.gcr if a.mode != "fp64" else 0
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    cand = (10, 12, 16, 20, 28, 32, 40, 48, 70) if cheb else (28, 32, 40, 48, 70)
    tune = (autotune_cases(cases, a.mode, candidates=cand, gcr=gcr, log=log, cheb=cheb)
            if (a.autotune and a.mode != "fp64") else None)
    res = run(cases, a.steps, a.mode, a.guess, w, a.threads, tune, log=log, graphs=a.graphs, gcr=gcr, cheb=cheb)
    res["cheb"] = list(cheb) if cheb else None
    res["graphs"] = a.graphs
    res["gcr"] = gcr
    if tune:
        res["tuned"] = tune
    res["cases"] = [{k: c[k] for k in ("nx", "nz", "RaT", "visT", "tpc")} for c in cases]
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
