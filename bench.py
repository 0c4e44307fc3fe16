"""Time-step throughput of the MI355X hot path (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config cfg2] [--mode mixed|dc64|fp64]

A step is one pass of the reference's time loop body (main.f90:174-205):
multigrid, compute_velocities, the strain rate, and energy_time with RK4 and
RK_update=1 -- four multigrid solves to a 1e-8 relative residual, four RK
tracer stages over all tracers, five tracer-to-node binnings with the C
filter, the T filter and the T-field RK stages.  Inputs are synthetic
(seeded) and resident in HBM when the timed region starts.

Modes: `mixed` (default) and `dc64` solve by defect correction with a
double-double streamfunction and FP32 / FP64 W-cycle corrections; `fp64` is
the reference's own cycle, which FP64 rounding keeps above a 1e-8 residual at
2048x1024 (it stops with "Too Many Multigrid Cycles").  The defect-correction
modes run GCR over the cycles (`--gcr 16`) and, before the warm-up, tune their
sweep counts with real time steps: the common count, one for levels 3 and
below, one for level 2, then the common count again (level 1's);
`--dc-sweeps` [`--level2-sweeps`, `--coarse-sweeps`] fix them instead.

Prints one JSON line.  `roofline` reports the level-1 Jacobi-Richardson sweep
(the dominant kernel): algorithmic bytes per interior point are the
reference's arrays read or written once -- S, eta, eta_x, eta_z, eta_xx,
eta_zz, eta_xz, RHS, dt_stream, ratio read, the new iterate written: 11 x 8 B
= 88 B in FP64, 44 B in FP32 -- over the HIP-event-timed average launch
duration on the library's stream.  The kernels themselves move less
(`kernel_bytes_per_launch`: eta_xx - eta_zz is one plane formed once per
solve, and the ratio plane is read only on the level where the reference's
ratio is not the reset value): 80 B / 36 B per point.  `step_roofline`
sums the same per-point bytes over every sweep of the step plus 304 B per
tracer for the RK stages and 20 B per tracer per binning.  `cpu_baseline`
times the CPU oracle's level-1 sweep on this host and converts it to
steps/s with the sweep counts the GPU executed per step.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SWEEP_BYTES_PER_POINT = {"fp64": 88.0, "dc64": 88.0, "mixed": 44.0}      # algorithmic
KERNEL_BYTES_PER_POINT = {"fp64": 80.0, "dc64": 80.0, "mixed": 36.0}     # what the kernels move
TRACER_BYTES_PER_STEP = 304.0 + 5 * 20.0  # RK4 stages + five binnings, per tracer
DC_SWEEP_CANDIDATES = (28, 32, 36, 40, 48, 56, 70)
# sweeps per visit on levels 3 and below, then on level 2 (0: the common count), tuned after it
COARSE_SWEEP_CANDIDATES = (4, 8, 12, 16, 24, 0)
LEVEL2_SWEEP_CANDIDATES = (8, 12, 16, 24, 0)
SWEEP_KERNEL = {"fp64": ("k_sweep", "sweep64"), "dc64": ("k_sweep", "sweep64"), "mixed": ("k_sweep32", "sweep32")}


def cpu_baseline(params, sweeps_per_step, budget_s=12.0):
    """Oracle (C restatement, OpenMP) level-1 sweep time on the host cores,
    scaled to steps/s with the GPU's per-level sweep counts (smaller levels
    scaled by point count, which favours the CPU)."""
    import numpy as np

    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    o = oracle.Oracle(params, 0)
    nx, nz = params.nx, params.nz
    for oid in (oracle.OR_VIS,):
        o.plane(oid, 1)[:] = 1.0
    rng = np.random.default_rng(0)
    o.plane(oracle.OR_ERROR, 1)[:] = rng.standard_normal((nz + 5, nx + 5))
    o.plane(oracle.OR_RHS, 1)[:] = rng.standard_normal((nz + 5, nx + 5))
    o.plane(oracle.OR_DTSTREAM, 1)[:] = 1e-9
    n = 0
    t0 = time.perf_counter()
    while True:
        o.sweep(1, 1 + (n % params.Nmulti), params.Nmulti, 0)
        n += 1
        el = time.perf_counter() - t0
        if (el > budget_s and n >= 2) or n >= 400:
            break
    t_sweep = el / n
    pts1 = (nx - 1) * (nz - 1)
    fe = 0.0
    for L, cnt in enumerate(sweeps_per_step, start=1):
        ptsL = (nx // 2 ** (L - 1) - 1) * (nz // 2 ** (L - 1) - 1)
        fe += cnt * ptsL / pts1
    o.close()
    step_s = fe * t_sweep
    return {
        "value": (1.0 / step_s) if step_s > 0 else None,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{n} level-1 Jacobi-Richardson sweeps of the CPU oracle at {nx}x{nz} "
                   f"({t_sweep * 1e3:.1f} ms/sweep, {threads} OpenMP threads), scaled by the "
                   f"{fe:.0f} fine-grid-equivalent sweeps per step the GPU executed; viscosity "
                   f"hierarchy, coarse solves, filters and tracers not counted (favours the CPU)"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--mode", default="mixed", choices=["mixed", "dc64", "fp64"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cycle", default="W", choices=["W", "F", "V"])
    ap.add_argument("--tolerance", type=float, default=None,
                    help="relative residual of every solve (default: the preset's 1e-8)")
    ap.add_argument("--nmulti", type=int, default=None, help="smoother sweeps per level visit (default: the preset's)")
    ap.add_argument("--dc-sweeps", type=int, default=None,
                    help="sweeps per visit of the defect-correction cycle (default: autotuned before the warm-up)")
    ap.add_argument("--coarse-sweeps", type=int, default=None,
                    help="sweeps per visit on levels 3 and below with --dc-sweeps (default: that count)")
    ap.add_argument("--level2-sweeps", type=int, default=None,
                    help="sweeps per visit on level 2 with --dc-sweeps (default: that count)")
    ap.add_argument("--guess", default="learned", choices=["plain", "extrap", "learned"],
                    help="initial-guess candidates of the defect-correction solves (learned: with the committed "
                         "network of this workload, if there is one)")
    ap.add_argument("--gcr", type=int, default=16,
                    help="GCR directions kept over the defect-correction cycles (0: plain cycles)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs the hot path on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding is a design note only)", file=sys.stderr)
        return 2

    from tracer_code_amd import Solver, _abi, config, ic

    ov = dict(cycle=args.cycle)
    if args.tolerance:
        ov["tolerance"] = args.tolerance
    if args.nmulti:
        ov["Nmulti"] = args.nmulti
    params, cfg = config.preset_params(args.config, **ov)
    t_setup = time.perf_counter()
    T, r, s, ty = ic.workload(cfg, seed=args.seed)
    ntr = len(r)
    print(f"[bench] {args.config}: {params.nx}x{params.nz}, {params.grids} levels, {ntr} tracers, "
          f"inputs generated in {time.perf_counter() - t_setup:.1f}s", file=sys.stderr, flush=True)
    solver = Solver(params, ntr)
    solver.set_mode(args.mode)
    solver.initialize(T, r, s, ty)
    del T, r, s, ty
    net = None
    guess = args.guess if args.mode != "fp64" else "plain"
    if guess == "learned":
        from tracer_code_amd import learn
        wpath = os.path.join(ROOT, "tracer_code_amd", "weights", f"unet_{args.config}_L4.npz")
        if os.path.exists(wpath):
            w, meta = learn.load_weights(wpath)
            net = learn.Network(*learn.level_shape(params.nx, params.nz, meta["level"]), trainable=False)
            net.set_weights(w)
            solver.learn_attach(net)
        else:
            guess = "extrap"
    gcr = args.gcr if args.mode != "fp64" else 0
    if args.mode != "fp64":
        solver.guess_enable(guess)
        solver.set_gcr(gcr)
    tune = None
    tiles = None
    coarse_sweeps, coarse_tune = 0, None
    level2_sweeps, level2_tune = 0, None
    dc_sweeps = params.Nmulti
    if args.mode != "fp64":
        if args.dc_sweeps is None:
            # untimed: one real time step per candidate
            t0 = time.perf_counter()
            dc_sweeps, tune = solver.autotune_dc_sweeps(DC_SWEEP_CANDIDATES, 1)
            print(f"[bench] autotuned sweeps per visit: {dc_sweeps} ({time.perf_counter() - t0:.1f}s; "
                  f"ms/step {tune})", file=sys.stderr, flush=True)
            if params.grids > 3:
                # levels 3 and below, then level 2: counts of their own
                t0 = time.perf_counter()
                coarse_sweeps, coarse_tune = solver.autotune_dc_level_sweeps(3, params.grids - 1,
                                                                             COARSE_SWEEP_CANDIDATES, 1)
                level2_sweeps, level2_tune = solver.autotune_dc_level_sweeps(2, 2, LEVEL2_SWEEP_CANDIDATES, 1)
                print(f"[bench] autotuned sweeps per visit on levels >= 3: {coarse_sweeps or dc_sweeps}, on level 2: "
                      f"{level2_sweeps or dc_sweeps} ({time.perf_counter() - t0:.1f}s; ms/step {coarse_tune}, "
                      f"{level2_tune})", file=sys.stderr, flush=True)
                # the common count again, now that the coarse levels have their own
                t0 = time.perf_counter()
                dc_sweeps, tune = solver.autotune_dc_sweeps(DC_SWEEP_CANDIDATES, 1)
                print(f"[bench] sweeps per visit on levels 1-2, tuned again: {dc_sweeps} "
                      f"({time.perf_counter() - t0:.1f}s; ms/step {tune})", file=sys.stderr, flush=True)
            if args.mode == "mixed":
                t0 = time.perf_counter()
                tiles = solver.autotune_tb_tiles(1)
                print(f"[bench] autotuned tiles by level: {tiles} ({time.perf_counter() - t0:.1f}s)",
                      file=sys.stderr, flush=True)
        else:
            dc_sweeps = args.dc_sweeps
            solver.set_dc_sweeps(dc_sweeps)
            if args.coarse_sweeps:
                coarse_sweeps = args.coarse_sweeps
                for L in range(3, params.grids):
                    solver.set_dc_level_sweeps(L, coarse_sweeps)
            if args.level2_sweeps:
                level2_sweeps = args.level2_sweeps
                solver.set_dc_level_sweeps(2, level2_sweeps)
    for w in range(args.warmup):
        t0 = time.perf_counter()
        solver.time_step(1)
        st = solver.stats()
        print(f"[bench] warmup step {w}: {time.perf_counter() - t0:.2f}s q_last={st.q_last} "
              f"dt={st.dt_last:.3e}", file=sys.stderr, flush=True)
    before = solver.stats()
    gcr0 = solver.gcr_counts()[0]
    sweeps0 = list(before.sweeps_exec)
    cycles0 = before.mg_cycles
    solver.set_profiling(True)
    solver.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        solver.time_step(1)
        print(f"[bench] step {k} done at {time.perf_counter() - t0:.2f}s", file=sys.stderr, flush=True)
    solver.sync()
    elapsed = time.perf_counter() - t0
    avg_us, nlaunch = solver.kernel_timing(_abi.TC_K_SWEEP_L1)
    solver.set_profiling(False)
    after = solver.stats()
    gcr_updates = solver.gcr_counts()[0] - gcr0
    sweeps_per_step = [(a - b) / args.steps for a, b in zip(after.sweeps_exec, sweeps0)][: params.grids]
    cycles_per_step = (after.mg_cycles - cycles0) / args.steps
    steps_per_s = args.steps / elapsed
    pts1 = (params.nx - 1) * (params.nz - 1)
    bytes_per_launch = SWEEP_BYTES_PER_POINT[args.mode] * pts1
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9 if avg_us > 0 else None
    traffic = None
    kname, kfile = SWEEP_KERNEL[args.mode]
    tfile = os.path.join(ROOT, "profiles", f"{kfile}_l1_{params.nx}x{params.nz}_hbm_bytes.txt")
    if os.path.exists(tfile):
        for line in open(tfile):
            if line.startswith("bytes_per_launch"):
                traffic = float(line.split("=")[1])
    # whole step: every sweep of every level at the per-point bytes, plus
    # the tracer stages and binnings (residuals, transfers, filters and the
    # viscosity set-up are not counted)
    step_bytes = TRACER_BYTES_PER_STEP * ntr
    for L, cnt in enumerate(sweeps_per_step, start=1):
        step_bytes += cnt * (params.nx // 2 ** (L - 1) - 1) * (params.nz // 2 ** (L - 1) - 1) * \
            SWEEP_BYTES_PER_POINT[args.mode]
    step_gbs = step_bytes / (elapsed / args.steps) / 1e9
    out = {
        "metric": "steps/s @2048x1024+42M tracers to 1e-8 resid: mixed+learned vs plain FP64",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "dc64": "f64 (double-double residual)",
                  "mixed": "f32 cycle, f64/double-double residual"}[args.mode],
        "data": "synthetic (seeded T0 with Fourier modes, jittered tracers)",
        "config": {
            "workload": f"{params.nx}x{params.nz} aspect {params.aspect:g}, Ra={params.RaT:g}, "
                        f"{ntr} tracers, multigrid to {params.tolerance:g} relative residual",
            "mode": args.mode,
            "cycle": args.cycle,
            "sweeps_per_visit": dc_sweeps,
            "sweeps_per_visit_level_2": level2_sweeps or dc_sweeps,
            "sweeps_per_visit_levels_3_up": coarse_sweeps or dc_sweeps,
            "coarse_autotune_ms_per_step": coarse_tune,
            "level2_autotune_ms_per_step": level2_tune,
            "reference_sweeps_per_visit": params.Nmulti,
            "autotune_ms_per_step": tune,
            "tb_tiles_by_level": tiles,
            "dc_retries": after.dc_retries,
            "gcr_directions": gcr,
            "gcr_updates": gcr_updates,
            "guess": guess,
            "starts": {k: after.guesses_used[i] - before.guesses_used[i]
                       for i, k in enumerate(("plain", "extrap", "learned"))},
            "levels": params.grids,
            "mg_cycles_per_step": cycles_per_step,
            "sweeps_per_step": sweeps_per_step,
        },
        "roofline": {
            "kernel": f"{kname} level 1 (Jacobi-Richardson smoother)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
            "traffic": traffic,
            "avg_launch_us": avg_us,
            "launches": nlaunch,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "kernel_bytes_per_launch": KERNEL_BYTES_PER_POINT[args.mode] * pts1,
        },
        "step_roofline": {"algorithmic_bytes_per_step": step_bytes, "achieved": step_gbs, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS},
    }
    if net is not None:
        solver.learn_attach(None)
        net.close()
    solver.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(params, sweeps_per_step)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
