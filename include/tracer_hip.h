/*
 * tracer_hip.h -- C-ABI of the MI355X time-step hot path of tracer_code.
 *
 * The reference (seantrim/tracer_code, Fortran 90 + OpenMP) has no FFI; its
 * hot path is a set of subroutines on module-global state.  Every entry point
 * below replaces the body of one of those subroutines (or one of the inline
 * OpenMP tracer loops of energy_time) and is meant to be bound from Fortran
 * with ISO_C_BINDING (the interface module is shown in INTEGRATION.md).
 *
 * Conventions
 *  - All functions return int: 0 = ok, >0 = the reference's own `stop`
 *    conditions, <0 = library / HIP errors.  tc_last_error() gives text.
 *  - Host arrays use the reference's Fortran layout: a ghosted grid field
 *    f(-2:nx+2,-2:nz+2) is passed as the address of f(-2,-2), i fastest,
 *    leading dimension nx+5.  Composition fields Cnew/Cvis/Cbuoy are
 *    type-fastest, C(1:ntype,-2:nx+2,-2:nz+2), and count_type is
 *    count_type(0:ntype,-2:nx+2,-2:nz+2) (basics.f90:42-98).
 *  - One host thread and one HIP stream per context; not re-entrant.
 *    Host arrays are valid when a call returns (synchronous at the boundary).
 *  - Device state (multigrid hierarchy, fields, tracers) is owned by the
 *    library; host arrays stay owned by the caller.
 *  - Only order=2 finite differences, iterate_order=1 (the Euler/Richardson
 *    smoother), uniform grids (gridX=gridZ=0), Newtonian rheology
 *    (non_Newtonian=0, ivis in {0,1,2}) and global time steps
 *    (local_time=0) are accepted; anything else returns TC_ERR_UNSUPPORTED.
 */
#ifndef TRACER_HIP_H
#define TRACER_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TC_MAX_TYPES 4
#define TC_MAX_LEVELS 6
#define TC_MAX_MULTI 52 /* multi(0:51), multigrid_cycle.f90:6 */

/* Status codes. */
enum {
    TC_OK = 0,
    TC_ERR_MAX_CYCLES = 1,    /* iterative_solver.f90:164-167 */
    TC_ERR_RESID_BLOWUP = 2,  /* iterative_solver.f90:275-278 */
    TC_ERR_TRACER_DOMAIN = 3, /* composition.f90:709-712 */
    TC_ERR_SINGULAR = 4,      /* DGBTRF INFO>0 on the coarsest grid */
    TC_ERR_ARG = -1,
    TC_ERR_HIP = -2,
    TC_ERR_UNSUPPORTED = -3,
    TC_ERR_NOMEM = -4
};

/* Solver precision modes (tc_set_mode). */
enum {
    TC_MODE_FP64 = 0,  /* plain path: the reference's Jacobi-Richardson W-cycle in FP64 */
    TC_MODE_MIXED = 1, /* defect correction: double-double psi and fine residual, FP32 W-cycle corrections */
    TC_MODE_DC64 = 2   /* defect correction as MIXED with FP64 W-cycle corrections */
};
/* In both defect-correction modes the correction cycle's coarsest-grid
 * solve uses the inverse of the previous solve's coarse operator while the
 * current one is factorised on a side stream (ready for the next solve); the
 * first solve waits for its own.  The outer residual is exact, so every solve
 * still ends at the same tolerance. */

/* Field identifiers for tc_put_field / tc_get_field. */
enum {
    TC_F_T = 0,      /* T(-2:nx+2,-2:nz+2) */
    TC_F_SF = 1,     /* SF (streamfunction) */
    TC_F_TVIS = 2,   /* Tvis (smoothed T for viscosity) */
    TC_F_TBUOY = 3,  /* Tbuoy (T for buoyancy) */
    TC_F_CNEW = 4,   /* Cnew(1:ntype,...) */
    TC_F_CVIS = 5,   /* Cvis(1:ntype,...) */
    TC_F_CBUOY = 6,  /* Cbuoy(1:ntype,...) */
    TC_F_U = 7,      /* u: the ghosted (-2:nx+2,-2:nz+2) frame; the reference's u(-1:nx+1,-1:nz+1) sits inside it */
    TC_F_W = 8,      /* w: same frame as u */
    TC_F_STRAIN = 9, /* strain */
    TC_F_VIS = 10,   /* vis_grid(1,...) = vis after a solve */
    TC_F_CONDUCT = 11,
    TC_F_NEMPTY = 12,     /* int32 nempty(-2:nx+2,-2:nz+2) */
    TC_F_COUNT_TYPE = 13, /* int32 count_type(0:ntype,...) */
    TC_F_T0 = 14,    /* energy_time stage fields T0/T1/T2 (device-resident driver) */
    TC_F_T1 = 15,
    TC_F_T2 = 16,
    TC_F_SF_LO = 17, /* get only: low part of the double-double streamfunction of the last
                        defect-correction solve (SF holds the high part) */
    TC_F_RHS = 18,   /* get only: RHS(1) of the last solve (compute_RHS, iterative_solver.f90:571-590) */
    TC_F_COUNT
};

/* Bits of tc_params.fixes. */
enum { TC_FIX_ZERO_COARSE = 1 };

/* Tracer position sets (energy.f90 arrays rtr/str, rtr0/str0, rtr1/str1, rtr2/str2). */
enum { TC_TR_X = 0, TC_TR_X0 = 1, TC_TR_X1 = 2, TC_TR_X2 = 3 };

/*
 * Run parameters: the compile-time inputs of main.f90:8-116 plus what the
 * Fortran host derives at start-up (FD_coefficients main.f90:350-469,
 * multigrid_setup multigrid_cycle.f90:1-99, derivative_eigenvalues
 * energy.f90:501-588, compute_smoothing_parameters smoother.f90:31-73).
 * Index j of a 5-point table D(-2:2) is stored at [j+2]; Dint[m-1][j+1]
 * holds D(m,j) of the interpolation table D(1:2,-1:1).
 */
typedef struct tc_params {
    int32_t nx, nz;       /* cells after optimize_grid */
    int32_t grids;        /* multigrid levels after optimize_grid */
    int32_t order;        /* must be 2 */
    int32_t iterate_order;/* must be 1 */
    int32_t Vbc, Tbot;    /* velocity / bottom thermal BC switches */
    int32_t Nmulti, vis_iter, max_cycles;
    int32_t nsteps;       /* W/V-cycle length */
    int32_t multi[TC_MAX_MULTI]; /* multi(0:nsteps+1) */
    int32_t comp, ntype, ivis, torder, RK_update;
    /* Bit mask of departures from the reference multigrid (0 = the
     * reference exactly).  TC_FIX_ZERO_COARSE: error(L), 1<L<grids, is
     * zeroed each time the cycle descends into level L (the standard
     * coarse-grid correction).  The reference zeroes it once (main.f90:162)
     * and otherwise starts each visit from the stale correction, which
     * diverges at large viscosity contrasts (1e5 from T x 100 from C). */
    int32_t fixes;
    double aspect;
    double tolerance, eig_ratio, eig_ratio_min;
    double courant, courantT, courantC, courant_stream;
    double RaT, visT, visP, H;
    double RaC[TC_MAX_TYPES], visC[TC_MAX_TYPES];
    double conduct_factor[TC_MAX_TYPES], Htype[TC_MAX_TYPES];
    double delta_T;                  /* T smoothing threshold */
    double delta_C[TC_MAX_TYPES];    /* C smoothing thresholds */
    double D1[5], D2[5], D3[5], D4[5];
    double Dint[2][3];
    /* Fourier symbol maxima (derivative_eigenvalues) */
    double Tr_base, Trr_base, Trrr_base, Trrrr_base, Trs_base, Trrs_base, Trrss_base;
} tc_params;

typedef struct tc_ctx tc_ctx;

/* Timing / counters for the last call(s); filled by tc_get_stats. */
typedef struct tc_stats {
    int32_t q_last;          /* multigrid cycles of the last solve */
    int32_t n_eta;           /* smoother_vis passes at level 1 of the last solve */
    int32_t n_temp;          /* smoother_time_T iterations of the last call */
    int32_t n_comp[TC_MAX_TYPES];
    int32_t guess_used;      /* start of the last solve: 0 plain, 1 stage increment, 2 network */
    int32_t pad;
    double residual_last;    /* residual_mag(1) of the last solve */
    double res_ref_last;     /* res_ref(1) of the last solve */
    double eig_fine;
    double dt_last;          /* dt of the last time step */
    double time;             /* model time */
    int64_t sweeps_exec[TC_MAX_LEVELS]; /* smoother sweeps executed per level (cumulative) */
    int64_t residuals_exec[TC_MAX_LEVELS];
    int64_t coarse_solves;
    int64_t vis_iters;       /* smoother_vis iterations (cumulative) */
    int64_t c_iters;         /* smoother_time iterations (cumulative) */
    int64_t t_iters;         /* smoother_time_T iterations (cumulative) */
    int64_t mg_cycles;       /* multigrid cycles (cumulative) */
    int64_t guesses_used[3]; /* solves started from each TC_GUESS_* candidate (cumulative) */
    /* Last solve: residual_mag(1) after the first level-1 smoothing from the
     * plain start, the stage increment and the learned guess (0: not judged). */
    double guess_res[3];
    int64_t dc_retries;      /* defect-correction solves redone with Nmulti sweeps (cumulative) */
    /* Cycles of each solve of the last tc_time_step, in call order: the solve
     * before energy_time (main.f90:181), then the RK_update solves of stages
     * 2, 3 and 4 (energy.f90:240, 331, 397); 0 where a stage had none. */
    int32_t q_step[4];
} tc_stats;

/* ---- lifetime ---------------------------------------------------------- */
/* Replaces storage (basics.f90:37-110) + the device half of compute_basics
 * (basics.f90:112-196): allocates the level hierarchy and ntr tracers. */
int tc_init(const tc_params* p, int64_t ntr, tc_ctx** out);
int tc_finalize(tc_ctx* ctx);
const char* tc_last_error(const tc_ctx* ctx);
int tc_set_mode(tc_ctx* ctx, int mode);
/* Jacobi-Richardson sweeps per level visit of the defect-correction modes'
 * correction cycle (0: the parameters' Nmulti, the reference's value).  A
 * solve that with it diverges or exceeds TC_DC_FAST_MAX_CYCLES cycles is
 * redone with Nmulti (counted in tc_stats.dc_retries). */
#define TC_DC_FAST_MAX_CYCLES 200
int tc_set_dc_sweeps(tc_ctx* ctx, int32_t sweeps);
/* Sweeps per visit of one level of the tuned defect-correction cycle (0:
 * the common count of tc_set_dc_sweeps); the Nmulti retry ignores it. */
int tc_set_dc_level_sweeps(tc_ctx* ctx, int32_t level, int32_t sweeps);
/* HIP graphs for the FP32 defect-correction cycle (no counterpart in the
 * reference): the part of a cycle below level 1 (restrictions, the coarse
 * levels' visits, the coarse solve, the prolongation into psi and the fine
 * residual) is the same sequence of launches in every cycle of a solve.  With
 * enable set it is captured once per solve and replayed as one graph launch
 * in the solve's other cycles (results bit-identical to the launches).  Meant
 * for many concurrent contexts (tc_ensemble_step), whose host threads then
 * issue one launch per cycle instead of hundreds.  Off by default. */
int tc_set_graphs(tc_ctx* ctx, int32_t enable);
/* Segment graphs captured and replayed so far. */
int tc_graph_counts(const tc_ctx* ctx, int64_t* captures, int64_t* replays);
/* GCR acceleration of the defect-correction cycles (no counterpart in the
 * reference).  Each cycle is used as a preconditioner: its change of psi is
 * a search direction, orthogonalised against up to `directions` (0..16) kept
 * ones, and psi moves to the point of least residual (2-norm) over them.
 * The residual is then recomputed exactly, so every solve stops on the same
 * test and tolerance as without it.  0 (the default) runs the plain cycles. */
int tc_set_gcr(tc_ctx* ctx, int32_t directions);
/* GCR updates made and restarts forced by a dependent direction so far. */
int tc_gcr_counts(const tc_ctx* ctx, int64_t* updates, int64_t* restarts);
/* Coarse-grid factorisation interval of the defect-correction modes (no
 * counterpart in the reference, whose Newtonian multigrid refactorises the
 * coarsest operator in every call, iterative_solver.f90:95-98): the banded
 * LU and explicit inverse of the coarsest operator are recomputed (on a side
 * stream, used from the next refresh on) every `solves`-th solve (>= 1; 1,
 * the default, is every solve).  The coarse solve is inside the correction
 * cycle, so the exact outer residual, the stop test and the tolerance are
 * unchanged. */
int tc_set_coarse_refresh(tc_ctx* ctx, int32_t solves);
/* smoother_time's C filter loop (smoother.f90:286-340) with several filter
 * iterations per kernel launch, staged in on-chip memory (enable=1, the
 * default), or one launch per iteration (enable=0).  Both give the same
 * iterates, stop test, iteration counts and bits. */
int tc_set_filter_blocking(tc_ctx* ctx, int32_t enable);
/* Level-1 smoother of the FP32 defect-correction cycle (mode TC_MODE_MIXED;
 * no counterpart in the reference, whose smoother is the Chebyshev-
 * accelerated Jacobi-Richardson sweep of iterate_stream,
 * iterative_solver.f90:860-912): with sweeps > 0 each level-1 visit runs
 * `sweeps` multicolour Gauss-Seidel sweeps (9 colours, tiles staged in LDS
 * and coupled between launches) instead of its Jacobi-Richardson sweeps,
 * `per_launch` (1..8) of them per kernel launch, relaxation factor omega
 * (0 < omega < 2).  sweeps = 0 (the default) keeps the Jacobi-Richardson
 * sweeps.  The outer residual, stop test and tolerance are unchanged. */
int tc_set_gs_smoother(tc_ctx* ctx, int32_t sweeps, int32_t per_launch, double omega);
/* Chebyshev interval of the level-1 visits of the FP32 defect-correction
 * cycle (mode TC_MODE_MIXED).  The reference's sweep factors
 * (iterate_stream, iterative_solver.f90:873-899) take the step of sweep n of
 * N as c = 2 / (-lb - la + (lb - la) cos((2(N-n+1)-1) pi / 2N)) with
 * lb = -2, la = -1 (the default here too).  A wider interval (la closer to
 * 0, lb <= la < 0 required) also damps the middle of the spectrum, so fewer
 * sweeps per visit do the same work.  It applies to the solves at a tuned
 * count (tc_set_dc_sweeps, other than Nmulti); the Nmulti retry of a solve
 * that diverged or stalled, the coarser levels, the reference form and the
 * dc64 mode keep the reference's factors.  The outer residual, stop test
 * and tolerance are unchanged. */
int tc_set_dc_cheb_interval(tc_ctx* ctx, double lb, double la);
/* Autotuner of the defect-correction sweep count: runs steps_each time
 * steps (real steps of the run) with each of the n candidates, and keeps
 * the largest candidate whose wall time per step is within 3% of the
 * fastest among those that needed no retry (0, i.e. Nmulti, if all did).
 * ms_per_step[n] (may be NULL) receives the measured times, *best the
 * choice, which is also set as with tc_set_dc_sweeps. */
/* Autotuner of one sweep count for the levels first_level..last_level
 * (tc_set_dc_level_sweeps on each; last_level < grids, the coarsest grid is
 * solved directly), after tc_autotune_dc_sweeps: runs steps_each time steps
 * per candidate (0 stands for the common count) and keeps the largest count
 * within 3% of the fastest among those that needed no retry, which it leaves
 * set; *best receives it.  With the cycles accelerated by tc_set_gcr the
 * coarse levels need fewer sweeps than the common count (measured at
 * 2048x1024). */
int tc_autotune_dc_level_sweeps(tc_ctx* ctx, int32_t first_level, int32_t last_level, const int32_t* candidates,
                                int32_t n, int32_t steps_each, int32_t* best, double* ms_per_step);
int tc_autotune_dc_sweeps(tc_ctx* ctx, const int32_t* candidates, int32_t n, int32_t steps_each, int32_t* best,
                          double* ms_per_step);
/* Tile of the mixed mode's temporally blocked sweeps on `level` (2 ..
 * grids-1): 16 or 32 (16x16 or 32x32 cells per workgroup, each staged with
 * an 8-cell halo), 0 = chosen by the level's size. */
int tc_set_tb_tile(tc_ctx* ctx, int32_t level, int32_t tile);
/* Autotuner of those tiles: level by level from 2, runs steps_each time
 * steps with each tile and keeps the faster (ties: the size-based choice);
 * tiles[TC_MAX_LEVELS] (may be NULL) receives the tile of each level
 * (0 where the level has no such sweeps). */
int tc_autotune_tb_tiles(tc_ctx* ctx, int32_t steps_each, int32_t* tiles);
int tc_get_stats(const tc_ctx* ctx, tc_stats* out);
int tc_sync(tc_ctx* ctx);

/* ---- data movement (the Fortran host keeps its own copies) ------------- */
int tc_put_field(tc_ctx* ctx, int field, const void* host);
int tc_get_field(tc_ctx* ctx, int field, void* host);
/* rtr/str/Ttype of ntr tracers (arrays.f90:3,8-9, ID order). */
int tc_put_tracers(tc_ctx* ctx, const double* rtr, const double* str, const int32_t* ttype);
/* which = TC_TR_X / _X0 / _X1 / _X2; returned in ID order. */
int tc_get_tracers(tc_ctx* ctx, int which, double* r, double* s);
/* Ttype of the tracers in ID order (equalize_tracers changes it). */
int tc_get_tracer_types(tc_ctx* ctx, int32_t* ttype);
/* Model time and dt (energy.f90 stability_RK4 writes dt). */
int tc_set_time(tc_ctx* ctx, double time, double dt);

/* ---- drop-in entry points ---------------------------------------------- */
/* multigrid (iterative_solver.f90:79-173): solves for SF from the device
 * Tbuoy/Cbuoy/Tvis/Cvis using SF as the warm start; q_out = cycles used,
 * resid_out = residual_mag(1), resref_out = res_ref(1). */
int tc_multigrid(tc_ctx* ctx, int32_t* q_out, double* resid_out, double* resref_out);
/* compute_velocities (momentum.f90:532-551) + velocityBCs. */
int tc_compute_velocities(tc_ctx* ctx);
/* compute_strain_rate_invariant (momentum.f90:32-62). */
int tc_compute_strain_rate_invariant(tc_ctx* ctx);
/* smoother_time_T(Tin) (smoother.f90:342-376); src = TC_F_T/_T0/_T1/_T2. */
int tc_smoother_time_T(tc_ctx* ctx, int src_field, int32_t* n_temp);
/* tracers_to_corners (composition.f90:669-790) on the position set `which`,
 * followed (smooth=1) by smoother_time and load_conductivity_array. */
int tc_tracers_to_corners(tc_ctx* ctx, int which, int smooth, int32_t* n_comp);
/* smoother_time (smoother.f90:286-340) alone. */
int tc_smoother_time(tc_ctx* ctx, int32_t* n_comp);
/* load_conductivity_array (energy.f90:31-58). */
int tc_load_conductivity(tc_ctx* ctx);
/* One RK stage of the tracer loops of energy_time (energy.f90:182-225,
 * 270-312, 355-378, 420-440), including compute_derivatives of u and w
 * (energy.f90:170-171 etc.); dt is the full time step. */
int tc_tracer_stage(tc_ctx* ctx, int stage, double dt);
/* stability_RK4 (energy.f90:600-646): returns dt from the device u,w. */
int tc_stability_RK4(tc_ctx* ctx, double* dt_out);
/* One T-field RK stage of energy_time (energy.f90:157-167, 248-260,
 * 339-351, 405-416) on the device fields. */
int tc_energy_stage(tc_ctx* ctx, int stage, double dt);
/* nsteps iterations of the body of the time loop (main.f90:174-205) with
 * tracer=0, comp=1, RK4 and RK_update=1: multigrid, compute_velocities,
 * strain, energy_time; everything stays on the device. */
int tc_time_step(tc_ctx* ctx, int nsteps);

/* ---- host start-up helper ----------------------------------------------- */
/* derivative_eigenvalues (energy.f90:501-598) for callers without the
 * Fortran host: Fourier-symbol maxima of the FD tables, single-precision
 * T_test products as in the reference.  base = Tr, Trr, Trrr, Trrrr, Trs,
 * Trrs, Trrss (the *_base scalars of tc_params).  No GPU needed. */
int tc_derivative_eigenvalues(const double D1[5], const double D2[5], const double D3[5], const double D4[5],
                              double base[7]);

/* ---- restart files ------------------------------------------------------ */
/* The reference's restart files (print_restart main.f90:301-332, read in
 * build_grid grid.f90:56-95): Fortran unformatted sequential, one record
 * framed by 4-byte byte counts.  Trestart holds xg(-3:nx+3), zg(-3:nz+3),
 * T (if has_T), Cnew(1:ntype_c,...) (if ntype_c > 0) and SF, the grid fields
 * ghosted (-2:nx+2,-2:nz+2) as at the rest of this interface;
 * restart_tracers holds rtr, str and Ttype (int32).  A record whose length
 * does not match the sizes is rejected with TC_ERR_ARG.  Host only. */
int tc_restart_read(const char* path, int32_t nx, int32_t nz, int32_t has_T, int32_t ntype_c, double* xg,
                    double* zg, double* T, double* Cnew, double* SF);
int tc_restart_write(const char* path, int32_t nx, int32_t nz, int32_t has_T, int32_t ntype_c, const double* xg,
                     const double* zg, const double* T, const double* Cnew, const double* SF);
int tc_restart_tracers_read(const char* path, int64_t ntr, double* rtr, double* str, int32_t* ttype);
int tc_restart_tracers_write(const char* path, int64_t ntr, const double* rtr, const double* str,
                             const int32_t* ttype);

/* ---- diagnostics -------------------------------------------------------- */
/* Replaces stats (basics.f90:236-324) on the device fields: recomputes the
 * strain rate invariant, then out = {vrms, Nu (top), Tavg, Cavg (type 1),
 * nodes without tracers, min and max tracers per node, Nu at the bottom,
 * max strain on the top row}, with the trapezoid integrals of integrate2D /
 * integrate1Dx (basics.f90:341-390).  The sums run in another order than
 * the reference's loops (equal to rounding); the integers are exact. */
int tc_diagnostics(tc_ctx* ctx, double out[9]);

/* One row of stats.dat (basics.f90:236-324) in its column order: the
 * quantities of tc_diagnostics plus RaT_avg (integrate2D(RaT/conduct/vis),
 * vis of level 1), the surface mobility, the corner heat fluxes q1..q4 and,
 * per type, Cavg and the entrainment (the mass above row kent over mass).
 * kent is entrainment_setup's row (composition.f90:582-590: the last k with
 * k*ds <= dent); mass[t] is Cavg of type t+1 at the first call (basics.f90:
 * 250-258); mass = NULL takes this call's Cavg, as the first call does.
 * Sums run in another order than the reference's loops (equal to rounding). */
typedef struct tc_stats_row {
    double time, Tavg, vrms, Nu, Nu_bot, RaT_avg, mobility, strain_max;
    double q[4];
    int32_t empty_cells, tracer_min, tracer_max, kent;
    double Cavg[TC_MAX_TYPES];
    double ent[TC_MAX_TYPES];
} tc_stats_row;
int tc_stats_record(tc_ctx* ctx, int32_t kent, const double* mass, tc_stats_row* out);

/* ---- snapshot staging ---------------------------------------------------- */
/* Replaces the reads of snapshots (main.f90:234-268): copies T, SF, u, w,
 * strain (recomputed, as stats does), vis, Tvis, nempty, Cnew and Cvis into
 * device staging planes on the compute stream, then to pinned host memory
 * on a stream of their own, and returns at once: the next time steps run
 * while the copy and the host's writing proceed.  A new stage waits for the
 * previous host copy. */
int tc_snapshot_stage(tc_ctx* ctx);
/* Waits for the staged copy and returns one staged field in tc_get_field's
 * layout (TC_F_T, _SF, _U, _W, _STRAIN, _VIS, _TVIS, _NEMPTY, _CNEW, _CVIS). */
int tc_snapshot_get(tc_ctx* ctx, int field, void* host);

/* ---- tracer equalisation ------------------------------------------------ */
/* Replaces equalize_tracers (composition.f90:83-580): tracers of nodes with
 * more than their type's share (nexcess > 0) are moved to the nodes that lack
 * the most (nexcess == minexcess), donors near the walls first, by the
 * reference's rules (the 3x3 single-type test, the type choice toward the
 * initial per-type totals Nideal, the lottery passes over shuffled row and
 * column orders, the second loop, the wall clamps), then the tracers are
 * binned again (tracers_to_corners without the C filter).  The reference
 * draws its random numbers from one mzran stream per OpenMP thread
 * (composition.f90:7-19), so its result depends on the thread count; here
 * every draw is a counter-based hash of (seed, call, cell or tracer index),
 * so the result is the same on any device and in any tracer order.
 * Positions stay inside the receiving node's control volume.  Temperature
 * tracers (tracer=1) are not supported: the temperature lives on the grid.
 * Fields of the info record are the columns of equalize.dat
 * (composition.f90:566-574). */
typedef struct tc_equalize_info {
    int32_t did_work;                     /* 0: every nexcess was 0, nothing moved */
    int32_t passes;                       /* lottery passes of both loops */
    int32_t tpc_type[TC_MAX_TYPES + 1];   /* tracers per node wanted for each type */
    int32_t minexcess[TC_MAX_TYPES + 1];  /* after the moves; 2147483647 = no node */
    int64_t deficit;                      /* sum of nexcess over nodes with nexcess <= 0, after */
    int64_t moved;                        /* tracers repositioned */
    int64_t nfree;                        /* donor tracers */
    int64_t count[TC_MAX_TYPES + 1];      /* tracers of each type after (bookkeeping) */
} tc_equalize_info;

/* Cinit[ntype]: the initial mean composition of each type (Cavg at time 0,
 * basics.f90:323); tpc: tracers per node for comp=0.  With in_time_step
 * set, tc_time_step equalizes after stability_RK4 (energy.f90:133-135). */
int tc_set_equalize(tc_ctx* ctx, int32_t tpc, const double* Cinit, uint64_t seed, int32_t in_time_step);
/* One equalisation now of the current tracers; call k (from 0) of a context
 * draws from stream (seed, k).  info may be NULL. */
int tc_equalize_tracers(tc_ctx* ctx, tc_equalize_info* info);

/* ---- ensembles (no counterpart in the reference) ------------------------ */
/* Runs nsteps time steps (tc_time_step) on each of n independent contexts
 * -- separate cases, any grids and parameters -- concurrently: `threads`
 * host threads (0: one per context) take the contexts from a shared queue,
 * each context on its own stream, so the GPU overlaps the cases' launches
 * (a single case leaves most of the chip idle on its coarse levels).  Each
 * case's results are those of running it alone.  Returns 0 or the first
 * nonzero status, with the index of that context in *failed (-1 if none). */
int tc_ensemble_step(tc_ctx* const* ctxs, int32_t n, int32_t nsteps, int32_t threads, int32_t* failed);

/* ---- tracer order (no counterpart in the reference) --------------------- */
/* Reorders the device tracer arrays by the node each tracer is binned to
 * (row-major node index; the order carries no meaning in the reference).
 * tc_get_tracers keeps returning the caller's order; tc_put_tracers sets
 * the caller's order again. */
int tc_sort_tracers(tc_ctx* ctx);
/* Sort automatically at the start of every `interval`-th tc_time_step step
 * (0: never, the default). */
int tc_set_tracer_sort(tc_ctx* ctx, int32_t interval);
/* caller_index[j] (ntr entries) = the caller's index of the tracer held in
 * device slot j (the identity before any sort). */
int tc_tracer_slots(tc_ctx* ctx, int64_t* caller_index);

/* ---- measurement hooks (no counterpart in the reference) --------------- */
/* Kernel ids for tc_kernel_timing. */
enum { TC_K_SWEEP_L1 = 0, TC_K_COUNT };
/* enable=1: time the level-1 sweeps with HIP events on the library stream
 * (the stream the kernels run on): one event pair around each visit's
 * back-to-back launches (around each launch where the reference form's wall
 * reset follows the first sweep); enable=0 stops. */
int tc_set_profiling(tc_ctx* ctx, int enable);
/* Average device duration (microseconds) and launch count of kernel `which`
 * since profiling was enabled. */
int tc_kernel_timing(tc_ctx* ctx, int which, double* avg_us, int64_t* count);

/* ---- learned initial guess (no counterpart in the reference) ----------- */
/*
 * A 4-level U-Net (channels 16/32/64/128, 3x3 convolutions, GELU; BF16
 * weights and activations, FP32 accumulation on MFMA) predicts the increment
 * of the streamfunction from (T, C, log10 eta, psi_prev, psi_prev - psi_prev2)
 * on the H x W = nz x nx node grid (psi_prev, psi_prev2: the last two
 * solutions).  The hook is the warm start error(1,:,:)=SF of
 * multigrid (iterative_solver.f90:87): in the defect-correction modes the
 * solver evaluates the exact residual of the guess and of the plain warm
 * start and iterates from the better one, to the same tolerance.
 */
typedef struct tc_nn tc_nn;
int tc_nn_create(int32_t H, int32_t W, int32_t trainable, tc_nn** out);
int tc_nn_destroy(tc_nn* nn);
const char* tc_nn_last_error(const tc_nn* nn);
int tc_nn_param_count(const tc_nn* nn, int64_t* count);
/* parameters: per layer W[cout][9][cin] then b[cout]; last the 1x1 head w[16], b */
int tc_nn_set_weights(tc_nn* nn, const float* w, int64_t count);
int tc_nn_get_weights(tc_nn* nn, float* w, int64_t count);
int tc_nn_get_grads(tc_nn* nn, float* g, int64_t count);
/* x: H*W*5 (pixel k*W+i, channel fastest); target, mask: H*W (may be NULL) */
int tc_nn_set_sample(tc_nn* nn, const float* x, const float* target, const float* mask);
int tc_nn_forward(tc_nn* nn, float* out);
int tc_nn_backward(tc_nn* nn, double* loss);
int tc_nn_adam(tc_nn* nn, float lr, float grad_scale);
/* Solver side: attach a network whose H x W are the nz x nx of a multigrid
 * level L < grids (the network then works on level L: inputs restricted
 * there, its predicted increment prolonged to the fine grid), record
 * training samples (inputs before / increments after each solve) for the
 * next `capacity` solves, train on them, and use the network as the
 * initial guess. */
int tc_learn_attach(tc_ctx* ctx, tc_nn* nn);
int tc_learn_record(tc_ctx* ctx, int32_t capacity);
int tc_learn_samples(const tc_ctx* ctx, int32_t* count);
int tc_learn_train(tc_ctx* ctx, int32_t epochs, float lr, int32_t batch, uint64_t seed, double* mean_loss);
/* Initial guess of the defect-correction modes' solves inside tc_time_step.
 * TC_GUESS_PLAIN: the last solution (the reference's warm start).
 * TC_GUESS_EXTRAP: also the stage increment: the last solution (the previous
 * RK stage's) plus the change between the same two stages in the last step.
 * TC_GUESS_LEARNED: also the stage increment, and the stage's extrapolation
 * in time from its solutions in the last two steps plus the attached
 * network's predicted correction of that extrapolation.  Every candidate is
 * judged by the residual after the cycle's first level-1 smoothing and the
 * smallest wins; the plain start is always a candidate, and every solve
 * iterates to the same tolerance.  A candidate that lost its last 8
 * judgements is built and judged only every 8th solve until it wins. */
enum { TC_GUESS_PLAIN = 0, TC_GUESS_EXTRAP = 1, TC_GUESS_LEARNED = 2 };
int tc_guess_enable(tc_ctx* ctx, int mode);

/* ---- binning without allocation ------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The binning of tc_tracers_to_corners -- nearest-node
 * counts per type, mirrored ghost counts, nempty, the tracer ratio Cnew
 * (empty nodes keep their value) and enforceBCs_comp (composition.f90:
 * 669-803), without the C filter -- on caller-owned DEVICE buffers, with no
 * context and no allocation.  Planes use the reference layout
 * f(-2:nx+2,-2:nz+2), i fastest, one plane after another per type:
 * count_type[t] (t = 0..ntype), Cnew[t-1] (t = 1..ntype).  count_type may be
 * NULL: the counts are then kept in the library's fixed per-GPU work array
 * of TC_BIN_NOALLOC_MAX_COUNTS int32 words, which (ntype+1)(nx+5)(nz+5) must
 * not exceed (else TC_ERR_NOMEM).  Runs on `stream` (a hipStream_t, NULL for
 * the default stream) and synchronizes it; returns TC_ERR_TRACER_DOMAIN
 * when a tracer is outside the domain.  Calls that overlap on one GPU share
 * that array and race. */
#define TC_BIN_NOALLOC_MAX_COUNTS (8u << 20)
int tc_bin_tracers_noalloc(int32_t nx, int32_t nz, double aspect, int32_t ntype, int32_t comp, const double* r,
                           const double* s, const int32_t* ttype, int64_t ntr, int32_t* count_type, int32_t* nempty,
                           double* Cnew, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* TRACER_HIP_H */
