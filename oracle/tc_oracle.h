/*
 * tc_oracle.h -- CPU oracle of the tracer_code time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Fortran (seantrim/tracer_code) used as the checker for the HIP library.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product path never does.
 *
 * Layout: every grid field is a ghosted plane f(-2:nx+2,-2:nz+2), i fastest,
 * leading dimension nx+5 (the reference's Fortran layout).  Multi-level
 * arrays of the reference (level-fastest, basics.f90:87-98) are stored as one
 * fine-size plane per level; the coarse levels use the low corner of their
 * plane exactly as the reference does.  Composition arrays are one plane
 * per type.
 */
#ifndef TC_ORACLE_H
#define TC_ORACLE_H
#include <stdint.h>
#include "../include/tracer_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_state or_state;

/* field ids for or_ptr (double planes unless noted) */
enum {
    OR_T = 0, OR_SF, OR_U, OR_W, OR_STRAIN, OR_TVIS, OR_TBUOY, OR_CONDUCT,
    OR_CONDUCT_R, OR_CONDUCT_S, OR_DTS, OR_DERR_U, OR_DERR_W, OR_T0, OR_T1, OR_T2,
    OR_RESIDUAL,
    /* per type (idx = type 1..ntype) */
    OR_CNEW = 32, OR_CVIS, OR_CBUOY, OR_C0,
    /* per level (idx = level 1..grids) */
    OR_ERROR = 64, OR_RHS, OR_VIS, OR_VISX, OR_VISZ, OR_VISXX, OR_VISZZ, OR_VISXZ,
    OR_DTSTREAM, OR_DTVIS, OR_RATIO,
    /* int planes (idx = type 0..ntype for COUNT) */
    OR_COUNT = 96, OR_NEMPTY,
    /* tracer arrays (length ntr) */
    OR_RTR = 128, OR_STR, OR_RTR0, OR_STR0, OR_RTR1, OR_STR1, OR_RTR2, OR_STR2, OR_TTYPE,
    /* coarse banded matrix AB(2kl+ku+1, ngrid_coarse) column-major, RowB, IPIV */
    OR_MCOARSE = 160, OR_ROWB, OR_IPIV, OR_BCOARSE
};

/* scalar ids for or_get_scalar / or_set_scalar */
enum {
    ORS_DT = 0, ORS_TIME, ORS_RESID1, ORS_RESREF1, ORS_EIG_FINE, ORS_DELTA_T,
    ORS_Q = 16, ORS_N_ETA, ORS_N_TEMP, ORS_N_COMP1, ORS_STATUS, ORS_KL, ORS_KU,
    ORS_NGC, ORS_TSTEP, ORS_SMOOTH_VIS, ORS_NVIS_STORE, ORS_NN_STORE,
    ORS_CYCLES_TOTAL, ORS_VIS_ITERS, ORS_C_ITERS, ORS_T_ITERS,
    /* cycles of each solve of the last or_time_step, in call order */
    ORS_Q_STEP0 = 32, ORS_Q_STEP1, ORS_Q_STEP2, ORS_Q_STEP3,
    /* n_comp of types 1..4 after the last smoother_time */
    ORS_N_COMP_T1 = 36, ORS_N_COMP_T2, ORS_N_COMP_T3, ORS_N_COMP_T4,
    /* interpolations of out-of-domain positions (cumulative) */
    ORS_INTERP_CLAMPS = 40,
    /* the cycle limit of or_multigrid (tc_params.max_cycles) */
    ORS_MAX_CYCLES = 41
};

or_state* or_create(const tc_params* p, int64_t ntr);
void or_destroy(or_state* s);
void* or_ptr(or_state* s, int which, int idx);
double or_get_scalar(const or_state* s, int which);
void or_set_scalar(or_state* s, int which, double v);
int or_ld(const or_state* s);
int or_nx_grid(const or_state* s, int level);
int or_nz_grid(const or_state* s, int level);
double or_residual_mag(const or_state* s, int p);
double or_res_ref(const or_state* s, int level);

/* host set-up helpers (init-time routines of the reference) */
void or_fd_coefficients(int order, double D1[5], double D2[5], double D3[5], double D4[5], double Dint[2][3]);
void or_optimize_grid(int* nx, int* nz, int* grids, int n_points_min);
int  or_multigrid_setup(int grids, int multi[TC_MAX_MULTI]);
void or_derivative_eigenvalues(const double D1[5], const double D2[5], const double D3[5],
                               const double D4[5], double base[7]);

/* hot-path routines */
void or_stability_vis_smoother(or_state* s, int level);
void or_stability_RK4_smoother(or_state* s);
void or_enforce_bcs(or_state* s, double* T0);
void or_enforce_bcs_stream(or_state* s, int level, double* S0);
void or_enforce_bcs_comp(or_state* s, double* C0);
double or_stream_space(or_state* s, int level, int i, int k, const double* S0);
void or_iterate_stream(or_state* s, int level, int Niter, int dir, int p);
void or_sweep(or_state* s, int level, int n, int Niter, int dir);
void or_compute_residual(or_state* s, int level, int p);
void or_compute_RHS(or_state* s, int level);
void or_compute_reference_residual(or_state* s, int level);
void or_restrict(or_state* s, int level, const double* in, double* out);
void or_prolong(or_state* s, int level_coarse);
void or_visf_plane(or_state* s, double* out);
void or_viscosity_gradients(or_state* s, int level);
void or_test_eigenvalues(or_state* s, int level);
void or_iterate_local_time_steps(or_state* s, int level);
void or_smoother_vis(or_state* s, int level, int interpolate);
void or_compute_viscosity(or_state* s, int level);
void or_viscosity_bcs(or_state* s, int level);
void or_viscosity_bcs_initialize(or_state* s, int level);
void or_compute_matrix_coarse(or_state* s);
int  or_dgbequ(int m, int n, int kl, int ku, const double* ab, int ldab, double* r, double* c,
               double* rowcnd, double* colcnd, double* amax);
int  or_dgbtrf(int m, int n, int kl, int ku, double* ab, int ldab, int* ipiv);
void or_dgbtrs(int n, int kl, int ku, const double* ab, int ldab, const int* ipiv, double* b);
void or_coarse_solve(or_state* s);
int  or_multigrid(or_state* s);
void or_compute_velocities(or_state* s);
void or_compute_strain_rate_invariant(or_state* s);
int  or_smoother_time_T(or_state* s, const double* Tin);
void or_smoother_time(or_state* s);
void or_load_conductivity_array(or_state* s);
int  or_tracers_to_corners(or_state* s, int which);
void or_compute_derivatives(or_state* s, const double* f, double* DERr);
double or_interpolate(const or_state* s, double r, double sv, const double* DERr, const double* f);
int  or_tracer_stage(or_state* s, int stage, double dt);
void or_energy_stage(or_state* s, int stage, double dt);
double or_stability_RK4(or_state* s);
int  or_energy_time(or_state* s);
int  or_time_step(or_state* s);
int  or_initialize_state(or_state* s);
/* diagnostics (basics.f90:236-324): out[0]=vrms, [1]=Nu, [2]=Tavg, [3]=Cavg(1),
 * [4]=empty_cells, [5]=tracer_min, [6]=tracer_max, [7]=Nu_bot, [8]=strain_max */
void or_stats(or_state* s, double out[9]);
/* the whole stats.dat row (basics.f90:236-324); mass NULL = this call's Cavg */
void or_stats_record(or_state* s, int kent, const double* mass, tc_stats_row* out);
/* equalize_tracers (composition.f90:83-580) in its counter-based form */
int  or_equalize_tracers(or_state* s, int tpc, const double* Cinit, uint64_t seed, uint64_t call,
                         tc_equalize_info* info);
/* with on set, or_energy_time equalizes after stability_RK4 (energy.f90:133-135),
 * call k of the run drawing from stream (seed, k) */
void or_set_equalize(or_state* s, int on, int tpc, const double* Cinit, uint64_t seed);
/* the pieces of the counter-based random stream, for the tests */
double or_eq_uniform(uint64_t seed, uint64_t call, unsigned stream, uint64_t a, uint64_t b);
uint64_t or_eq_perm40(uint64_t seed, uint64_t call, uint64_t id);
void or_eq_shuffle(int32_t* a, int n, uint64_t seed, uint64_t call, unsigned stream, uint64_t id);

#ifdef __cplusplus
}
#endif
#endif
