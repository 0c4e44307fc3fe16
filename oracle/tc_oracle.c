/*
 * tc_oracle.c -- CPU oracle (test infrastructure only) for the tracer_code
 * time-step hot path.  A line-by-line restatement in C of the reference
 * Fortran; every routine cites the reference file:line it follows.  It is
 * compiled with -O2 -ffp-contract=off so that each floating-point operation
 * is evaluated in the order the Fortran source writes it.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this code.  The reference needs Intel MKL (DGBEQU/DGBTRF/DGBTRS) which the
 * image lacks, so the banded LAPACK routines are restated from the published
 * LAPACK reference algorithm (dgbequ/dgbtf2/dgbtrs); MKL's blocked DGBTRF
 * rounds differently, so that part is pinned against an independent dense
 * solve, not bit for bit.
 */
#include "tc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <stddef.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SPAN 2
#define PII 3.1415926535897932

struct or_state {
    tc_params p;
    int nx, nz, grids, ld, ldz, ntype;
    size_t plane;
    int nxg[TC_MAX_LEVELS + 1], nzg[TC_MAX_LEVELS + 1];
    double drg[TC_MAX_LEVELS + 1], dsg[TC_MAX_LEVELS + 1], dr2g[TC_MAX_LEVELS + 1], ds2g[TC_MAX_LEVELS + 1];
    double dr3g[TC_MAX_LEVELS + 1], ds3g[TC_MAX_LEVELS + 1], dr4g[TC_MAX_LEVELS + 1], ds4g[TC_MAX_LEVELS + 1];
    double drdsg[TC_MAX_LEVELS + 1], drds2g[TC_MAX_LEVELS + 1], dr2dsg[TC_MAX_LEVELS + 1], dr2ds2g[TC_MAX_LEVELS + 1];
    double dr, ds, dr2, ds2, dro2, dso2, dist_r, dist_s;
    double D1[5], D2[5], D3[5], D4[5];
    double Drs[5][5], Drss[5][5], Drrs[5][5], Drrss[5][5]; /* [i+2][k+2] */
    double Dint[2][3];
    /* grid metrics (uniform grids: xr=zs=1, others 0), index offset SPAN */
    double *xr, *xrr, *xrrr, *xrrrr, *zs, *zss, *zsss, *zssss;
    double *zg; /* zg(-3:nz+3) stored with offset 3 */
    /* grid fields */
    double *T, *SF, *u, *w, *strain, *Tvis, *Tbuoy, *conduct, *conduct_r, *conduct_s, *dts;
    double *DERr_u, *DERr_w, *T0, *T1, *T2, *residual, *tmp, *tmp2;
    double *Cnew[TC_MAX_TYPES + 1], *Cvis[TC_MAX_TYPES + 1], *Cbuoy[TC_MAX_TYPES + 1], *C0[TC_MAX_TYPES + 1];
    int *count[TC_MAX_TYPES + 1], *nempty;
    /* multigrid */
    double *error[TC_MAX_LEVELS + 1], *RHS[TC_MAX_LEVELS + 1], *vis[TC_MAX_LEVELS + 1];
    double *visx[TC_MAX_LEVELS + 1], *visz[TC_MAX_LEVELS + 1], *visxx[TC_MAX_LEVELS + 1];
    double *viszz[TC_MAX_LEVELS + 1], *visxz[TC_MAX_LEVELS + 1];
    double *visf[TC_MAX_LEVELS + 1], *visxf[TC_MAX_LEVELS + 1], *viszf[TC_MAX_LEVELS + 1];
    double *visxxf[TC_MAX_LEVELS + 1], *viszzf[TC_MAX_LEVELS + 1], *visxzf[TC_MAX_LEVELS + 1];
    double *dt_stream[TC_MAX_LEVELS + 1], *dt_vis[TC_MAX_LEVELS + 1], *ratio[TC_MAX_LEVELS + 1];
    double res_ref[TC_MAX_LEVELS + 1], residual_mag[TC_MAX_MULTI];
    /* eigenvalue scalars (module variables of basics.f90:14-15) */
    double Tr_max, Trr_max, Trrr_max, Trrrr_max, Ts_max, Tss_max, Tsss_max, Tssss_max;
    double Trs_max, Trrs_max, Trss_max, Trrss_max;
    /* coarse direct solve */
    int ngc, kl, ku;
    double *Mc, *Bc, *RowB, *ColB;
    int *ipiv;
    /* tracers */
    int64_t ntr;
    double *rtr, *str, *rtr0, *str0, *rtr1, *str1, *rtr2, *str2;
    int32_t *Ttype;
    /* scalars */
    double dt, time, eig_fine;
    int tstep, q, n_eta, n_temp, n_comp[TC_MAX_TYPES + 1], status, smooth_vis;
    int Nvis_store, nn_store;
    int64_t cycles_total, vis_iters, c_iters, t_iters;
    /* interpolations whose cell index needed the clamp (positions outside
     * the domain; the reference would read outside its arrays there) */
    int64_t interp_clamps;
    /* cycles of each solve of the running or_time_step (q_step_n = solves
     * recorded so far; -1 outside a time step) */
    int q_step[4], q_step_n;
    /* equalize_tracers in the time step */
    int eq_on, eq_tpc;
    double eq_Cinit[TC_MAX_TYPES];   /* type 1..ntype at [0..ntype-1] */
    uint64_t eq_seed, eq_calls;
};

#define IX(s, i, k) ((size_t)((k) + SPAN) * (size_t)(s)->ld + (size_t)((i) + SPAN))

static double* alloc_plane(const or_state* s) { return (double*)calloc(s->plane, sizeof(double)); }
static int* alloc_iplane(const or_state* s) { return (int*)calloc(s->plane, sizeof(int)); }

/* ------------------------------------------------------------------------ */
/* Start-up routines (host side in the reference).                          */
/* ------------------------------------------------------------------------ */

/* FD_coefficients (main.f90:350-469): Fornberg weights on the nodes
 * a = 0, 1, -1, 2, -2, ...; D1/D2 on order+1 points, D3/D4 on order+3,
 * interpolation table D(m,.) on order+m-1 points, then the enforced
 * (anti)symmetries.  Only order=2 is used by the hot path. */
void or_fd_coefficients(int order, double D1[5], double D2[5], double D3[5], double D4[5], double Dint[2][3])
{
    const int Dnumber = order;
    const int bigM = Dnumber > 4 ? Dnumber : 4;
    const int bigN = order + bigM - 2;
    const int span = order / 2 + 1, span1 = order / 2, span_interp = (order + Dnumber - 2) / 2;
    double a[16];
    /* delta(-1:bigM, 0:bigN, 0:bigN) */
    static double delta[8][16][16];
#define DL(m, n, i) delta[(m) + 1][n][i]
    memset(delta, 0, sizeof(delta));
    a[0] = 0.0;
    int k = 1;
    for (int i = 1; i <= bigN; i += 2) {
        a[i] = (double)k;
        a[i + 1] = -(double)k;
        k++;
    }
    const double x0 = 0.0;
    DL(0, 0, 0) = 1.0;
    double c1 = 1.0;
    for (int n = 1; n <= bigN; ++n) {
        double c2 = 1.0;
        for (int i = 0; i <= n - 1; ++i) {
            double c3 = a[n] - a[i];
            c2 = c2 * c3;
            if (n <= bigM) DL(n, n - 1, i) = 0.0;
            for (int m = 0; m <= (n < bigM ? n : bigM); ++m)
                DL(m, n, i) = ((a[n] - x0) * DL(m, n - 1, i) - (double)m * DL(m - 1, n - 1, i)) / c3;
        }
        for (int m = 0; m <= (n < bigM ? n : bigM); ++m)
            DL(m, n, n) = (c1 / c2) * ((double)m * DL(m - 1, n - 1, n - 1) - (a[n - 1] - x0) * DL(m, n - 1, n - 1));
        c1 = c2;
    }
    for (int j = 0; j < 5; ++j) D1[j] = D2[j] = D3[j] = D4[j] = 0.0;
    D1[2] = DL(1, order, 0);
    D2[2] = DL(2, order, 0);
    k = 1;
    for (int i = 1; i <= order; i += 2) {
        D1[2 + k] = DL(1, order, i);
        D1[2 - k] = DL(1, order, i + 1);
        D2[2 + k] = DL(2, order, i);
        D2[2 - k] = DL(2, order, i + 1);
        k++;
    }
    /* D(1:Dnumber,-span_interp:span_interp) (main.f90:403-422) */
    for (int m = 0; m < 2; ++m)
        for (int j = 0; j < 3; ++j) Dint[m][j] = 0.0;
    for (int m = 2; m <= Dnumber; m += 2) {
        Dint[m - 2][1] = DL(m - 1, order + m - 2, 0);
        Dint[m - 1][1] = DL(m, order + m - 2, 0);
        k = 1;
        for (int i = 1; i <= order + m - 2; i += 2) {
            Dint[m - 2][1 + k] = DL(m - 1, order + m - 2, i);
            Dint[m - 2][1 - k] = DL(m - 1, order + m - 2, i + 1);
            Dint[m - 1][1 + k] = DL(m, order + m - 2, i);
            Dint[m - 1][1 - k] = DL(m, order + m - 2, i + 1);
            k++;
        }
    }
    for (int m = 1; m <= Dnumber; m += 2) {
        Dint[m - 1][1] = 0.0;
        for (int i = 1; i <= span_interp; ++i) {
            Dint[m - 1][1 + i] = -Dint[m - 1][1 - i];
            Dint[m][1 + i] = Dint[m][1 - i];
        }
    }
    D3[2] = DL(3, order + 2, 0);
    D4[2] = DL(4, order + 2, 0);
    k = 1;
    for (int i = 1; i <= order + 2; i += 2) {
        D3[2 + k] = DL(3, order + 2, i);
        D3[2 - k] = DL(3, order + 2, i + 1);
        D4[2 + k] = DL(4, order + 2, i);
        D4[2 - k] = DL(4, order + 2, i + 1);
        k++;
    }
    D1[2] = 0.0;
    D3[2] = 0.0;
    for (int i = 1; i <= span; ++i) {
        D1[2 + i] = -D1[2 - i];
        D2[2 + i] = D2[2 - i];
        D3[2 + i] = -D3[2 - i];
        D4[2 + i] = D4[2 - i];
    }
    (void)span1;
#undef DL
}

/* optimize_grid (iterative_solver.f90:1-61). */
void or_optimize_grid(int* nx, int* nz, int* grids, int n_points_min)
{
    int n_points_max = *nx < *nz ? *nx : *nz;
    int n_coarse_min = 1, n_coarse_max = *grids - 1;
    int n_max = n_points_max * (1 << n_coarse_max);
    int nxbig, n_in;
    if (*nz <= *nx) { nxbig = 1; n_in = *nz; } else { nxbig = 0; n_in = *nx; }
    int n_temp = n_max, n_coarse_temp = 0, n_points_temp = 0;
    for (int n_points = n_points_min; n_points <= n_points_max; ++n_points)
        for (int n_coarse = n_coarse_min; n_coarse <= n_coarse_max; ++n_coarse) {
            int n = n_points * (1 << n_coarse);
            if (n >= n_in && n < n_temp) {
                n_temp = n;
                n_coarse_temp = n_coarse;
                n_points_temp = n_points;
            }
        }
    *grids = n_coarse_temp + 1;
    if (nxbig == 0) *nx = n_temp; else *nz = n_temp;
    n_in = (*nz <= *nx) ? *nx : *nz;
    n_points_temp = (int)lround((double)n_in / (double)(1 << n_coarse_temp));
    n_temp = n_points_temp * (1 << n_coarse_temp);
    if (nxbig == 0) *nz = n_temp; else *nx = n_temp;
}

/* multigrid_setup (multigrid_cycle.f90:1-99): the active W-cycle (V for two
 * levels) visit sequences and the periodic wrap multi(0), multi(nsteps+1). */
int or_multigrid_setup(int grids, int multi[TC_MAX_MULTI])
{
    static const int w2[] = {1, 2};
    static const int w3[] = {1, 2, 3, 2, 3, 2};
    static const int w4[] = {1, 2, 3, 4, 3, 4, 3, 2, 3, 4, 3, 4, 3, 2};
    static const int w5[] = {1, 2, 3, 4, 5, 4, 5, 4, 3, 4, 5, 4, 3, 2, 3, 4, 5, 4, 3, 4, 5, 4, 5, 4, 3, 2};
    static const int w6[] = {1, 2, 3, 4, 5, 6, 5, 6, 5, 4, 5, 6, 5, 4, 3, 4, 5, 6, 5, 4, 3,
                             2, 3, 4, 5, 6, 5, 4, 3, 4, 5, 6, 5, 4, 5, 6, 5, 6, 5, 4, 3, 2};
    const int* seq = NULL;
    int nsteps = 0;
    switch (grids) {
    case 2: seq = w2; nsteps = 2; break;
    case 3: seq = w3; nsteps = 6; break;
    case 4: seq = w4; nsteps = 14; break;
    case 5: seq = w5; nsteps = 26; break;
    case 6: seq = w6; nsteps = 42; break;
    default: return -1;
    }
    for (int i = 0; i < TC_MAX_MULTI; ++i) multi[i] = 0;
    for (int i = 0; i < nsteps; ++i) multi[i + 1] = seq[i];
    multi[0] = multi[nsteps];
    multi[nsteps + 1] = multi[1];
    return nsteps;
}

/* T_test (energy.f90:590-598): cmplx() without a KIND is default (single)
 * precision, so each factor is rounded to float and multiplied in float. */
static void t_test(int i, int k, double kr, double ks, double* re, double* im)
{
    double r = (double)i, sv = (double)k;
    volatile float a = (float)cos(kr * r), b = (float)sin(kr * r);
    volatile float c = (float)cos(ks * sv), d = (float)sin(ks * sv);
    volatile float ac = a * c, bd = b * d, ad = a * d, bc = b * c;
    volatile float x = ac - bd, y = ad + bc;
    *re = (double)x;
    *im = (double)y;
}

/* derivative_eigenvalues (energy.f90:501-588): maxima of the Fourier symbols
 * of the FD operators.  base[] = Tr, Trr, Trrr, Trrrr, Trs, Trrs, Trrss. */
void or_derivative_eigenvalues(const double D1[5], const double D2[5], const double D3[5],
                               const double D4[5], double base[7])
{
    const int n_angle = (1 << 11) + 1;
    const double d_angle = 2.0 * PII / (double)(n_angle - 1);
    double Drs[3][3], Drrs[3][3], Drrss[3][3];
    for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b) {
            Drs[a + 1][b + 1] = D1[2 + b] * D1[2 + a];
            Drrs[a + 1][b + 1] = D1[2 + b] * D2[2 + a];
            Drrss[a + 1][b + 1] = D2[2 + b] * D2[2 + a];
        }
    double Tr_max = 0, Trr_max = 0, Trrr_max = 0, Trrrr_max = 0, Trs_max = 0, Trrs_max = 0, Trrss_max = 0;
#pragma omp parallel for reduction(max : Tr_max, Trr_max, Trrr_max, Trrrr_max, Trs_max, Trrs_max, Trrss_max)
    for (int angle = 1; angle <= n_angle; ++angle) {
        double ks = 0.0, kr = (double)(angle - 1) * d_angle;
        double tr_re = 0, tr_im = 0, trr_re = 0, trr_im = 0, t3_re = 0, t3_im = 0, t4_re = 0, t4_im = 0;
        double re, im;
        for (int j = -1; j <= 1; ++j) {
            t_test(j, 0, kr, ks, &re, &im);
            tr_re += D1[2 + j] * re; tr_im += D1[2 + j] * im;
            trr_re += D2[2 + j] * re; trr_im += D2[2 + j] * im;
        }
        for (int j = -2; j <= 2; ++j) {
            t_test(j, 0, kr, ks, &re, &im);
            t3_re += D3[2 + j] * re; t3_im += D3[2 + j] * im;
            t4_re += D4[2 + j] * re; t4_im += D4[2 + j] * im;
        }
        /* division by T_test(0,0,kr,ks) = (1,0) leaves the values unchanged */
        if (fabs(tr_im) > Tr_max) Tr_max = fabs(tr_im);
        if (fabs(trr_re) > Trr_max) Trr_max = fabs(trr_re);
        if (fabs(t3_im) > Trrr_max) Trrr_max = fabs(t3_im);
        if (fabs(t4_re) > Trrrr_max) Trrrr_max = fabs(t4_re);
        for (int angle2 = 1; angle2 <= n_angle; ++angle2) {
            ks = (double)(angle2 - 1) * d_angle;
            double trs_re = 0, trrs_im = 0, trrss_re = 0;
            for (int j = -1; j <= 1; ++j)
                for (int jj = -1; jj <= 1; ++jj) {
                    t_test(j, jj, kr, ks, &re, &im);
                    trs_re += Drs[j + 1][jj + 1] * re;
                    trrs_im += Drrs[j + 1][jj + 1] * im;
                    trrss_re += Drrss[j + 1][jj + 1] * re;
                }
            if (fabs(trs_re) > Trs_max) Trs_max = fabs(trs_re);
            if (fabs(trrs_im) > Trrs_max) Trrs_max = fabs(trrs_im);
            if (fabs(trrss_re) > Trrss_max) Trrss_max = fabs(trrss_re);
        }
    }
    base[0] = Tr_max; base[1] = Trr_max; base[2] = Trrr_max; base[3] = Trrrr_max;
    base[4] = Trs_max; base[5] = Trrs_max; base[6] = Trrss_max;
}

/* ------------------------------------------------------------------------ */
/* State                                                                     */
/* ------------------------------------------------------------------------ */

or_state* or_create(const tc_params* p, int64_t ntr)
{
    or_state* s = (or_state*)calloc(1, sizeof(or_state));
    s->q_step_n = -1;
    s->p = *p;
    s->nx = p->nx; s->nz = p->nz; s->grids = p->grids; s->ntype = p->ntype;
    s->ld = p->nx + 2 * SPAN + 1;
    s->ldz = p->nz + 2 * SPAN + 1;
    s->plane = (size_t)s->ld * (size_t)s->ldz;
    /* compute_basics (basics.f90:112-183) */
    s->dr = p->aspect / (double)p->nx;
    s->ds = 1.0 / (double)p->nz;
    s->dr2 = s->dr * s->dr;
    s->ds2 = s->ds * s->ds;
    s->dro2 = 0.5 * s->dr;
    s->dso2 = 0.5 * s->ds;
    s->dist_r = 1e-4 * s->dr;
    s->dist_s = 1e-4 * s->ds;
    for (int j = 1; j <= s->grids; ++j) {
        s->nxg[j] = p->nx / (1 << (j - 1));
        s->nzg[j] = p->nz / (1 << (j - 1));
        s->drg[j] = p->aspect / (double)s->nxg[j];
        s->dsg[j] = 1.0 / (double)s->nzg[j];
        s->dr2g[j] = s->drg[j] * s->drg[j];
        s->ds2g[j] = s->dsg[j] * s->dsg[j];
        s->dr3g[j] = pow(s->drg[j], 3.0);
        s->ds3g[j] = pow(s->dsg[j], 3.0);
        s->dr4g[j] = pow(s->drg[j], 4.0);
        s->ds4g[j] = pow(s->dsg[j], 4.0);
        s->drdsg[j] = s->drg[j] * s->dsg[j];
        s->drds2g[j] = s->drg[j] * s->ds2g[j];
        s->dr2dsg[j] = s->dr2g[j] * s->dsg[j];
        s->dr2ds2g[j] = s->dr2g[j] * s->ds2g[j];
    }
    memcpy(s->D1, p->D1, sizeof(s->D1));
    memcpy(s->D2, p->D2, sizeof(s->D2));
    memcpy(s->D3, p->D3, sizeof(s->D3));
    memcpy(s->D4, p->D4, sizeof(s->D4));
    memcpy(s->Dint, p->Dint, sizeof(s->Dint));
    /* Drs(i,k)=D1(k)*D1(i) ... (main.f90:448-459) */
    for (int i = -2; i <= 2; ++i)
        for (int k = -2; k <= 2; ++k) {
            int in1 = (i >= -1 && i <= 1 && k >= -1 && k <= 1);
            s->Drs[i + 2][k + 2] = in1 ? s->D1[k + 2] * s->D1[i + 2] : 0.0;
            s->Drss[i + 2][k + 2] = in1 ? s->D2[k + 2] * s->D1[i + 2] : 0.0;
            s->Drrs[i + 2][k + 2] = in1 ? s->D1[k + 2] * s->D2[i + 2] : 0.0;
            s->Drrss[i + 2][k + 2] = in1 ? s->D2[k + 2] * s->D2[i + 2] : 0.0;
        }
    /* uniform-grid metrics (grid.f90:110-145 with gridX=gridZ=0) */
    int nmx = p->nx + 2 * SPAN + 1, nmz = p->nz + 2 * SPAN + 1;
    s->xr = (double*)calloc(nmx, sizeof(double)); s->xrr = (double*)calloc(nmx, sizeof(double));
    s->xrrr = (double*)calloc(nmx, sizeof(double)); s->xrrrr = (double*)calloc(nmx, sizeof(double));
    s->zs = (double*)calloc(nmz, sizeof(double)); s->zss = (double*)calloc(nmz, sizeof(double));
    s->zsss = (double*)calloc(nmz, sizeof(double)); s->zssss = (double*)calloc(nmz, sizeof(double));
    for (int i = 0; i < nmx; ++i) s->xr[i] = 1.0;
    for (int k = 0; k < nmz; ++k) s->zs[k] = 1.0;
    /* zg(k) = k*ds with the mirrored ghosts of build_grid (grid.f90:51-54, 101-108) */
    s->zg = (double*)calloc(p->nz + 7, sizeof(double));
    for (int k = 0; k <= p->nz; ++k) s->zg[k + 3] = s->ds * (double)k;
    for (int k = 1; k <= 3; ++k) {
        s->zg[-k + 3] = -s->zg[k + 3];
        s->zg[p->nz + k + 3] = 1.0 - s->zg[p->nz - k + 3] + 1.0;
    }
    s->zg[3] = 0.0;
    s->zg[p->nz + 3] = 1.0;

    s->T = alloc_plane(s); s->SF = alloc_plane(s); s->u = alloc_plane(s); s->w = alloc_plane(s);
    s->strain = alloc_plane(s); s->Tvis = alloc_plane(s); s->Tbuoy = alloc_plane(s);
    s->conduct = alloc_plane(s); s->conduct_r = alloc_plane(s); s->conduct_s = alloc_plane(s);
    s->dts = alloc_plane(s); s->DERr_u = alloc_plane(s); s->DERr_w = alloc_plane(s);
    s->T0 = alloc_plane(s); s->T1 = alloc_plane(s); s->T2 = alloc_plane(s);
    s->residual = alloc_plane(s); s->tmp = alloc_plane(s); s->tmp2 = alloc_plane(s);
    for (int t = 0; t <= s->ntype; ++t) s->count[t] = alloc_iplane(s);
    s->nempty = alloc_iplane(s);
    for (int t = 1; t <= s->ntype; ++t) {
        s->Cnew[t] = alloc_plane(s); s->Cvis[t] = alloc_plane(s);
        s->Cbuoy[t] = alloc_plane(s); s->C0[t] = alloc_plane(s);
    }
    for (int L = 1; L <= s->grids; ++L) {
        s->error[L] = alloc_plane(s); s->RHS[L] = alloc_plane(s); s->vis[L] = alloc_plane(s);
        s->visx[L] = alloc_plane(s); s->visz[L] = alloc_plane(s); s->visxx[L] = alloc_plane(s);
        s->viszz[L] = alloc_plane(s); s->visxz[L] = alloc_plane(s);
        s->visf[L] = alloc_plane(s); s->visxf[L] = alloc_plane(s); s->viszf[L] = alloc_plane(s);
        s->visxxf[L] = alloc_plane(s); s->viszzf[L] = alloc_plane(s); s->visxzf[L] = alloc_plane(s);
        s->dt_stream[L] = alloc_plane(s); s->dt_vis[L] = alloc_plane(s); s->ratio[L] = alloc_plane(s);
    }
    /* coarse matrix bandwidth analysis (Compute_Matrix_coarse analysis=1,
     * iterative_solver.f90:1048-1070, Database=2 ordering) */
    {
        int nxc = s->nxg[s->grids], nzc = s->nzg[s->grids];
        s->ngc = (nxc + 1) * (nzc + 1);
        int limit[5] = {0, 1, 2, 1, 0};
        int ku = 0, kl = 0;
        for (int pp = 1; pp <= s->ngc; ++pp) {
            int kp = pp - (nzc + 1) * ((pp - 1) / (nzc + 1)) - 1, ip = (pp - 1) / (nzc + 1);
            for (int j = -2; j <= 2; ++j) {
                int iq = j + ip;
                if (iq > 0 && iq < nxc)
                    for (int kq = -limit[j + 2] + kp; kq <= kp + limit[j + 2]; ++kq)
                        if (kq > 0 && kq < nzc) {
                            int qq = (nzc + 1) * iq + kq + 1;
                            int t = qq - pp;
                            if (t > ku) ku = t;
                            if (-t > kl) kl = -t;
                        }
            }
        }
        s->kl = kl; s->ku = ku;
        s->Mc = (double*)calloc((size_t)(2 * kl + ku + 1) * (size_t)s->ngc, sizeof(double));
        s->Bc = (double*)calloc(s->ngc, sizeof(double));
        s->RowB = (double*)calloc(s->ngc, sizeof(double));
        s->ColB = (double*)calloc(s->ngc, sizeof(double));
        s->ipiv = (int*)calloc(s->ngc, sizeof(int));
    }
    s->ntr = ntr;
    if (ntr > 0) {
        s->rtr = (double*)calloc(ntr, sizeof(double)); s->str = (double*)calloc(ntr, sizeof(double));
        s->rtr0 = (double*)calloc(ntr, sizeof(double)); s->str0 = (double*)calloc(ntr, sizeof(double));
        s->rtr1 = (double*)calloc(ntr, sizeof(double)); s->str1 = (double*)calloc(ntr, sizeof(double));
        s->rtr2 = (double*)calloc(ntr, sizeof(double)); s->str2 = (double*)calloc(ntr, sizeof(double));
        s->Ttype = (int32_t*)calloc(ntr, sizeof(int32_t));
    }
    s->smooth_vis = 1;
    s->tstep = 0;
    return s;
}

void or_destroy(or_state* s)
{
    if (!s) return;
    double* d[] = {s->xr, s->xrr, s->xrrr, s->xrrrr, s->zs, s->zss, s->zsss, s->zssss, s->zg, s->T, s->SF, s->u, s->w,
                   s->strain, s->Tvis, s->Tbuoy, s->conduct, s->conduct_r, s->conduct_s, s->dts, s->DERr_u,
                   s->DERr_w, s->T0, s->T1, s->T2, s->residual, s->tmp, s->tmp2, s->Mc, s->Bc, s->RowB, s->ColB,
                   s->rtr, s->str, s->rtr0, s->str0, s->rtr1, s->str1, s->rtr2, s->str2};
    for (size_t i = 0; i < sizeof(d) / sizeof(d[0]); ++i) free(d[i]);
    free(s->ipiv); free(s->Ttype); free(s->nempty);
    for (int t = 0; t <= s->ntype; ++t) free(s->count[t]);
    for (int t = 1; t <= s->ntype; ++t) { free(s->Cnew[t]); free(s->Cvis[t]); free(s->Cbuoy[t]); free(s->C0[t]); }
    for (int L = 1; L <= s->grids; ++L) {
        double* e[] = {s->error[L], s->RHS[L], s->vis[L], s->visx[L], s->visz[L], s->visxx[L], s->viszz[L],
                       s->visxz[L], s->visf[L], s->visxf[L], s->viszf[L], s->visxxf[L], s->viszzf[L], s->visxzf[L],
                       s->dt_stream[L], s->dt_vis[L], s->ratio[L]};
        for (size_t i = 0; i < sizeof(e) / sizeof(e[0]); ++i) free(e[i]);
    }
    free(s);
}

void* or_ptr(or_state* s, int which, int idx)
{
    switch (which) {
    case OR_T: return s->T; case OR_SF: return s->SF; case OR_U: return s->u; case OR_W: return s->w;
    case OR_STRAIN: return s->strain; case OR_TVIS: return s->Tvis; case OR_TBUOY: return s->Tbuoy;
    case OR_CONDUCT: return s->conduct; case OR_CONDUCT_R: return s->conduct_r; case OR_CONDUCT_S: return s->conduct_s;
    case OR_DTS: return s->dts; case OR_DERR_U: return s->DERr_u; case OR_DERR_W: return s->DERr_w;
    case OR_T0: return s->T0; case OR_T1: return s->T1; case OR_T2: return s->T2; case OR_RESIDUAL: return s->residual;
    case OR_CNEW: return s->Cnew[idx]; case OR_CVIS: return s->Cvis[idx]; case OR_CBUOY: return s->Cbuoy[idx];
    case OR_C0: return s->C0[idx];
    case OR_ERROR: return s->error[idx]; case OR_RHS: return s->RHS[idx]; case OR_VIS: return s->vis[idx];
    case OR_VISX: return s->visx[idx]; case OR_VISZ: return s->visz[idx]; case OR_VISXX: return s->visxx[idx];
    case OR_VISZZ: return s->viszz[idx]; case OR_VISXZ: return s->visxz[idx];
    case OR_DTSTREAM: return s->dt_stream[idx]; case OR_DTVIS: return s->dt_vis[idx]; case OR_RATIO: return s->ratio[idx];
    case OR_COUNT: return s->count[idx]; case OR_NEMPTY: return s->nempty;
    case OR_RTR: return s->rtr; case OR_STR: return s->str; case OR_RTR0: return s->rtr0; case OR_STR0: return s->str0;
    case OR_RTR1: return s->rtr1; case OR_STR1: return s->str1; case OR_RTR2: return s->rtr2; case OR_STR2: return s->str2;
    case OR_TTYPE: return s->Ttype;
    case OR_MCOARSE: return s->Mc; case OR_ROWB: return s->RowB; case OR_IPIV: return s->ipiv; case OR_BCOARSE: return s->Bc;
    default: return NULL;
    }
}

double or_get_scalar(const or_state* s, int which)
{
    switch (which) {
    case ORS_DT: return s->dt; case ORS_TIME: return s->time; case ORS_RESID1: return s->residual_mag[1];
    case ORS_RESREF1: return s->res_ref[1]; case ORS_EIG_FINE: return s->eig_fine; case ORS_DELTA_T: return s->p.delta_T;
    case ORS_Q: return s->q; case ORS_N_ETA: return s->n_eta; case ORS_N_TEMP: return s->n_temp;
    case ORS_N_COMP1: return s->n_comp[1]; case ORS_STATUS: return s->status; case ORS_KL: return s->kl;
    case ORS_KU: return s->ku; case ORS_NGC: return s->ngc; case ORS_TSTEP: return s->tstep;
    case ORS_SMOOTH_VIS: return s->smooth_vis; case ORS_NVIS_STORE: return s->Nvis_store;
    case ORS_NN_STORE: return s->nn_store; case ORS_CYCLES_TOTAL: return (double)s->cycles_total;
    case ORS_VIS_ITERS: return (double)s->vis_iters; case ORS_C_ITERS: return (double)s->c_iters;
    case ORS_T_ITERS: return (double)s->t_iters;
    case ORS_Q_STEP0: case ORS_Q_STEP1: case ORS_Q_STEP2: case ORS_Q_STEP3: return s->q_step[which - ORS_Q_STEP0];
    case ORS_N_COMP_T1: case ORS_N_COMP_T2: case ORS_N_COMP_T3: case ORS_N_COMP_T4:
        return s->n_comp[1 + which - ORS_N_COMP_T1];
    case ORS_INTERP_CLAMPS: return (double)s->interp_clamps;
    case ORS_MAX_CYCLES: return s->p.max_cycles;
    default: return NAN;
    }
}

void or_set_scalar(or_state* s, int which, double v)
{
    switch (which) {
    case ORS_DT: s->dt = v; break; case ORS_TIME: s->time = v; break;
    case ORS_TSTEP: s->tstep = (int)v; break; case ORS_SMOOTH_VIS: s->smooth_vis = (int)v; break;
    case ORS_STATUS: s->status = (int)v; break;
    case ORS_NVIS_STORE: s->Nvis_store = (int)v; break; case ORS_NN_STORE: s->nn_store = (int)v; break;
    case ORS_MAX_CYCLES: s->p.max_cycles = (int)v; break;
    default: break;
    }
}

int or_ld(const or_state* s) { return s->ld; }
int or_nx_grid(const or_state* s, int L) { return s->nxg[L]; }
int or_nz_grid(const or_state* s, int L) { return s->nzg[L]; }
double or_residual_mag(const or_state* s, int p) { return s->residual_mag[p]; }
double or_res_ref(const or_state* s, int L) { return s->res_ref[L]; }

/* ------------------------------------------------------------------------ */
/* Small helpers: dot products in the Fortran source's summation order.      */
/* ------------------------------------------------------------------------ */

/* dot_product(D(-1:1), f(i-1:i+1)) along a stride */
static inline double dot3(const double* D /* D[-1..1] at D[1..3] of a 5-table */, const double* f, ptrdiff_t st)
{
    double acc = 0.0;
    acc += D[1] * f[-st];
    acc += D[2] * f[0];
    acc += D[3] * f[st];
    return acc;
}
static inline double dot5(const double* D, const double* f, ptrdiff_t st)
{
    double acc = 0.0;
    acc += D[0] * f[-2 * st];
    acc += D[1] * f[-st];
    acc += D[2] * f[0];
    acc += D[3] * f[st];
    acc += D[4] * f[2 * st];
    return acc;
}
/* sum(M(-1:1,-1:1)*f(i-1:i+1,k-1:k+1)) in column-major order */
static inline double sum9(const double M[5][5], const double* f, ptrdiff_t ld)
{
    double acc = 0.0;
    for (int dk = -1; dk <= 1; ++dk)
        for (int di = -1; di <= 1; ++di) acc += M[di + 2][dk + 2] * f[dk * ld + di];
    return acc;
}

/* ------------------------------------------------------------------------ */
/* Boundary conditions                                                       */
/* ------------------------------------------------------------------------ */

/* enforceBCs (energy.f90:455-475): insulating sides, isothermal bottom
 * (Tbot=0) or insulating (Tbot=1), isothermal top. */
void or_enforce_bcs(or_state* s, double* T0)
{
    const int nx = s->nx, nz = s->nz;
    for (int i = 1; i <= SPAN; ++i)
        for (int k = 0; k <= nz; ++k) {
            T0[IX(s, -i, k)] = T0[IX(s, i, k)];
            T0[IX(s, nx + i, k)] = T0[IX(s, nx - i, k)];
        }
    for (int k = 1; k <= SPAN; ++k)
        for (int i = -SPAN; i <= nx + SPAN; ++i) {
            if (s->p.Tbot == 0) T0[IX(s, i, -k)] = -(T0[IX(s, i, k)] - 1.0) + 1.0;
            else T0[IX(s, i, -k)] = T0[IX(s, i, k)];
            T0[IX(s, i, nz + k)] = -T0[IX(s, i, nz - k)];
        }
    for (int i = -SPAN; i <= nx + SPAN; ++i) {
        if (s->p.Tbot == 0) T0[IX(s, i, 0)] = 1.0;
        T0[IX(s, i, nz)] = 0.0;
    }
}

/* enforceBCs_stream (iterative_solver.f90:965-989). */
void or_enforce_bcs_stream(or_state* s, int level, double* S0)
{
    const int nx = s->nxg[level], nz = s->nzg[level];
    for (int k = -SPAN; k <= nz + SPAN; ++k) { S0[IX(s, 0, k)] = 0.0; S0[IX(s, nx, k)] = 0.0; }
    for (int i = -SPAN; i <= nx + SPAN; ++i) { S0[IX(s, i, 0)] = 0.0; S0[IX(s, i, nz)] = 0.0; }
    for (int i = 1; i <= SPAN; ++i)
        for (int k = 0; k <= nz; ++k) {
            S0[IX(s, -i, k)] = -S0[IX(s, i, k)];
            S0[IX(s, nx + i, k)] = -S0[IX(s, nx - i, k)];
        }
    for (int i = 1; i <= SPAN; ++i)
        for (int ii = -SPAN; ii <= nx + SPAN; ++ii) {
            if (s->p.Vbc == 1) {
                S0[IX(s, ii, -i)] = S0[IX(s, ii, i)];
                S0[IX(s, ii, nz + i)] = S0[IX(s, ii, nz - i)];
            } else {
                S0[IX(s, ii, -i)] = -S0[IX(s, ii, i)];
                S0[IX(s, ii, nz + i)] = -S0[IX(s, ii, nz - i)];
            }
        }
}

/* enforceBCs_comp (composition.f90:792-803).  The loop interleaves sides and
 * rows per ghost index, so the row ghosts at i=1 read column -2/nx+2 before
 * it is refreshed at i=2; this order is kept. */
void or_enforce_bcs_comp(or_state* s, double* C0)
{
    const int nx = s->nx, nz = s->nz;
    for (int i = 1; i <= SPAN; ++i) {
        for (int k = 0; k <= nz; ++k) {
            C0[IX(s, -i, k)] = C0[IX(s, i, k)];
            C0[IX(s, nx + i, k)] = C0[IX(s, nx - i, k)];
        }
        for (int ii = -SPAN; ii <= nx + SPAN; ++ii) {
            C0[IX(s, ii, -i)] = C0[IX(s, ii, i)];
            C0[IX(s, ii, nz + i)] = C0[IX(s, ii, nz - i)];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Streamfunction operator and smoother                                      */
/* ------------------------------------------------------------------------ */

/* stream_space with build=0, scaling=0 (iterative_solver.f90:456-561):
 * returns RHS - L(S0) at interior point (i,k) of `level`. */
double or_stream_space(or_state* s, int level, int i, int k, const double* S0)
{
    const ptrdiff_t ld = s->ld;
    const double* P = S0 + IX(s, i, k);
    const int i1 = i * (1 << (level - 1)), k1 = k * (1 << (level - 1));
    double Sr = dot3(s->D1, P, 1) / s->drg[level];
    double Ss = dot3(s->D1, P, ld) / s->dsg[level];
    double Srr = dot3(s->D2, P, 1) / s->dr2g[level];
    double Sss = dot3(s->D2, P, ld) / s->ds2g[level];
    double Srrr = dot5(s->D3, P, 1) / s->dr3g[level];
    double Ssss = dot5(s->D3, P, ld) / s->ds3g[level];
    double Srrrr = dot5(s->D4, P, 1) / s->dr4g[level];
    double Sssss = dot5(s->D4, P, ld) / s->ds4g[level];
    double Srs = sum9(s->Drs, P, ld) / s->drdsg[level];
    double Srss = sum9(s->Drss, P, ld) / s->drds2g[level];
    double Srrs = sum9(s->Drrs, P, ld) / s->dr2dsg[level];
    double Srrss = sum9(s->Drrss, P, ld) / s->dr2ds2g[level];
    const double xr = s->xr[i1 + SPAN], xrr = s->xrr[i1 + SPAN], xrrr = s->xrrr[i1 + SPAN], xrrrr = s->xrrrr[i1 + SPAN];
    const double zs = s->zs[k1 + SPAN], zss = s->zss[k1 + SPAN], zsss = s->zsss[k1 + SPAN], zssss = s->zssss[k1 + SPAN];
    double Sx = Sr / xr;
    double Sz = Ss / zs;
    double Sxx = (Srr - Sx * xrr) / (xr * xr);
    double Szz = (Sss - Sz * zss) / (zs * zs);
    double Sxxx = (Srrr - 3.0 * Sxx * xr * xrr - Sx * xrrr) / pow(xr, 3.0);
    double Szzz = (Ssss - 3.0 * Szz * zs * zss - Sz * zsss) / pow(zs, 3.0);
    double Sxxxx = (Srrrr - 6.0 * Sxxx * xrr * (xr * xr) - (3.0 * (xrr * xrr) + 4.0 * xr * xrrr) * Sxx - Sx * xrrrr) / pow(xr, 4.0);
    double Szzzz = (Sssss - 6.0 * Szzz * zss * (zs * zs) - (3.0 * (zss * zss) + 4.0 * zs * zsss) * Szz - Sz * zssss) / pow(zs, 4.0);
    double Sxz = Srs / xr / zs;
    double Sxxz = (Srrs - Sxz * zs * xrr) / zs / (xr * xr);
    double Sxzz = (Srss - Sxz * xr * zss) / xr / (zs * zs);
    double xz = xr * zs;
    double Sxxzz = (Srrss - Sxzz * xrr * (zs * zs) - Sxxz * zss * (xr * xr) - Sxz * xrr * zss) / (xz * xz);
    const size_t c = IX(s, i, k);
    double t1, t2, t3, t4, t5, t6;
    const double vis = s->vis[level][c];
    if (s->p.ivis == 0) {
        t1 = 0.0;
        t2 = vis * (Sxxxx + Szzzz);
        t3 = 2.0 * vis * Sxxzz;
        t4 = 0.0; t5 = 0.0; t6 = 0.0;
    } else {
        t1 = (s->visxx[level][c] - s->viszz[level][c]) * (Sxx - Szz);
        t2 = vis * (Sxxxx + Szzzz);
        t3 = 2.0 * vis * Sxxzz;
        t4 = 4.0 * s->visxz[level][c] * Sxz;
        t5 = 2.0 * s->visx[level][c] * (Sxxx + Sxzz);
        t6 = 2.0 * s->visz[level][c] * (Szzz + Sxxz);
    }
    double ss = t1 + t2 + t3 + t4 + t5 + t6 - s->RHS[level][c];
    return -ss;
}

/* One Euler sweep n (1..Niter) of iterate_stream with iterate_order=1
 * (iterative_solver.f90:871-912): Chebyshev factor c from the ratio, c=1 on
 * the upward leg, Jacobi update from a full copy S0, then enforceBCs_stream. */
void or_sweep(or_state* s, int level, int n, int Niter, int dir)
{
    const int nxl = s->nxg[level], nzl = s->nzg[level];
    double ntemp = (double)(Niter - n + 1);
    const double la = -1.0;
    double temp = cos((2.0 * ntemp - 1.0) * PII / (double)(2 * Niter));
    double* S0 = s->tmp;
    memcpy(S0, s->error[level], s->plane * sizeof(double));
    double* E = s->error[level];
    const double* ratio = s->ratio[level];
    const double* dts = s->dt_stream[level];
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nzl - 1; ++k)
        for (int i = 1; i <= nxl - 1; ++i) {
            size_t c0 = IX(s, i, k);
            double cc;
            if (dir == 1) {
                cc = 1.0;
            } else {
                double lb;
                double r = ratio[c0];
                if (r > 1e99) lb = -2.0;
                else lb = -2.0 * (r * r) / (1.0 + r * r);
                cc = 2.0 / (-lb - la + (lb - la) * temp);
            }
            double t = or_stream_space(s, level, i, k, S0);
            E[c0] = S0[c0] + cc * dts[c0] * t;
        }
    or_enforce_bcs_stream(s, level, E);
}

/* compute_residual (iterative_solver.f90:258-279).  The L1 norm is summed in
 * the Fortran column-major order. */
void or_compute_residual(or_state* s, int level, int p)
{
    const int nxl = s->nxg[level], nzl = s->nzg[level];
    double* R = s->residual;
    const double* E = s->error[level];
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nzl - 1; ++k)
        for (int i = 1; i <= nxl - 1; ++i) R[IX(s, i, k)] = or_stream_space(s, level, i, k, E);
    double acc = 0.0;
    for (int k = 1; k <= nzl - 1; ++k)
        for (int i = 1; i <= nxl - 1; ++i) acc += fabs(R[IX(s, i, k)]);
    s->residual_mag[p] = acc / (double)(nxl - 1) / (double)(nzl - 1);
    if (level == 1 && s->residual_mag[p] > 1e10 * s->res_ref[level]) s->status = TC_ERR_RESID_BLOWUP;
}

/* iterate_stream (iterative_solver.f90:860-963), iterate_order=1. */
void or_iterate_stream(or_state* s, int level, int Niter, int dir, int p)
{
    for (int n = 1; n <= Niter; ++n) or_sweep(s, level, n, Niter, dir);
    or_compute_residual(s, level, p);
}

/* compute_reference_residual (iterative_solver.f90:281-288). */
void or_compute_reference_residual(or_state* s, int level)
{
    const int nxl = s->nxg[level], nzl = s->nzg[level];
    double acc = 0.0;
    for (int k = 1; k <= nzl - 1; ++k)
        for (int i = 1; i <= nxl - 1; ++i) acc += fabs(s->RHS[level][IX(s, i, k)]);
    s->res_ref[level] = acc / (double)(nxl - 2) / (double)(nzl - 2);
}

/* restrict / restrict_coefficients (iterative_solver.f90:596-626, 628-664):
 * full weighting whose corner sum repeats (ii-1,kk+1) and omits (ii+1,kk-1). */
void or_restrict(or_state* s, int level, const double* in, double* out)
{
    const int nxl = s->nxg[level], nzl = s->nzg[level];
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nzl - 1; ++k)
        for (int i = 1; i <= nxl - 1; ++i) {
            int ii = 2 * i, kk = 2 * k;
            double corners = (in[IX(s, ii - 1, kk - 1)] + in[IX(s, ii - 1, kk + 1)] + in[IX(s, ii + 1, kk + 1)] +
                              in[IX(s, ii - 1, kk + 1)]) / 16.0;
            double sides = (in[IX(s, ii - 1, kk)] + in[IX(s, ii, kk - 1)] + in[IX(s, ii + 1, kk)] + in[IX(s, ii, kk + 1)]) / 8.0;
            double centre = in[IX(s, ii, kk)] / 4.0;
            out[IX(s, i, k)] = corners + sides + centre;
        }
}

/* compute_RHS (iterative_solver.f90:563-594). */
void or_compute_RHS(or_state* s, int level)
{
    if (level == 1) {
        const int nxl = s->nxg[1], nzl = s->nzg[1];
#pragma omp parallel for schedule(static)
        for (int k = 1; k <= nzl - 1; ++k)
            for (int i = 1; i <= nxl - 1; ++i) {
                size_t c = IX(s, i, k);
                double rhs = 0.0;
                if (s->p.RaT != 0.0) {
                    double Tr = dot3(s->D1, s->Tbuoy + c, 1) / s->dr;
                    double Tx = Tr / s->xr[i + SPAN];
                    rhs = s->p.RaT * Tx;
                }
                if (s->p.comp == 1)
                    for (int j = 1; j <= s->ntype; ++j) {
                        double Cr = dot3(s->D1, s->Cbuoy[j] + c, 1) / s->dr;
                        double Cx = Cr / s->xr[i + SPAN];
                        rhs = rhs - s->p.RaC[j - 1] * Cx;
                    }
                s->RHS[1][c] = rhs;
            }
    } else {
        or_restrict(s, level, s->residual, s->RHS[level]);
    }
}

/* prolong (iterative_solver.f90:175-233): bilinear interpolation of the
 * coarse correction added into the finer one, then enforceBCs_stream. */
void or_prolong(or_state* s, int lc)
{
    const int lf = lc - 1;
    double* Ef = s->error[lf];
    const double* Ec = s->error[lc];
    const int nxc = s->nxg[lc], nzc = s->nzg[lc], nxf = s->nxg[lf], nzf = s->nzg[lf];
    for (int i = 1; i <= nxc - 1; ++i)
        for (int k = 1; k <= nzc - 1; ++k) Ef[IX(s, 2 * i, 2 * k)] = Ef[IX(s, 2 * i, 2 * k)] + Ec[IX(s, i, k)];
    for (int i = 2; i <= nxf - 2; i += 2)
        for (int k = 1; k <= nzf - 1; k += 2)
            Ef[IX(s, i, k)] = Ef[IX(s, i, k)] + (Ec[IX(s, i / 2, (k - 1) / 2)] + Ec[IX(s, i / 2, (k + 1) / 2)]) / 2.0;
    for (int k = 2; k <= nzf - 2; k += 2)
        for (int i = 1; i <= nxf - 1; i += 2)
            Ef[IX(s, i, k)] = Ef[IX(s, i, k)] + (Ec[IX(s, (i - 1) / 2, k / 2)] + Ec[IX(s, (i + 1) / 2, k / 2)]) / 2.0;
    for (int i = 1; i <= nxf - 1; i += 2)
        for (int k = 1; k <= nzf - 1; k += 2)
            Ef[IX(s, i, k)] = Ef[IX(s, i, k)] + (Ec[IX(s, (i - 1) / 2, (k - 1) / 2)] + Ec[IX(s, (i + 1) / 2, (k - 1) / 2)] +
                                                 Ec[IX(s, (i + 1) / 2, (k + 1) / 2)] + Ec[IX(s, (i - 1) / 2, (k + 1) / 2)]) / 4.0;
    or_enforce_bcs_stream(s, lf, Ef);
}

/* ------------------------------------------------------------------------ */
/* Viscosity hierarchy                                                       */
/* ------------------------------------------------------------------------ */

/* visf (momentum.f90:1-30) for ivis = 0, 1, 2. */
static double visf(const or_state* s, int i, int k)
{
    const size_t c = IX(s, i, k);
    const double zgk = s->zg[k + 3];
    if (s->p.ivis == 1) {
        return exp(-s->Tvis[c] * log(s->p.visT) + (1.0 - zgk) * log(s->p.visP));
    } else if (s->p.ivis == 2 && s->p.comp == 1) {
        double dp = 0.0;
        for (int t = 1; t <= s->ntype; ++t) dp += s->Cvis[t][c] * log(s->p.visC[t - 1]);
        if (s->p.RaT != 0.0) return exp(-s->Tvis[c] * log(s->p.visT) - dp + (1.0 - zgk) * log(s->p.visP));
        return exp(-dp + (1.0 - zgk) * log(s->p.visP));
    }
    return 1.0;
}

/* visf over the whole ghosted plane of level 1 (the first loop of
 * compute_viscosity, iterative_solver.f90:340-345), without the filter that
 * follows it there. */
void or_visf_plane(or_state* s, double* out)
{
    for (int k = -SPAN; k <= s->nz + SPAN; ++k)
        for (int i = -SPAN; i <= s->nx + SPAN; ++i) out[IX(s, i, k)] = visf(s, i, k);
}

/* viscosity_gradients (iterative_solver.f90:290-329). */
void or_viscosity_gradients(or_state* s, int level)
{
    const int nxl = s->nxg[level], nzl = s->nzg[level];
    const ptrdiff_t ld = s->ld;
    double *vx = s->visx[level], *vz = s->visz[level], *vxx = s->visxx[level], *vzz = s->viszz[level], *vxz = s->visxz[level];
    const double* V = s->vis[level];
    if (s->p.ivis == 0) {
        for (size_t c = 0; c < s->plane; ++c) vx[c] = vz[c] = vxx[c] = vzz[c] = vxz[c] = 0.0;
        return;
    }
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nzl - 1; ++k)
        for (int i = 1; i <= nxl - 1; ++i) {
            int ii = i * (1 << (level - 1)), kk = k * (1 << (level - 1));
            size_t c = IX(s, i, k);
            double vis_r = dot3(s->D1, V + c, 1) / s->drg[level];
            double vis_rr = dot3(s->D2, V + c, 1) / s->dr2g[level];
            double vis_s = dot3(s->D1, V + c, ld) / s->dsg[level];
            double vis_ss = dot3(s->D2, V + c, ld) / s->ds2g[level];
            double vis_rs = sum9(s->Drs, V + c, ld) / s->drdsg[level];
            double xr = s->xr[ii + SPAN], zs = s->zs[kk + SPAN];
            vx[c] = vis_r / xr;
            vz[c] = vis_s / zs;
            vxx[c] = (vis_rr - vx[c] * s->xrr[ii + SPAN]) / (xr * xr);
            vzz[c] = (vis_ss - vz[c] * s->zss[kk + SPAN]) / (zs * zs);
            vxz[c] = vis_rs / xr / zs;
        }
}

static void set_eig_scalars(or_state* s, int level)
{
    /* (iterative_solver.f90:674-685, 709-720) */
    s->Ts_max = s->p.Tr_base / s->dsg[level];
    s->Tss_max = -s->p.Trr_base / s->ds2g[level];
    s->Tsss_max = s->p.Trrr_base / s->ds3g[level];
    s->Tssss_max = s->p.Trrrr_base / s->ds4g[level];
    s->Tr_max = s->p.Tr_base / s->drg[level];
    s->Trr_max = -s->p.Trr_base / s->dr2g[level];
    s->Trrr_max = s->p.Trrr_base / s->dr3g[level];
    s->Trrrr_max = s->p.Trrrr_base / s->dr4g[level];
    s->Trs_max = -s->p.Trs_base / s->drdsg[level];
    s->Trrss_max = s->p.Trrss_base / s->dr2ds2g[level];
    s->Trss_max = s->p.Trrs_base / s->drds2g[level];
    s->Trrs_max = s->p.Trrs_base / s->dr2dsg[level];
}

/* eigI (iterative_solver.f90:750-803). */
static double eigI(const or_state* s, int level, int i, int k)
{
    const int i1 = i * (1 << (level - 1)), k1 = k * (1 << (level - 1));
    const double xr = s->xr[i1 + SPAN], xrr = s->xrr[i1 + SPAN], xrrr = s->xrrr[i1 + SPAN], xrrrr = s->xrrrr[i1 + SPAN];
    const double zs = s->zs[k1 + SPAN], zss = s->zss[k1 + SPAN], zsss = s->zsss[k1 + SPAN], zssss = s->zssss[k1 + SPAN];
    double Sr = s->Tr_max, Ss = s->Ts_max, Srr = 0.0, Sss = 0.0, Srrr = s->Trrr_max, Ssss = s->Tsss_max;
    double Srrrr = 0.0, Sssss = 0.0, Srs = 0.0, Srrs = s->Trrs_max, Srss = s->Trss_max, Srrss = 0.0;
    double Sx = Sr / xr, Sz = Ss / zs;
    double Sxx = (Srr + fabs(Sx * xrr)) / (xr * xr);
    double Szz = (Sss + fabs(Sz * zss)) / (zs * zs);
    double Sxxx = (Srrr + fabs(3.0 * Sxx * xr * xrr) + fabs(Sx * xrrr)) / fabs(pow(xr, 3.0));
    double Szzz = (Ssss + fabs(3.0 * Szz * zs * zss) + fabs(Sz * zsss)) / fabs(pow(zs, 3.0));
    double Sxxxx = (Srrrr + fabs(6.0 * Sxxx * xrr * (xr * xr)) + (fabs(3.0 * (xrr * xrr) + 4.0 * xr * xrrr) * fabs(Sxx)) +
                    fabs(Sx * xrrrr)) / pow(xr, 4.0);
    double Szzzz = (Sssss + fabs(6.0 * Szzz * zss * (zs * zs)) + (fabs(3.0 * (zss * zss) + 4.0 * zs * zsss) * fabs(Szz)) +
                    fabs(Sz * zssss)) / pow(zs, 4.0);
    double Sxz = Srs / xr / zs;
    double Sxxz = (Srrs + fabs(Sxz * zs * xrr)) / fabs(zs) / (xr * xr);
    double Sxzz = (Srss + fabs(Sxz * xr * zss)) / fabs(xr) / (zs * zs);
    double xz = xr * zs;
    double Sxxzz = (Srrss + fabs(Sxzz * xrr) * (zs * zs) + fabs(Sxxz * zss) * (xr * xr) + fabs(Sxz * xrr * zss)) / (xz * xz);
    const size_t c = IX(s, i, k);
    double t1 = (s->visxx[level][c] + s->viszz[level][c]) * (Sxx + Szz);
    double t2 = s->vis[level][c] * (Sxxxx + Szzzz);
    double t3 = 2.0 * s->vis[level][c] * Sxxzz;
    double t4 = 4.0 * s->visxz[level][c] * Sxz;
    double t5 = 2.0 * s->visx[level][c] * (Sxxx + Sxzz);
    double t6 = 2.0 * s->visz[level][c] * (Szzz + Sxxz);
    return fabs(t1) + fabs(t2) + fabs(t3) + fabs(t4) + fabs(t5) + fabs(t6);
}

/* eigR (iterative_solver.f90:806-857). */
static double eigR(const or_state* s, int level, int i, int k)
{
    const int i1 = i * (1 << (level - 1)), k1 = k * (1 << (level - 1));
    const double xr = s->xr[i1 + SPAN], xrr = s->xrr[i1 + SPAN], xrrr = s->xrrr[i1 + SPAN], xrrrr = s->xrrrr[i1 + SPAN];
    const double zs = s->zs[k1 + SPAN], zss = s->zss[k1 + SPAN], zsss = s->zsss[k1 + SPAN], zssss = s->zssss[k1 + SPAN];
    double Sr = 0.0, Ss = 0.0, Srr = s->Trr_max, Sss = s->Tss_max, Srrr = 0.0, Ssss = 0.0;
    double Srrrr = s->Trrrr_max, Sssss = s->Tssss_max, Srs = s->Trs_max, Srrs = 0.0, Srss = 0.0, Srrss = s->Trrss_max;
    double Sx = Sr / xr, Sz = Ss / zs;
    double Sxx = (Srr - Sx * xrr) / (xr * xr);
    double Szz = (Sss - Sz * zss) / (zs * zs);
    double Sxxx = (Srrr - 3.0 * Sxx * xr * xrr - Sx * xrrr) / pow(xr, 3.0);
    double Szzz = (Ssss - 3.0 * Szz * zs * zss - Sz * zsss) / pow(zs, 3.0);
    double Sxxxx = (Srrrr - 6.0 * Sxxx * xrr * (xr * xr) - (3.0 * (xrr * xrr) + 4.0 * xr * xrrr) * Sxx - Sx * xrrrr) / pow(xr, 4.0);
    double Szzzz = (Sssss - 6.0 * Szzz * zss * (zs * zs) - (3.0 * (zss * zss) + 4.0 * zs * zsss) * Szz - Sz * zssss) / pow(zs, 4.0);
    double Sxz = Srs / xr / zs;
    double Sxxz = (Srrs - Sxz * zs * xrr) / zs / (xr * xr);
    double Sxzz = (Srss - Sxz * xr * zss) / xr / (zs * zs);
    double xz = xr * zs;
    double Sxxzz = (Srrss - Sxzz * xrr * (zs * zs) - Sxxz * zss * (xr * xr) - Sxz * xrr * zss) / (xz * xz);
    const size_t c = IX(s, i, k);
    double t1 = (s->visxx[level][c] - s->viszz[level][c]) * (Sxx - Szz);
    double t2 = s->vis[level][c] * (Sxxxx + Szzzz);
    double t3 = 2.0 * s->vis[level][c] * Sxxzz;
    double t4 = 4.0 * s->visxz[level][c] * Sxz;
    double t5 = 2.0 * s->visx[level][c] * (Sxxx + Sxzz);
    double t6 = 2.0 * s->visz[level][c] * (Szzz + Sxxz);
    return -(t1 + t2 + t3 + t4 + t5 + t6);
}

/* test_eigenvalues (iterative_solver.f90:666-698): resets the whole ratio
 * array of every level to 1e99, then fills `level`. */
void or_test_eigenvalues(or_state* s, int level)
{
    set_eig_scalars(s, level);
    for (int L = 1; L <= s->grids - 1; ++L)
        for (size_t c = 0; c < s->plane; ++c) s->ratio[L][c] = 1e99;
    if (s->p.ivis == 0) return;
    const int nxl = s->nxg[level], nzl = s->nzg[level];
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nzl - 1; ++k)
        for (int i = 1; i <= nxl - 1; ++i) {
            double eI = eigI(s, level, i, k);
            double eR = eigR(s, level, i, k);
            s->ratio[level][IX(s, i, k)] = fabs(eR / eI);
        }
}

/* iterate_local_time_steps (iterative_solver.f90:700-748), iterate_order=1. */
void or_iterate_local_time_steps(or_state* s, int level)
{
    set_eig_scalars(s, level);
    const int nxl = s->nxg[level], nzl = s->nzg[level];
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nzl - 1; ++k)
        for (int i = 1; i <= nxl - 1; ++i) {
            double eI = eigI(s, level, i, k);
            double eR = eigR(s, level, i, k);
            s->dt_stream[level][IX(s, i, k)] = -s->p.courant_stream * 2.0 * eR / (eR * eR + eI * eI);
        }
}

/* stability_vis_smoother (smoother.f90:465-508): dt_vis of `level`. */
void or_stability_vis_smoother(or_state* s, int level)
{
    double Ts_max = s->p.Tr_base / s->dsg[level];
    double Tss_max = -s->p.Trr_base / s->ds2g[level];
    double Tr_max = s->p.Tr_base / s->drg[level];
    double Trr_max = -s->p.Trr_base / s->dr2g[level];
    for (int i = 0; i <= s->nxg[level]; ++i)
        for (int k = 1; k <= s->nzg[level] - 1; ++k) {
            int i1 = i * (1 << (level - 1)), k1 = k * (1 << (level - 1));
            double xr = s->xr[i1 + SPAN], zs = s->zs[k1 + SPAN];
            double eR = fabs(Trr_max / (xr * xr)) + fabs(Tss_max / (zs * zs));
            double eI = fabs(s->xrr[i1 + SPAN] / pow(xr, 3.0)) * Tr_max + fabs(s->zss[k1 + SPAN] / pow(zs, 3.0)) * Ts_max;
            s->dt_vis[level][IX(s, i, k)] = 2.0 * eR / (eR * eR + eI * eI);
        }
}

/* stability_RK4_smoother (smoother.f90:423-444): dts_array. */
void or_stability_RK4_smoother(or_state* s)
{
    double Ts_max = s->p.Tr_base / s->ds;
    double Tss_max = -s->p.Trr_base / s->ds2;
    double Tr_max = s->p.Tr_base / s->dr;
    double Trr_max = -s->p.Trr_base / s->dr2;
    for (int i = 0; i <= s->nx; ++i)
        for (int k = 0; k <= s->nz; ++k) {
            double xr = s->xr[i + SPAN], zs = s->zs[k + SPAN];
            double eR = fabs(Trr_max / (xr * xr)) + fabs(Tss_max / (zs * zs));
            double eI = fabs(s->xrr[i + SPAN] / pow(xr, 3.0)) * Tr_max + fabs(s->zss[k + SPAN] / pow(zs, 3.0)) * Ts_max;
            s->dts[IX(s, i, k)] = 2.0 * eR / (eR * eR + eI * eI);
        }
}

/* vis_space (smoother.f90:141-163). */
static double vis_space(const or_state* s, int i, int k, const double* T0, int level)
{
    const int i1 = i * (1 << (level - 1)), k1 = k * (1 << (level - 1));
    const double* P = T0 + IX(s, i, k);
    double Tr = dot3(s->D1, P, 1) / s->drg[level];
    double Ts = dot3(s->D1, P, s->ld) / s->dsg[level];
    double Trr = dot3(s->D2, P, 1) / s->dr2g[level];
    double Tss = dot3(s->D2, P, s->ld) / s->ds2g[level];
    double xr = s->xr[i1 + SPAN], zs = s->zs[k1 + SPAN];
    double Tx = Tr / xr, Tz = Ts / zs;
    double Txx = (Trr - Tx * s->xrr[i1 + SPAN]) / (xr * xr);
    double Tzz = (Tss - Tz * s->zss[k1 + SPAN]) / (zs * zs);
    return Txx + Tzz;
}

/* viscosityBCs_initialize / viscosityBCs (iterative_solver.f90:379-453). */
static void visc_bcs(or_state* s, int level, int initialize)
{
    double* V = s->vis[level];
    const int nxl = s->nxg[level], nzl = s->nzg[level];
    if (level > 1) {
        const double* Vf = s->vis[level - 1];
        if (initialize)
            for (int k = 1; k <= nzl - 1; ++k) {
                V[IX(s, 0, k)] = Vf[IX(s, 0, 2 * k)];
                V[IX(s, nxl, k)] = Vf[IX(s, s->nxg[level - 1], 2 * k)];
            }
        for (int i = 0; i <= nxl; ++i) {
            V[IX(s, i, 0)] = Vf[IX(s, 2 * i, 0)];
            V[IX(s, i, nzl)] = Vf[IX(s, 2 * i, s->nzg[level - 1])];
        }
    }
    for (int i = 1; i <= SPAN; ++i)
        for (int k = 0; k <= nzl; ++k) {
            V[IX(s, -i, k)] = V[IX(s, i, k)];
            V[IX(s, nxl + i, k)] = V[IX(s, nxl - i, k)];
        }
    const double visT2 = s->p.visT * s->p.visT;
    for (int k = 1; k <= SPAN; ++k)
        for (int i = -SPAN; i <= nxl + SPAN; ++i) {
            if (s->p.RaT != 0.0) {                V[IX(s, i, -k)] = 1.0 / V[IX(s, i, k)] / visT2;
                V[IX(s, i, nzl + k)] = 1.0 / V[IX(s, i, nzl - k)];
            } else {
                V[IX(s, i, -k)] = V[IX(s, i, k)];
                V[IX(s, i, nzl + k)] = V[IX(s, i, nzl - k)];
            }
        }
}
void or_viscosity_bcs(or_state* s, int level) { visc_bcs(s, level, 0); }
void or_viscosity_bcs_initialize(or_state* s, int level) { visc_bcs(s, level, 1); }

static double ratio_min(const or_state* s, int level)
{
    double m = INFINITY;
    for (int k = 1; k <= s->nzg[level] - 1; ++k)
        for (int i = 1; i <= s->nxg[level] - 1; ++i) {
            double r = s->ratio[level][IX(s, i, k)];
            if (r < m) m = r;
        }
    return m;
}

/* smoother_vis (smoother.f90:165-263): Richardson/Chebyshev filtering of the
 * viscosity of `level` with the data-dependent exits on the minimum
 * eigenvalue ratio.  Nvis_store/nn_store are module variables and carry
 * over between calls exactly as in the reference. */
void or_smoother_vis(or_state* s, int level, int interpolate)
{
    int option, Niter;
    double lb, la;
    if (level == 1 && interpolate == 0) { option = 1; Niter = 6; lb = -5.0; la = -2.0; }
    else { option = 0; Niter = s->p.vis_iter; lb = -2.0; la = -1.0; }
    if (interpolate == 0 || level == 1) { s->Nvis_store = 0; s->nn_store = 1; }
    int kmin, kmax;
    if (s->p.RaT == 0.0) { kmin = 0; kmax = s->nzg[level]; }
    else { kmin = 1; kmax = s->nzg[level] - 1; }
    int n = 0, trigger = 0;
    double* T0 = s->tmp2;
    for (;;) {
        for (int nn = s->nn_store; nn <= Niter; ++nn) {
            or_viscosity_gradients(s, level);
            or_test_eigenvalues(s, level);
            double rmin = ratio_min(s, level);
            s->vis_iters++;
            if (option == 1) {
                if ((s->Nvis_store >= 1 && rmin > s->p.eig_ratio_min) || s->p.ivis == 0 || n > 20) {
                    s->eig_fine = rmin;
                    s->n_eta = n;
                    trigger = 1;
                    break;
                }
            } else {
                if (interpolate == 0) {
                    if (rmin > s->p.eig_ratio || s->p.ivis == 0) { trigger = 1; s->nn_store = nn; break; }
                } else {
                    if (s->Nvis_store >= 1 || s->p.ivis == 0) { trigger = 1; break; }
                }
            }
            int ntemp = Niter - nn + 1;
            double c = 2.0 / (-lb - la + (lb - la) * cos((double)(2 * ntemp - 1) * PII / (double)(2 * Niter)));
            memcpy(T0, s->vis[level], s->plane * sizeof(double));
            double* V = s->vis[level];
            const double* dtv = s->dt_vis[level];
#pragma omp parallel for schedule(static)
            for (int k = kmin; k <= kmax; ++k)
                for (int i = 0; i <= s->nxg[level]; ++i) {
                    size_t cc = IX(s, i, k);
                    double temp = vis_space(s, i, k, T0, level);
                    double dts = dtv[cc] * 0.99;
                    V[cc] = T0[cc] + c * dts * temp;
                }
            or_viscosity_bcs(s, level);
        }
        if (trigger == 1) break;
        s->Nvis_store++;
        s->nn_store = 1;
        if (option == 1) { s->nn_store = 1; Niter = 6; lb = -2.0; la = -1.0; }
        n++;
    }
}

/* compute_viscosity (iterative_solver.f90:331-377). */
void or_compute_viscosity(or_state* s, int level)
{
    if (level == 1) {
        double* V = s->vis[1];
#pragma omp parallel for schedule(static)
        for (int k = -SPAN; k <= s->nzg[1] + SPAN; ++k)
            for (int i = -SPAN; i <= s->nxg[1] + SPAN; ++i) V[IX(s, i, k)] = visf(s, i, k);
        or_smoother_vis(s, 1, 0);
    } else {
        const int L1 = level - 1;
        memcpy(s->visf[L1], s->vis[L1], s->plane * sizeof(double));
        memcpy(s->visxf[L1], s->visx[L1], s->plane * sizeof(double));
        memcpy(s->viszf[L1], s->visz[L1], s->plane * sizeof(double));
        memcpy(s->visxxf[L1], s->visxx[L1], s->plane * sizeof(double));
        memcpy(s->viszzf[L1], s->viszz[L1], s->plane * sizeof(double));
        memcpy(s->visxzf[L1], s->visxz[L1], s->plane * sizeof(double));
        or_smoother_vis(s, L1, 1);
        or_restrict(s, level, s->vis[L1], s->vis[level]);
        or_viscosity_bcs_initialize(s, level);
        if (level < s->grids) or_smoother_vis(s, level, 0);
        else or_viscosity_gradients(s, level);
        memcpy(s->vis[L1], s->visf[L1], s->plane * sizeof(double));
        memcpy(s->visx[L1], s->visxf[L1], s->plane * sizeof(double));
        memcpy(s->visz[L1], s->viszf[L1], s->plane * sizeof(double));
        memcpy(s->visxx[L1], s->visxxf[L1], s->plane * sizeof(double));
        memcpy(s->viszz[L1], s->viszzf[L1], s->plane * sizeof(double));
        memcpy(s->visxz[L1], s->visxzf[L1], s->plane * sizeof(double));
    }
}

/* ------------------------------------------------------------------------ */
/* Coarsest grid: banded matrix and LAPACK-semantics LU                      */
/* ------------------------------------------------------------------------ */

/* stream_space with build=1 (iterative_solver.f90:487-519, 521-560):
 * coefficient of S(iq,kq) (times Sval) in row (i,k) on the coarsest level. */
static double stream_coeff(const or_state* s, int level, int i, int k, int iq, int kq, double Sval)
{
    int di = iq - i, dk = kq - k;
    double Sr = 0, Srr = 0, Srrr = 0, Srrrr = 0, Ss = 0, Sss = 0, Ssss = 0, Sssss = 0;
    if (k == kq) {
        Sr = s->D1[di + 2] * Sval / s->drg[level];
        Srr = s->D2[di + 2] * Sval / s->dr2g[level];
        Srrr = s->D3[di + 2] * Sval / s->dr3g[level];
        Srrrr = s->D4[di + 2] * Sval / s->dr4g[level];
    }
    if (i == iq) {
        Ss = s->D1[dk + 2] * Sval / s->dsg[level];
        Sss = s->D2[dk + 2] * Sval / s->ds2g[level];
        Ssss = s->D3[dk + 2] * Sval / s->ds3g[level];
        Sssss = s->D4[dk + 2] * Sval / s->ds4g[level];
    }
    double Srs = s->Drs[di + 2][dk + 2] * Sval / s->drdsg[level];
    double Srss = s->Drss[di + 2][dk + 2] * Sval / s->drds2g[level];
    double Srrs = s->Drrs[di + 2][dk + 2] * Sval / s->dr2dsg[level];
    double Srrss = s->Drrss[di + 2][dk + 2] * Sval / s->dr2ds2g[level];
    const int i1 = i * (1 << (level - 1)), k1 = k * (1 << (level - 1));
    const double xr = s->xr[i1 + SPAN], xrr = s->xrr[i1 + SPAN], xrrr = s->xrrr[i1 + SPAN], xrrrr = s->xrrrr[i1 + SPAN];
    const double zs = s->zs[k1 + SPAN], zss = s->zss[k1 + SPAN], zsss = s->zsss[k1 + SPAN], zssss = s->zssss[k1 + SPAN];
    double Sx = Sr / xr, Sz = Ss / zs;
    double Sxx = (Srr - Sx * xrr) / (xr * xr);
    double Szz = (Sss - Sz * zss) / (zs * zs);
    double Sxxx = (Srrr - 3.0 * Sxx * xr * xrr - Sx * xrrr) / pow(xr, 3.0);
    double Szzz = (Ssss - 3.0 * Szz * zs * zss - Sz * zsss) / pow(zs, 3.0);
    double Sxxxx = (Srrrr - 6.0 * Sxxx * xrr * (xr * xr) - (3.0 * (xrr * xrr) + 4.0 * xr * xrrr) * Sxx - Sx * xrrrr) / pow(xr, 4.0);
    double Szzzz = (Sssss - 6.0 * Szzz * zss * (zs * zs) - (3.0 * (zss * zss) + 4.0 * zs * zsss) * Szz - Sz * zssss) / pow(zs, 4.0);
    double Sxz = Srs / xr / zs;
    double Sxxz = (Srrs - Sxz * zs * xrr) / zs / (xr * xr);
    double Sxzz = (Srss - Sxz * xr * zss) / xr / (zs * zs);
    double xz = xr * zs;
    double Sxxzz = (Srrss - Sxzz * xrr * (zs * zs) - Sxxz * zss * (xr * xr) - Sxz * xrr * zss) / (xz * xz);
    const size_t c = IX(s, i, k);
    const double vis = s->vis[level][c];
    double t1, t2, t3, t4, t5, t6;
    if (s->p.ivis == 0) {
        t1 = 0.0; t2 = vis * (Sxxxx + Szzzz); t3 = 2.0 * vis * Sxxzz; t4 = 0.0; t5 = 0.0; t6 = 0.0;
    } else {
        t1 = (s->visxx[level][c] - s->viszz[level][c]) * (Sxx - Szz);
        t2 = vis * (Sxxxx + Szzzz);
        t3 = 2.0 * vis * Sxxzz;
        t4 = 4.0 * s->visxz[level][c] * Sxz;
        t5 = 2.0 * s->visx[level][c] * (Sxxx + Sxzz);
        t6 = 2.0 * s->visz[level][c] * (Szzz + Sxxz);
    }
    return t1 + t2 + t3 + t4 + t5 + t6;
}

/* DGBEQU (LAPACK): row scale factors R(i) = 1/max_j |A(i,j)| (and column
 * factors, unused by the reference).  ab holds A(i,j) at ab[(ku+i-j) + j*ldab]
 * (0-based i,j). */
int or_dgbequ(int m, int n, int kl, int ku, const double* ab, int ldab, double* r, double* c,
              double* rowcnd, double* colcnd, double* amax)
{
    const double smlnum = DBL_MIN, bignum = 1.0 / DBL_MIN;
    for (int i = 0; i < m; ++i) r[i] = 0.0;
    for (int j = 0; j < n; ++j) {
        int i0 = j - ku > 0 ? j - ku : 0, i1 = j + kl < m - 1 ? j + kl : m - 1;
        for (int i = i0; i <= i1; ++i) {
            double v = fabs(ab[(ku + i - j) + (size_t)j * ldab]);
            if (v > r[i]) r[i] = v;
        }
    }
    double rcmin = bignum, rcmax = 0.0;
    for (int i = 0; i < m; ++i) {
        if (r[i] > rcmax) rcmax = r[i];
        if (r[i] < rcmin) rcmin = r[i];
    }
    *amax = rcmax;
    if (rcmin == 0.0) {
        for (int i = 0; i < m; ++i) if (r[i] == 0.0) return i + 1;
    }
    for (int i = 0; i < m; ++i) {
        double v = r[i] > smlnum ? r[i] : smlnum;
        v = v < bignum ? v : bignum;
        r[i] = 1.0 / v;
    }
    *rowcnd = (rcmin > smlnum ? rcmin : smlnum) / (rcmax < bignum ? rcmax : bignum);
    for (int j = 0; j < n; ++j) c[j] = 0.0;
    for (int j = 0; j < n; ++j) {
        int i0 = j - ku > 0 ? j - ku : 0, i1 = j + kl < m - 1 ? j + kl : m - 1;
        for (int i = i0; i <= i1; ++i) {
            double v = fabs(ab[(ku + i - j) + (size_t)j * ldab]) * r[i];
            if (v > c[j]) c[j] = v;
        }
    }
    double cmin = bignum, cmax = 0.0;
    for (int j = 0; j < n; ++j) {
        if (c[j] < cmin) cmin = c[j];
        if (c[j] > cmax) cmax = c[j];
    }
    if (cmin == 0.0) {
        for (int j = 0; j < n; ++j) if (c[j] == 0.0) return m + j + 1;
    }
    for (int j = 0; j < n; ++j) {
        double v = c[j] > smlnum ? c[j] : smlnum;
        v = v < bignum ? v : bignum;
        c[j] = 1.0 / v;
    }
    *colcnd = (cmin > smlnum ? cmin : smlnum) / (cmax < bignum ? cmax : bignum);
    return 0;
}

/* DGBTF2 (LAPACK reference, unblocked banded LU with partial pivoting).
 * ab: (2kl+ku+1) x n, column-major, A(i,j) at ab[(kl+ku+i-j) + j*ldab]. */
int or_dgbtrf(int m, int n, int kl, int ku, double* ab, int ldab, int* ipiv)
{
#define AB(i, j) ab[(size_t)(i) + (size_t)(j) * (size_t)ldab] /* 0-based row, col */
    const int kv = ku + kl;
    int info = 0;
    for (int j = ku + 1; j < (kv < n ? kv : n); ++j)
        for (int i = kv - j; i < kl; ++i) AB(i, j) = 0.0;
    int ju = 0;
    int mn = m < n ? m : n;
    for (int j = 0; j < mn; ++j) {
        if (j + kv < n)
            for (int i = 0; i < kl; ++i) AB(i, j + kv) = 0.0;
        int km = kl < m - 1 - j ? kl : m - 1 - j;
        /* IDAMAX over AB(kv..kv+km, j): first index of the largest |x| */
        int jp = 0;
        double best = fabs(AB(kv, j));
        for (int t = 1; t <= km; ++t) {
            double v = fabs(AB(kv + t, j));
            if (v > best) { best = v; jp = t; }
        }
        ipiv[j] = jp + j + 1; /* 1-based like LAPACK */
        if (AB(kv + jp, j) != 0.0) {
            int jt = j + ku + jp;
            if (jt > n - 1) jt = n - 1;
            if (jt > ju) ju = jt;
            if (jp != 0)
                for (int t = 0; t <= ju - j; ++t) {
                    double tmp = AB(kv + jp - t, j + t);
                    AB(kv + jp - t, j + t) = AB(kv - t, j + t);
                    AB(kv - t, j + t) = tmp;
                }
            if (km > 0) {
                double rp = 1.0 / AB(kv, j);
                for (int t = 1; t <= km; ++t) AB(kv + t, j) *= rp;
                if (ju > j)
                    for (int jj = 1; jj <= ju - j; ++jj) {
                        double y = AB(kv - jj, j + jj);
                        if (y != 0.0) {
                            double temp = -1.0 * y;
                            for (int t = 1; t <= km; ++t) AB(kv + t - jj, j + jj) += AB(kv + t, j) * temp;
                        }
                    }
            }
        } else if (info == 0) {
            info = j + 1;
        }
    }
    return info;
#undef AB
}

/* DGBTRS('N', NRHS=1) (LAPACK reference): forward pass with the stored
 * interchanges and multipliers, then DTBSV upper non-unit. */
void or_dgbtrs(int n, int kl, int ku, const double* ab, int ldab, const int* ipiv, double* b)
{
#define AB(i, j) ab[(size_t)(i) + (size_t)(j) * (size_t)ldab]
    const int kd = ku + kl; /* 0-based row of the diagonal */
    if (kl > 0)
        for (int j = 0; j < n - 1; ++j) {
            int lm = kl < n - 1 - j ? kl : n - 1 - j;
            int l = ipiv[j] - 1;
            if (l != j) { double t = b[l]; b[l] = b[j]; b[j] = t; }
            if (b[j] != 0.0) {
                double temp = -1.0 * b[j];
                for (int t = 1; t <= lm; ++t) b[j + t] += AB(kd + t, j) * temp;
            }
        }
    const int k = kl + ku;
    for (int j = n - 1; j >= 0; --j) {
        if (b[j] != 0.0) {
            b[j] = b[j] / AB(k, j);
            double temp = b[j];
            int i0 = j - k > 0 ? j - k : 0;
            for (int i = j - 1; i >= i0; --i) b[i] -= temp * AB(k + i - j, j);
        }
    }
#undef AB
}

/* Compute_Matrix_coarse with analysis=0 (iterative_solver.f90:1072-1176):
 * assemble, fold ghost columns with the antisymmetric/symmetric signs, wall
 * identity rows, DGBEQU row scaling. */
void or_compute_matrix_coarse(or_state* s)
{
    const int G = s->grids, nxc = s->nxg[G], nzc = s->nzg[G];
    const int kl = s->kl, ku = s->ku, ldab = 2 * kl + ku + 1, kl_ku_1 = kl + ku + 1;
    const int N = nzc + 1;
    double* M = s->Mc;
#define MC(r, q) M[(size_t)((r) - 1) + (size_t)((q) - 1) * (size_t)ldab] /* 1-based as in the reference */
    memset(M, 0, sizeof(double) * (size_t)ldab * (size_t)s->ngc);
    const int limit[5] = {0, 1, 2, 1, 0};
    for (int p = 1; p <= s->ngc; ++p) {
        int kp = p - N * ((p - 1) / N) - 1, ip = (p - 1) / N;
        if (ip > 0 && ip < nxc && kp > 0 && kp < nzc)
            for (int j = -2; j <= 2; ++j) {
                int iq = j + ip;
                if (iq >= 0 && iq <= nxc)
                    for (int kq = -limit[j + 2] + kp; kq <= kp + limit[j + 2]; ++kq)
                        if (kq >= 0 && kq <= nzc) {
                            int q = N * iq + kq + 1;
                            MC(kl_ku_1 + p - q, q) = stream_coeff(s, G, ip, kp, iq, kq, 1.0);
                        }
            }
    }
    for (int p = 1; p <= s->ngc; ++p) {
        int kp = p - N * ((p - 1) / N) - 1, ip = (p - 1) / N;
        if (ip > 0 && ip < nxc && kp > 0 && kp < nzc)
            for (int j = -2; j <= 2; ++j) {
                int iq = j + ip;
                for (int kq = -limit[j + 2] + kp; kq <= kp + limit[j + 2]; ++kq) {
                    if (iq < 0 || iq > nxc || kq < 0 || kq > nzc) {
                        int iiq, kkq;
                        double xval, zval;
                        if (iq < 0) { iiq = -iq; xval = -1.0; }
                        else if (iq > nxc) { iiq = 2 * nxc - iq; xval = -1.0; }
                        else { iiq = iq; xval = 1.0; }
                        if (kq < 0) { kkq = -kq; zval = (s->p.Vbc == 0) ? -1.0 : 1.0; }
                        else if (kq > nzc) { kkq = 2 * nzc - kq; zval = (s->p.Vbc == 0) ? -1.0 : 1.0; }
                        else { kkq = kq; zval = 1.0; }
                        double Sval = xval * zval;
                        int q = N * iiq + kkq + 1;
                        MC(kl_ku_1 + p - q, q) = MC(kl_ku_1 + p - q, q) + stream_coeff(s, G, ip, kp, iq, kq, Sval);
                    }
                }
            }
    }
    for (int kp = 0; kp <= nzc; ++kp) {
        MC(kl_ku_1, N * 0 + kp + 1) = 1.0;
        MC(kl_ku_1, N * nxc + kp + 1) = 1.0;
    }
    for (int ip = 0; ip <= nxc; ++ip) {
        MC(kl_ku_1, N * ip + 0 + 1) = 1.0;
        MC(kl_ku_1, N * ip + nzc + 1) = 1.0;
    }
    double rowcnd, colcnd, amax;
    /* DGBEQU on rows kl+1..2kl+ku+1 with LDAB = kl+ku+1 (copy-in of the section) */
    {
        const int ld2 = kl + ku + 1;
        double* sec = (double*)malloc(sizeof(double) * (size_t)ld2 * (size_t)s->ngc);
        for (int q = 0; q < s->ngc; ++q)
            for (int r = 0; r < ld2; ++r) sec[r + (size_t)q * ld2] = M[(size_t)(kl + r) + (size_t)q * ldab];
        or_dgbequ(s->ngc, s->ngc, kl, ku, sec, ld2, s->RowB, s->ColB, &rowcnd, &colcnd, &amax);
        free(sec);
    }
    for (int q = 1; q <= s->ngc; ++q) {
        int p0 = q - kl > 1 ? q - kl : 1, p1 = q + ku < s->ngc ? q + ku : s->ngc;
        for (int p = p0; p <= p1; ++p) MC(kl_ku_1 + p - q, q) = s->RowB[p - 1] * MC(kl_ku_1 + p - q, q);
    }
#undef MC
}

/* Compute_coarse_vector + DGBTRS + Stream_grid_coarse + enforceBCs_stream
 * (iterative_solver.f90:139-142, 1181-1204). */
void or_coarse_solve(or_state* s)
{
    const int G = s->grids, nxc = s->nxg[G], nzc = s->nzg[G], N = nzc + 1;
    for (int p = 0; p < s->ngc; ++p) s->Bc[p] = 0.0;
    for (int i = 1; i <= nxc - 1; ++i)
        for (int k = 1; k <= nzc - 1; ++k) {
            int p = N * i + k + 1;
            s->Bc[p - 1] = s->RHS[G][IX(s, i, k)] * s->RowB[p - 1];
        }
    or_dgbtrs(s->ngc, s->kl, s->ku, s->Mc, 2 * s->kl + s->ku + 1, s->ipiv, s->Bc);
    for (int p = 1; p <= s->ngc; ++p) {
        int kp = p - N * ((p - 1) / N) - 1, ip = (p - 1) / N;
        s->error[G][IX(s, ip, kp)] = s->Bc[p - 1];
    }
    or_enforce_bcs_stream(s, G, s->error[G]);
}

/* ------------------------------------------------------------------------ */
/* multigrid (iterative_solver.f90:79-173), Newtonian branch.               */
/* ------------------------------------------------------------------------ */
int or_multigrid(or_state* s)
{
    const int G = s->grids;
    memcpy(s->error[1], s->SF, s->plane * sizeof(double));
    or_compute_RHS(s, 1);
    or_compute_reference_residual(s, 1);
    for (int L = 1; L <= G; ++L) or_compute_viscosity(s, L);
    if (s->time == 0.0 || s->p.ivis > 0) {
        or_compute_matrix_coarse(s);
        int info = or_dgbtrf(s->ngc, s->ngc, s->kl, s->ku, s->Mc, 2 * s->kl + s->ku + 1, s->ipiv);
        if (info != 0) { s->status = TC_ERR_SINGULAR; return s->status; }
    }
    for (int L = 1; L <= G; ++L) or_iterate_local_time_steps(s, L);
    const int* multi = s->p.multi;
    const int nsteps = s->p.nsteps;
    int q = 1;
    s->status = 0;
    for (;;) {
        int done = 0;
        for (int p = 1; p <= nsteps; ++p) {
            int level = multi[p];
            int dir = (multi[p + 1] > multi[p]) ? 0 : 1;
            if (multi[p - 1] < multi[p]) {
                or_compute_RHS(s, level);
                if (p <= G && q == 1) or_compute_reference_residual(s, level);
                /* optional standard coarse-grid correction (not the reference) */
                if (s->p.fixes & TC_FIX_ZERO_COARSE) memset(s->error[level], 0, s->plane * sizeof(double));
            }
            if (level < G) {
                or_iterate_stream(s, level, s->p.Nmulti, dir, p);
                if (s->status) return s->status;
            } else {
                or_coarse_solve(s);
            }
            if (level == 1 && s->residual_mag[1] / s->res_ref[1] <= s->p.tolerance) { done = 1; break; }
            if (multi[p + 1] < multi[p]) or_prolong(s, level);
        }
        s->cycles_total++;
        if (done) break;
        if (s->residual_mag[1] / s->res_ref[1] <= s->p.tolerance) break;
        q++;
        if (q > s->p.max_cycles) { s->q = q; s->status = TC_ERR_MAX_CYCLES; return s->status; }
    }
    memcpy(s->SF, s->error[1], s->plane * sizeof(double));
    s->q = q;
    if (s->q_step_n >= 0 && s->q_step_n < 4) s->q_step[s->q_step_n++] = q;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Velocities, strain rate                                                   */
/* ------------------------------------------------------------------------ */

/* compute_velocities (momentum.f90:532-551) + velocityBCs
 * (composition.f90:806-836).  u,w live in the ghosted frame; only
 * (-1:nx+1,-1:nz+1) is meaningful, the outer ring stays 0. */
void or_compute_velocities(or_state* s)
{
    const int nx = s->nx, nz = s->nz;
    double *u = s->u, *w = s->w;
    memset(u, 0, s->plane * sizeof(double));
    memset(w, 0, s->plane * sizeof(double));
    for (int k = 0; k <= nz; ++k)
        for (int i = 0; i <= nx; ++i) {
            size_t c = IX(s, i, k);
            double SFr = dot3(s->D1, s->SF + c, 1) / s->dr;
            double SFs = dot3(s->D1, s->SF + c, s->ld) / s->ds;
            double SFx = SFr / s->xr[i + SPAN];
            double SFz = SFs / s->zs[k + SPAN];
            u[c] = SFz;
            w[c] = -SFx;
        }
    /* velocityBCs, span_interp = 1 */
    for (int k = 0; k <= nz; ++k) {
        w[IX(s, -1, k)] = w[IX(s, 1, k)];
        w[IX(s, nx + 1, k)] = w[IX(s, nx - 1, k)];
    }
    for (int i = 0; i <= nx; ++i) {
        if (s->p.Vbc == 0) {
            u[IX(s, i, -1)] = u[IX(s, i, 1)];
            u[IX(s, i, nz + 1)] = u[IX(s, i, nz - 1)];
        } else {
            u[IX(s, i, -1)] = -u[IX(s, i, 1)];
            u[IX(s, i, nz + 1)] = -u[IX(s, i, nz - 1)];
        }
    }
    for (int i = -1; i <= nx + 1; ++i) {
        w[IX(s, i, -1)] = -w[IX(s, i, 1)];
        w[IX(s, i, nz + 1)] = -w[IX(s, i, nz - 1)];
    }
    for (int k = -1; k <= nz + 1; ++k) {
        u[IX(s, -1, k)] = -u[IX(s, 1, k)];
        u[IX(s, nx + 1, k)] = -u[IX(s, nx - 1, k)];
    }
    for (int k = -1; k <= nz + 1; ++k) { u[IX(s, 0, k)] = 0.0; u[IX(s, nx, k)] = 0.0; }
    for (int i = -1; i <= nx + 1; ++i) { w[IX(s, i, 0)] = 0.0; w[IX(s, i, nz)] = 0.0; }
    if (s->p.Vbc == 1)
        for (int i = -1; i <= nx + 1; ++i) { u[IX(s, i, 0)] = 0.0; u[IX(s, i, nz)] = 0.0; }
}

/* compute_strain_rate_invariant (momentum.f90:32-62). */
void or_compute_strain_rate_invariant(or_state* s)
{
    const int nx = s->nx, nz = s->nz;
    double* st = s->strain;
    for (int k = 0; k <= nz; ++k)
        for (int i = 0; i <= nx; ++i) {
            size_t c = IX(s, i, k);
            double ur = dot3(s->D1, s->u + c, 1) / s->dr;
            double us = dot3(s->D1, s->u + c, s->ld) / s->ds;
            double wr = dot3(s->D1, s->w + c, 1) / s->dr;
            double ws = dot3(s->D1, s->w + c, s->ld) / s->ds;
            double ux = ur / s->xr[i + SPAN], uz = us / s->zs[k + SPAN];
            double wx = wr / s->xr[i + SPAN], wz = ws / s->zs[k + SPAN];
            double a = uz + wx;
            st[c] = sqrt(0.5 * (2.0 * (ux * ux + wz * wz) + a * a));
        }
    for (int i = 1; i <= SPAN; ++i)
        for (int k = 0; k <= nz; ++k) {
            st[IX(s, -i, k)] = st[IX(s, i, k)];
            st[IX(s, nx + i, k)] = st[IX(s, nx - i, k)];
        }
    for (int k = 1; k <= SPAN; ++k)
        for (int i = -SPAN; i <= nx + SPAN; ++i) {
            st[IX(s, i, -k)] = st[IX(s, i, k)];
            st[IX(s, i, nz + k)] = st[IX(s, i, nz - k)];
        }
}

/* ------------------------------------------------------------------------ */
/* T and C filters                                                           */
/* ------------------------------------------------------------------------ */

/* smoothness (smoother.f90:75-95): largest jump to the 5 forward neighbours.
   max(a1,...,a5) is gfortran's MAX, which returns the non-NaN argument (fmax);
   a NaN result fails "a > diff".  So NaN jumps are ignored. */
static double smoothness(const or_state* s, const double* T0)
{
    double diff = 0.0;
    for (int i = 0; i <= s->nx - 1; ++i)
        for (int k = 1; k <= s->nz - 1; ++k) {
            double c = T0[IX(s, i, k)];
            double a1 = fabs(T0[IX(s, i, k - 1)] - c);
            double a2 = fabs(T0[IX(s, i + 1, k - 1)] - c);
            double a3 = fabs(T0[IX(s, i + 1, k)] - c);
            double a4 = fabs(T0[IX(s, i, k + 1)] - c);
            double a5 = fabs(T0[IX(s, i + 1, k + 1)] - c);
            double a = fmax(fmax(fmax(fmax(a1, a2), a3), a4), a5);
            if (a > diff) diff = a;
        }
    return diff;
}

/* smoother_space (smoother.f90:121-139). */
static double smoother_space(const or_state* s, int i, int k, const double* T0)
{
    const double* P = T0 + IX(s, i, k);
    double Tr = dot3(s->D1, P, 1) / s->dr;
    double Ts = dot3(s->D1, P, s->ld) / s->ds;
    double Trr = dot3(s->D2, P, 1) / s->dr2;
    double Tss = dot3(s->D2, P, s->ld) / s->ds2;
    double xr = s->xr[i + SPAN], zs = s->zs[k + SPAN];
    double Tx = Tr / xr, Tz = Ts / zs;
    double Txx = (Trr - Tx * s->xrr[i + SPAN]) / (xr * xr);
    double Tzz = (Tss - Tz * s->zss[k + SPAN]) / (zs * zs);
    return Txx + Tzz;
}

/* smoother_time_T (smoother.f90:342-376). */
int or_smoother_time_T(or_state* s, const double* Tin)
{
    if (Tin != s->Tvis) memcpy(s->Tvis, Tin, s->plane * sizeof(double));
    memcpy(s->Tbuoy, s->Tvis, s->plane * sizeof(double));
    if (s->p.RaT == 0.0) return 0;
    int n = 0;
    double* T0 = s->tmp2;
    for (;;) {
        double diff = smoothness(s, s->Tvis);
        if (diff < s->p.delta_T) { s->n_temp = n; break; }
        memcpy(T0, s->Tvis, s->plane * sizeof(double));
#pragma omp parallel for schedule(static)
        for (int k = 0; k <= s->nz; ++k)
            for (int i = 0; i <= s->nx; ++i) {
                size_t c = IX(s, i, k);
                double temp = smoother_space(s, i, k, T0);
                double dts = s->dts[c] * s->p.courantT;
                s->Tvis[c] = T0[c] + dts * temp;
            }
        or_enforce_bcs(s, s->Tvis);
        n++;
        s->t_iters++;
    }
    return n;
}

/* smoother_time (smoother.f90:286-340).  Niter is not reset between types
 * (as in the reference). */
void or_smoother_time(or_state* s)
{
    int Niter = 10;
    const double lb = -5.0, la = -1.0;
    for (int t = 1; t <= s->ntype; ++t) {
        memcpy(s->Cvis[t], s->Cnew[t], s->plane * sizeof(double));
        memcpy(s->Cbuoy[t], s->Cnew[t], s->plane * sizeof(double));
    }
    double* T0 = s->tmp2;
    for (int t = 1; t <= s->ntype; ++t) {
        int nn = 0, trigger = 0;
        double* C = s->Cvis[t];
        for (;;) {
            for (int n = 1; n <= Niter; ++n) {
                double diff = smoothness(s, C);
                if (diff < s->p.delta_C[t - 1]) { trigger = 1; break; }
                double c;
                if (nn == 0) {
                    int ntemp = Niter - n + 1;
                    c = 2.0 / (-lb - la + (lb - la) * cos((double)(2 * ntemp - 1) * PII / (double)(2 * Niter)));
                } else {
                    c = 1.0;
                }
                memcpy(T0, C, s->plane * sizeof(double));
#pragma omp parallel for schedule(static)
                for (int k = 0; k <= s->nz; ++k)
                    for (int i = 0; i <= s->nx; ++i) {
                        size_t cc = IX(s, i, k);
                        double temp = smoother_space(s, i, k, T0);
                        double dts = s->dts[cc] * s->p.courantC;
                        double v = T0[cc] + c * dts * temp;
                        if (v < 0.0) v = 0.0;
                        if (v > 1.0) v = 1.0;
                        C[cc] = v;
                    }
                or_enforce_bcs_comp(s, C);
                s->c_iters++;
            }
            if (trigger == 1) break;
            Niter = 1;
            nn++;
        }
        s->n_comp[t] = nn;
    }
}

/* conductivity + load_conductivity_array (energy.f90:16-58). */
void or_load_conductivity_array(or_state* s)
{
    const int nx = s->nx, nz = s->nz;
    double* K = s->conduct;
    for (int k = 0; k <= nz; ++k)
        for (int i = 0; i <= nx; ++i) {
            size_t c = IX(s, i, k);
            double v = 1.0;
            if (s->p.comp != 0)
                for (int j = 1; j <= s->ntype; ++j) v = exp(-s->Cvis[j][c] * log(1.0 / s->p.conduct_factor[j - 1]));
            K[c] = v;
        }
    for (int i = 1; i <= 2; ++i)
        for (int k = 0; k <= nz; ++k) {
            K[IX(s, -i, k)] = K[IX(s, i, k)];
            K[IX(s, nx + i, k)] = K[IX(s, nx - i, k)];
        }
    for (int k = 1; k <= 2; ++k)
        for (int i = -2; i <= nx + 2; ++i) {
            K[IX(s, i, -k)] = K[IX(s, i, k)];
            K[IX(s, i, nz + k)] = K[IX(s, i, nz - k)];
        }
    for (int k = -1; k <= nz + 1; ++k)
        for (int i = -1; i <= nx + 1; ++i) {
            size_t c = IX(s, i, k);
            s->conduct_r[c] = dot3(s->D1, K + c, 1) / s->dr;
            s->conduct_s[c] = dot3(s->D1, K + c, s->ld) / s->ds;
        }
}

/* ------------------------------------------------------------------------ */
/* Tracers                                                                   */
/* ------------------------------------------------------------------------ */

static void tracer_set(or_state* s, int which, double** r, double** sv)
{
    switch (which) {
    case TC_TR_X0: *r = s->rtr0; *sv = s->str0; break;
    case TC_TR_X1: *r = s->rtr1; *sv = s->str1; break;
    case TC_TR_X2: *r = s->rtr2; *sv = s->str2; break;
    default: *r = s->rtr; *sv = s->str; break;
    }
}

/* tracers_to_corners (composition.f90:669-790) with tracer=0, comp=1:
 * nearest-node counts per type, mirrored ghost counts, nempty, the tracer
 * ratio Cnew (empty nodes keep the previous value), enforceBCs_comp, then
 * smoother_time (if smooth_vis) and load_conductivity_array. */
int or_tracers_to_corners(or_state* s, int which)
{
    const int nx = s->nx, nz = s->nz;
    double *R, *S;
    tracer_set(s, which, &R, &S);
    for (int t = 0; t <= s->ntype; ++t) memset(s->count[t], 0, s->plane * sizeof(int));
    for (int64_t id = 0; id < s->ntr; ++id) {
        double r = R[id], sv = S[id];
        int i = (int)lround(r / s->dr);
        int k = (int)lround(sv / s->ds);
        if (r <= 0.0 || r >= s->p.aspect || sv <= 0.0 || sv >= 1.0) { s->status = TC_ERR_TRACER_DOMAIN; return s->status; }
        s->count[s->Ttype[id]][IX(s, i, k)] += 1;
    }
    for (int t = 0; t <= s->ntype; ++t) {
        int* C = s->count[t];
        for (int i = 1; i <= SPAN; ++i)
            for (int k = 0; k <= nz; ++k) {
                C[IX(s, -i, k)] = C[IX(s, i, k)];
                C[IX(s, nx + i, k)] = C[IX(s, nx - i, k)];
            }
        for (int k = 1; k <= SPAN; ++k)
            for (int i = 0; i <= nx; ++i) {
                C[IX(s, i, -k)] = C[IX(s, i, k)];
                C[IX(s, i, nz + k)] = C[IX(s, i, nz - k)];
            }
    }
    for (size_t c = 0; c < s->plane; ++c) {
        int acc = 0;
        for (int t = 0; t <= s->ntype; ++t) acc += s->count[t][c];
        s->nempty[c] = acc;
    }
    for (int k = 0; k <= nz; ++k)
        for (int i = 0; i <= nx; ++i) {
            size_t c = IX(s, i, k);
            int ne = s->nempty[c];
            if (ne > 0 && s->p.comp == 1)
                for (int t = 1; t <= s->ntype; ++t) {
                    int ct = s->count[t][c];
                    if (ct == ne) s->Cnew[t][c] = 1.0;
                    else if (ct == 0) s->Cnew[t][c] = 0.0;
                    else s->Cnew[t][c] = (double)ct / (double)ne;
                }
        }
    if (s->p.comp == 1) {
        for (int t = 1; t <= s->ntype; ++t) or_enforce_bcs_comp(s, s->Cnew[t]);
        if (s->smooth_vis == 1) or_smoother_time(s);
    }
    or_load_conductivity_array(s);
    return 0;
}

/* compute_derivatives (composition.f90:900-919), order=2: the central first
 * derivative along r on k=-1..nz+1, i=0..nx. */
void or_compute_derivatives(or_state* s, const double* f, double* DERr)
{
    for (int k = -1; k <= s->nz + 1; ++k)
        for (int i = 0; i <= s->nx; ++i) {
            size_t c = IX(s, i, k);
            double acc = 0.0;
            acc += s->Dint[0][0] * f[c - 1];
            acc += s->Dint[0][1] * f[c];
            acc += s->Dint[0][2] * f[c + 1];
            DERr[c] = acc / s->dr; /* dr_power(1) = dr**1.0 = dr */
        }
}

/* interpolate (composition.f90:838-898), order=2: first-order Taylor
 * expansion in r from the nearer node (forward if hr <= dr/2, else backward
 * from i+1) on rows k-1..k+2, then in s with the central D(1,.). */
double or_interpolate(const or_state* s, double r, double sv, const double* DERr, const double* f)
{
    int i = (int)(r / s->dr);
    int k = (int)(sv / s->ds);
    /* In-domain positions give 0<=i<=nx-1, 0<=k<=nz-1.  The reference reads
     * outside its arrays otherwise; the clamp (same as the HIP kernel) only
     * keeps invalid input inside the ghost ring, and is counted. */
    if (i < 0 || i > s->nx - 1 || k < 0 || k > s->nz - 1)
        __atomic_fetch_add(&((or_state*)s)->interp_clamps, 1, __ATOMIC_RELAXED);
    i = i < 0 ? 0 : (i > s->nx - 1 ? s->nx - 1 : i);
    k = k < 0 ? 0 : (k > s->nz - 1 ? s->nz - 1 : k);
    double hr = r - (double)i * s->dr;
    double fR[4]; /* jj = -1..2 */
    if (hr <= s->dro2) {
        double a1 = hr;
        for (int jj = -1; jj <= 2; ++jj) {
            size_t c = IX(s, i, k + jj);
            fR[jj + 1] = f[c] + DERr[c] * a1;
        }
    } else {
        hr = s->dr - hr;
        double a1 = -hr;
        for (int jj = -1; jj <= 2; ++jj) {
            size_t c = IX(s, i + 1, k + jj);
            fR[jj + 1] = f[c] + DERr[c] * a1;
        }
    }
    double hs = sv - (double)k * s->ds;
    double fi;
    if (hs <= s->dso2) {
        double acc = 0.0;
        acc += s->Dint[0][0] * fR[0];
        acc += s->Dint[0][1] * fR[1];
        acc += s->Dint[0][2] * fR[2];
        double DERs = acc / s->ds;
        fi = fR[1] + DERs * hs;
    } else {
        hs = s->ds - hs;
        double acc = 0.0;
        acc += s->Dint[0][0] * fR[1];
        acc += s->Dint[0][1] * fR[2];
        acc += s->Dint[0][2] * fR[3];
        double DERs = acc / s->ds;
        fi = fR[2] + DERs * (-hs);
    }
    return fi;
}

/* One RK stage of the tracer loops in energy_time (energy.f90:169-228,
 * 243-246 + 261-315, 352-381, 417-443) for tracer=0, local_time=0, uniform
 * grids.  Stage 1 also copies rtr0=rtr (energy.f90:146-149). */
int or_tracer_stage(or_state* s, int stage, double dt)
{
    or_compute_derivatives(s, s->u, s->DERr_u);
    or_compute_derivatives(s, s->w, s->DERr_w);
    const int64_t ntr = s->ntr;
    const int torder = s->p.torder;
    if (stage == 1) {
        memcpy(s->rtr0, s->rtr, ntr * sizeof(double));
        memcpy(s->str0, s->str, ntr * sizeof(double));
        double dtstage = (torder == 2) ? dt : 0.5 * dt;
#pragma omp parallel for schedule(static)
        for (int64_t id = 0; id < ntr; ++id) {
            double tr = or_interpolate(s, s->rtr0[id], s->str0[id], s->DERr_u, s->u);
            double ts = or_interpolate(s, s->rtr0[id], s->str0[id], s->DERr_w, s->w);
            s->rtr[id] = tr;
            s->str[id] = ts;
            s->rtr1[id] = s->rtr0[id] + dtstage * tr;
            s->str1[id] = s->str0[id] + dtstage * ts;
        }
    } else if (stage == 2) {
        double dtstage = 0.5 * dt;
#pragma omp parallel for schedule(static)
        for (int64_t id = 0; id < ntr; ++id) {
            double tr = or_interpolate(s, s->rtr1[id], s->str1[id], s->DERr_u, s->u);
            double ts = or_interpolate(s, s->rtr1[id], s->str1[id], s->DERr_w, s->w);
            if (torder == 2) {
                s->rtr[id] = s->rtr[id] + tr;
                s->str[id] = s->str[id] + ts;
                s->rtr[id] = s->rtr0[id] + dtstage * s->rtr[id];
                s->str[id] = s->str0[id] + dtstage * s->str[id];
            } else {
                s->rtr[id] = s->rtr[id] + 2.0 * tr;
                s->str[id] = s->str[id] + 2.0 * ts;
                s->rtr2[id] = s->rtr0[id] + dtstage * tr;
                s->str2[id] = s->str0[id] + dtstage * ts;
            }
        }
    } else if (stage == 3) {
        double dtt = dt;
#pragma omp parallel for schedule(static)
        for (int64_t id = 0; id < ntr; ++id) {
            double tr = or_interpolate(s, s->rtr2[id], s->str2[id], s->DERr_u, s->u);
            s->rtr[id] = s->rtr[id] + 2.0 * tr;
            s->rtr1[id] = s->rtr0[id] + dtt * tr;
            double ts = or_interpolate(s, s->rtr2[id], s->str2[id], s->DERr_w, s->w);
            s->str[id] = s->str[id] + 2.0 * ts;
            s->str1[id] = s->str0[id] + dtt * ts;
        }
    } else if (stage == 4) {
        double dto6 = dt / 6.0;
#pragma omp parallel for schedule(static)
        for (int64_t id = 0; id < ntr; ++id) {
            double tr = or_interpolate(s, s->rtr1[id], s->str1[id], s->DERr_u, s->u);
            s->rtr[id] = s->rtr0[id] + dto6 * (s->rtr[id] + tr);
            double ts = or_interpolate(s, s->rtr1[id], s->str1[id], s->DERr_w, s->w);
            s->str[id] = s->str0[id] + dto6 * (s->str[id] + ts);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Energy equation (T field) -- energy.f90 stays on the host in the drop-in; */
/* restated here so that a whole time step can be checked.                    */
/* ------------------------------------------------------------------------ */

/* heating (energy.f90:1-14). */
static double heating(const or_state* s, size_t c)
{
    if (s->p.comp == 0) return s->p.H;
    double sum = 0.0;
    for (int j = 1; j <= s->ntype; ++j) sum += s->Cvis[j][c];
    double h = s->p.H * (1.0 - sum);
    for (int j = 1; j <= s->ntype; ++j) h = s->p.Htype[j - 1] * s->Cvis[j][c] + h;
    return h;
}

/* energy_space (energy.f90:60-80). */
static double energy_space(const or_state* s, int i, int k, const double* T0)
{
    const size_t c = IX(s, i, k);
    const double* P = T0 + c;
    double Tr = dot3(s->D1, P, 1) / s->dr;
    double Ts = dot3(s->D1, P, s->ld) / s->ds;
    double Trr = dot3(s->D2, P, 1) / s->dr2;
    double Tss = dot3(s->D2, P, s->ld) / s->ds2;
    double xr = s->xr[i + SPAN], zs = s->zs[k + SPAN];
    double kx = s->conduct_r[c] / xr;
    double kz = s->conduct_s[c] / zs;
    double Tx = Tr / xr, Tz = Ts / zs;
    double Txx = (Trr - Tx * s->xrr[i + SPAN]) / (xr * xr);
    double Tzz = (Tss - Tz * s->zss[k + SPAN]) / (zs * zs);
    return s->conduct[c] * (Txx + Tzz) + (kx * Tx + kz * Tz) - s->u[c] * Tx - s->w[c] * Tz + heating(s, c);
}

/* T-field part of RK stage `stage` of energy_time with tracer=0
 * (energy.f90:157-167, 248-260, 339-351, 405-416). */
void or_energy_stage(or_state* s, int stage, double dt)
{
    const int nx = s->nx, nz = s->nz;
    if (s->p.RaT == 0.0) return;
    if (stage == 1) {
#pragma omp parallel for schedule(static)
        for (int k = 0; k <= nz; ++k)
            for (int i = 0; i <= nx; ++i) {
                size_t c = IX(s, i, k);
                double temp = energy_space(s, i, k, s->T0);
                s->T[c] = temp;
                s->T1[c] = s->T0[c] + 0.5 * dt * temp;
            }
        or_enforce_bcs(s, s->T1);
    } else if (stage == 2) {
#pragma omp parallel for schedule(static)
        for (int k = 0; k <= nz; ++k)
            for (int i = 0; i <= nx; ++i) {
                size_t c = IX(s, i, k);
                double temp = energy_space(s, i, k, s->T1);
                s->T[c] = s->T[c] + 2.0 * temp;
                s->T2[c] = s->T0[c] + 0.5 * dt * temp;
            }
        or_enforce_bcs(s, s->T2);
    } else if (stage == 3) {
#pragma omp parallel for schedule(static)
        for (int k = 0; k <= nz; ++k)
            for (int i = 0; i <= nx; ++i) {
                size_t c = IX(s, i, k);
                double temp = energy_space(s, i, k, s->T2);
                s->T[c] = s->T[c] + 2.0 * temp;
                s->T1[c] = s->T0[c] + dt * temp;
            }
        or_enforce_bcs(s, s->T1);
    } else {
#pragma omp parallel for schedule(static)
        for (int k = 0; k <= nz; ++k)
            for (int i = 0; i <= nx; ++i) {
                size_t c = IX(s, i, k);
                double temp = energy_space(s, i, k, s->T1);
                s->T[c] = s->T0[c] + (dt / 6.0) * (s->T[c] + temp);
            }
        or_enforce_bcs(s, s->T);
    }
}

/* stability_RK4 (energy.f90:600-646) with tracer=0 and the eigA/eigD of
 * energy.f90:686-706. */
double or_stability_RK4(or_state* s)
{
    double Ts_max = s->p.Tr_base / s->ds;
    double Tss_max = -s->p.Trr_base / s->ds2;
    double Tr_max = s->p.Tr_base / s->dr;
    double Trr_max = -s->p.Trr_base / s->dr2;
    double dt;
    if (s->p.RaT != 0.0) {
        double eA_max = 0.0, eD_max = 0.0;
        for (int i = 0; i <= s->nx; ++i)
            for (int k = 0; k <= s->nz; ++k) {
                size_t c = IX(s, i, k);
                double xr = s->xr[i + SPAN], zs = s->zs[k + SPAN];
                double K = s->conduct[c];
                double eA = fabs(-(K * s->xrr[i + SPAN] / pow(xr, 3.0)) - s->u[c] / xr + s->conduct_r[c] / (xr * xr)) * Tr_max +
                            fabs(-(K * s->zss[k + SPAN] / pow(zs, 3.0)) - s->w[c] / zs + s->conduct_s[c] / (zs * zs)) * Ts_max;
                double eD = K * (fabs(Trr_max / (xr * xr)) + fabs(Tss_max / (zs * zs)));
                if (eA > eA_max) eA_max = eA;
                if (eD > eD_max) eD_max = eD;
            }
        double dtA, dtD;
        if (s->p.torder == 2) { dtA = 1.0 / eA_max; dtD = 2.0 / eD_max; }
        else { dtA = 2.8 / eA_max; dtD = 2.8 / eD_max; }
        if (eA_max == 0.0) dt = dtD;
        else dt = 1.0 / (1.0 / dtA + 1.0 / dtD);
    } else {
        dt = 0.5; /* dt_manual */
    }
    return s->p.courant * dt;
}

/* energy_time (energy.f90:117-453) with tracer=0, comp=1, local_time=0,
 * equalize=0, RK_update=1: the RK4 (or RK2) stages interleaving the T-field
 * update, tracers_to_corners, smoother_time_T, the multigrid solves and the
 * tracer loops. */
int or_energy_time(or_state* s)
{
    int st;
    s->dt = or_stability_RK4(s);
    if (s->eq_on) {
        tc_equalize_info info;
        if ((st = or_equalize_tracers(s, s->eq_tpc, s->eq_Cinit, s->eq_seed, s->eq_calls++, &info))) return st;
    }
    const double dt = s->dt;
    const int torder = s->p.torder;
    for (int t = 1; t <= s->ntype; ++t) memcpy(s->C0[t], s->Cnew[t], s->plane * sizeof(double));
    memcpy(s->T0, s->T, s->plane * sizeof(double));
    /* stage 1 */
    or_energy_stage(s, 1, dt);
    or_tracer_stage(s, 1, dt); /* also rtr0=rtr */
    if (torder >= 2) {
        if (s->p.RK_update == 1) {
            if ((st = or_tracers_to_corners(s, TC_TR_X1))) return st;
            or_smoother_time_T(s, s->T1);
            if ((st = or_multigrid(s))) return st;
            or_compute_velocities(s);
        }
        or_energy_stage(s, 2, dt);
        or_tracer_stage(s, 2, dt);
    }
    if (torder >= 3) {
        if (s->p.RK_update == 1) {
            for (int t = 1; t <= s->ntype; ++t) memcpy(s->Cnew[t], s->C0[t], s->plane * sizeof(double));
            if ((st = or_tracers_to_corners(s, TC_TR_X2))) return st;
            or_smoother_time_T(s, s->T2);
            if ((st = or_multigrid(s))) return st;
            or_compute_velocities(s);
        }
        or_energy_stage(s, 3, dt);
        or_tracer_stage(s, 3, dt);
    }
    if (torder >= 4) {
        if (s->p.RK_update == 1) {
            for (int t = 1; t <= s->ntype; ++t) memcpy(s->Cnew[t], s->C0[t], s->plane * sizeof(double));
            if ((st = or_tracers_to_corners(s, TC_TR_X1))) return st;
            or_smoother_time_T(s, s->T1);
            if ((st = or_multigrid(s))) return st;
            or_compute_velocities(s);
        }
        or_energy_stage(s, 4, dt);
        or_tracer_stage(s, 4, dt);
    }
    for (int t = 1; t <= s->ntype; ++t) memcpy(s->Cnew[t], s->C0[t], s->plane * sizeof(double));
    if ((st = or_tracers_to_corners(s, TC_TR_X))) return st;
    or_smoother_time_T(s, s->T);
    return 0;
}

/* Body of the time loop (main.f90:174-205) with iterate=1, local_time=0:
 * multigrid, compute_velocities, the strain rate of stats, energy_time,
 * time += dt. */
int or_time_step(or_state* s)
{
    int st;
    memset(s->q_step, 0, sizeof(s->q_step));
    s->q_step_n = 0;
    st = or_multigrid(s);
    if (!st) {
        or_compute_velocities(s);
        or_compute_strain_rate_invariant(s);
        st = or_energy_time(s);
    }
    s->q_step_n = -1;
    if (st) return st;
    s->time += s->dt;
    s->tstep++;
    return 0;
}

/* Start-up after T, SF and the tracers are loaded (the restart path of
 * main.f90:121-163 and grid.f90:56-95): enforceBCs(T), enforceBCs_stream(SF),
 * the grid-dependent time-step tables, tracers_to_corners without smoothing,
 * smoother_time_T(T), smoother_time, load_conductivity_array, error(2:)=0. */
int or_initialize_state(or_state* s)
{
    or_enforce_bcs(s, s->T);
    or_enforce_bcs_stream(s, 1, s->SF);
    or_stability_RK4_smoother(s);
    for (int L = 1; L <= s->grids; ++L) or_stability_vis_smoother(s, L);
    s->smooth_vis = 0;
    int st = or_tracers_to_corners(s, TC_TR_X);
    s->smooth_vis = 1;
    if (st) return st;
    or_smoother_time_T(s, s->T);
    if (s->p.comp != 0) or_smoother_time(s);
    or_load_conductivity_array(s);
    for (int L = 2; L <= s->grids; ++L) memset(s->error[L], 0, s->plane * sizeof(double));
    s->time = 0.0;
    s->tstep = 0;
    return 0;
}

/* stats (basics.f90:236-324): trapezoid integrals of integrate2D /
 * integrate1Dx (basics.f90:341-390). */
static double integrate1Dx(const or_state* s, const double* f, ptrdiff_t stride)
{
    double fint = 0.0;
    for (int i = 0; i <= s->nx - 1; ++i)
        fint = fint + 0.5 * (f[i * stride] * s->xr[i + SPAN] + f[(i + 1) * stride] * s->xr[i + 1 + SPAN]) * s->dr;
    return fint;
}
static double integrate2D_fn(const or_state* s, double (*g)(const or_state*, size_t, const void*), const void* ctx)
{
    double* fintx = (double*)malloc(sizeof(double) * (s->nz + 1));
    double* row = (double*)malloc(sizeof(double) * (s->nx + 1));
    for (int k = 0; k <= s->nz; ++k) {
        for (int i = 0; i <= s->nx; ++i) row[i] = g(s, IX(s, i, k), ctx);
        fintx[k] = integrate1Dx(s, row, 1);
    }
    double fint = 0.0;
    for (int k = 0; k <= s->nz - 1; ++k)
        fint = fint + 0.5 * (fintx[k] * s->zs[k + SPAN] + fintx[k + 1] * s->zs[k + 1 + SPAN]) * s->ds;
    free(fintx);
    free(row);
    return fint;
}
static double g_vel2(const or_state* s, size_t c, const void* ctx) { (void)ctx; return s->u[c] * s->u[c] + s->w[c] * s->w[c]; }
static double g_field(const or_state* s, size_t c, const void* ctx) { (void)s; return ((const double*)ctx)[c]; }

void or_stats(or_state* s, double out[9])
{
    const int nx = s->nx, nz = s->nz;
    or_compute_strain_rate_invariant(s);
    double strain_max = -INFINITY;
    for (int i = 0; i <= nx; ++i) if (s->strain[IX(s, i, nz)] > strain_max) strain_max = s->strain[IX(s, i, nz)];
    double Cavg = 0.0;
    if (s->p.comp == 1) Cavg = integrate2D_fn(s, g_field, s->Cnew[1]) / s->p.aspect;
    int empty = 0, tmin = 1 << 30, tmax = -1;
    for (int k = 0; k <= nz; ++k)
        for (int i = 0; i <= nx; ++i) {
            int v = s->nempty[IX(s, i, k)];
            if (v == 0) empty++;
            if (v < tmin) tmin = v;
            if (v > tmax) tmax = v;
        }
    double vrms = sqrt(integrate2D_fn(s, g_vel2, NULL) / s->p.aspect);
    double Tavg = integrate2D_fn(s, g_field, s->T) / s->p.aspect;
    double* Tz_surf = (double*)malloc(sizeof(double) * (nx + 1));
    double* Tz_bot = (double*)malloc(sizeof(double) * (nx + 1));
    double* a = (double*)malloc(sizeof(double) * (nx + 1));
    double* b = (double*)malloc(sizeof(double) * (nx + 1));
    for (int i = 0; i <= nx; ++i) {
        Tz_surf[i] = dot3(s->D1, s->T + IX(s, i, nz), s->ld) / s->ds / s->zs[nz + SPAN];
        Tz_bot[i] = dot3(s->D1, s->T + IX(s, i, 0), s->ld) / s->ds / s->zs[0 + SPAN];
    }
    for (int i = 0; i <= nx; ++i) { a[i] = -s->conduct[IX(s, i, nz)] * Tz_surf[i]; b[i] = s->T[IX(s, i, 0)]; }
    double Nu = integrate1Dx(s, a, 1) / integrate1Dx(s, b, 1);
    for (int i = 0; i <= nx; ++i) a[i] = -s->conduct[IX(s, i, 0)] * Tz_bot[i];
    double Nu_bot = integrate1Dx(s, a, 1) / s->p.aspect;
    free(Tz_surf); free(Tz_bot); free(a); free(b);
    out[0] = vrms; out[1] = Nu; out[2] = Tavg; out[3] = Cavg;
    out[4] = empty; out[5] = tmin; out[6] = tmax; out[7] = Nu_bot; out[8] = strain_max;
}

static double g_ratavg(const or_state* s, size_t c, const void* ctx)
{
    (void)ctx;
    return s->p.RaT / s->conduct[c] / s->vis[1][c];
}

/* integrate1Dz_ent (basics.f90:393-404): the z trapezoid from row kent */
static double integrate_ent(const or_state* s, const double* f, int kent)
{
    double* fintx = (double*)malloc(sizeof(double) * (s->nz + 1));
    for (int k = 0; k <= s->nz; ++k) fintx[k] = integrate1Dx(s, f + IX(s, 0, k), 1);
    double fint = 0.0;
    for (int k = kent; k <= s->nz - 1; ++k)
        fint = fint + 0.5 * (fintx[k] * s->zs[k + SPAN] + fintx[k + 1] * s->zs[k + 1 + SPAN]) * s->ds;
    free(fintx);
    return fint;
}

void or_stats_record(or_state* s, int kent, const double* mass, tc_stats_row* out)
{
    const int nx = s->nx, nz = s->nz;
    double d[9];
    or_stats(s, d);
    memset(out, 0, sizeof(*out));
    out->time = s->time;
    out->vrms = d[0]; out->Nu = d[1]; out->Tavg = d[2]; out->Nu_bot = d[7]; out->strain_max = d[8];
    out->empty_cells = (int32_t)d[4]; out->tracer_min = (int32_t)d[5]; out->tracer_max = (int32_t)d[6];
    out->kent = kent;
    out->RaT_avg = integrate2D_fn(s, g_ratavg, NULL) / s->p.aspect;
    double* top = (double*)malloc(sizeof(double) * (nx + 1));
    for (int i = 0; i <= nx; ++i) {
        const size_t c = IX(s, i, nz);
        top[i] = s->u[c] * s->u[c] + s->w[c] * s->w[c];
    }
    out->mobility = sqrt(integrate1Dx(s, top, 1) / s->p.aspect) / d[0];
    free(top);
    const int ci[4] = {0, nx, nx, 0}, ck[4] = {nz, nz, 0, 0};
    for (int j = 0; j < 4; ++j) {
        const size_t c = IX(s, ci[j], ck[j]);
        out->q[j] = -s->conduct[c] * dot3(s->D1, s->T + c, s->ld) / s->ds / s->zs[ck[j] + SPAN];
    }
    if (s->p.comp == 1)
        for (int t = 1; t <= s->ntype; ++t) {
            out->Cavg[t - 1] = integrate2D_fn(s, g_field, s->Cnew[t]) / s->p.aspect;
            const double m = mass ? mass[t - 1] : out->Cavg[t - 1];
            out->ent[t - 1] = integrate_ent(s, s->Cnew[t], kent) / s->p.aspect / m;
        }
}

/* ------------------------------------------------------------------------ */
/* equalize_tracers (composition.f90:83-580)                                 */
/* ------------------------------------------------------------------------ */
/* The reference draws from one Marsaglia mzran stream per OpenMP thread
 * (composition.f90:7-19, seeds basics.f90:198-213) and splits its loops by
 * thread, so its result depends on the thread count.  This restatement keeps
 * every rule of the routine and replaces each draw by a counter-based one:
 * u = hash(seed, call, stream, a, b) with the SplitMix64 finalizer (Steele,
 * Lea and Flood, "Fast splittable pseudorandom number generators", 2014) as
 * the mixing function.  Streams: per-node tracer shuffles (a = node), donor
 * placement (a = tracer index), row and column visit orders (a = pass).  The
 * two donor shuffles (composition.f90:307-315, 353-359) become an order by a
 * keyed bijection of the tracer index; the per-thread donor lists become one
 * list per type (a single thread of the reference).  The node index is
 * c = k*(nx+1) + i, the tracer index the caller's (0-based). */

#define EQ_GOLD 0x9e3779b97f4a7c15ULL
enum { EQ_S_NODE = 1, EQ_S_PLACE = 2, EQ_S_ROWS = 3, EQ_S_COLS = 4 };

static inline uint64_t eq_mix(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}
static inline uint64_t eq_key(uint64_t seed, uint64_t call) { return eq_mix(seed + EQ_GOLD * (call + 1)); }
static inline double eq_u(uint64_t key, unsigned stream, uint64_t a, uint64_t b)
{
    uint64_t x = eq_mix(key ^ ((uint64_t)stream << 48) ^ a);
    x = eq_mix(x + EQ_GOLD * (b + 1));
    return (double)(x >> 11) * 0x1.0p-53;
}
/* a bijection of [0, 2^40): xor, odd multiplications and right xorshifts,
 * each invertible modulo 2^40 */
static inline uint64_t eq_p40(uint64_t key, uint64_t id)
{
    const uint64_t M = (1ULL << 40) - 1;
    uint64_t x = (id ^ key) & M;
    x = (x * 0x9e3779b97f4a7c15ULL) & M;
    x ^= x >> 20;
    x = (x * 0xbf58476d1ce4e5b9ULL) & M;
    x ^= x >> 17;
    x = (x * 0x94d049bb133111ebULL) & M;
    x ^= x >> 21;
    return x;
}
/* shuffle_serial (composition.f90:22-33): element j swaps with a uniformly
 * drawn later element, never itself (Sattolo's algorithm: one cycle) */
static void eq_shuffle(int32_t* a, int n, uint64_t key, unsigned stream, uint64_t id)
{
    for (int j = 0; j + 1 < n; ++j) {
        const double u = eq_u(key, stream, id, (uint64_t)j);
        const int r = (int)floor(u * (double)(n - 1 - j)) + j + 1;
        const int32_t t = a[j];
        a[j] = a[r];
        a[r] = t;
    }
}

double or_eq_uniform(uint64_t seed, uint64_t call, unsigned stream, uint64_t a, uint64_t b)
{
    return eq_u(eq_key(seed, call), stream, a, b);
}
uint64_t or_eq_perm40(uint64_t seed, uint64_t call, uint64_t id) { return eq_p40(eq_key(seed, call), id); }
void or_eq_shuffle(int32_t* a, int n, uint64_t seed, uint64_t call, unsigned stream, uint64_t id)
{
    eq_shuffle(a, n, eq_key(seed, call), stream, id);
}

void or_set_equalize(or_state* s, int on, int tpc, const double* Cinit, uint64_t seed)
{
    s->eq_on = on;
    s->eq_tpc = tpc;
    for (int t = 0; t < s->ntype; ++t) s->eq_Cinit[t] = Cinit ? Cinit[t] : 0.0;
    s->eq_seed = seed;
    s->eq_calls = 0;
}

/* minexcess of the lottery loop (composition.f90:387-423, mode 0), of the
 * second loop (:484-500, mode 1: its enriched types test tpc_type(0)), and
 * after the moves (:547-554, mode 2: no tpc_type test).  Fortran's minval of
 * an empty selection is huge(0). */
static void eq_minexcess(const or_state* s, int nt, const int* ne, const int* tpc_type, int mode, int* mx)
{
    const int nx = s->nx, nz = s->nz;
    for (int p = 0; p <= nt; ++p) {
        const int gate = mode == 2 ? 1 : (mode == 1 ? tpc_type[0] > 0 : tpc_type[p] > 0);
        const int gate0 = mode == 2 ? 1 : tpc_type[0] > 0;
        if (!(p == 0 ? gate0 : gate)) { mx[p] = 0; continue; }
        int m = 2147483647;
        for (int k = 0; k <= nz; ++k)
            for (int i = 0; i <= nx; ++i) {
                const int v = ne[k * (nx + 1) + i];
                if (v > 0) continue;
                const size_t c = IX(s, i, k);
                int sel;
                if (s->p.comp == 0) sel = 1;
                else if (p == 0) {
                    double acc = 0.0;
                    for (int t = 1; t <= s->ntype; ++t) acc += s->Cnew[t][c];
                    sel = acc == 0.0;
                } else sel = s->Cnew[p][c] == 1.0;
                if (sel && v < m) m = v;
            }
        mx[p] = m;
    }
}

/* first type of largest count at node c (maxloc, composition.f90:438) */
static int eq_type_save(const or_state* s, int nt, size_t c)
{
    int best = 0;
    for (int t = 1; t <= nt; ++t)
        if (s->count[t][c] > s->count[best][c]) best = t;
    return best;
}

typedef struct { uint64_t key; int32_t id; } eq_donor;
static int eq_cmp(const void* a, const void* b)
{
    const uint64_t x = ((const eq_donor*)a)->key, y = ((const eq_donor*)b)->key;
    return x < y ? -1 : (x > y);
}

int or_equalize_tracers(or_state* s, int tpc, const double* Cinit, uint64_t seed, uint64_t call,
                        tc_equalize_info* info)
{
    const int nx = s->nx, nz = s->nz, nt = s->p.comp == 1 ? s->ntype : 0;
    const int64_t ntr = s->ntr;
    const size_t V = (size_t)(nx + 1) * (size_t)(nz + 1);
    const uint64_t key = eq_key(seed, call);
    memset(info, 0, sizeof(*info));
    /* Nideal (composition.f90:103-111) */
    int64_t Nideal[TC_MAX_TYPES + 1] = {0};
    if (s->p.comp == 1) {
        int64_t acc = 0;
        for (int t = 1; t <= nt; ++t) { Nideal[t] = llround(Cinit[t - 1] * (double)ntr); acc += Nideal[t]; }
        Nideal[0] = ntr - acc;
    } else Nideal[0] = ntr;
    /* tpc_type (composition.f90:114-140) */
    int tpc_type[TC_MAX_TYPES + 1] = {0};
    if (s->p.comp == 0) tpc_type[0] = tpc;
    else
        for (int t = 0; t <= nt; ++t) {
            int64_t n = 0, sum = 0;
            for (int k = 0; k <= nz; ++k)
                for (int i = 0; i <= nx; ++i) {
                    const size_t c = IX(s, i, k);
                    int sel;
                    if (t == 0) {
                        double acc = 0.0;
                        for (int q = 1; q <= s->ntype; ++q) acc += s->Cnew[q][c];
                        sel = acc == 0.0;
                    } else sel = s->Cnew[t][c] == 1.0;
                    if (sel) { n++; sum += s->count[t][c]; }
                }
            tpc_type[t] = n > 0 ? (int)lround((double)sum / (double)n) : 0;
        }
    for (int t = 0; t <= nt; ++t) info->tpc_type[t] = tpc_type[t];
    /* nexcess (composition.f90:118-161) */
    int* ne = (int*)calloc(V, sizeof(int));
    int64_t nfree = 0;
    int any = 0;
    for (int k = 0; k <= nz; ++k)
        for (int i = 0; i <= nx; ++i) {
            const size_t c = IX(s, i, k);
            const int n0 = s->nempty[c];
            int v = 0;
            if (s->p.comp == 0) {
                if (n0 > 0) v = n0 - tpc;
            } else if (n0 > 0) {
                for (int t = 0; t <= nt; ++t) {
                    int all = 1;
                    for (int dk = -1; dk <= 1 && all; ++dk)
                        for (int di = -1; di <= 1; ++di)
                            if (s->count[t][IX(s, i + di, k + dk)] != s->nempty[IX(s, i + di, k + dk)]) { all = 0; break; }
                    if (all) { v = n0 - tpc_type[t]; break; }
                }
            }
            ne[k * (nx + 1) + i] = v;
            if (v > 0) nfree += v;
            if (v != 0) any = 1;
        }
    info->nfree = nfree;
    if (!any) { free(ne); return 0; }   /* composition.f90:166-171 */
    info->did_work = 1;
    /* tracers of each node in index order (composition.f90:195-205), then
     * shuffled (:209-217) */
    int32_t* cnt = (int32_t*)calloc(V + 1, sizeof(int32_t));
    int32_t* list = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ntr > 0 ? ntr : 1));
    int32_t* node = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ntr > 0 ? ntr : 1));
    for (int64_t id = 0; id < ntr; ++id) {
        const double r = s->rtr[id], sv = s->str[id];
        if (!(r > 0.0 && r < s->p.aspect && sv > 0.0 && sv < 1.0)) {   /* composition.f90:709-712 */
            free(ne); free(cnt); free(list); free(node);
            s->status = TC_ERR_TRACER_DOMAIN;
            return s->status;
        }
        const int i = (int)lround(r / s->dr), k = (int)lround(sv / s->ds);
        node[id] = k * (nx + 1) + i;
        cnt[node[id] + 1]++;
    }
    for (size_t c = 0; c < V; ++c) cnt[c + 1] += cnt[c];
    {
        int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * V);
        memcpy(fill, cnt, sizeof(int32_t) * V);
        for (int64_t id = 0; id < ntr; ++id) list[fill[node[id]]++] = (int32_t)id;
        free(fill);
    }
    /* donors (composition.f90:229-335): wall nodes give their tracers within
     * dist of a wall first (segment 0), then the rest (segment 1); interior
     * nodes their first nexcess tracers */
    eq_donor* dn = (eq_donor*)malloc(sizeof(eq_donor) * (size_t)(nfree > 0 ? nfree : 1));
    int64_t nd = 0;
    const double dr_ = s->dist_r, ds_ = s->dist_s, asp = s->p.aspect;
    for (int k = 0; k <= nz; ++k)
        for (int i = 0; i <= nx; ++i) {
            const size_t c = (size_t)k * (nx + 1) + i;
            int want = ne[c];
            if (want <= 0) continue;
            int32_t* L = list + cnt[c];
            const int n = cnt[c + 1] - cnt[c];
            eq_shuffle(L, n, key, EQ_S_NODE, c);
            const int edge = i == 0 || i == nx || k == 0 || k == nz;
            for (int pass = edge ? 0 : 1; pass < 2 && want > 0; ++pass)
                for (int q = 0; q < n && want > 0; ++q) {
                    const int32_t id = L[q];
                    const double r = s->rtr[id], sv = s->str[id];
                    const int near = sv < ds_ || sv > 1.0 - ds_ || r < dr_ || r > asp - dr_;
                    if (edge && near != (pass == 0)) continue;
                    const uint64_t bucket = (uint64_t)s->Ttype[id] * 2 + (uint64_t)pass;
                    dn[nd].key = (bucket << 40) | eq_p40(key, (uint64_t)id);
                    dn[nd].id = id;
                    nd++;
                    want--;
                }
        }
    free(cnt); free(list); free(node);
    if (nd != nfree) {   /* composition.f90:344-347 */
        free(ne); free(dn);
        s->status = TC_ERR_ARG;
        return TC_ERR_ARG;
    }
    qsort(dn, (size_t)nd, sizeof(eq_donor), eq_cmp);
    int64_t nfree_t[TC_MAX_TYPES + 1] = {0}, off[TC_MAX_TYPES + 2] = {0};
    for (int64_t j = 0; j < nd; ++j) nfree_t[(dn[j].key >> 40) / 2]++;
    for (int t = 0; t <= nt; ++t) off[t + 1] = off[t] + nfree_t[t];
    /* lottery (composition.f90:362-463), then the second loop (:466-541) */
    int32_t* rows = (int32_t*)malloc(sizeof(int32_t) * (nx + 1));
    int32_t* cols = (int32_t*)malloc(sizeof(int32_t) * (nz + 1));
    for (int i = 0; i <= nx; ++i) rows[i] = i;
    for (int k = 0; k <= nz; ++k) cols[k] = k;
    int64_t used[TC_MAX_TYPES + 1] = {0}, tcount[TC_MAX_TYPES + 1];
    int mx[TC_MAX_TYPES + 1];
    uint64_t pass = 0;
    int escape = 0;
#define EQ_TCOUNT()                                                                  \
    do {                                                                             \
        escape = 1;                                                                  \
        for (int t = 0; t <= nt; ++t) {                                              \
            int64_t a = 0;                                                           \
            for (int k = 0; k <= nz; ++k)                                            \
                for (int i = 0; i <= nx; ++i) a += s->count[t][IX(s, i, k)];         \
            tcount[t] = a;                                                           \
            if (a != Nideal[t]) escape = 0;                                          \
        }                                                                            \
    } while (0)
    EQ_TCOUNT();
    for (int loop = 0; loop < 2; ++loop)
        for (;;) {
            if (loop == 0 && !escape) EQ_TCOUNT();
            eq_minexcess(s, nt, ne, tpc_type, loop, mx);
            int allz = 1;
            int64_t su = 0;
            for (int t = 0; t <= nt; ++t) { if (mx[t] != 0) allz = 0; su += used[t]; }
            if (allz || su >= nfree) break;
            eq_shuffle(rows, nx + 1, key, EQ_S_ROWS, pass);
            eq_shuffle(cols, nz + 1, key, EQ_S_COLS, pass);
            pass++;
            int tg = 0;
            if (loop == 0 && !escape)
                for (int p = 0; p <= nt; ++p)
                    if (tcount[p] > Nideal[p]) { tg = p; break; }
            int moved = 0;
            for (int ii = 0; ii <= nx; ++ii)
                for (int kk = 0; kk <= nz; ++kk) {
                    const int i = rows[ii], k = cols[kk];
                    const size_t c = IX(s, i, k), cc = (size_t)k * (nx + 1) + i;
                    const int ts = eq_type_save(s, nt, c);
                    const int tt = (loop == 0 && !escape) ? tg : ts;
                    if (!(ne[cc] == mx[ts] && used[tt] < nfree_t[tt] && ne[cc] < 0)) continue;
                    const int32_t id = dn[off[tt] + used[tt]].id;
                    used[tt]++;
                    moved = 1;
                    const double pr = (eq_u(key, EQ_S_PLACE, (uint64_t)id, 0) - 0.5) * s->dr;
                    const double ps = (eq_u(key, EQ_S_PLACE, (uint64_t)id, 1) - 0.5) * s->ds;
                    double r = s->dr * (double)i + pr, sv = s->ds * (double)k + ps;
                    if (r <= 0.0) r = dr_;
                    if (r >= asp) r = asp - dr_;
                    if (sv <= 0.0) sv = ds_;
                    if (sv >= 1.0) sv = 1.0 - ds_;
                    s->rtr[id] = r;
                    s->str[id] = sv;
                    if (s->p.comp == 1) s->Ttype[id] = ts;
                    if (loop == 0) {
                        s->count[tt][c] -= 1;
                        s->count[s->Ttype[id]][c] += 1;
                    }
                    ne[cc] += 1;
                }
            if (!moved) break;
        }
#undef EQ_TCOUNT
    info->passes = (int32_t)pass;
    eq_minexcess(s, nt, ne, tpc_type, 2, mx);
    for (int t = 0; t <= nt; ++t) {
        info->minexcess[t] = mx[t];
        info->moved += used[t];
        int64_t a = 0;
        for (int k = 0; k <= nz; ++k)
            for (int i = 0; i <= nx; ++i) a += s->count[t][IX(s, i, k)];
        info->count[t] = a;
    }
    for (size_t c = 0; c < V; ++c)
        if (ne[c] <= 0) info->deficit += ne[c];
    free(rows); free(cols); free(ne); free(dn);
    /* tracers_to_corners without the C filter (composition.f90:576-578) */
    const int sv_save = s->smooth_vis;
    s->smooth_vis = 0;
    const int st = or_tracers_to_corners(s, TC_TR_X);
    s->smooth_vis = sv_save;
    return st;
}
