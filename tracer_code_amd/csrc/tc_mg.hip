// Multigrid kernels of the plain FP64 path: the Jacobi-Richardson sweep of
// iterate_stream, compute_residual, restrict, prolong, compute_RHS and the
// viscosity hierarchy (compute_viscosity, smoother_vis, viscosity_gradients,
// test_eigenvalues, iterate_local_time_steps).  Reference:
// iterative_solver.f90 and smoother.f90 of seantrim/tracer_code.
#include "tc_kernels.h"

namespace tc {

// ---------------------------------------------------------------------------
// iterate_stream sweep (iterative_solver.f90:871-912) fused with the ghost
// part of enforceBCs_stream.  RATIO selects the per-point Chebyshev factor
// computed from ratio(level) (iterative_solver.f90:880-885); otherwise the
// factor is the scalar cconst (1 on the upward leg, 2/(3-temp) when the
// level's ratio is the reset value 1e99).
// ---------------------------------------------------------------------------
template <bool POW2, bool RATIO>
__global__ __launch_bounds__(256) void k_sweep(const double* __restrict__ S0, double* __restrict__ E,
                                               const double* __restrict__ vis, const double* __restrict__ vx,
                                               const double* __restrict__ vz, const double* __restrict__ vxx,
                                               const double* __restrict__ vzz, const double* __restrict__ vxz,
                                               const double* __restrict__ rhs, const double* __restrict__ dts,
                                               const double* __restrict__ ratio, Grid g, FD F, Spacing h, int ivis,
                                               double temp, double cconst, double zsg)
{
    const int2 tile = xcd_tile();
    const int i = tile.x * blockDim.x + threadIdx.x + 1;
    const int k = tile.y * blockDim.y + threadIdx.y + 1;
    if (i > g.nx - 1 || k > g.nz - 1) return;
    const size_t c = g.at(i, k);
    double t = stream_space<POW2>(F, h, ivis, S0, c, g.ld, vis[c], vx[c], vz[c], vxx[c], vzz[c], vxz[c], rhs[c]);
    double cc;
    if constexpr (RATIO) {
        const double la = -1.0;
        double r = ratio[c], lb;
        if (r > 1e99) lb = -2.0;
        else lb = -2.0 * (r * r) / (1.0 + r * r);
        cc = 2.0 / (-lb - la + (lb - la) * temp);
    } else {
        cc = cconst;
    }
    double v = S0[c] + cc * dts[c] * t;
    E[c] = v;
    stream_ghosts(E, g, i, k, v, zsg);
}

// Full enforceBCs_stream on one level, one workgroup (boundary work only).
__global__ void k_bc_stream(double* __restrict__ S, Grid g, int Vbc)
{
    const int nx = g.nx, nz = g.nz;
    for (int k = -SPAN + (int)threadIdx.x; k <= nz + SPAN; k += blockDim.x) { S[g.at(0, k)] = 0.0; S[g.at(nx, k)] = 0.0; }
    for (int i = -SPAN + (int)threadIdx.x; i <= nx + SPAN; i += blockDim.x) { S[g.at(i, 0)] = 0.0; S[g.at(i, nz)] = 0.0; }
    __syncthreads();
    for (int k = threadIdx.x; k <= nz; k += blockDim.x)
        for (int i = 1; i <= SPAN; ++i) {
            S[g.at(-i, k)] = -S[g.at(i, k)];
            S[g.at(nx + i, k)] = -S[g.at(nx - i, k)];
        }
    __syncthreads();
    for (int ii = -SPAN + (int)threadIdx.x; ii <= nx + SPAN; ii += blockDim.x)
        for (int i = 1; i <= SPAN; ++i) {
            if (Vbc == 1) {
                S[g.at(ii, -i)] = S[g.at(ii, i)];
                S[g.at(ii, nz + i)] = S[g.at(ii, nz - i)];
            } else {
                S[g.at(ii, -i)] = -S[g.at(ii, i)];
                S[g.at(ii, nz + i)] = -S[g.at(ii, nz - i)];
            }
        }
}

// compute_residual (iterative_solver.f90:258-279): residual = RHS - L(S);
// optionally per-block partial sums of |residual| for residual_mag.
template <bool POW2>
__global__ __launch_bounds__(256) void k_residual(const double* __restrict__ S, double* __restrict__ R,
                                                  const double* __restrict__ vis, const double* __restrict__ vx,
                                                  const double* __restrict__ vz, const double* __restrict__ vxx,
                                                  const double* __restrict__ vzz, const double* __restrict__ vxz,
                                                  const double* __restrict__ rhs, Grid g, FD F, Spacing h, int ivis,
                                                  double* __restrict__ partial)
{
    const int2 tile = xcd_tile();
    const int i = tile.x * blockDim.x + threadIdx.x + 1;
    const int k = tile.y * blockDim.y + threadIdx.y + 1;
    double a = 0.0;
    if (i <= g.nx - 1 && k <= g.nz - 1) {
        const size_t c = g.at(i, k);
        double r = stream_space<POW2>(F, h, ivis, S, c, g.ld, vis[c], vx[c], vz[c], vxx[c], vzz[c], vxz[c], rhs[c]);
        R[c] = r;
        a = fabs(r);
    }
    if (partial) {
        a = block_sum(a);
        if (threadIdx.x == 0 && threadIdx.y == 0) partial[tile.y * gridDim.x + tile.x] = a;
    }
}

// restrict / restrict_coefficients (iterative_solver.f90:596-664), with the
// reference's corner weights (ii-1,kk+1 twice, ii+1,kk-1 absent).
__global__ __launch_bounds__(256) void k_restrict(const double* __restrict__ in, Grid gi, double* __restrict__ out, Grid go)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > go.nx - 1 || k > go.nz - 1) return;
    const int ii = 2 * i, kk = 2 * k;
    double corners = (in[gi.at(ii - 1, kk - 1)] + in[gi.at(ii - 1, kk + 1)] + in[gi.at(ii + 1, kk + 1)] +
                      in[gi.at(ii - 1, kk + 1)]) / 16.0;
    double sides = (in[gi.at(ii - 1, kk)] + in[gi.at(ii, kk - 1)] + in[gi.at(ii + 1, kk)] + in[gi.at(ii, kk + 1)]) / 8.0;
    double centre = in[gi.at(ii, kk)] / 4.0;
    out[go.at(i, k)] = corners + sides + centre;
}

// prolong (iterative_solver.f90:175-233): one thread per fine interior
// point; the four parity classes of the reference's loops are disjoint.
// Fused with the ghost part of enforceBCs_stream(lf).
__global__ __launch_bounds__(256) void k_prolong(double* __restrict__ Ef, Grid gf, const double* __restrict__ Ec, Grid gc,
                                                 double zsg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > gf.nx - 1 || k > gf.nz - 1) return;
    const size_t c = gf.at(i, k);
    double add;
    const bool ie = (i % 2) == 0, ke = (k % 2) == 0;
    if (ie && ke) add = Ec[gc.at(i / 2, k / 2)];
    else if (ie) add = (Ec[gc.at(i / 2, (k - 1) / 2)] + Ec[gc.at(i / 2, (k + 1) / 2)]) / 2.0;
    else if (ke) add = (Ec[gc.at((i - 1) / 2, k / 2)] + Ec[gc.at((i + 1) / 2, k / 2)]) / 2.0;
    else
        add = (Ec[gc.at((i - 1) / 2, (k - 1) / 2)] + Ec[gc.at((i + 1) / 2, (k - 1) / 2)] +
               Ec[gc.at((i + 1) / 2, (k + 1) / 2)] + Ec[gc.at((i - 1) / 2, (k + 1) / 2)]) / 4.0;
    double v = Ef[c] + add;
    Ef[c] = v;
    stream_ghosts(Ef, gf, i, k, v, zsg);
}

// compute_RHS(1) (iterative_solver.f90:571-590) + per-block |RHS| sums
// for compute_reference_residual.
__global__ __launch_bounds__(256) void k_rhs1(const double* __restrict__ Tb, PtrArr Cb,
                                              double* __restrict__ rhs, Grid g, FD F, double dr, double RaT, int comp,
                                              int ntype, RaCArr RaC, double* __restrict__ partial)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
    double a = 0.0;
    if (i <= g.nx - 1 && k <= g.nz - 1) {
        const size_t c = g.at(i, k);
        double r = 0.0;
        if (RaT != 0.0) {
            double Tr = dot3(F.D1, Tb, c, 1) / dr;
            r = RaT * Tr;
        }
        if (comp == 1)
            for (int j = 0; j < ntype; ++j) {
                double Cr = dot3(F.D1, Cb.v[j], c, 1) / dr;
                r = r - RaC.v[j] * Cr;
            }
        rhs[c] = r;
        a = fabs(r);
    }
    a = block_sum(a);
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = a;
}

// Sum of |f| over the interior of a level (compute_reference_residual).
__global__ __launch_bounds__(256) void k_abs_partial(const double* __restrict__ f, Grid g, double* __restrict__ partial)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
    double a = 0.0;
    if (i <= g.nx - 1 && k <= g.nz - 1) a = fabs(f[g.at(i, k)]);
    a = block_sum(a);
    if (threadIdx.x == 0 && threadIdx.y == 0) partial[blockIdx.y * gridDim.x + blockIdx.x] = a;
}

// Deterministic final reduction of n partials by one workgroup:
// op 0 = sum, 1 = min, 2 = max.  Writes out[slot].
__global__ __launch_bounds__(256) void k_reduce_final(const double* __restrict__ partial, int n, int op,
                                                      double* __restrict__ out, int slot)
{
    double a = (op == 0) ? 0.0 : (op == 1 ? INFINITY : -INFINITY);
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        double v = partial[t];
        if (op == 0) a += v;
        else if (op == 1) a = v < a ? v : a;
        else a = v > a ? v : a;
    }
    if (op == 0) a = block_sum(a);
    else if (op == 1) a = block_min(a);
    else a = block_max(a);
    if (threadIdx.x == 0) out[slot] = a;
}

// ---------------------------------------------------------------------------
// Viscosity hierarchy
// ---------------------------------------------------------------------------

// visf (momentum.f90:1-30) on the ghosted level-1 grid
// (compute_viscosity, iterative_solver.f90:340-346).
__global__ __launch_bounds__(256) void k_visf(double* __restrict__ vis, const double* __restrict__ Tvis,
                                              PtrArr Cvis, const double* __restrict__ zg,
                                              Grid g, int ivis, int comp, int ntype, double RaT, double lvisT,
                                              double lvisP, RaCArr lvisC)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x - SPAN;
    const int k = blockIdx.y * blockDim.y + threadIdx.y - SPAN;
    if (i > g.nx + SPAN || k > g.nz + SPAN) return;
    const size_t c = g.at(i, k);
    const double zgk = zg[k + 3];
    double v = 1.0;
    if (ivis == 1) {
        v = exp(-Tvis[c] * lvisT + (1.0 - zgk) * lvisP);
    } else if (ivis == 2 && comp == 1) {
        double dp = 0.0;
        for (int t = 0; t < ntype; ++t) dp += Cvis.v[t][c] * lvisC.v[t];
        if (RaT != 0.0) v = exp(-Tvis[c] * lvisT - dp + (1.0 - zgk) * lvisP);
        else v = exp(-dp + (1.0 - zgk) * lvisP);
    }
    vis[c] = v;
}

// The loop control of smoother_vis (smoother.f90:165-263) for one iteration
// of a loop that has not stopped, from the minimum a of the eigenvalue
// ratios: the exit test of the pass, and either the stop (ctl[VC_DONE],
// with no sweep: ctl[VC_NOSWEEP]) or the Chebyshev factor of this
// iteration's sweep (dctl[VD_C]) and the advance to the next iteration, with
// the pass bookkeeping at the end of a pass.
__device__ void vis_decide(double a, const VisDecide& P, int32_t* __restrict__ ctl, double* __restrict__ dctl)
{
    const double rmin = (P.ivis == 0) ? 1e99 : a;
    ctl[VC_ITERS] += 1;
    const int nn = ctl[VC_NN], n = ctl[VC_N], Nvs = ctl[VC_NVIS];
    bool trig;
    if (P.option == 1) trig = (Nvs >= 1 && rmin > P.eig_ratio_min) || P.ivis == 0 || n > 20;
    else if (P.interpolate == 0) trig = rmin > P.eig_ratio || P.ivis == 0;
    else trig = Nvs >= 1 || P.ivis == 0;
    if (trig) {
        ctl[VC_DONE] = 1;
        ctl[VC_NOSWEEP] = 1;
        if (P.option == 1) {
            dctl[VD_EIG] = rmin;
            ctl[VC_NETA] = n;
        } else if (P.interpolate == 0) {
            ctl[VC_NNSTORE] = nn;
        }
        return;
    }
    dctl[VD_C] = (ctl[VC_TAB] == 0) ? P.c0[nn - 1] : P.c1[nn - 1];
    ctl[VC_NOSWEEP] = 0;
    int next = nn + 1;
    if (next > ctl[VC_NITER]) {  // end of the pass
        ctl[VC_NVIS] = Nvs + 1;
        ctl[VC_NNSTORE] = 1;
        if (P.option == 1) {
            ctl[VC_NITER] = 6;
            ctl[VC_TAB] = 1;
        }
        ctl[VC_N] = n + 1;
        next = 1;
        if (n + 1 > 100000) {
            ctl[VC_ERR] = 1;
            ctl[VC_DONE] = 1;
        }
    }
    ctl[VC_NN] = next;
}

// viscosity_gradients + test_eigenvalues (iterative_solver.f90:290-329,
// 666-698, eigI 750-803, eigR 806-857) on uniform grids, with per-block
// minima of the ratio for smoother_vis (smoother.f90:209).  gate: the
// device-controlled smoother_vis loop has stopped; copy_dst: the plane V is
// also copied there (the source of the iteration's sweep, which the loop's
// decision may still switch off; the copy is then unused); store = 0: the gradient and
// ratio planes are not written (inside the loop only the block minima are
// read, and the planes are written from the final viscosity after it);
// slots: each block's minimum also goes into slot blk % VIS_SLOTS by atomicMin
// (the loop's decision reads the minimum of the slots in the next launch).
template <bool POW2>
__global__ __launch_bounds__(256) void k_vis_grad_eig(const double* __restrict__ V, double* __restrict__ vx,
                                                      double* __restrict__ vz, double* __restrict__ vxx,
                                                      double* __restrict__ vzz, double* __restrict__ vxz,
                                                      double* __restrict__ ratio, Grid g, FD F, Spacing h, int ivis,
                                                      EigScalars e, int want_ratio, double* __restrict__ partial,
                                                      const int32_t* __restrict__ gate,
                                                      double* __restrict__ copy_dst, int store,
                                                      unsigned long long* __restrict__ slots)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const bool in = i <= g.nx - 1 && k <= g.nz - 1;
    // the point's neighbourhood, loaded first so that it arrives with the
    // gate and the copy's values (every thread: one beyond the interior
    // reads the neighbourhood of point (1,1) and discards it)
    const Reg9 S = load9(V, g.at(in ? i : 1, in ? k : 1), g.ld);
    if (gate && gate[0]) return;  // the device-controlled loop has ended
    if (copy_dst) {
        const size_t nt = (size_t)gridDim.x * gridDim.y * blockDim.x * blockDim.y;
        for (size_t j = ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * (blockDim.x * blockDim.y) +
                        threadIdx.y * blockDim.x + threadIdx.x;
             j < g.elems; j += nt)
            copy_dst[j] = V[j];
    }
    double m = INFINITY;
    if (in) {
        const size_t c = g.at(i, k);
        double gx = 0.0, gz = 0.0, gxx = 0.0, gzz = 0.0, gxz = 0.0;
        if (ivis != 0) {
            gx = sdiv<POW2>(dot3g<0>(F.D1, S), h.dr, h.idr);
            gxx = sdiv<POW2>(dot3g<0>(F.D2, S), h.dr2, h.idr2);
            gz = sdiv<POW2>(dot3g<1>(F.D1, S), h.ds, h.ids);
            gzz = sdiv<POW2>(dot3g<1>(F.D2, S), h.ds2, h.ids2);
            gxz = sdiv<POW2>(sum9g(F.Drs, S), h.drds, h.idrds);
        }
        if (store) { vx[c] = gx; vz[c] = gz; vxx[c] = gxx; vzz[c] = gzz; vxz[c] = gxz; }
        if (want_ratio && ivis != 0) {
            const double vis = S(0, 0);
            // eigI with xr=zs=1, xrr=..=0: only the odd-derivative terms survive
            double t1 = (gxx + gzz) * (0.0 + 0.0);
            double t2 = vis * (0.0 + 0.0);
            double t3 = 2.0 * vis * 0.0;
            double t4 = 4.0 * gxz * 0.0;
            double t5 = 2.0 * gx * (e.Trrr_max + e.Trss_max);
            double t6 = 2.0 * gz * (e.Tsss_max + e.Trrs_max);
            double eI = fabs(t1) + fabs(t2) + fabs(t3) + fabs(t4) + fabs(t5) + fabs(t6);
            double r1 = (gxx - gzz) * (e.Trr_max - e.Tss_max);
            double r2 = vis * (e.Trrrr_max + e.Tssss_max);
            double r3 = 2.0 * vis * e.Trrss_max;
            double r4 = 4.0 * gxz * e.Trs_max;
            double r5 = 2.0 * gx * (0.0 + 0.0);
            double r6 = 2.0 * gz * (0.0 + 0.0);
            double eR = -(r1 + r2 + r3 + r4 + r5 + r6);
            double r = fabs(eR / eI);
            if (store) ratio[c] = r;
            m = r;
        }
    }
    if (partial || slots) {
        m = block_min(m);
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            const unsigned blk = blockIdx.y * gridDim.x + blockIdx.x;
            if (partial) partial[blk] = m;
            // the ratio is >= 0 (or NaN, skipped as by the comparison minimum):
            // its bit pattern orders as the value
            if (slots) atomicMin(slots + (size_t)(blk % VIS_SLOTS) * VIS_SLOT_STRIDE, (unsigned long long)__double_as_longlong(m));
        }
    }
}

// iterate_local_time_steps (iterative_solver.f90:700-748), iterate_order=1.
__global__ __launch_bounds__(256) void k_local_dt(const double* __restrict__ V, const double* __restrict__ vx,
                                                  const double* __restrict__ vz, const double* __restrict__ vxx,
                                                  const double* __restrict__ vzz, const double* __restrict__ vxz,
                                                  double* __restrict__ dts, Grid g, EigScalars e, double courant_stream)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > g.nx - 1 || k > g.nz - 1) return;
    const size_t c = g.at(i, k);
    const double vis = V[c], gx = vx[c], gz = vz[c], gxx = vxx[c], gzz = vzz[c], gxz = vxz[c];
    double t1 = (gxx + gzz) * (0.0 + 0.0);
    double t2 = vis * (0.0 + 0.0);
    double t3 = 2.0 * vis * 0.0;
    double t4 = 4.0 * gxz * 0.0;
    double t5 = 2.0 * gx * (e.Trrr_max + e.Trss_max);
    double t6 = 2.0 * gz * (e.Tsss_max + e.Trrs_max);
    double eI = fabs(t1) + fabs(t2) + fabs(t3) + fabs(t4) + fabs(t5) + fabs(t6);
    double r1 = (gxx - gzz) * (e.Trr_max - e.Tss_max);
    double r2 = vis * (e.Trrrr_max + e.Tssss_max);
    double r3 = 2.0 * vis * e.Trrss_max;
    double r4 = 4.0 * gxz * e.Trs_max;
    double r5 = 2.0 * gx * (0.0 + 0.0);
    double r6 = 2.0 * gz * (0.0 + 0.0);
    double eR = -(r1 + r2 + r3 + r4 + r5 + r6);
    dts[c] = -courant_stream * 2.0 * eR / (eR * eR + eI * eI);
}

// One smoother_vis sweep (smoother.f90:239-248) from the copy T0 into V on
// i=0..nx, k=kmin..kmax.
template <bool POW2>
__global__ __launch_bounds__(256) void k_vis_sweep(const double* __restrict__ T0, double* __restrict__ V,
                                                   const double* __restrict__ dtv, Grid g, FD F, Spacing h, int kmin,
                                                   int kmax, double c_arg, const double* __restrict__ cptr,
                                                   const int32_t* __restrict__ gate)
{
    if (gate && gate[0]) return;
    const double c = cptr ? cptr[0] : c_arg;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y * blockDim.y + threadIdx.y + kmin;
    if (i > g.nx || k > kmax) return;
    const size_t cc = g.at(i, k);
    double Trr = sdiv<POW2>(dot3(F.D2, T0, cc, 1), h.dr2, h.idr2);
    double Tss = sdiv<POW2>(dot3(F.D2, T0, cc, g.ld), h.ds2, h.ids2);
    double temp = Trr + Tss;
    double dts = dtv[cc] * 0.99;
    V[cc] = T0[cc] + c * dts * temp;
}

// viscosityBCs / viscosityBCs_initialize (iterative_solver.f90:379-453).
// The routine runs three dependent phases: (1) on coarser levels, rows 0
// and nz (and with initialize the columns 0 and nx) from the finer level;
// (2) the side ghost columns mirrored for k=0..nz; (3) the bottom/top ghost
// rows from rows 1..SPAN and nz-1..nz-SPAN over all columns.  Each output
// cell here is computed straight from the state before the call, by
// following its chain of sources back through the phases (val1 = the value
// after phase 1), so one thread per boundary cell needs no barrier; the
// arithmetic is the reference's.
__device__ __forceinline__ double visc_val1(const double* __restrict__ V, const Grid& g,
                                            const double* __restrict__ Vf, const Grid& gf, int level,
                                            int initialize, int i, int k)
{
    if (level > 1) {
        if ((k == 0 || k == g.nz) && i >= 0 && i <= g.nx) return Vf[gf.at(2 * i, k == 0 ? 0 : gf.nz)];
        if (initialize && (i == 0 || i == g.nx) && k >= 1 && k <= g.nz - 1)
            return Vf[gf.at(i == 0 ? 0 : gf.nx, 2 * k)];
    }
    return V[g.at(i, k)];
}

__global__ __launch_bounds__(256) void k_visc_bcs(double* __restrict__ V, Grid g, const double* __restrict__ Vf,
                                                  Grid gf, int level, int initialize, double RaT, double visT2,
                                                  const int32_t* __restrict__ gate)
{
    if (gate && gate[0]) return;
    const int nx = g.nx, nz = g.nz;
    const int nA = (level > 1) ? 2 * (nx + 1) + (initialize ? 2 * (nz - 1) : 0) : 0;  // phase 1 cells
    const int nB = 2 * SPAN * (nz + 1);                                                // side ghosts
    const int nC = 2 * SPAN * (nx + 2 * SPAN + 1);                                     // row ghosts
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    auto v1 = [&](int i, int k) { return visc_val1(V, g, Vf, gf, level, initialize, i, k); };
    if (e < nA) {
        int i, k;
        if (e < 2 * (nx + 1)) { i = e % (nx + 1); k = (e < nx + 1) ? 0 : nz; }
        else { const int f = e - 2 * (nx + 1); k = 1 + f % (nz - 1); i = (f < nz - 1) ? 0 : nx; }
        V[g.at(i, k)] = v1(i, k);
        return;
    }
    e -= nA;
    if (e < nB) {
        const int k = e % (nz + 1), f = e / (nz + 1);  // f = 0..2*SPAN-1
        const int j = 1 + f % SPAN;
        if (f < SPAN) V[g.at(-j, k)] = v1(j, k);
        else V[g.at(nx + j, k)] = v1(nx - j, k);
        return;
    }
    e -= nB;
    if (e < nC) {
        const int w = nx + 2 * SPAN + 1;
        const int i = -SPAN + e % w, f = e / w;  // f = 0..2*SPAN-1
        const int kk = 1 + f % SPAN;
        // the value after phase 2 at (i, kr): ghost columns are mirrors
        auto v2 = [&](int kr) {
            const int im = (i < 0) ? -i : (i > nx ? 2 * nx - i : i);
            return v1(im, kr);
        };
        if (f < SPAN) V[g.at(i, -kk)] = (RaT != 0.0) ? 1.0 / v2(kk) / visT2 : v2(kk);
        else V[g.at(i, nz + kk)] = (RaT != 0.0) ? 1.0 / v2(nz - kk) : v2(nz - kk);
    }
}

// One iteration's sweep of smoother_vis (k_vis_sweep) with the
// viscosityBCs that follow it (k_visc_bcs with initialize = 0) in one launch.
// Thread (i,k), i = 0..nx, k = 0..nz: its value after the sweep (rows
// kmin..kmax; the other rows keep theirs), after phase 1 of the boundary
// routine (rows 0 and nz of a coarser level from the finer level), and then
// the cells the routine fills from it: the side ghosts of columns 1..SPAN
// and nx-SPAN..nx-1 (phase 2) and, from rows 1..SPAN and nz-SPAN..nz-1, the
// bottom and top ghost rows in its own column and in its mirror ghost
// column (phase 3).  Every cell k_visc_bcs writes is written here by the one
// thread that owns its source, with the same arithmetic.
//
// The iteration's decision is made here too (VisLoop): every block takes the
// minimum of the slots k_vis_grad_eig filled and runs vis_decide on a copy of
// the loop state of this iteration (ctl_in, dctl_in); all blocks reach the
// same decision.  Block 0 writes the new state to the other buffer (ctl_out,
// dctl_out: the buffers alternate with the iteration's parity, so no block
// reads a state another block writes) and resets the slot set of the next
// iteration.  Once the loop has stopped, the state is only carried over.
template <bool POW2>
__global__ __launch_bounds__(256) void k_vis_sweep_bcs(const double* __restrict__ T0, double* __restrict__ V,
                                                       const double* __restrict__ dtv, Grid g, FD F, Spacing h,
                                                       int kmin, int kmax, VisLoopArgs lp,
                                                       const double* __restrict__ Vf, Grid gf, int level, double RaT,
                                                       double visT2)
{
    __shared__ int s_sweep;
    __shared__ double s_c;
    static_assert(VIS_SLOTS == 64, "one slot per lane of the first wave (blockDim.x = 64)");
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y * blockDim.y + threadIdx.y;
    const int nx = g.nx, nz = g.nz;
    const bool act = i <= nx && k <= nz;
    // the sweep's values, loaded before the decision so that they arrive with
    // the slots (every thread: one beyond the grid reads around point (0,0)
    // and discards it)
    const size_t cc = g.at(act ? i : 0, act ? k : 0);
    const double t0 = T0[cc], tm = T0[cc - 1], tp = T0[cc + 1], tdm = T0[cc - g.ld], tdp = T0[cc + g.ld];
    const double dv = dtv[cc];
    // rows 0 and nz of a coarser level take the finer level's value
    const bool fromf = act && level > 1 && (k == 0 || k == nz);
    const double* __restrict__ fsrc = Vf ? Vf : T0;
    const double vf = fsrc[fromf ? gf.at(2 * i, k == 0 ? 0 : gf.nz) : cc];
    if (threadIdx.y == 0) {
        // the minimum of the slots, a slot each lane; an unset slot (NaN bit
        // pattern) and a NaN ratio count as +inf, as the comparison minimum
        // of the block minima skips them
        double a = __longlong_as_double((long long)lp.slots_in[(size_t)threadIdx.x * VIS_SLOT_STRIDE]);
        if (!(a < INFINITY)) a = INFINITY;
        for (int o = 32; o > 0; o >>= 1) {
            const double w = __shfl_xor(a, o, 64);
            a = w < a ? w : a;
        }
        if (threadIdx.x == 0) {
            int32_t st[VC_COUNT];
            double dst[VD_COUNT];
            for (int q = 0; q < VC_COUNT; ++q) st[q] = lp.ctl_in[q];
            for (int q = 0; q < VD_COUNT; ++q) dst[q] = lp.dctl_in[q];
            if (st[VC_DONE]) {
                st[VC_NOSWEEP] = 1;
            } else {
                vis_decide(a, lp.P, st, dst);
            }
            s_sweep = !st[VC_NOSWEEP];
            s_c = dst[VD_C];
            if (blockIdx.x == 0 && blockIdx.y == 0) {
                for (int q = 0; q < VC_COUNT; ++q) lp.ctl_out[q] = st[q];
                for (int q = 0; q < VD_COUNT; ++q) lp.dctl_out[q] = dst[q];
                for (int q = 0; q < VIS_SLOTS; ++q) lp.slots_reset[(size_t)q * VIS_SLOT_STRIDE] = ~0ull;
            }
        }
    }
    __syncthreads();
    if (!s_sweep) return;
    if (!act) return;
    double v;
    if (k >= kmin && k <= kmax) {
        const double c = s_c;
        // dot3 along each axis, on the loaded values
        double ar = F.D2[1] * tm;
        ar += F.D2[2] * t0;
        ar += F.D2[3] * tp;
        double as = F.D2[1] * tdm;
        as += F.D2[2] * t0;
        as += F.D2[3] * tdp;
        double Trr = sdiv<POW2>(ar, h.dr2, h.idr2);
        double Tss = sdiv<POW2>(as, h.ds2, h.ids2);
        double temp = Trr + Tss;
        double dts = dv * 0.99;
        v = t0 + c * dts * temp;
    } else {
        v = t0;
    }
    if (fromf) v = vf;
    V[cc] = v;
    const bool lo_i = i >= 1 && i <= SPAN, hi_i = i >= nx - SPAN && i <= nx - 1;
    if (lo_i) V[g.at(-i, k)] = v;
    if (hi_i) V[g.at(2 * nx - i, k)] = v;
    if (k >= 1 && k <= SPAN) {
        const double b = (RaT != 0.0) ? 1.0 / v / visT2 : v;
        V[g.at(i, -k)] = b;
        if (lo_i) V[g.at(-i, -k)] = b;
        if (hi_i) V[g.at(2 * nx - i, -k)] = b;
    }
    if (k >= nz - SPAN && k <= nz - 1) {
        const double t = (RaT != 0.0) ? 1.0 / v : v;
        V[g.at(i, 2 * nz - k)] = t;
        if (lo_i) V[g.at(-i, 2 * nz - k)] = t;
        if (hi_i) V[g.at(2 * nx - i, 2 * nz - k)] = t;
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
static inline dim3 grid2(int nx, int ny, dim3 b) { return dim3((nx + b.x - 1) / b.x, (ny + b.y - 1) / b.y); }

void launch_sweep(hipStream_t st, bool pow2, bool use_ratio, const Level& L, const double* S0, double* E, const FD& F,
                  const Spacing& h, int ivis, double temp, double cconst, double zsg)
{
    dim3 b(64, 4);
    dim3 gr = grid2(L.g.nx - 1, L.g.nz - 1, b);
    if (pow2) {
        if (use_ratio)
            hipLaunchKernelGGL((k_sweep<true, true>), gr, b, 0, st, S0, E, L.vis, L.visx, L.visz, L.visxx, L.viszz, L.visxz,
                               L.rhs, L.dtstream, L.ratio, L.g, F, h, ivis, temp, cconst, zsg);
        else
            hipLaunchKernelGGL((k_sweep<true, false>), gr, b, 0, st, S0, E, L.vis, L.visx, L.visz, L.visxx, L.viszz,
                               L.visxz, L.rhs, L.dtstream, L.ratio, L.g, F, h, ivis, temp, cconst, zsg);
    } else {
        if (use_ratio)
            hipLaunchKernelGGL((k_sweep<false, true>), gr, b, 0, st, S0, E, L.vis, L.visx, L.visz, L.visxx, L.viszz,
                               L.visxz, L.rhs, L.dtstream, L.ratio, L.g, F, h, ivis, temp, cconst, zsg);
        else
            hipLaunchKernelGGL((k_sweep<false, false>), gr, b, 0, st, S0, E, L.vis, L.visx, L.visz, L.visxx, L.viszz,
                               L.visxz, L.rhs, L.dtstream, L.ratio, L.g, F, h, ivis, temp, cconst, zsg);
    }
}

void launch_bc_stream(hipStream_t st, double* S, const Grid& g, int Vbc)
{
    hipLaunchKernelGGL(k_bc_stream, dim3(1), dim3(256), 0, st, S, g, Vbc);
}

int launch_residual(hipStream_t st, bool pow2, const Level& L, const double* S, double* R, const FD& F,
                    const Spacing& h, int ivis, double* partial)
{
    dim3 b(64, 4);
    dim3 gr = grid2(L.g.nx - 1, L.g.nz - 1, b);
    if (pow2)
        hipLaunchKernelGGL((k_residual<true>), gr, b, 0, st, S, R, L.vis, L.visx, L.visz, L.visxx, L.viszz, L.visxz,
                           L.rhs, L.g, F, h, ivis, partial);
    else
        hipLaunchKernelGGL((k_residual<false>), gr, b, 0, st, S, R, L.vis, L.visx, L.visz, L.visxx, L.viszz, L.visxz,
                           L.rhs, L.g, F, h, ivis, partial);
    return (int)(gr.x * gr.y);
}

void launch_restrict(hipStream_t st, const double* in, const Grid& gi, double* out, const Grid& go)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_restrict, grid2(go.nx - 1, go.nz - 1, b), b, 0, st, in, gi, out, go);
}

void launch_prolong(hipStream_t st, double* Ef, const Grid& gf, const double* Ec, const Grid& gc, double zsg)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_prolong, grid2(gf.nx - 1, gf.nz - 1, b), b, 0, st, Ef, gf, Ec, gc, zsg);
}

int launch_rhs1(hipStream_t st, const double* Tb, const PtrArr& Cb, double* rhs, const Grid& g, const FD& F,
                double dr, double RaT, int comp, int ntype, const RaCArr& RaC, double* partial)
{
    dim3 b(64, 4);
    dim3 gr = grid2(g.nx - 1, g.nz - 1, b);
    hipLaunchKernelGGL(k_rhs1, gr, b, 0, st, Tb, Cb, rhs, g, F, dr, RaT, comp, ntype, RaC, partial);
    return (int)(gr.x * gr.y);
}

int launch_abs_partial(hipStream_t st, const double* f, const Grid& g, double* partial)
{
    dim3 b(64, 4);
    dim3 gr = grid2(g.nx - 1, g.nz - 1, b);
    hipLaunchKernelGGL(k_abs_partial, gr, b, 0, st, f, g, partial);
    return (int)(gr.x * gr.y);
}

void launch_reduce_final(hipStream_t st, const double* partial, int n, int op, double* out, int slot)
{
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, partial, n, op, out, slot);
}

void launch_visf(hipStream_t st, double* vis, const double* Tvis, const PtrArr& Cvis, const double* zg,
                 const Grid& g, int ivis, int comp, int ntype, double RaT, double lvisT, double lvisP, const RaCArr& lvisC)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_visf, grid2(g.nx + 2 * SPAN + 1, g.nz + 2 * SPAN + 1, b), b, 0, st, vis, Tvis, Cvis, zg, g, ivis,
                       comp, ntype, RaT, lvisT, lvisP, lvisC);
}

int launch_vis_grad_eig(hipStream_t st, bool pow2, Level& L, const FD& F, const Spacing& h, int ivis,
                        const EigScalars& e, int want_ratio, double* partial, const int32_t* gate,
                        double* copy_dst, bool store, unsigned long long* slots)
{
    dim3 b(64, 4);
    dim3 gr = grid2(L.g.nx - 1, L.g.nz - 1, b);
    if (pow2)
        hipLaunchKernelGGL((k_vis_grad_eig<true>), gr, b, 0, st, L.vis, L.visx, L.visz, L.visxx, L.viszz, L.visxz,
                           L.ratio, L.g, F, h, ivis, e, want_ratio, partial, gate, copy_dst, store ? 1 : 0, slots);
    else
        hipLaunchKernelGGL((k_vis_grad_eig<false>), gr, b, 0, st, L.vis, L.visx, L.visz, L.visxx, L.viszz, L.visxz,
                           L.ratio, L.g, F, h, ivis, e, want_ratio, partial, gate, copy_dst, store ? 1 : 0, slots);
    return (int)(gr.x * gr.y);
}

void launch_local_dt(hipStream_t st, Level& L, const EigScalars& e, double courant_stream)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_local_dt, grid2(L.g.nx - 1, L.g.nz - 1, b), b, 0, st, L.vis, L.visx, L.visz, L.visxx, L.viszz,
                       L.visxz, L.dtstream, L.g, e, courant_stream);
}

void launch_vis_sweep(hipStream_t st, bool pow2, const double* T0, double* V, const double* dtv, const Grid& g,
                      const FD& F, const Spacing& h, int kmin, int kmax, double c, const double* cptr,
                      const int32_t* gate)
{
    dim3 b(64, 4);
    dim3 gr = grid2(g.nx + 1, kmax - kmin + 1, b);
    if (pow2)
        hipLaunchKernelGGL((k_vis_sweep<true>), gr, b, 0, st, T0, V, dtv, g, F, h, kmin, kmax, c, cptr, gate);
    else
        hipLaunchKernelGGL((k_vis_sweep<false>), gr, b, 0, st, T0, V, dtv, g, F, h, kmin, kmax, c, cptr, gate);
}

void launch_vis_sweep_bcs(hipStream_t st, bool pow2, const double* T0, double* V, const double* dtv, const Grid& g,
                          const FD& F, const Spacing& h, int kmin, int kmax, const VisLoopArgs& lp,
                          const double* Vf, const Grid& gf, int level, double RaT, double visT2)
{
    dim3 b(64, 4);
    dim3 gr = grid2(g.nx + 1, g.nz + 1, b);
    if (pow2)
        hipLaunchKernelGGL((k_vis_sweep_bcs<true>), gr, b, 0, st, T0, V, dtv, g, F, h, kmin, kmax, lp, Vf, gf, level,
                           RaT, visT2);
    else
        hipLaunchKernelGGL((k_vis_sweep_bcs<false>), gr, b, 0, st, T0, V, dtv, g, F, h, kmin, kmax, lp, Vf, gf, level,
                           RaT, visT2);
}

void launch_visc_bcs(hipStream_t st, double* V, const Grid& g, const double* Vf, const Grid& gf, int level, int init,
                     double RaT, double visT2, const int32_t* gate)
{
    const int n = ((level > 1) ? 2 * (g.nx + 1) + (init ? 2 * (g.nz - 1) : 0) : 0) + 2 * SPAN * (g.nz + 1) +
                  2 * SPAN * (g.nx + 2 * SPAN + 1);
    hipLaunchKernelGGL(k_visc_bcs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, V, g, Vf, gf, level, init, RaT,
                       visT2, gate);
}




}  // namespace tc
