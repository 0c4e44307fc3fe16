// Defect-correction multigrid (TC_MODE_MIXED and TC_MODE_DC64).
//
// At 2048x1024 the FP64 rounding of the fourth differences of an FP64
// streamfunction leaves a relative residual of ~8e-8, above the 1e-8
// tolerance.  The defect-correction form keeps psi as a double-double
// (hi + lo), evaluates the fine-grid residual RHS - L(psi) with error-free
// transformations, and computes each correction with one W-cycle of the
// reference's schedule in FP32 (MIXED) or FP64 (DC64) started from zero.  In
// exact arithmetic a cycle is the same affine map as the reference's cycle
// with zeroed coarse corrections; only the rounding differs.  The coarsest
// grid is solved in FP64 in both forms.
#include "tc_kernels.h"
#include "tc_stencil32.h"

namespace tc {

// 1/h^n of the stencil terms as double-double reciprocals: hi = fl(1/x),
// lo = fl((1 - hi*x) / x) with the residual 1 - hi*x exact by FMA
enum { IH_R2, IH_S2, IH_R3, IH_S3, IH_R4, IH_S4, IH_RS, IH_RS2, IH_R2S, IH_R2S2, IH_COUNT };
struct InvSpacing {
    double hi[IH_COUNT], lo[IH_COUNT];
};
static InvSpacing inv_spacing(const Spacing& h)
{
    const double x[IH_COUNT] = {h.dr2, h.ds2, h.dr3, h.ds3, h.dr4, h.ds4, h.drds, h.drds2, h.dr2ds, h.dr2ds2};
    InvSpacing r;
    for (int q = 0; q < IH_COUNT; ++q) {
        r.hi[q] = 1.0 / x[q];
        r.lo[q] = std::fma(-r.hi[q], x[q], 1.0) / x[q];
    }
    return r;
}

// ---- double-double helpers (Knuth TwoSum, FMA TwoProd) ---------------------
struct dd {
    double hi, lo;
};
__device__ __forceinline__ dd two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
__device__ __forceinline__ dd quick_two_sum(double a, double b)
{
    double s = a + b;
    return {s, b - (s - a)};
}
__device__ __forceinline__ dd dd_add(dd a, dd b)
{
    dd s = two_sum(a.hi, b.hi);
    dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ dd dd_mul_d(dd a, double b)
{
    double p = a.hi * b;
    double e = fma(a.hi, b, -p);
    e = fma(a.lo, b, e);
    return quick_two_sum(p, e);
}
__device__ __forceinline__ dd dd_neg(dd a) { return {-a.hi, -a.lo}; }
// Compensated stencil sums.  A finite difference of psi = hi + lo cancels
// to ~h^4 of |psi| (the FP64 floor of the reference form); here every tap's
// product w*hi is split exactly (FMA TwoProd), the products are summed with
// TwoSum, and the rounding errors, the product errors and the w*lo terms
// are gathered in a second double.  The result (s, e) is the difference to
// ~eps^2 * sum|w*hi|, i.e. double-double accuracy, at a third of the
// instructions of a double-double accumulation.
struct CSum {
    double s = 0.0, e = 0.0;
};
__device__ __forceinline__ void csum_tap(CSum& a, double w, double h, double l)
{
    const double p = w * h;
    const double pe = fma(w, h, -p);
    const double s = a.s + p;
    const double bb = s - a.s;
    const double t = (a.s - (s - bb)) + (p - bb);
    a.s = s;
    a.e = fma(w, l, a.e + (t + pe));
}
__device__ __forceinline__ dd csum_dot(const double* D, int j0, int j1, const double* __restrict__ hi,
                                       const double* __restrict__ lo, size_t c, ptrdiff_t st)
{
    CSum a;
    for (int j = j0; j <= j1; ++j) {
        const double w = D[j + 2];
        if (w == 0.0) continue;
        const size_t q = c + j * st;
        csum_tap(a, w, hi[q], lo[q]);
    }
    return quick_two_sum(a.s, a.e);
}
__device__ __forceinline__ dd csum_9(const double* M, const double* __restrict__ hi, const double* __restrict__ lo,
                                     size_t c, ptrdiff_t ld)
{
    CSum a;
    for (int dk = -1; dk <= 1; ++dk)
        for (int di = -1; di <= 1; ++di) {
            const double w = M[(dk + 1) * 3 + (di + 1)];
            if (w == 0.0) continue;
            const size_t q = c + dk * ld + di;
            csum_tap(a, w, hi[q], lo[q]);
        }
    return quick_two_sum(a.s, a.e);
}
// a * b for double-doubles (to ~eps^2)
__device__ __forceinline__ dd dd_mul(dd a, dd b)
{
    const double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p);
    e = fma(a.hi, b.lo, e);
    e = fma(a.lo, b.hi, e);
    return quick_two_sum(p, e);
}

// Fine-grid residual r = RHS - L(hi+lo) in double-double; writes r (FP64)
// and its FP32 rounding, and per-block sums of |r|.  The 1/h^n factors are
// double-double reciprocals (InvSpacing).
constexpr int RD_BX = 64, RD_BY = 4, RD_H = 2;              // block shape (launch_residual_dd), halo
constexpr int RD_W = RD_BX + 2 * RD_H, RD_CELLS = RD_W * (RD_BY + 2 * RD_H);
constexpr int RD_NST = (RD_CELLS + RD_BX * RD_BY - 1) / (RD_BX * RD_BY);  // staged cells per thread

__global__ __launch_bounds__(256) void k_residual_dd(const double* __restrict__ hi, const double* __restrict__ lo,
                                                     const double* __restrict__ vis, const double* __restrict__ vx,
                                                     const double* __restrict__ vz, const double* __restrict__ vxx,
                                                     const double* __restrict__ vzz, const double* __restrict__ vxz,
                                                     const double* __restrict__ rhs, double* __restrict__ r64,
                                                     float* __restrict__ r32, Grid g, FD F, InvSpacing ih, int ivis,
                                                     double* __restrict__ partial)
{
    // the block's 64x4 points and a 2-cell halo of hi and lo staged in LDS:
    // the 62 stencil taps of each point read LDS instead of the L1/TA path
    __shared__ double sh[RD_CELLS], sl[RD_CELLS];
    const int2 tile = xcd_tile();
    const int i = tile.x * RD_BX + threadIdx.x + 1;
    const int k = tile.y * RD_BY + threadIdx.y + 1;
    const bool in = i <= g.nx - 1 && k <= g.nz - 1;
    // the point's coefficients and right-hand side, loaded with the staged
    // cells (every thread, branch-free, as in k_absorb_inc)
    const size_t c = g.at(in ? i : 1, in ? k : 1);
    const double v = vis[c], cvx = vx[c], cvz = vz[c], cvxx = vxx[c], cvzz = vzz[c], cvxz = vxz[c], crhs = rhs[c];
    {
        const int i0 = tile.x * RD_BX + 1 - RD_H, k0 = tile.y * RD_BY + 1 - RD_H;
        double h[RD_NST], o[RD_NST];
#pragma unroll
        for (int u = 0; u < RD_NST; ++u) {
            const int l = threadIdx.y * RD_BX + threadIdx.x + u * RD_BX * RD_BY;
            const int ii = i0 + l % RD_W, kk = k0 + l / RD_W;
            const bool ok = l < RD_CELLS && ii >= -SPAN && ii <= g.nx + SPAN && kk >= -SPAN && kk <= g.nz + SPAN;
            const size_t q = g.at(ok ? ii : 0, ok ? kk : 0);
            const double hv = hi[q], ov = lo[q];
            h[u] = ok ? hv : 0.0;
            o[u] = ok ? ov : 0.0;
        }
#pragma unroll
        for (int u = 0; u < RD_NST; ++u) {
            const int l = threadIdx.y * RD_BX + threadIdx.x + u * RD_BX * RD_BY;
            if (l < RD_CELLS) {
                sh[l] = h[u];
                sl[l] = o[u];
            }
        }
    }
    __syncthreads();
    double a = 0.0;
    if (in) {
        const size_t cl = (size_t)(threadIdx.y + RD_H) * RD_W + threadIdx.x + RD_H;
        const ptrdiff_t ld = RD_W;
        auto sc = [&](dd x, int q) { return dd_mul(x, dd{ih.hi[q], ih.lo[q]}); };
        dd Srrrr = sc(csum_dot(F.D4, -2, 2, sh, sl, cl, 1), IH_R4);
        dd Sssss = sc(csum_dot(F.D4, -2, 2, sh, sl, cl, ld), IH_S4);
        dd Srrss = sc(csum_9(F.Drrss, sh, sl, cl, ld), IH_R2S2);
        dd sum = dd_mul_d(dd_add(Srrrr, Sssss), v);
        sum = dd_add(sum, dd_mul_d(Srrss, 2.0 * v));
        if (ivis != 0) {
            dd Srr = sc(csum_dot(F.D2, -1, 1, sh, sl, cl, 1), IH_R2);
            dd Sss = sc(csum_dot(F.D2, -1, 1, sh, sl, cl, ld), IH_S2);
            dd Srrr = sc(csum_dot(F.D3, -2, 2, sh, sl, cl, 1), IH_R3);
            dd Ssss = sc(csum_dot(F.D3, -2, 2, sh, sl, cl, ld), IH_S3);
            dd Srs = sc(csum_9(F.Drs, sh, sl, cl, ld), IH_RS);
            dd Srss = sc(csum_9(F.Drss, sh, sl, cl, ld), IH_RS2);
            dd Srrs = sc(csum_9(F.Drrs, sh, sl, cl, ld), IH_R2S);
            sum = dd_add(sum, dd_mul_d(dd_add(Srr, dd_neg(Sss)), cvxx - cvzz));
            sum = dd_add(sum, dd_mul_d(Srs, 4.0 * cvxz));
            sum = dd_add(sum, dd_mul_d(dd_add(Srrr, Srss), 2.0 * cvx));
            sum = dd_add(sum, dd_mul_d(dd_add(Ssss, Srrs), 2.0 * cvz));
        }
        dd r = dd_add({crhs, 0.0}, dd_neg(sum));
        const double rv = r.hi + r.lo;
        r64[c] = rv;
        if (r32) r32[c] = (float)rv;
        a = fabs(rv);
    }
    if (partial) {
        a = block_sum(a);
        if (threadIdx.x == 0 && threadIdx.y == 0) partial[tile.y * gridDim.x + tile.x] = a;
    }
}

// psi(hi,lo) += e on the interior with the enforceBCs_stream ghost mirrors;
// the correction e is cleared at the same points (its walls stay zero), so
// the next visit of level 1 starts from a zero correction.
template <typename T>
__global__ __launch_bounds__(256) void k_dd_absorb(double* __restrict__ hi, double* __restrict__ lo, T* __restrict__ e,
                                                   Grid g, double zsg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > g.nx - 1 || k > g.nz - 1) return;
    const size_t c = g.at(i, k);
    dd s = dd_add({hi[c], lo[c]}, {(double)e[c], 0.0});
    hi[c] = s.hi;
    lo[c] = s.lo;
    stream_ghosts(hi, g, i, k, s.hi, zsg);
    stream_ghosts(lo, g, i, k, s.lo, zsg);
    e[c] = T(0);
    stream_ghosts(e, g, i, k, T(0), T(zsg));
}

// The bilinear prolongation of a coarse correction at fine interior (i,k)
// (prolong, iterative_solver.f90:175-233), in FP64 as k_prolong_dd adds it.
template <typename T>
__device__ __forceinline__ double prolong_val(const T* __restrict__ Ec, const Grid& gc, int i, int k)
{
    const bool ie = (i % 2) == 0, ke = (k % 2) == 0;
    if (ie && ke) return (double)Ec[gc.at(i / 2, k / 2)];
    if (ie) return ((double)Ec[gc.at(i / 2, (k - 1) / 2)] + (double)Ec[gc.at(i / 2, (k + 1) / 2)]) / 2.0;
    if (ke) return ((double)Ec[gc.at((i - 1) / 2, k / 2)] + (double)Ec[gc.at((i + 1) / 2, k / 2)]) / 2.0;
    return ((double)Ec[gc.at((i - 1) / 2, (k - 1) / 2)] + (double)Ec[gc.at((i + 1) / 2, (k - 1) / 2)] +
            (double)Ec[gc.at((i + 1) / 2, (k + 1) / 2)] + (double)Ec[gc.at((i - 1) / 2, (k + 1) / 2)]) / 4.0;
}

// prolong_val with its loads outside the branches: the four corners' values
// are loaded at every point (one address four times at an even-even point)
// and combined by the same case formulas, so the result has the same bits
// and a caller's loads can all be in flight together.  (i, k) interior.
template <typename T>
__device__ __forceinline__ double prolong_val_nb(const T* __restrict__ Ec, const Grid& gc, int i, int k)
{
    const bool ie = (i % 2) == 0, ke = (k % 2) == 0;
    const int ia = ie ? i / 2 : (i - 1) / 2, ib = ie ? i / 2 : (i + 1) / 2;
    const int ka = ke ? k / 2 : (k - 1) / 2, kb = ke ? k / 2 : (k + 1) / 2;
    const double aa = (double)Ec[gc.at(ia, ka)], ba = (double)Ec[gc.at(ib, ka)];
    const double ab = (double)Ec[gc.at(ia, kb)], bb = (double)Ec[gc.at(ib, kb)];
    // every case formed and one selected (branches would take the loads
    // into them)
    const double ve = (aa + ab) / 2.0, vk = (aa + ba) / 2.0, vo = (aa + ba + bb + ab) / 4.0;
    return (ie && ke) ? aa : (ie ? ve : (ke ? vk : vo));
}

// psi(hi,lo) += bilinear prolongation of the level-2 correction
// (prolong, iterative_solver.f90:175-233) with the ghost mirrors.  Folding
// the coarse-grid correction into psi before the level-1 post-smoothing keeps
// the smooth part out of the FP32 level-1 correction, whose rounding would
// otherwise leave a high-frequency defect proportional to its magnitude.
template <typename T>
__global__ __launch_bounds__(256) void k_prolong_dd(double* __restrict__ hi, double* __restrict__ lo, Grid gf,
                                                    const T* __restrict__ Ec, Grid gc, double zsg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > gf.nx - 1 || k > gf.nz - 1) return;
    const size_t c = gf.at(i, k);
    const double add = prolong_val(Ec, gc, i, k);
    dd s = dd_add({hi[c], lo[c]}, {add, 0.0});
    hi[c] = s.hi;
    lo[c] = s.lo;
    stream_ghosts(hi, gf, i, k, s.hi, zsg);
    stream_ghosts(lo, gf, i, k, s.lo, zsg);
}

// The defect kept up to date by increments: psi(hi,lo) += e (as k_dd_absorb
// or k_prolong_dd) and r -= L(e), so r stays RHS - L(psi) without the
// double-double residual of the whole psi.  L(e) is the reference-form FP64
// operator (stream_space, as k_residual, with the FP64 coefficients): its
// rounding is that of the FP64 residual of psi scaled by |e|/|psi|, far
// below the tolerance for a correction; the solve confirms convergence on
// the exact residual (k_residual_dd) before it stops.  e is staged in LDS
// with the residual's 2-cell halo.  PRO = false: e is the level-1 correction
// (with the ghosts the sweeps wrote); PRO = true: the bilinear prolongation of
// the level-2 correction Ec, formed at the staged interior cells, 0 on the
// walls and the enforceBCs_stream mirrors at the ghosts.  The kernel does not
// clear e: other workgroups read it as their halo.  With partial, the block
// sums of |r| for the stop test.
template <typename T, bool PRO, bool POW2>
__global__ __launch_bounds__(256) void k_absorb_inc(double* __restrict__ hi, double* __restrict__ lo,
                                                    const T* __restrict__ e, Grid gc, const double* __restrict__ vis,
                                                    const double* __restrict__ vx, const double* __restrict__ vz,
                                                    const double* __restrict__ vxx, const double* __restrict__ vzz,
                                                    const double* __restrict__ vxz, double* __restrict__ r64,
                                                    float* __restrict__ r32, Grid g, FD F, Spacing h, int ivis,
                                                    double zsg, double* __restrict__ partial)
{
    __shared__ double se[RD_CELLS];
    const int2 tile = xcd_tile();
    const int i = tile.x * RD_BX + threadIdx.x + 1;
    const int k = tile.y * RD_BY + threadIdx.y + 1;
    const bool in = i <= g.nx - 1 && k <= g.nz - 1;
    // the point's own values, loaded before the staging so that their
    // latency overlaps the halo's.  Every thread loads (a thread beyond the
    // interior reads point (1,1) and discards it): behind a branch the
    // compiler folds the operator's coefficient products into it and waits
    // for the loads there.
    const size_t c = g.at(in ? i : 1, in ? k : 1);
    const double ph = hi[c], pl = lo[c], pr = r64[c];
    const double cv[6] = {vis[c], vx[c], vz[c], vxx[c], vzz[c], vxz[c]};
    {
        // the staged cells of this thread, all loaded before the first is
        // stored (a rolled loop waits for each load in turn)
        const int i0 = tile.x * RD_BX + 1 - RD_H, k0 = tile.y * RD_BY + 1 - RD_H;
        double v[RD_NST];
#pragma unroll
        for (int u = 0; u < RD_NST; ++u) {
            const int l = threadIdx.y * RD_BX + threadIdx.x + u * RD_BX * RD_BY;
            const int ii = i0 + l % RD_W, kk = k0 + l / RD_W;
            // branch-free loads: a cell that takes no value reads a valid
            // address and is set to zero after
            if constexpr (PRO) {
                int mi = ii, mk = kk;
                double sg = 1.0;
                if (mi < 0) { mi = -mi; sg = -sg; } else if (mi > g.nx) { mi = 2 * g.nx - mi; sg = -sg; }
                if (mk < 0) { mk = -mk; sg = sg * zsg; } else if (mk > g.nz) { mk = 2 * g.nz - mk; sg = sg * zsg; }
                const bool ok = l < RD_CELLS && mi >= 1 && mi <= g.nx - 1 && mk >= 1 && mk <= g.nz - 1;
                const double pv = prolong_val_nb(e, gc, ok ? mi : 1, ok ? mk : 1);
                v[u] = ok ? sg * pv : 0.0;
            } else {
                const bool ok = l < RD_CELLS && ii >= -SPAN && ii <= g.nx + SPAN && kk >= -SPAN && kk <= g.nz + SPAN;
                const double ev = (double)e[g.at(ok ? ii : 0, ok ? kk : 0)];
                v[u] = ok ? ev : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < RD_NST; ++u) {
            const int l = threadIdx.y * RD_BX + threadIdx.x + u * RD_BX * RD_BY;
            if (l < RD_CELLS) se[l] = v[u];
        }
    }
    __syncthreads();
    double a = 0.0;
    if (in) {
        const size_t cl = (size_t)(threadIdx.y + RD_H) * RD_W + threadIdx.x + RD_H;
        const dd s = dd_add({ph, pl}, {se[cl], 0.0});
        hi[c] = s.hi;
        lo[c] = s.lo;
        stream_ghosts(hi, g, i, k, s.hi, zsg);
        stream_ghosts(lo, g, i, k, s.lo, zsg);
        // 0 - L(e)
        const double le =
            stream_space_g<POW2>(F, h, ivis, PlaneAt<double>{se, cl, RD_W}, cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], 0.0);
        const double rv = pr + le;
        r64[c] = rv;
        if (r32) r32[c] = (float)rv;
        a = fabs(rv);
    }
    if (partial) {
        a = block_sum(a);
        if (threadIdx.x == 0 && threadIdx.y == 0) partial[tile.y * gridDim.x + tile.x] = a;
    }
}

// ---- FP32 inner cycle (stencil in tc_stencil32.h) ---------------------------
#pragma clang fp contract(fast)
template <bool RATIO>
__global__ __launch_bounds__(256) void k_sweep32(const float* __restrict__ S0, float* __restrict__ E,
                                                 const float* __restrict__ vis, const float* __restrict__ vx,
                                                 const float* __restrict__ vz, const float* __restrict__ a1,
                                                 const float* __restrict__ vxz, const float* __restrict__ rhs,
                                                 const float* __restrict__ dts, const float* __restrict__ lb32, Grid g,
                                                 FD32 F, H32 h, float temp, float cconst, float zsg)
{
    const int2 tile = xcd_tile();
    const int i = tile.x * blockDim.x + threadIdx.x + 1;
    const int k = tile.y * blockDim.y + threadIdx.y + 1;
    if (i > g.nx - 1 || k > g.nz - 1) return;
    const size_t c = g.at(i, k);
    float t = stream_space32<float>(F, h, S0, c, g.ld, vis[c], vx[c], vz[c], a1[c], vxz[c], rhs[c]);
    float cc = cconst;
    if constexpr (RATIO) {
        const float lb = lb32[c];
        cc = 2.0f / (-lb + 1.0f + (lb + 1.0f) * temp);
    }
    float v = S0[c] + cc * dts[c] * t;
    E[c] = v;
    stream_ghosts(E, g, i, k, v, zsg);
}

__global__ __launch_bounds__(256) void k_residual32(const float* __restrict__ S, float* __restrict__ R,
                                                    const float* __restrict__ vis, const float* __restrict__ vx,
                                                    const float* __restrict__ vz, const float* __restrict__ a1,
                                                    const float* __restrict__ vxz, const float* __restrict__ rhs,
                                                    Grid g, FD32 F, H32 h)
{
    const int2 tile = xcd_tile();
    const int i = tile.x * blockDim.x + threadIdx.x + 1;
    const int k = tile.y * blockDim.y + threadIdx.y + 1;
    if (i > g.nx - 1 || k > g.nz - 1) return;
    const size_t c = g.at(i, k);
    R[c] = (float)stream_space32<double>(F, h, S, c, g.ld, vis[c], vx[c], vz[c], a1[c], vxz[c], rhs[c]);
}

__global__ __launch_bounds__(256) void k_restrict32(const float* __restrict__ in, Grid gi, float* __restrict__ out,
                                                    Grid go)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > go.nx - 1 || k > go.nz - 1) return;
    const int ii = 2 * i, kk = 2 * k;
    float corners = (in[gi.at(ii - 1, kk - 1)] + in[gi.at(ii - 1, kk + 1)] + in[gi.at(ii + 1, kk + 1)] +
                     in[gi.at(ii - 1, kk + 1)]) * (1.0f / 16.0f);
    float sides = (in[gi.at(ii - 1, kk)] + in[gi.at(ii, kk - 1)] + in[gi.at(ii + 1, kk)] + in[gi.at(ii, kk + 1)]) *
                  (1.0f / 8.0f);
    out[go.at(i, k)] = corners + sides + in[gi.at(ii, kk)] * 0.25f;
}

// compute_residual of fine level Lf fused with restrict into level Lf+1:
// one thread per coarse interior point evaluates the 8 fine residuals the
// restriction weights (the reference's corner set, iterative_solver.f90:
// 604-609, reads (ii-1,kk+1) twice and not (ii+1,kk-1)), each rounded to FP32
// exactly as k_residual32 stores it, and combines them in k_restrict32's
// order -- the same values as the two kernels, without the residual plane.
// With bout (the coarsest level), also the equilibrated right-hand side of
// the coarse solve at the interior unknowns, as k_coarse_rhs32 forms it.
__global__ __launch_bounds__(256) void k_residual_restrict32(const float* __restrict__ S, const float* __restrict__ vis,
                                                             const float* __restrict__ vx, const float* __restrict__ vz,
                                                             const float* __restrict__ a1,
                                                             const float* __restrict__ vxz,
                                                             const float* __restrict__ rhs, Grid gi, FD32 F, H32 h,
                                                             float* __restrict__ out, Grid go,
                                                             const double* __restrict__ RowB, double* __restrict__ bout)
{
    const int2 tile = xcd_tile();
    const int i = tile.x * blockDim.x + threadIdx.x + 1;
    const int k = tile.y * blockDim.y + threadIdx.y + 1;
    const bool in = i <= go.nx - 1 && k <= go.nz - 1;
    // every value the eight residuals read, loaded before the first of them
    // (computed one after another from memory, each waited for its own
    // loads): the 7x7 fine neighbourhood and the residual points'
    // coefficients.  A thread beyond the interior loads around coarse point
    // (1,1) and stores nothing.
    const int ii = 2 * (in ? i : 1), kk = 2 * (in ? k : 1);
    float P[49];
#pragma unroll
    for (int dk = -3; dk <= 3; ++dk)
#pragma unroll
        for (int di = -3; di <= 3; ++di) P[(dk + 3) * 7 + di + 3] = S[gi.at(ii + di, kk + dk)];
    // the residual points, in the order the restriction takes them
    constexpr int RA[8] = {-1, -1, 1, -1, 0, 1, 0, 0}, RB[8] = {-1, 1, 1, 0, -1, 0, 1, 0};
    float cf[8][6];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const size_t c = gi.at(ii + RA[t], kk + RB[t]);
        cf[t][0] = vis[c]; cf[t][1] = vx[c]; cf[t][2] = vz[c]; cf[t][3] = a1[c]; cf[t][4] = vxz[c]; cf[t][5] = rhs[c];
    }
    if (!in) return;
    float rv[8];
#pragma unroll
    for (int t = 0; t < 8; ++t)
        rv[t] = (float)stream_space32_g<double>(F, h, Patch7At{P, RA[t], RB[t]}, cf[t][0], cf[t][1], cf[t][2], cf[t][3],
                                                cf[t][4], cf[t][5]);
    const float v = restrict_fw32(rv[0], rv[1], rv[2], rv[3], rv[4], rv[5], rv[6], rv[7]);
    out[go.at(i, k)] = v;
    if (bout) {
        const int q = i * (go.nz + 1) + k;
        bout[q] = (double)v * RowB[q];
    }
}

__global__ __launch_bounds__(256) void k_prolong32(float* __restrict__ Ef, Grid gf, const float* __restrict__ Ec,
                                                   Grid gc, float zsg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > gf.nx - 1 || k > gf.nz - 1) return;
    const size_t c = gf.at(i, k);
    const float v = __fadd_rn(Ef[c], prolong_add32(Ec, gc, i, k));
    Ef[c] = v;
    stream_ghosts(Ef, gf, i, k, v, zsg);
}
#pragma clang fp contract(off)

// The equilibrated right-hand side of a coarse solve, b[q] = rhs * RowB at
// the interior unknowns (q = i*(nz+1) + k) and 0 at the walls, formed once so
// that the GEMV below streams only the inverse.
__global__ __launch_bounds__(256) void k_coarse_rhs32(const double* __restrict__ RowB, const float* __restrict__ rhs,
                                                      Grid g, int n, double* __restrict__ b)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int N = g.nz + 1, kq = q % N, iq = q / N;
    b[q] = (iq > 0 && iq < g.nx && kq > 0 && kq < g.nz) ? (double)rhs[g.at(iq, kq)] * RowB[q] : 0.0;
}

// Coarse solve of the FP32 correction cycle: the explicit inverse rounded to
// FP32 (half the bytes of the GEMV; the cycle's corrections are FP32), FP64
// right-hand side and accumulation, result rounded to FP32: one workgroup
// per unknown, its row of the inverse times b.  J > 0: n <= 256 J, and a
// thread loads its J chunks' values before the first product (a rolled loop
// waits for each chunk's loads in turn); the products are summed in the
// loop's order, so the result has the same bits.  J = 0: the loop.
template <int J>
__global__ __launch_bounds__(256) void k_coarse_apply32(const float* __restrict__ Ainv, const double* __restrict__ b,
                                                        float* __restrict__ E, Grid g, int n, float zsg)
{
    const int N = g.nz + 1;
    const int ip = blockIdx.x + 1, kp = blockIdx.y + 1;
    const float* row = Ainv + (size_t)(N * ip + kp) * (size_t)n;
    double acc = 0.0;
    if constexpr (J > 0) {
        float a[J];
        double x[J];
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int q = threadIdx.x + 256 * j;
            a[j] = row[q < n ? q : 0];
            x[j] = b[q < n ? q : 0];
        }
#pragma unroll
        for (int j = 0; j < J; ++j)
            if (threadIdx.x + 256 * j < n) acc += (double)a[j] * x[j];
    } else {
        for (int q = threadIdx.x; q < n; q += blockDim.x) acc += (double)row[q] * b[q];
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) {
        const float v = (float)acc;
        E[g.at(ip, kp)] = v;
        stream_ghosts(E, g, ip, kp, v, zsg);
    }
}

// FP64 -> FP32 coefficient copies (a1 = vis_xx - vis_zz; lb from the ratio
// as in iterate_stream, iterative_solver.f90:880-884).
__global__ __launch_bounds__(256) void k_coeffs_to_f32(const double* __restrict__ vis, const double* __restrict__ vx,
                                                       const double* __restrict__ vz, const double* __restrict__ vxx,
                                                       const double* __restrict__ vzz, const double* __restrict__ vxz,
                                                       const double* __restrict__ dts, const double* __restrict__ ratio,
                                                       float* __restrict__ vis32, float* __restrict__ vx32,
                                                       float* __restrict__ vz32, float* __restrict__ a132,
                                                       float* __restrict__ vxz32, float* __restrict__ dts32,
                                                       float* __restrict__ lb32, Grid g, int ivis)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > g.nx - 1 || k > g.nz - 1) return;
    const size_t c = g.at(i, k);
    vis32[c] = (float)vis[c];
    vx32[c] = ivis ? (float)vx[c] : 0.0f;
    vz32[c] = ivis ? (float)vz[c] : 0.0f;
    a132[c] = ivis ? (float)(vxx[c] - vzz[c]) : 0.0f;
    vxz32[c] = ivis ? (float)vxz[c] : 0.0f;
    dts32[c] = (float)dts[c];
    if (lb32) {
        const double r = ratio[c];
        lb32[c] = (float)((r > 1e99) ? -2.0 : -2.0 * (r * r) / (1.0 + r * r));
    }
}

// out = a + c * (a - b) over whole planes: linear extrapolation in time of
// two streamfunctions; the ghost mirrors of enforceBCs_stream are linear,
// so the result keeps them
__global__ __launch_bounds__(256) void k_extrap(const double* __restrict__ a, const double* __restrict__ b, double c,
                                                double* __restrict__ out, size_t n)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = fma(c, a[e] - b[e], a[e]);
}

// out = a + (b - c) over whole planes: the stage-increment guess (plain start
// plus a difference of two solutions); linear, so it keeps the ghost mirrors
__global__ __launch_bounds__(256) void k_add_diff(const double* __restrict__ a, const double* __restrict__ b,
                                                  const double* __restrict__ c, double* __restrict__ out, size_t n)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = a[e] + (b[e] - c[e]);
}

// ---- GCR over the defect-correction cycles ----------------------------------
// One cycle maps psi_s (with residual r_s) to psi' (residual r'); its
// direction is z = psi' - psi_s and the operator applied to it is
// q = r_s - r', from the two residuals the cycles compute anyway.  Both are
// kept in FP64: the fourth-order operator would turn an FP32 rounding of z
// into a high-frequency defect far above the tolerance, and q enters the
// orthogonalisation, whose cancellation would amplify an FP32 rounding of it.
// Interior points only: the update writes psi's ghosts by the mirrors of the
// boundary routine.
//
// k_gcr_dots: the per-block partial sums over the interior of (q, Q_j) and
// (r_s, Q_j) for the NB kept directions, then (q, q) and (r_s, q), with
// q = r_s - r'; block b takes rows 1+b, 1+b+gridDim.x, ...  It only reads
// (z and q are stored by k_gcr_update), so the loads of successive points
// are free to overlap.  Each value is summed over the block as block_sum
// does (wave shuffles, then the waves in order), with one barrier for all.
template <int NB>
__global__ __launch_bounds__(256) void k_gcr_dots(const double* __restrict__ rs, const double* __restrict__ r,
                                                  GcrSet Q, Grid g, double* __restrict__ part)
{
    constexpr int ND = 2 * NB + 2;
    double acc[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) acc[j] = 0.0;
    for (int k = 1 + blockIdx.x; k <= g.nz - 1; k += gridDim.x)
        for (int i = 1 + threadIdx.x; i <= g.nx - 1; i += blockDim.x) {
            const size_t c = g.at(i, k);
            // every load of the point first: issued back to back, they
            // overlap (loaded one per use, each waited for the last)
            double b[NB > 0 ? NB : 1];
#pragma unroll
            for (int j = 0; j < NB; ++j) b[j] = Q.v[j][c];
            const double rv = rs[c], qv = rv - r[c];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                acc[2 * j] += qv * b[j];
                acc[2 * j + 1] += rv * b[j];
            }
            acc[2 * NB] += qv * qv;
            acc[2 * NB + 1] += rv * qv;
        }
    __shared__ double sh[4][ND];
    static_assert(ND <= 256, "one thread per value");
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        double a = acc[j];
        for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
        if (lane == 0) sh[w][j] = a;
    }
    __syncthreads();
    if (threadIdx.x < ND) {
        double t = sh[0][threadIdx.x];
        for (int q = 1; q < 4; ++q) t += sh[q][threadIdx.x];
        part[(size_t)blockIdx.x * GCR_NDOT + threadIdx.x] = t;
    }
}

// value j of the dots: the sum of the blocks' partials (one workgroup per
// value, a fixed order)
__global__ __launch_bounds__(256) void k_gcr_final(const double* __restrict__ part, int nblk, double* __restrict__ out)
{
    double a = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) a += part[(size_t)b * GCR_NDOT + blockIdx.x];
    a = block_sum(a);
    if (threadIdx.x == 0) out[blockIdx.x] = a;
}

// The step's coefficients from the dots (one thread): beta_j against the
// kept directions (squared norms qq per slot), alpha = (r, q^)/(q^, q^).
// flag 1: the new direction is (nearly) in the span of the kept ones
// (|q^|^2 <= 1e-6 |q|^2), so it is used alone; flag 2: no usable direction
// (the cycle left the residual unchanged), so alpha = 1, beta = 0 keep the
// cycle's result.  The new slot's squared norm is stored for later steps.
__global__ void k_gcr_coef(const double* __restrict__ out, int nb, GcrIdx idx, double* __restrict__ qq,
                           GcrCoef* __restrict__ cf, int32_t* __restrict__ flag)
{
    if (threadIdx.x != 0) return;
    const double qq0 = out[2 * nb], rq0 = out[2 * nb + 1];
    double q2 = qq0, rq = rq0;
    double beta[GCR_MAX];
    for (int j = 0; j < nb; ++j) {
        beta[j] = out[2 * j] / qq[idx.v[j]];
        q2 -= beta[j] * out[2 * j];
        rq -= beta[j] * out[2 * j + 1];
    }
    int f = 0;
    if (nb > 0 && !(q2 > 1e-6 * qq0)) {
        for (int j = 0; j < nb; ++j) beta[j] = 0.0;
        q2 = qq0;
        rq = rq0;
        f = 1;
    }
    double alpha = rq / q2;
    if (!(q2 > 0.0) || !isfinite(alpha)) {
        for (int j = 0; j < nb; ++j) beta[j] = 0.0;
        alpha = 1.0;
        f = 2;
    }
    for (int j = 0; j < nb; ++j) cf->beta[j] = beta[j];
    cf->alpha = alpha;
    qq[idx.v[nb]] = q2;
    *flag = f;
}

// the new direction z = (hi - hs) + (lo - ls) and its image q = r_s - r,
// orthogonalised against the kept ones and stored at position NB of Z and Q;
// psi = psi_s + alpha z in double-double, written to psi (with its ghosts)
// and to psi_s; its residual r_s - alpha q written to r (FP64 and, for the
// FP32 cycle, FP32) and to r_s
template <int NB>
__global__ __launch_bounds__(256) void k_gcr_update(GcrSetW Z, GcrSetW Q, const GcrCoef* __restrict__ cfp,
                                                    double* __restrict__ hi,
                                                    double* __restrict__ lo, double* __restrict__ hs,
                                                    double* __restrict__ ls, double* __restrict__ rs,
                                                    double* __restrict__ r, float* __restrict__ r32, Grid g,
                                                    double zsg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > g.nx - 1 || k > g.nz - 1) return;
    const size_t c = g.at(i, k);
    const GcrCoef& cf = *cfp;
    double z = (hi[c] - hs[c]) + (lo[c] - ls[c]), q = rs[c] - r[c];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        z -= cf.beta[j] * Z.v[j][c];
        q -= cf.beta[j] * Q.v[j][c];
    }
    Z.v[NB][c] = z;
    Q.v[NB][c] = q;
    const dd s = dd_add({hs[c], ls[c]}, {cf.alpha * z, 0.0});
    hi[c] = s.hi;
    lo[c] = s.lo;
    hs[c] = s.hi;
    ls[c] = s.lo;
    stream_ghosts(hi, g, i, k, s.hi, zsg);
    stream_ghosts(lo, g, i, k, s.lo, zsg);
    const double rn = rs[c] - cf.alpha * q;
    rs[c] = rn;
    r[c] = rn;
    if (r32) r32[c] = (float)rn;
}

// ---- launchers -------------------------------------------------------------
static inline dim3 grid2(int nx, int ny, dim3 b) { return dim3((nx + b.x - 1) / b.x, (ny + b.y - 1) / b.y); }

int gcr_blocks(const Grid& g) { return std::min(g.nz - 1, 1024); }

void launch_gcr_dots(hipStream_t st, int nb, const double* rs, const double* r, const GcrSet& Q, const Grid& g,
                     double* part, double* out)
{
    const int nblk = gcr_blocks(g);
    const dim3 gr(nblk), b(256);
#define TC_GCR_DOTS(N) hipLaunchKernelGGL((k_gcr_dots<N>), gr, b, 0, st, rs, r, Q, g, part)
    switch (nb) {
    case 0: TC_GCR_DOTS(0); break;
    case 1: TC_GCR_DOTS(1); break;
    case 2: TC_GCR_DOTS(2); break;
    case 3: TC_GCR_DOTS(3); break;
    case 4: TC_GCR_DOTS(4); break;
    case 5: TC_GCR_DOTS(5); break;
    case 6: TC_GCR_DOTS(6); break;
    case 7: TC_GCR_DOTS(7); break;
    case 8: TC_GCR_DOTS(8); break;
    case 9: TC_GCR_DOTS(9); break;
    case 10: TC_GCR_DOTS(10); break;
    case 11: TC_GCR_DOTS(11); break;
    case 12: TC_GCR_DOTS(12); break;
    case 13: TC_GCR_DOTS(13); break;
    case 14: TC_GCR_DOTS(14); break;
    default: TC_GCR_DOTS(15); break;
    }
#undef TC_GCR_DOTS
    hipLaunchKernelGGL(k_gcr_final, dim3(2 * nb + 2), dim3(256), 0, st, part, nblk, out);
}

void launch_gcr_coef(hipStream_t st, int nb, const double* out, const GcrIdx& idx, double* qq, GcrCoef* cf,
                     int32_t* flag)
{
    hipLaunchKernelGGL(k_gcr_coef, dim3(1), dim3(64), 0, st, out, nb, idx, qq, cf, flag);
}

void launch_gcr_update(hipStream_t st, int nb, const GcrSetW& Z, const GcrSetW& Q, const GcrCoef* cf, double* hi,
                       double* lo, double* hs, double* ls, double* rs, double* r, float* r32, const Grid& g,
                       double zsg)
{
    dim3 b(64, 4);
    const dim3 gr = grid2(g.nx - 1, g.nz - 1, b);
#define TC_GCR_UPD(N)                                                                                              \
    hipLaunchKernelGGL((k_gcr_update<N>), gr, b, 0, st, Z, Q, cf, hi, lo, hs, ls, rs, r, r32, g, zsg)
    switch (nb) {
    case 0: TC_GCR_UPD(0); break;
    case 1: TC_GCR_UPD(1); break;
    case 2: TC_GCR_UPD(2); break;
    case 3: TC_GCR_UPD(3); break;
    case 4: TC_GCR_UPD(4); break;
    case 5: TC_GCR_UPD(5); break;
    case 6: TC_GCR_UPD(6); break;
    case 7: TC_GCR_UPD(7); break;
    case 8: TC_GCR_UPD(8); break;
    case 9: TC_GCR_UPD(9); break;
    case 10: TC_GCR_UPD(10); break;
    case 11: TC_GCR_UPD(11); break;
    case 12: TC_GCR_UPD(12); break;
    case 13: TC_GCR_UPD(13); break;
    case 14: TC_GCR_UPD(14); break;
    default: TC_GCR_UPD(15); break;
    }
#undef TC_GCR_UPD
}

void launch_add_diff(hipStream_t st, const double* a, const double* b, const double* c, double* out, size_t n)
{
    hipLaunchKernelGGL(k_add_diff, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, c, out, n);
}

void launch_extrap(hipStream_t st, const double* a, const double* b, double c, double* out, size_t n)
{
    hipLaunchKernelGGL(k_extrap, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, c, out, n);
}

int launch_residual_dd(hipStream_t st, const double* hi, const double* lo, const Level& L, const double* rhs,
                       double* r64, float* r32, const FD& F, const Spacing& h, int ivis, double* partial)
{
    dim3 b(RD_BX, RD_BY);
    dim3 gr = grid2(L.g.nx - 1, L.g.nz - 1, b);
    hipLaunchKernelGGL(k_residual_dd, gr, b, 0, st, hi, lo, L.vis, L.visx, L.visz, L.visxx, L.viszz, L.visxz, rhs, r64,
                       r32, L.g, F, inv_spacing(h), ivis, partial);
    return (int)(gr.x * gr.y);
}

void launch_dd_absorb(hipStream_t st, double* hi, double* lo, float* e32, double* e64, const Grid& g, double zsg)
{
    dim3 b(64, 4);
    if (e32) hipLaunchKernelGGL((k_dd_absorb<float>), grid2(g.nx - 1, g.nz - 1, b), b, 0, st, hi, lo, e32, g, zsg);
    else hipLaunchKernelGGL((k_dd_absorb<double>), grid2(g.nx - 1, g.nz - 1, b), b, 0, st, hi, lo, e64, g, zsg);
}

template <typename T, bool PRO>
static void absorb_inc_t(hipStream_t st, bool pow2, dim3 gr, dim3 b, double* hi, double* lo, const T* e, const Grid& gc,
                         const Level& L, double* r64, float* r32, const FD& F, const Spacing& h, int ivis, double zsg,
                         double* partial)
{
    if (pow2)
        hipLaunchKernelGGL((k_absorb_inc<T, PRO, true>), gr, b, 0, st, hi, lo, e, gc, L.vis, L.visx, L.visz, L.visxx,
                           L.viszz, L.visxz, r64, r32, L.g, F, h, ivis, zsg, partial);
    else
        hipLaunchKernelGGL((k_absorb_inc<T, PRO, false>), gr, b, 0, st, hi, lo, e, gc, L.vis, L.visx, L.visz, L.visxx,
                           L.viszz, L.visxz, r64, r32, L.g, F, h, ivis, zsg, partial);
}

int launch_absorb_inc(hipStream_t st, double* hi, double* lo, const float* e32, const double* e64, const Grid* gc,
                      const Level& L, double* r64, float* r32, const FD& F, const Spacing& h, int ivis, bool pow2,
                      double zsg, double* partial)
{
    dim3 b(RD_BX, RD_BY);
    dim3 gr = grid2(L.g.nx - 1, L.g.nz - 1, b);
    const Grid gcv = gc ? *gc : Grid{};
    if (e32) {
        if (gc) absorb_inc_t<float, true>(st, pow2, gr, b, hi, lo, e32, gcv, L, r64, r32, F, h, ivis, zsg, partial);
        else absorb_inc_t<float, false>(st, pow2, gr, b, hi, lo, e32, gcv, L, r64, r32, F, h, ivis, zsg, partial);
    } else {
        if (gc) absorb_inc_t<double, true>(st, pow2, gr, b, hi, lo, e64, gcv, L, r64, r32, F, h, ivis, zsg, partial);
        else absorb_inc_t<double, false>(st, pow2, gr, b, hi, lo, e64, gcv, L, r64, r32, F, h, ivis, zsg, partial);
    }
    return (int)(gr.x * gr.y);
}

void launch_prolong_dd(hipStream_t st, double* hi, double* lo, const Grid& gf, const float* e32, const double* e64,
                       const Grid& gc, double zsg)
{
    dim3 b(64, 4);
    if (e32)
        hipLaunchKernelGGL((k_prolong_dd<float>), grid2(gf.nx - 1, gf.nz - 1, b), b, 0, st, hi, lo, gf, e32, gc, zsg);
    else
        hipLaunchKernelGGL((k_prolong_dd<double>), grid2(gf.nx - 1, gf.nz - 1, b), b, 0, st, hi, lo, gf, e64, gc, zsg);
}

void launch_sweep32(hipStream_t st, bool use_ratio, const Level& L, const float* S0, float* E, const float* rhs,
                    const FD& F, const Spacing& h, double temp, double cconst, double zsg)
{
    dim3 b(64, 4);
    dim3 gr = grid2(L.g.nx - 1, L.g.nz - 1, b);
    const FD32 f = fd32_from(F);
    const H32 hh = h32_from(h);
    if (use_ratio)
        hipLaunchKernelGGL((k_sweep32<true>), gr, b, 0, st, S0, E, L.vis32, L.visx32, L.visz32, L.a132, L.visxz32, rhs,
                           L.dts32, L.lb32, L.g, f, hh, (float)temp, (float)cconst, (float)zsg);
    else
        hipLaunchKernelGGL((k_sweep32<false>), gr, b, 0, st, S0, E, L.vis32, L.visx32, L.visz32, L.a132, L.visxz32, rhs,
                           L.dts32, L.lb32, L.g, f, hh, (float)temp, (float)cconst, (float)zsg);
}

void launch_residual32(hipStream_t st, const Level& L, const float* S, float* R, const float* rhs, const FD& F,
                       const Spacing& h)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_residual32, grid2(L.g.nx - 1, L.g.nz - 1, b), b, 0, st, S, R, L.vis32, L.visx32, L.visz32,
                       L.a132, L.visxz32, rhs, L.g, fd32_from(F), h32_from(h));
}

void launch_restrict32(hipStream_t st, const float* in, const Grid& gi, float* out, const Grid& go)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_restrict32, grid2(go.nx - 1, go.nz - 1, b), b, 0, st, in, gi, out, go);
}

void launch_residual_restrict32(hipStream_t st, const Level& Lf, const float* S, const float* rhs, const FD& F,
                                const Spacing& h, float* out, const Grid& go, const double* RowB, double* bout)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_residual_restrict32, grid2(go.nx - 1, go.nz - 1, b), b, 0, st, S, Lf.vis32, Lf.visx32,
                       Lf.visz32, Lf.a132, Lf.visxz32, rhs, Lf.g, fd32_from(F), h32_from(h), out, go, RowB, bout);
}

void launch_prolong32(hipStream_t st, float* Ef, const Grid& gf, const float* Ec, const Grid& gc, double zsg)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_prolong32, grid2(gf.nx - 1, gf.nz - 1, b), b, 0, st, Ef, gf, Ec, gc, (float)zsg);
}

void launch_coarse_apply32(hipStream_t st, const float* Ainv, const double* RowB, const Level& L, const float* rhs,
                           float* E, int ngc, double zsg, double* b, bool b_ready)
{
    if (!b_ready)
        hipLaunchKernelGGL(k_coarse_rhs32, dim3((unsigned)((ngc + 255) / 256)), dim3(256), 0, st, RowB, rhs, L.g, ngc,
                           b);
    const dim3 gr(L.g.nx - 1, L.g.nz - 1), bl(256);
    const int chunks = (ngc + 255) / 256;
    if (chunks <= 2) hipLaunchKernelGGL(k_coarse_apply32<2>, gr, bl, 0, st, Ainv, b, E, L.g, ngc, (float)zsg);
    else if (chunks <= 4) hipLaunchKernelGGL(k_coarse_apply32<4>, gr, bl, 0, st, Ainv, b, E, L.g, ngc, (float)zsg);
    else if (chunks <= 8) hipLaunchKernelGGL(k_coarse_apply32<8>, gr, bl, 0, st, Ainv, b, E, L.g, ngc, (float)zsg);
    else if (chunks <= 12) hipLaunchKernelGGL(k_coarse_apply32<12>, gr, bl, 0, st, Ainv, b, E, L.g, ngc, (float)zsg);
    else if (chunks <= 16) hipLaunchKernelGGL(k_coarse_apply32<16>, gr, bl, 0, st, Ainv, b, E, L.g, ngc, (float)zsg);
    else hipLaunchKernelGGL(k_coarse_apply32<0>, gr, bl, 0, st, Ainv, b, E, L.g, ngc, (float)zsg);
}

__global__ __launch_bounds__(256) void k_to_f32(const double* __restrict__ in, float* __restrict__ out, size_t n)
{
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (size_t)gridDim.x * blockDim.x)
        out[j] = (float)in[j];
}

void launch_to_f32(hipStream_t st, const double* in, float* out, size_t n)
{
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_to_f32, dim3(blocks), dim3(256), 0, st, in, out, n);
}

void launch_coeffs_to_f32(hipStream_t st, Level& L, bool with_lb, int ivis)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_coeffs_to_f32, grid2(L.g.nx - 1, L.g.nz - 1, b), b, 0, st, L.vis, L.visx, L.visz, L.visxx,
                       L.viszz, L.visxz, L.dtstream, L.ratio, L.vis32, L.visx32, L.visz32, L.a132, L.visxz32, L.dts32,
                       with_lb ? L.lb32 : nullptr, L.g, ivis);
}

}  // namespace tc
