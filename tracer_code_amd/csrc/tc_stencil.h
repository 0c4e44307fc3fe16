// Device stencil helpers shared by the multigrid kernels.
//
// Every expression keeps the operand order of the reference Fortran so that,
// compiled with -ffp-contract=off, the FP64 path reproduces the CPU result of
// the same operation sequence.  On uniform grids the metric chain of
// stream_space (iterative_solver.f90:521-533) divides by xr=zs=1 and
// subtracts terms multiplied by xrr=...=0; both are exact no-ops for finite
// values, so the uniform-grid path omits them.  Divisions by a grid spacing
// are replaced by a multiplication with its reciprocal only when the spacing
// is a power of two (then the two are bitwise identical); POW2 selects that.
#pragma once
#include "tc_internal.h"

namespace tc {

template <bool POW2>
__device__ __forceinline__ double sdiv(double a, double d, double inv)
{
    if constexpr (POW2) return a * inv; else return a / d;
}

struct Spacing {
    double dr, ds, dr2, ds2, dr3, ds3, dr4, ds4, drds, drds2, dr2ds, dr2ds2;
    double idr, ids, idr2, ids2, idr3, ids3, idr4, ids4, idrds, idrds2, idr2ds, idr2ds2;
};

// dot_product(D(-1:1), f(c-st:c+st)) with D given as the 5-table D[-2..2].
__device__ __forceinline__ double dot3(const double* D, const double* __restrict__ f, size_t c, ptrdiff_t st)
{
    double acc = D[1] * f[c - st];
    acc += D[2] * f[c];
    acc += D[3] * f[c + st];
    return acc;
}
__device__ __forceinline__ double dot5(const double* D, const double* __restrict__ f, size_t c, ptrdiff_t st)
{
    double acc = D[0] * f[c - 2 * st];
    acc += D[1] * f[c - st];
    acc += D[2] * f[c];
    acc += D[3] * f[c + st];
    acc += D[4] * f[c + 2 * st];
    return acc;
}
// sum(M(-1:1,-1:1)*f(...)) in column-major order; M[(dk+1)*3+(di+1)].
__device__ __forceinline__ double sum9(const double* M, const double* __restrict__ f, size_t c, ptrdiff_t ld)
{
    double acc = M[0] * f[c - ld - 1];
    acc += M[1] * f[c - ld];
    acc += M[2] * f[c - ld + 1];
    acc += M[3] * f[c - 1];
    acc += M[4] * f[c];
    acc += M[5] * f[c + 1];
    acc += M[6] * f[c + ld - 1];
    acc += M[7] * f[c + ld];
    acc += M[8] * f[c + ld + 1];
    return acc;
}

struct FD {
    double D1[5], D2[5], D3[5], D4[5];
    double Drs[9], Drss[9], Drrs[9], Drrss[9];
};

// L(S) - RHS evaluated as in stream_space (build=0, scaling=0), negated:
// returns RHS - L(S).  The six terms are summed left to right.
template <bool POW2>
__device__ __forceinline__ double stream_space(const FD& F, const Spacing& h, int ivis, const double* __restrict__ S,
                                               size_t c, ptrdiff_t ld, double vis, double vx, double vz, double vxx,
                                               double vzz, double vxz, double rhs)
{
    double Sr = sdiv<POW2>(dot3(F.D1, S, c, 1), h.dr, h.idr);
    double Ss = sdiv<POW2>(dot3(F.D1, S, c, ld), h.ds, h.ids);
    double Srr = sdiv<POW2>(dot3(F.D2, S, c, 1), h.dr2, h.idr2);
    double Sss = sdiv<POW2>(dot3(F.D2, S, c, ld), h.ds2, h.ids2);
    double Srrr = sdiv<POW2>(dot5(F.D3, S, c, 1), h.dr3, h.idr3);
    double Ssss = sdiv<POW2>(dot5(F.D3, S, c, ld), h.ds3, h.ids3);
    double Srrrr = sdiv<POW2>(dot5(F.D4, S, c, 1), h.dr4, h.idr4);
    double Sssss = sdiv<POW2>(dot5(F.D4, S, c, ld), h.ds4, h.ids4);
    double Srs = sdiv<POW2>(sum9(F.Drs, S, c, ld), h.drds, h.idrds);
    double Srss = sdiv<POW2>(sum9(F.Drss, S, c, ld), h.drds2, h.idrds2);
    double Srrs = sdiv<POW2>(sum9(F.Drrs, S, c, ld), h.dr2ds, h.idr2ds);
    double Srrss = sdiv<POW2>(sum9(F.Drrss, S, c, ld), h.dr2ds2, h.idr2ds2);
    double t1, t2, t3, t4, t5, t6;
    if (ivis == 0) {
        t1 = 0.0;
        t2 = vis * (Srrrr + Sssss);
        t3 = 2.0 * vis * Srrss;
        t4 = 0.0; t5 = 0.0; t6 = 0.0;
    } else {
        t1 = (vxx - vzz) * (Srr - Sss);
        t2 = vis * (Srrrr + Sssss);
        t3 = 2.0 * vis * Srrss;
        t4 = 4.0 * vxz * Srs;
        t5 = 2.0 * vx * (Srrr + Srss);
        t6 = 2.0 * vz * (Ssss + Srrs);
    }
    double ss = t1 + t2 + t3 + t4 + t5 + t6 - rhs;
    return -ss;
}

// Ghost mirrors of enforceBCs_stream (iterative_solver.f90:965-989) for an
// interior point (i,k) with value v: antisymmetric sides, antisymmetric
// (Vbc=0) or symmetric (Vbc=1) top/bottom, corners by composition.
template <typename T>
__device__ __forceinline__ void stream_ghosts(T* __restrict__ E, const Grid& g, int i, int k, T v, T zsg)
{
    const int nx = g.nx, nz = g.nz;
    const bool lo_i = i <= SPAN, hi_i = i >= nx - SPAN, lo_k = k <= SPAN, hi_k = k >= nz - SPAN;
    if (lo_i) E[g.at(-i, k)] = -v;
    if (hi_i) E[g.at(2 * nx - i, k)] = -v;
    if (lo_k) E[g.at(i, -k)] = zsg * v;
    if (hi_k) E[g.at(i, 2 * nz - k)] = zsg * v;
    if (lo_i && lo_k) E[g.at(-i, -k)] = zsg * (-v);
    if (lo_i && hi_k) E[g.at(-i, 2 * nz - k)] = zsg * (-v);
    if (hi_i && lo_k) E[g.at(2 * nx - i, -k)] = zsg * (-v);
    if (hi_i && hi_k) E[g.at(2 * nx - i, 2 * nz - k)] = zsg * (-v);
}

}  // namespace tc
