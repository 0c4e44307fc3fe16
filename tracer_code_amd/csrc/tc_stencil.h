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

// The stencil sums read the field through an accessor S(di, dk) that
// returns S(i+di, k+dk); the pointer forms below wrap a strided plane.
// dot_product(D(-1:1), S along one axis) with D given as the 5-table D[-2..2].
template <int AX, typename G>
__device__ __forceinline__ double dot3g(const double* D, const G& S)
{
    double acc = D[1] * S(AX == 0 ? -1 : 0, AX == 0 ? 0 : -1);
    acc += D[2] * S(0, 0);
    acc += D[3] * S(AX == 0 ? 1 : 0, AX == 0 ? 0 : 1);
    return acc;
}
template <int AX, typename G>
__device__ __forceinline__ double dot5g(const double* D, const G& S)
{
    double acc = D[0] * S(AX == 0 ? -2 : 0, AX == 0 ? 0 : -2);
    acc += D[1] * S(AX == 0 ? -1 : 0, AX == 0 ? 0 : -1);
    acc += D[2] * S(0, 0);
    acc += D[3] * S(AX == 0 ? 1 : 0, AX == 0 ? 0 : 1);
    acc += D[4] * S(AX == 0 ? 2 : 0, AX == 0 ? 0 : 2);
    return acc;
}
// sum(M(-1:1,-1:1)*S(...)) in column-major order; M[(dk+1)*3+(di+1)].
template <typename G>
__device__ __forceinline__ double sum9g(const double* M, const G& S)
{
    double acc = M[0] * S(-1, -1);
    acc += M[1] * S(0, -1);
    acc += M[2] * S(1, -1);
    acc += M[3] * S(-1, 0);
    acc += M[4] * S(0, 0);
    acc += M[5] * S(1, 0);
    acc += M[6] * S(-1, 1);
    acc += M[7] * S(0, 1);
    acc += M[8] * S(1, 1);
    return acc;
}

// accessor of a strided plane around element c
template <typename T>
struct PlaneAt {
    const T* __restrict__ f;
    size_t c;
    ptrdiff_t ld;
    __device__ __forceinline__ T operator()(int di, int dk) const { return f[c + (ptrdiff_t)dk * ld + di]; }
};

// a 3x3 neighbourhood of element c held in registers, for the accessor
// forms: its nine loads are issued together, ahead of other work
struct Reg9 {
    double v[9];
    __device__ __forceinline__ double operator()(int di, int dk) const { return v[(dk + 1) * 3 + di + 1]; }
};
__device__ __forceinline__ Reg9 load9(const double* __restrict__ f, size_t c, ptrdiff_t ld)
{
    Reg9 r;
#pragma unroll
    for (int dk = -1; dk <= 1; ++dk)
#pragma unroll
        for (int di = -1; di <= 1; ++di) r.v[(dk + 1) * 3 + di + 1] = f[c + (ptrdiff_t)dk * ld + di];
    return r;
}

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
__device__ __forceinline__ double sum9(const double* M, const double* __restrict__ f, size_t c, ptrdiff_t ld)
{
    return sum9g(M, PlaneAt<double>{f, c, ld});
}

struct FD {
    double D1[5], D2[5], D3[5], D4[5];
    double Drs[9], Drss[9], Drrs[9], Drrss[9];
};

// L(S) - RHS evaluated as in stream_space (build=0, scaling=0), negated:
// returns RHS - L(S).  The six terms are summed left to right.
template <bool POW2, typename G>
__device__ __forceinline__ double stream_space_g(const FD& F, const Spacing& h, int ivis, const G& S, double vis,
                                                 double vx, double vz, double vxx, double vzz, double vxz, double rhs)
{
    double Srr = sdiv<POW2>(dot3g<0>(F.D2, S), h.dr2, h.idr2);
    double Sss = sdiv<POW2>(dot3g<1>(F.D2, S), h.ds2, h.ids2);
    double Srrr = sdiv<POW2>(dot5g<0>(F.D3, S), h.dr3, h.idr3);
    double Ssss = sdiv<POW2>(dot5g<1>(F.D3, S), h.ds3, h.ids3);
    double Srrrr = sdiv<POW2>(dot5g<0>(F.D4, S), h.dr4, h.idr4);
    double Sssss = sdiv<POW2>(dot5g<1>(F.D4, S), h.ds4, h.ids4);
    double Srs = sdiv<POW2>(sum9g(F.Drs, S), h.drds, h.idrds);
    double Srss = sdiv<POW2>(sum9g(F.Drss, S), h.drds2, h.idrds2);
    double Srrs = sdiv<POW2>(sum9g(F.Drrs, S), h.dr2ds, h.idr2ds);
    double Srrss = sdiv<POW2>(sum9g(F.Drrss, S), h.dr2ds2, h.idr2ds2);
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
template <bool POW2>
__device__ __forceinline__ double stream_space(const FD& F, const Spacing& h, int ivis, const double* __restrict__ S,
                                               size_t c, ptrdiff_t ld, double vis, double vx, double vz, double vxx,
                                               double vzz, double vxz, double rhs)
{
    return stream_space_g<POW2>(F, h, ivis, PlaneAt<double>{S, c, ld}, vis, vx, vz, vxx, vzz, vxz, rhs);
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

// Mirror writes of enforceBCs_stream for interior (i,k) into the staged
// region with global origin (gi0, gk0); only cells inside it.
template <int R, typename T>
__device__ __forceinline__ void tb_ghosts(T* __restrict__ B, const Grid& g, int i, int k, int gi0, int gk0, T v, T zsg)
{
    const int nx = g.nx, nz = g.nz;
    auto put = [&](int ii, int kk, T x) {
        const int li = ii - gi0, lk = kk - gk0;
        if (li >= 0 && li < R && lk >= 0 && lk < R) B[lk * R + li] = x;
    };
    const bool lo_i = i <= SPAN, hi_i = i >= nx - SPAN, lo_k = k <= SPAN, hi_k = k >= nz - SPAN;
    if (lo_i) put(-i, k, -v);
    if (hi_i) put(2 * nx - i, k, -v);
    if (lo_k) put(i, -k, zsg * v);
    if (hi_k) put(i, 2 * nz - k, zsg * v);
    if (lo_i && lo_k) put(-i, -k, zsg * (-v));
    if (lo_i && hi_k) put(-i, 2 * nz - k, zsg * (-v));
    if (hi_i && lo_k) put(2 * nx - i, -k, zsg * (-v));
    if (hi_i && hi_k) put(2 * nx - i, 2 * nz - k, zsg * (-v));
}

}  // namespace tc
