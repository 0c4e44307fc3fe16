// FP32 form of the stream_space stencil (the mixed-precision cycle).
#pragma once
#include "tc_stencil.h"

namespace tc {

struct FD32 {
    float D1[5], D2[5], D3[5], D4[5];
    float Drs[9], Drss[9], Drrs[9], Drrss[9];
};
struct H32 {
    float idr, ids, idr2, ids2, idr3, ids3, idr4, ids4, idrds, idrds2, idr2ds, idr2ds2;
};

inline FD32 fd32_from(const FD& F)
{
    FD32 o;
    for (int j = 0; j < 5; ++j) { o.D1[j] = (float)F.D1[j]; o.D2[j] = (float)F.D2[j]; o.D3[j] = (float)F.D3[j]; o.D4[j] = (float)F.D4[j]; }
    for (int j = 0; j < 9; ++j) { o.Drs[j] = (float)F.Drs[j]; o.Drss[j] = (float)F.Drss[j]; o.Drrs[j] = (float)F.Drrs[j]; o.Drrss[j] = (float)F.Drrss[j]; }
    return o;
}
inline H32 h32_from(const Spacing& h)
{
    return H32{(float)h.idr, (float)h.ids, (float)h.idr2, (float)h.ids2, (float)h.idr3, (float)h.ids3,
               (float)h.idr4, (float)h.ids4, (float)h.idrds, (float)h.idrds2, (float)h.idr2ds, (float)h.idr2ds2};
}

#pragma clang fp contract(fast)
// FP32 fields and weights; A is the arithmetic type.  With A = double every
// product of two FP32 values is exact and only the sums round, which keeps
// the residual passed to the next level accurate relative to the (small)
// defect even where rhs and L(S) nearly cancel.
template <typename A, int AX, typename G>
__device__ __forceinline__ A dot3gf(const float* D, const G& S)
{
    return A(D[1]) * A(S(AX == 0 ? -1 : 0, AX == 0 ? 0 : -1)) + A(D[2]) * A(S(0, 0)) +
           A(D[3]) * A(S(AX == 0 ? 1 : 0, AX == 0 ? 0 : 1));
}
template <typename A, int AX, typename G>
__device__ __forceinline__ A dot5gf(const float* D, const G& S)
{
    return A(D[0]) * A(S(AX == 0 ? -2 : 0, AX == 0 ? 0 : -2)) + A(D[1]) * A(S(AX == 0 ? -1 : 0, AX == 0 ? 0 : -1)) +
           A(D[2]) * A(S(0, 0)) + A(D[3]) * A(S(AX == 0 ? 1 : 0, AX == 0 ? 0 : 1)) +
           A(D[4]) * A(S(AX == 0 ? 2 : 0, AX == 0 ? 0 : 2));
}
template <typename A, typename G>
__device__ __forceinline__ A sum9gf(const float* M, const G& S)
{
    return A(M[0]) * A(S(-1, -1)) + A(M[1]) * A(S(0, -1)) + A(M[2]) * A(S(1, -1)) + A(M[3]) * A(S(-1, 0)) +
           A(M[4]) * A(S(0, 0)) + A(M[5]) * A(S(1, 0)) + A(M[6]) * A(S(-1, 1)) + A(M[7]) * A(S(0, 1)) +
           A(M[8]) * A(S(1, 1));
}
// rhs - L(S); a1 = vis_xx - vis_zz
template <typename A, typename G>
__device__ __forceinline__ A stream_space32_g(const FD32& F, const H32& h, const G& S, float vis, float vx, float vz,
                                              float a1, float vxz, float rhs)
{
    A Srr = dot3gf<A, 0>(F.D2, S) * A(h.idr2);
    A Sss = dot3gf<A, 1>(F.D2, S) * A(h.ids2);
    A Srrr = dot5gf<A, 0>(F.D3, S) * A(h.idr3);
    A Ssss = dot5gf<A, 1>(F.D3, S) * A(h.ids3);
    A Srrrr = dot5gf<A, 0>(F.D4, S) * A(h.idr4);
    A Sssss = dot5gf<A, 1>(F.D4, S) * A(h.ids4);
    A Srs = sum9gf<A>(F.Drs, S) * A(h.idrds);
    A Srss = sum9gf<A>(F.Drss, S) * A(h.idrds2);
    A Srrs = sum9gf<A>(F.Drrs, S) * A(h.idr2ds);
    A Srrss = sum9gf<A>(F.Drrss, S) * A(h.idr2ds2);
    A L = A(a1) * (Srr - Sss) + A(vis) * (Srrrr + Sssss) + A(2) * A(vis) * Srrss + A(4) * A(vxz) * Srs +
          A(2) * A(vx) * (Srrr + Srss) + A(2) * A(vz) * (Ssss + Srrs);
    return A(rhs) - L;
}
template <typename A>
__device__ __forceinline__ A stream_space32(const FD32& F, const H32& h, const float* __restrict__ S, size_t c,
                                            ptrdiff_t ld, float vis, float vx, float vz, float a1, float vxz,
                                            float rhs)
{
    return stream_space32_g<A>(F, h, PlaneAt<float>{S, c, ld}, vis, vx, vz, a1, vxz, rhs);
}

// S(i+di, k+dk) of a 7x7 neighbourhood held in registers (row-major, centre
// at index 24) for the point (a, b) cells from its centre
struct Patch7At {
    const float* p;
    int a, b;
    __device__ __forceinline__ float operator()(int di, int dk) const { return p[(b + dk + 3) * 7 + a + di + 3]; }
};

// S(i+di, k+dk) of a tile staged in LDS with row pitch ld
struct LdsAt {
    const float* f;
    int c;
    int ld;
    __device__ __forceinline__ float operator()(int di, int dk) const { return f[c + dk * ld + di]; }
};

// restrict (iterative_solver.f90:596-626) at one coarse point from the fine
// residuals around its fine centre: the reference's weights, whose corner sum
// counts (ii-1,kk+1) twice and (ii+1,kk-1) not at all
__device__ __forceinline__ float restrict_fw32(float rmm, float rmp, float rpp, float rm0, float r0m, float rp0,
                                               float r0p, float r00)
{
    const float corners = (rmm + rmp + rpp + rmp) * (1.0f / 16.0f);
    const float sides = (rm0 + r0m + rp0 + r0p) * (1.0f / 8.0f);
    return corners + sides + r00 * 0.25f;
}
#pragma clang fp contract(off)

}  // namespace tc
