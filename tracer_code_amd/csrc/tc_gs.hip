// Multicolour Gauss-Seidel smoothing of the FP32 correction of the
// defect-correction cycle, on LDS-staged tiles (an accelerated form; the
// reference smooths with Chebyshev-accelerated Jacobi-Richardson sweeps,
// iterate_stream, iterative_solver.f90:860-912).
//
// The operator is the 13-tap form of the sweeps on the coarser levels: per
// cell the dt_stream-scaled weights of S(i+di,k+dk) (k_tb_weights32, once
// per solve).  Its taps reach 2 cells along the axes and 1 diagonally, so
// with the colour (i mod 3, k mod 3) no two cells of one colour are in each
// other's stencil: the cells of a colour can be relaxed at once, and nine
// colour phases make one Gauss-Seidel sweep,
//     S(p) <- S(p) + (dts rhs - sum_q w_q S(q)) / w_p.
// A workgroup owns a GS_T x GS_T tile and stages it with a 2-cell halo; each
// thread owns one 3x3 block of the tile (one cell per colour) and keeps its
// nine cells' weights in registers.  The halo holds the previous launch's
// values (the tiles are coupled between launches, as in a hybrid
// Jacobi/Gauss-Seidel smoother), so the result does not depend on the order
// in which workgroups run.  The ghost mirrors of enforceBCs_stream
// (:965-989) are refreshed in LDS as their interior cell is relaxed, and
// written out with the tile.
#include "tc_kernels.h"

namespace tc {

constexpr int GS_T = 48, GS_H = 2, GS_R = GS_T + 2 * GS_H, GS_THREADS = 256;
static_assert(GS_T == 3 * 16 && GS_THREADS == 16 * 16, "one 3x3 block of the tile per thread");

#pragma clang fp contract(fast)
__global__ __launch_bounds__(GS_THREADS) void k_gs32(const float* __restrict__ Ein, float* __restrict__ Eout,
                                                     const float* __restrict__ W, const float* __restrict__ rhs,
                                                     const float* __restrict__ dts, Grid g, int nk, float omega,
                                                     float zsg)
{
    __shared__ float S[GS_R * GS_R];
    const int tid = threadIdx.x;
    const int2 tile = xcd_tile();
    const int gi0 = 1 + tile.x * GS_T - GS_H, gk0 = 1 + tile.y * GS_T - GS_H;  // staged region origin
    for (int l = tid; l < GS_R * GS_R; l += GS_THREADS) {
        const int i = gi0 + l % GS_R, k = gk0 + l / GS_R;
        S[l] = (i >= -SPAN && i <= g.nx + SPAN && k >= -SPAN && k <= g.nz + SPAN) ? Ein[g.at(i, k)] : 0.0f;
    }
    // this thread's 3x3 block; cell c has colour c
    const int bi = GS_H + 3 * (tid % 16), bk = GS_H + 3 * (tid / 16);
    const size_t n = g.elems;
    float w[9][13], rd[9], inv[9];
    bool act[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        const int i = gi0 + bi + c % 3, k = gk0 + bk + c / 3;
        act[c] = i >= 1 && i <= g.nx - 1 && k >= 1 && k <= g.nz - 1;
        const size_t p = act[c] ? g.at(i, k) : g.at(1, 1);
#pragma unroll
        for (int q = 0; q < 13; ++q) w[c][q] = W[q * n + p];
        rd[c] = dts[p] * rhs[p];
        inv[c] = omega / w[c][0];
    }
    __syncthreads();
    for (int s = 0; s < nk; ++s) {
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            if (act[c]) {
                const int l = (bk + c / 3) * GS_R + bi + c % 3;
                const float* q = w[c];
                const float s0 = S[l];
                // taps in W13 order: c, xm1, xp1, xm2, xp2, zm1, zp1, zm2, zp2, mm, pm, mp, pp
                float acc = q[0] * s0;
                acc += q[1] * S[l - 1] + q[2] * S[l + 1] + q[3] * S[l - 2] + q[4] * S[l + 2];
                acc += q[5] * S[l - GS_R] + q[6] * S[l + GS_R] + q[7] * S[l - 2 * GS_R] + q[8] * S[l + 2 * GS_R];
                acc += q[9] * S[l - GS_R - 1] + q[10] * S[l - GS_R + 1] + q[11] * S[l + GS_R - 1] +
                       q[12] * S[l + GS_R + 1];
                const float nv = s0 + (rd[c] - acc) * inv[c];
                S[l] = nv;
                tb_ghosts<GS_R>(S, g, gi0 + bi + c % 3, gk0 + bk + c / 3, gi0, gk0, nv, zsg);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        if (!act[c]) continue;
        const int i = gi0 + bi + c % 3, k = gk0 + bk + c / 3;
        const float v = S[(bk + c / 3) * GS_R + bi + c % 3];
        Eout[g.at(i, k)] = v;
        stream_ghosts(Eout, g, i, k, v, zsg);
    }
}
#pragma clang fp contract(off)

void launch_gs32(hipStream_t st, const Level& L, const float* rhs, const float* Ein, float* Eout, int nk, double omega,
                 double zsg)
{
    const dim3 gr((unsigned)((L.g.nx - 1 + GS_T - 1) / GS_T), (unsigned)((L.g.nz - 1 + GS_T - 1) / GS_T));
    hipLaunchKernelGGL(k_gs32, gr, dim3(GS_THREADS), 0, st, Ein, Eout, L.w13, rhs, L.dts32, L.g, nk, (float)omega,
                       (float)zsg);
}

}  // namespace tc
