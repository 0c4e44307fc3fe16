// Temporally blocked Jacobi-Richardson sweeps for the correction levels of
// the defect-correction cycle.  One launch performs up to K consecutive
// sweeps of a visit (iterate_stream, iterative_solver.f90:871-912, with the
// ghost part of enforceBCs_stream :965-989).
//
// Each workgroup stages its TB_T x TB_T tile plus a halo of 2K cells in
// LDS, one to three cells per thread, and sweeps the staged region K times.  The
// band of still-valid cells shrinks by the stencil reach of 2 per sweep, so
// the tile itself stays exact.  It then writes only its tile and the tile's
// ghost mirrors.
//
// The operator is applied in its 13-tap form: per cell the weights of
// S(i+di,k+dk) are combined once per launch from eta, eta_x, eta_z,
// eta_xx-eta_zz and eta_xz (pre-scaled by dt_stream), so a sweep costs 13
// LDS reads and 13 FMAs.  This is the same operator with a different
// rounding order, so these kernels serve the defect-correction modes; the
// reference-form FP64 path keeps its one-sweep kernels.
#include "tc_kernels.h"
#include "tc_stencil32.h"

namespace tc {

// Tile TT x TT; K sweeps per launch need a halo of 2K cells, so the
// staged region is (TT + 4K)^2: 32^2 for TT=16, K=4 (4x the tile), 48^2 for
// TT=32, K=4 (2.25x) -- the larger tile stages less halo per updated cell
// and serves levels with enough tiles to fill the chip.  A launch that also
// forms the next coarser level's right-hand side stages X = 3 more cells
// (below).
constexpr int TB_KMAX = 8;
constexpr int TB_THREADS = 1024;
template <int TT, int K, int X = 0>
struct TBGeom {
    static constexpr int H = 2 * K + X;
    static constexpr int R = TT + 2 * H;
    static constexpr int CELLS = R * R;
    static constexpr int NPT = (CELLS + TB_THREADS - 1) / TB_THREADS;  // staged cells per thread
};

// a coarse correction to prolong into the level while staging (FP32 only)
template <typename T>
struct Prolong {
    const T* Ec = nullptr;
    Grid gc;
};

template <typename T>
struct SweepFactors {
    T temp[TB_KMAX];
    T cconst[TB_KMAX];
};

// FD tables and reciprocal spacings of one level in the arithmetic type.
template <typename T>
struct Tables {
    T D1[5], D2[5], D3[5], D4[5];
    T idr2, ids2, idr3, ids3, idr4, ids4, idrds, idrds2, idr2ds, idr2ds2;
};

// Per-level coefficient arrays; a1 = eta_xx - eta_zz for FP32, or eta_xx
// and eta_zz separately for FP64.
template <typename T>
struct Coefs {
    const T *vis, *vx, *vz, *vxx, *vzz, *a1, *vxz, *rhs, *dts, *ratio, *lb;
};

// The residual+restriction into the next coarser level, done by the last
// launch of a visit before a descent (RR): out (the coarse right-hand side),
// its grid, and for the coarsest level RowB/bout (the equilibrated
// right-hand side of the coarse solve); F, h the FP32 stencil tables.
struct RRArgs {
    float* out;
    Grid go;
    const double* RowB;
    double* bout;
    FD32 F;
    H32 h;
};
struct NoRR {};

// dt_stream-scaled weights of the 13 taps and the scaled RHS of one cell.
template <typename T>
struct W13 {
    T c, xm1, xp1, xm2, xp2, zm1, zp1, zm2, zp2, mm, pm, mp, pp, rd, lb;
};

template <typename T, bool RATIO>
__device__ __forceinline__ void tb_weights(const Coefs<T>& C, const Tables<T>& Tb, size_t c, W13<T>& w)
{
    const T vis = C.vis[c], vx = C.vx[c], vz = C.vz[c], vxz = C.vxz[c], dts = C.dts[c];
    T a1;
    if constexpr (sizeof(T) == 4) a1 = C.a1[c];
    else a1 = C.vxx[c] - C.vzz[c];
    const T two = T(2);
    // axis parts: a1 (S_rr - S_ss) + 2 vx S_rrr + 2 vz S_sss + vis (S_rrrr + S_ssss)
    auto wx = [&](int d) { return a1 * Tb.D2[d + 2] * Tb.idr2 + two * vx * Tb.D3[d + 2] * Tb.idr3 + vis * Tb.D4[d + 2] * Tb.idr4; };
    auto wz = [&](int d) { return -a1 * Tb.D2[d + 2] * Tb.ids2 + two * vz * Tb.D3[d + 2] * Tb.ids3 + vis * Tb.D4[d + 2] * Tb.ids4; };
    // 3x3 part: 4 vxz S_rs + 2 vx S_rss + 2 vz S_rrs + 2 vis S_rrss
    auto wb = [&](int di, int dk) {
        return T(4) * vxz * Tb.D1[dk + 2] * Tb.D1[di + 2] * Tb.idrds + two * vx * Tb.D2[dk + 2] * Tb.D1[di + 2] * Tb.idrds2 +
               two * vz * Tb.D1[dk + 2] * Tb.D2[di + 2] * Tb.idr2ds + two * vis * Tb.D2[dk + 2] * Tb.D2[di + 2] * Tb.idr2ds2;
    };
    w.c = dts * (wx(0) + wz(0) + wb(0, 0));
    w.xm1 = dts * (wx(-1) + wb(-1, 0));
    w.xp1 = dts * (wx(1) + wb(1, 0));
    w.xm2 = dts * wx(-2);
    w.xp2 = dts * wx(2);
    w.zm1 = dts * (wz(-1) + wb(0, -1));
    w.zp1 = dts * (wz(1) + wb(0, 1));
    w.zm2 = dts * wz(-2);
    w.zp2 = dts * wz(2);
    w.mm = dts * wb(-1, -1);
    w.pm = dts * wb(1, -1);
    w.mp = dts * wb(-1, 1);
    w.pp = dts * wb(1, 1);
    w.rd = dts * C.rhs[c];
    w.lb = T(-2);
    if constexpr (RATIO) {
        if constexpr (sizeof(T) == 4) {
            w.lb = C.lb[c];
        } else {
            const T r = C.ratio[c];
            w.lb = (r > T(1e99)) ? T(-2) : T(-2) * (r * r) / (T(1) + r * r);
        }
    }
}

#pragma clang fp contract(fast)
// The 13 dt-scaled tap weights of every interior cell of a level, once per
// solve (the operator's coefficients are fixed during a solve): plane q of
// W holds tap q in W13 order (c, xm1, xp1, xm2, xp2, zm1, zp1, zm2, zp2, mm,
// pm, mp, pp).  The same arithmetic as tb_weights, so the sweeps that read
// them give the same values.
__global__ __launch_bounds__(256) void k_tb_weights32(Coefs<float> C, Tables<float> Tb, Grid g,
                                                      float* __restrict__ W)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > g.nx - 1 || k > g.nz - 1) return;
    const size_t c = g.at(i, k), n = g.elems;
    W13<float> w;
    tb_weights<float, false>(C, Tb, c, w);
    const float v[13] = {w.c, w.xm1, w.xp1, w.xm2, w.xp2, w.zm1, w.zp1, w.zm2, w.zp2, w.mm, w.pm, w.mp, w.pp};
#pragma unroll
    for (int q = 0; q < 13; ++q) W[q * n + c] = v[q];
}

// RR: after the nk <= K sweeps, the fine residual rhs - L(S) of the cells
// [TT*tx+1, TT*tx+TT+1]^2 (the tile and one more row and column: the
// restriction footprint of the tile's coarse points, coarse i in
// [TT/2*tx+1, TT/2*tx+TT/2] with fine centre 2i) from the staged values, and
// their restriction.  The residual reads S 2 cells out, so the staged values
// must be exact from 2 cells below the tile to 3 above: the halo is 2K+3,
// and after nk sweeps every cell at least 2nk from the region's edge is.
// The arithmetic is k_residual_restrict32's, so the values are the same.
template <typename T, bool RATIO, int K, int TT, bool PRE = false, bool RR = false>
__global__ __launch_bounds__(TB_THREADS) void k_sweep_tb(const T* __restrict__ Ein, T* __restrict__ Eout, Coefs<T> C,
                                                        Tables<T> Tb, Grid g, SweepFactors<T> sf, int nk, T zsg,
                                                        const T* __restrict__ W, Prolong<T> pr,
                                                        std::conditional_t<RR, RRArgs, NoRR> rr)
{
    static_assert(!RR || (sizeof(T) == 4 && PRE), "the fused restriction serves the FP32 cycle");
    using G = TBGeom<TT, K, RR ? 3 : 0>;
    constexpr int R = G::R, H = G::H, NPT = G::NPT;
    __shared__ T buf[2][G::CELLS];
    const int2 tile = xcd_tile();
    const int gi0 = 1 + tile.x * TT - H, gk0 = 1 + tile.y * TT - H;  // staged region origin
    W13<T> w[NPT];
    int band[NPT];
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
        const int l = threadIdx.x + u * TB_THREADS;
        band[u] = 0;
        if (l >= G::CELLS) continue;
        const int li = l % R, lk = l / R;
        const int i = gi0 + li, k = gk0 + lk;
        T v = T(0);
        // Ein == nullptr: a visit that starts from a zero correction
        if (Ein && i >= -SPAN && i <= g.nx + SPAN && k >= -SPAN && k <= g.nz + SPAN) {
            v = Ein[g.at(i, k)];
            if constexpr (sizeof(T) == 4) {
                // the coarse-grid correction prolonged into this level (what
                // k_prolong32 would have added): at interior points, and at
                // the ghosts the mirror of their interior point's increment
                if (pr.Ec) {
                    int mi = i, mk = k;
                    T sg = T(1);
                    if (i < 0) { mi = -i; sg = -sg; } else if (i > g.nx) { mi = 2 * g.nx - i; sg = -sg; }
                    if (k < 0) { mk = -k; sg = sg * zsg; } else if (k > g.nz) { mk = 2 * g.nz - k; sg = sg * zsg; }
                    if (mi >= 1 && mi <= g.nx - 1 && mk >= 1 && mk <= g.nz - 1)
                        v = __fadd_rn(v, sg * prolong_add32(pr.Ec, pr.gc, mi, mk));
                }
            }
        }
        buf[0][l] = v;
        buf[1][l] = v;
        // sweep s updates the cells at least 2(s+1) from the region edge
        const int m = min(min(li, R - 1 - li), min(lk, R - 1 - lk));
        const bool interior = i >= 1 && i <= g.nx - 1 && k >= 1 && k <= g.nz - 1;
        if (interior && m >= 2) {
            band[u] = m / 2;
            const size_t c = g.at(i, k);
            if constexpr (PRE) {
                const size_t n = g.elems;
                W13<T>& q = w[u];
                q.c = W[c]; q.xm1 = W[n + c]; q.xp1 = W[2 * n + c]; q.xm2 = W[3 * n + c]; q.xp2 = W[4 * n + c];
                q.zm1 = W[5 * n + c]; q.zp1 = W[6 * n + c]; q.zm2 = W[7 * n + c]; q.zp2 = W[8 * n + c];
                q.mm = W[9 * n + c]; q.pm = W[10 * n + c]; q.mp = W[11 * n + c]; q.pp = W[12 * n + c];
                q.rd = C.dts[c] * C.rhs[c];
                q.lb = T(-2);
                if constexpr (RATIO) q.lb = C.lb[c];
            } else {
                tb_weights<T, RATIO>(C, Tb, c, w[u]);
            }
        }
    }
    // RR: this thread's fine residual cell and its coefficients, loaded now so
    // that their latency overlaps the sweeps
    constexpr int NF = TT + 1;
    const int rl = threadIdx.x, ri = 1 + TT * tile.x + rl % NF, rk = 1 + TT * tile.y + rl / NF;
    const bool rv = RR && rl < NF * NF && ri <= g.nx - 1 && rk <= g.nz - 1;
    float rc[6] = {};
    if constexpr (RR) {
        if (rv) {
            const size_t c = g.at(ri, rk);
            rc[0] = C.vis[c]; rc[1] = C.vx[c]; rc[2] = C.vz[c]; rc[3] = C.a1[c]; rc[4] = C.vxz[c]; rc[5] = C.rhs[c];
        }
    }
    __syncthreads();
    int cur = 0;
    for (int s = 0; s < nk; ++s) {
        const T* __restrict__ A = buf[cur];
        T* __restrict__ B = buf[1 - cur];
#pragma unroll
        for (int u = 0; u < NPT; ++u) {
            if (band[u] <= s) continue;
            const int l = threadIdx.x + u * TB_THREADS;
            const W13<T>& q = w[u];
            const T s0 = A[l];
            T acc = q.c * s0;
            acc += q.xm1 * A[l - 1] + q.xp1 * A[l + 1] + q.xm2 * A[l - 2] + q.xp2 * A[l + 2];
            acc += q.zm1 * A[l - R] + q.zp1 * A[l + R] + q.zm2 * A[l - 2 * R] + q.zp2 * A[l + 2 * R];
            acc += q.mm * A[l - R - 1] + q.pm * A[l - R + 1] + q.mp * A[l + R - 1] + q.pp * A[l + R + 1];
            T cc = sf.cconst[s];
            if constexpr (RATIO) cc = T(2) / (-q.lb + T(1) + (q.lb + T(1)) * sf.temp[s]);
            const T nv = s0 + cc * (q.rd - acc);
            B[l] = nv;
            tb_ghosts<R>(B, g, gi0 + l % R, gk0 + l / R, gi0, gk0, nv, zsg);
        }
        __syncthreads();
        cur = 1 - cur;
    }
    // the tile (interior cells) and its ghost mirrors
#pragma unroll
    for (int u = 0; u < NPT; ++u) {
        const int l = threadIdx.x + u * TB_THREADS;
        if (l >= G::CELLS) continue;
        const int li = l % R, lk = l / R, i = gi0 + li, k = gk0 + lk;
        if (li >= H && li < H + TT && lk >= H && lk < H + TT && i >= 1 && i <= g.nx - 1 && k >= 1 &&
            k <= g.nz - 1) {
            const T r = buf[cur][l];
            Eout[g.at(i, k)] = r;
            stream_ghosts(Eout, g, i, k, r, zsg);
        }
    }
    if constexpr (RR) {
        constexpr int NC = TT / 2;
        __shared__ float rres[NF * NF];
        const int fi0 = 1 + TT * tile.x, fk0 = 1 + TT * tile.y;
        static_assert(NF * NF <= TB_THREADS, "one fine residual per thread");
        if (rl < NF * NF)
            rres[rl] = rv ? (float)stream_space32_g<double>(rr.F, rr.h, LdsAt{buf[cur], (rk - gk0) * R + (ri - gi0), R},
                                                            rc[0], rc[1], rc[2], rc[3], rc[4], rc[5])
                          : 0.0f;
        __syncthreads();
        for (int l = threadIdx.x; l < NC * NC; l += TB_THREADS) {
            const int ic = NC * tile.x + 1 + l % NC, kc = NC * tile.y + 1 + l / NC;
            if (ic > rr.go.nx - 1 || kc > rr.go.nz - 1) continue;
            const int c0 = (2 * kc - fk0) * NF + (2 * ic - fi0);  // the fine centre in rres
            const float v = restrict_fw32(rres[c0 - NF - 1], rres[c0 + NF - 1], rres[c0 + NF + 1], rres[c0 - 1],
                                          rres[c0 - NF], rres[c0 + 1], rres[c0 + NF], rres[c0]);
            rr.out[rr.go.at(ic, kc)] = v;
            if (rr.bout) {
                const int q = ic * (rr.go.nz + 1) + kc;
                rr.bout[q] = (double)v * rr.RowB[q];
            }
        }
    }
}
#pragma clang fp contract(off)

// ---- launchers -------------------------------------------------------------
template <typename T>
static Tables<T> make_tables(const FD& F, const Spacing& h)
{
    Tables<T> t;
    for (int j = 0; j < 5; ++j) { t.D1[j] = (T)F.D1[j]; t.D2[j] = (T)F.D2[j]; t.D3[j] = (T)F.D3[j]; t.D4[j] = (T)F.D4[j]; }
    t.idr2 = (T)h.idr2; t.ids2 = (T)h.ids2; t.idr3 = (T)h.idr3; t.ids3 = (T)h.ids3; t.idr4 = (T)h.idr4; t.ids4 = (T)h.ids4;
    t.idrds = (T)h.idrds; t.idrds2 = (T)h.idrds2; t.idr2ds = (T)h.idr2ds; t.idr2ds2 = (T)h.idr2ds2;
    return t;
}

static inline dim3 tb_grid(const Grid& g, int TT) { return dim3((g.nx - 1 + TT - 1) / TT, (g.nz - 1 + TT - 1) / TT); }

int tb_max_sweeps(int K) { return K; }

// sweeps per launch of a level's FP32 blocked sweeps: 8 where the level has
// at most 16 tiles of 16x16 (a launch there is all latency: twice the sweeps
// for ~1.2x the launch time), else 4 (8 measured slower on larger levels)
int tb_sweeps_per_launch(const Grid& g)
{
    const dim3 t = tb_grid(g, 16);
    return (t.x * t.y <= 16) ? 8 : 4;
}

// the 32x32 tile where it still gives at least 2 workgroups per CU (256 CUs)
// and the arithmetic is FP32 (three staged cells of FP64 weights per thread
// would spill)
template <typename T>
static int tb_tile(const Grid& g)
{
    if (sizeof(T) != 4) return 16;
    const dim3 g32 = tb_grid(g, 32);
    return (g32.x * g32.y >= 512) ? 32 : 16;
}

template <typename T, int TT>
static void tb_launch_t(hipStream_t st, bool use_ratio, int K, const Grid& g, const T* S0, T* E, const Coefs<T>& C,
                        const Tables<T>& tb, const SweepFactors<T>& sf, int nk, T zsg, const T* W, const Prolong<T>& pr)
{
    const dim3 gr = tb_grid(g, TT), b(TB_THREADS);
    const NoRR no{};
    if constexpr (sizeof(T) == 4) {
        if (W && K == 4) {
            if (use_ratio)
                hipLaunchKernelGGL((k_sweep_tb<T, true, 4, TT, true>), gr, b, 0, st, S0, E, C, tb, g, sf, nk, zsg, W, pr, no);
            else
                hipLaunchKernelGGL((k_sweep_tb<T, false, 4, TT, true>), gr, b, 0, st, S0, E, C, tb, g, sf, nk, zsg, W, pr,
                                   no);
            return;
        }
        if constexpr (TT == 16) {
            if (W && K == 8) {
                if (use_ratio)
                    hipLaunchKernelGGL((k_sweep_tb<T, true, 8, 16, true>), gr, b, 0, st, S0, E, C, tb, g, sf, nk, zsg, W, pr,
                                       no);
                else
                    hipLaunchKernelGGL((k_sweep_tb<T, false, 8, 16, true>), gr, b, 0, st, S0, E, C, tb, g, sf, nk, zsg,
                                       W, pr, no);
                return;
            }
        }
    }
    if (K == 8) {
        if (use_ratio) hipLaunchKernelGGL((k_sweep_tb<T, true, 8, TT>), gr, b, 0, st, S0, E, C, tb, g, sf, nk, zsg, W, pr, no);
        else hipLaunchKernelGGL((k_sweep_tb<T, false, 8, TT>), gr, b, 0, st, S0, E, C, tb, g, sf, nk, zsg, W, pr, no);
    } else {
        if (use_ratio) hipLaunchKernelGGL((k_sweep_tb<T, true, 4, TT>), gr, b, 0, st, S0, E, C, tb, g, sf, nk, zsg, W, pr, no);
        else hipLaunchKernelGGL((k_sweep_tb<T, false, 4, TT>), gr, b, 0, st, S0, E, C, tb, g, sf, nk, zsg, W, pr, no);
    }
}

// the FP32 launch that also forms the next coarser level's right-hand side
template <int TT, int K>
static void tb_launch_rr(hipStream_t st, bool use_ratio, const Grid& g, const float* S0, float* E,
                         const Coefs<float>& C, const Tables<float>& tb, const SweepFactors<float>& sf, int nk,
                         float zsg, const float* W, const Prolong<float>& pr, const RRArgs& rr)
{
    const dim3 gr = tb_grid(g, TT), b(TB_THREADS);
    if (use_ratio)
        hipLaunchKernelGGL((k_sweep_tb<float, true, K, TT, true, true>), gr, b, 0, st, S0, E, C, tb, g, sf, nk, zsg, W,
                           pr, rr);
    else
        hipLaunchKernelGGL((k_sweep_tb<float, false, K, TT, true, true>), gr, b, 0, st, S0, E, C, tb, g, sf, nk, zsg, W,
                           pr, rr);
}

template <typename T>
static void tb_launch(hipStream_t st, bool use_ratio, int K, const Grid& g, const T* S0, T* E, const Coefs<T>& C,
                      const Tables<T>& tb, const SweepFactors<T>& sf, int nk, T zsg, const T* W = nullptr,
                      int tile = 0, const Prolong<T>& pr = Prolong<T>{})
{
    if constexpr (sizeof(T) == 4) {
        // 48x48 tiles (1.78x staged cells per tile cell against 2.25x for 32x32), precomputed weights only
        if (tile == 48 && K == 4 && W) {
            const dim3 gr = tb_grid(g, 48), b(TB_THREADS);
            const NoRR no{};
            if (use_ratio)
                hipLaunchKernelGGL((k_sweep_tb<T, true, 4, 48, true>), gr, b, 0, st, S0, E, C, tb, g, sf, nk, zsg, W, pr,
                                   no);
            else
                hipLaunchKernelGGL((k_sweep_tb<T, false, 4, 48, true>), gr, b, 0, st, S0, E, C, tb, g, sf, nk, zsg, W, pr,
                                   no);
            return;
        }
        if ((tile ? tile : tb_tile<T>(g)) == 32 && K == 4) {
            tb_launch_t<T, 32>(st, use_ratio, K, g, S0, E, C, tb, sf, nk, zsg, W, pr);
            return;
        }
    }
    tb_launch_t<T, 16>(st, use_ratio, K, g, S0, E, C, tb, sf, nk, zsg, W, pr);
}

// precomputed weights on every FP32 level: they save each launch the weight
// arithmetic of its staged cells (three per thread with 32x32 tiles), which
// bounds the launches more than the extra planes they read
bool tb_uses_w13(const Level& L) { (void)L; return true; }

void launch_tb_weights32(hipStream_t st, Level& L, const FD& F, const Spacing& h)
{
    if (!L.w13) {
        if (hipMalloc((void**)&L.w13, 13 * L.g.elems * sizeof(float)) != hipSuccess) {
            L.w13 = nullptr;
            return;
        }
    }
    // rhs: any plane (tb_weights also forms dts*rhs, unused here)
    Coefs<float> C{L.vis32, L.visx32, L.visz32, nullptr, nullptr, L.a132, L.visxz32, L.rhs32, L.dts32, nullptr, nullptr};
    dim3 b(64, 4);
    dim3 gr((L.g.nx - 1 + 63) / 64, (L.g.nz - 1 + 3) / 4);
    hipLaunchKernelGGL(k_tb_weights32, gr, b, 0, st, C, make_tables<float>(F, h), L.g, L.w13);
}

void launch_sweep_tb64(hipStream_t st, bool use_ratio, int K, const Level& L, const double* rhs, const double* S0,
                       double* E, const FD& F, const Spacing& h, const double* temp, const double* cconst, int nk,
                       double zsg)
{
    Coefs<double> C{L.vis, L.visx, L.visz, L.visxx, L.viszz, nullptr, L.visxz, rhs, L.dtstream, L.ratio, nullptr};
    SweepFactors<double> sf{};
    for (int s = 0; s < nk; ++s) { sf.temp[s] = temp[s]; sf.cconst[s] = cconst[s]; }
    tb_launch<double>(st, use_ratio, K, L.g, S0, E, C, make_tables<double>(F, h), sf, nk, zsg);
}

bool launch_sweep_tb32(hipStream_t st, bool use_ratio, int K, const Level& L, const float* rhs, const float* S0,
                       float* E, const FD& F, const Spacing& h, const double* temp, const double* cconst, int nk,
                       double zsg, const float* Pc, const Grid* gc, const RRTarget* rr)
{
    Coefs<float> C{L.vis32, L.visx32, L.visz32, nullptr, nullptr, L.a132, L.visxz32, rhs, L.dts32, nullptr, L.lb32};
    SweepFactors<float> sf{};
    for (int s = 0; s < nk; ++s) { sf.temp[s] = (float)temp[s]; sf.cconst[s] = (float)cconst[s]; }
    Prolong<float> pr;
    if (Pc && gc) {
        pr.Ec = Pc;
        pr.gc = *gc;
    }
    const Tables<float> tbl = make_tables<float>(F, h);
    const float* W = tb_uses_w13(L) ? L.w13 : nullptr;
    if (rr && rr->out && W) {
        // 16x16 tiles only: a 32x32 tile's 33^2 residuals take two per thread,
        // and its FP64 residual arithmetic then spills.  The wider staging
        // (1444 cells for K=4) needs ~85 VGPRs, so one workgroup fits a CU where
        // the plain launch fits two: on a level of more tiles than CUs the
        // fused launch runs in two rounds and costs more than the launch it
        // saves (16.0 against 8.3 + 6.3 us at 512 tiles; 8.6 against 5.7 + 6.3
        // at 128 tiles).
        const int TT = L.tb_tile ? L.tb_tile : tb_tile<float>(L.g);
        const dim3 t16 = tb_grid(L.g, 16);
        const RRArgs a{rr->out, rr->go, rr->RowB, rr->bout, fd32_from(F), h32_from(h)};
        if (TT == 16 && (K == 4 || K == 8) && (int)(t16.x * t16.y) <= rr->cus) {
            if (K == 4) tb_launch_rr<16, 4>(st, use_ratio, L.g, S0, E, C, tbl, sf, nk, (float)zsg, W, pr, a);
            else tb_launch_rr<16, 8>(st, use_ratio, L.g, S0, E, C, tbl, sf, nk, (float)zsg, W, pr, a);
            return true;
        }
    }
    tb_launch<float>(st, use_ratio, K, L.g, S0, E, C, tbl, sf, nk, (float)zsg, W, L.tb_tile, pr);
    return false;
}

}  // namespace tc
