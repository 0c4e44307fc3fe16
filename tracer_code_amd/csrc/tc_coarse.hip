// Coarsest-grid direct solve of the multigrid W-cycle
// (Compute_Matrix_coarse iterative_solver.f90:1026-1179, DGBEQU/DGBTRF at
// :1165/:97, Compute_coarse_vector/DGBTRS/Stream_grid_coarse :139-142,
// :1181-1204).
//
// The reference factors the row-equilibrated banded matrix with MKL and
// back-substitutes 8 times per 6-level W-cycle.  Sequential banded
// substitution is latency-bound on a GPU, so here the matrix is factored
// once per solve (LAPACK DGBTF2 semantics, one workgroup), its explicit
// inverse is formed column-parallel from those factors (DGBTRS applied to
// each unit vector), and every coarse visit becomes one dense GEMV that
// streams the inverse from HBM / Infinity Cache.
#include "tc_kernels.h"

namespace tc {

// Band storage: A(p,q) (1-based) at ab[(kl+ku+p-q) + (q-1)*ldab] (0-based row).
#define ABx(r, q) ab[(size_t)(r) + (size_t)(q) * (size_t)ldab]

// stream_space with build=1 (iterative_solver.f90:487-560), uniform grid:
// coefficient of S(iq,kq)*Sval in the row of interior point (i,k).
__device__ double stream_coeff(const FD& F, const Spacing& h, int ivis, double vis, double vx, double vz, double vxx,
                               double vzz, double vxz, int i, int k, int iq, int kq, double Sval)
{
    const int di = iq - i, dk = kq - k;
    double Sr = 0, Srr = 0, Srrr = 0, Srrrr = 0, Ss = 0, Sss = 0, Ssss = 0, Sssss = 0;
    if (k == kq) {
        Sr = F.D1[di + 2] * Sval / h.dr;
        Srr = F.D2[di + 2] * Sval / h.dr2;
        Srrr = F.D3[di + 2] * Sval / h.dr3;
        Srrrr = F.D4[di + 2] * Sval / h.dr4;
    }
    if (i == iq) {
        Ss = F.D1[dk + 2] * Sval / h.ds;
        Sss = F.D2[dk + 2] * Sval / h.ds2;
        Ssss = F.D3[dk + 2] * Sval / h.ds3;
        Sssss = F.D4[dk + 2] * Sval / h.ds4;
    }
    double m9 = 0.0, m9s = 0.0, m9r = 0.0, m9rs = 0.0;
    if (di >= -1 && di <= 1 && dk >= -1 && dk <= 1) {
        const int t = (dk + 1) * 3 + (di + 1);
        m9 = F.Drs[t]; m9s = F.Drss[t]; m9r = F.Drrs[t]; m9rs = F.Drrss[t];
    }
    double Srs = m9 * Sval / h.drds;
    double Srss = m9s * Sval / h.drds2;
    double Srrs = m9r * Sval / h.dr2ds;
    double Srrss = m9rs * Sval / h.dr2ds2;
    (void)Sr; (void)Ss;
    double t1, t2, t3, t4, t5, t6;
    if (ivis == 0) {
        t1 = 0.0; t2 = vis * (Srrrr + Sssss); t3 = 2.0 * vis * Srrss; t4 = 0.0; t5 = 0.0; t6 = 0.0;
    } else {
        t1 = (vxx - vzz) * (Srr - Sss);
        t2 = vis * (Srrrr + Sssss);
        t3 = 2.0 * vis * Srrss;
        t4 = 4.0 * vxz * Srs;
        t5 = 2.0 * vx * (Srrr + Srss);
        t6 = 2.0 * vz * (Ssss + Srrs);
    }
    return t1 + t2 + t3 + t4 + t5 + t6;
}

// One thread per matrix row p (Database=2 ordering p = (nz_c+1)*ip + kp + 1):
// set the in-domain stencil entries, add the folded ghost entries, identity
// rows on the walls, then DGBEQU's row factor R(p) = 1/max|A(p,:)| and the
// row scaling.  The band must be zero on entry.
__global__ __launch_bounds__(256) void k_coarse_assemble(double* __restrict__ ab, double* __restrict__ RowB,
                                                         const double* __restrict__ V, const double* __restrict__ vx,
                                                         const double* __restrict__ vz, const double* __restrict__ vxx,
                                                         const double* __restrict__ vzz, const double* __restrict__ vxz,
                                                         Grid g, FD F, Spacing h, int ivis, int Vbc, int kl, int ku, int n)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (p > n) return;
    const int ldab = 2 * kl + ku + 1, kl_ku_1 = kl + ku + 1;
    const int nxc = g.nx, nzc = g.nz, N = nzc + 1;
    const int kp = p - N * ((p - 1) / N) - 1, ip = (p - 1) / N;
    const int limit[5] = {0, 1, 2, 1, 0};
    if (ip > 0 && ip < nxc && kp > 0 && kp < nzc) {
        const size_t c = g.at(ip, kp);
        const double vis = V[c], gx = vx[c], gz = vz[c], gxx = vxx[c], gzz = vzz[c], gxz = vxz[c];
        for (int j = -2; j <= 2; ++j) {
            int iq = j + ip;
            if (iq >= 0 && iq <= nxc)
                for (int kq = -limit[j + 2] + kp; kq <= kp + limit[j + 2]; ++kq)
                    if (kq >= 0 && kq <= nzc) {
                        int q = N * iq + kq + 1;
                        ABx(kl_ku_1 + p - q - 1, q - 1) = stream_coeff(F, h, ivis, vis, gx, gz, gxx, gzz, gxz, ip, kp, iq, kq, 1.0);
                    }
        }
        for (int j = -2; j <= 2; ++j) {
            int iq = j + ip;
            for (int kq = -limit[j + 2] + kp; kq <= kp + limit[j + 2]; ++kq) {
                if (iq < 0 || iq > nxc || kq < 0 || kq > nzc) {
                    int iiq, kkq;
                    double xval, zval;
                    if (iq < 0) { iiq = -iq; xval = -1.0; }
                    else if (iq > nxc) { iiq = 2 * nxc - iq; xval = -1.0; }
                    else { iiq = iq; xval = 1.0; }
                    if (kq < 0) { kkq = -kq; zval = (Vbc == 0) ? -1.0 : 1.0; }
                    else if (kq > nzc) { kkq = 2 * nzc - kq; zval = (Vbc == 0) ? -1.0 : 1.0; }
                    else { kkq = kq; zval = 1.0; }
                    double Sval = xval * zval;
                    int q = N * iiq + kkq + 1;
                    double& a = ABx(kl_ku_1 + p - q - 1, q - 1);
                    a = a + stream_coeff(F, h, ivis, vis, gx, gz, gxx, gzz, gxz, ip, kp, iq, kq, Sval);
                }
            }
        }
    } else {
        ABx(kl_ku_1 - 1, p - 1) = 1.0;
    }
    // DGBEQU row factor over the band of row p, then the row scaling.
    const double smlnum = 2.2250738585072014e-308, bignum = 1.0 / smlnum;
    double rmax = 0.0;
    int q0 = p - kl > 1 ? p - kl : 1, q1 = p + ku < n ? p + ku : n;
    for (int q = q0; q <= q1; ++q) {
        double v = fabs(ABx(kl_ku_1 + p - q - 1, q - 1));
        if (v > rmax) rmax = v;
    }
    double rv = rmax > smlnum ? rmax : smlnum;
    rv = rv < bignum ? rv : bignum;
    double R = 1.0 / rv;
    RowB[p - 1] = R;
    // rows p with p-q in [-kl, ku] are scaled (iterative_solver.f90:1168-1172)
    int s0 = p - ku > 1 ? p - ku : 1, s1 = p + kl < n ? p + kl : n;
    for (int q = s0; q <= s1; ++q) {
        double& a = ABx(kl_ku_1 + p - q - 1, q - 1);
        a = R * a;
    }
}

// DGBTF2 (LAPACK, unblocked banded LU with partial pivoting) by one
// workgroup; same operations per element as the reference algorithm.
__global__ __launch_bounds__(256) void k_dgbtf2(double* __restrict__ ab, int ldab, int n, int kl, int ku,
                                                int32_t* __restrict__ ipiv, int32_t* __restrict__ info_out)
{
    const int kv = ku + kl;
    const int tid = threadIdx.x, nt = blockDim.x;
    __shared__ int sh_jp;
    __shared__ int sh_ju;
    __shared__ double sh_piv, sh_old;
    __shared__ int sh_info;
    for (int j = ku + 1; j < (kv < n ? kv : n); ++j)
        for (int i = kv - j + tid; i < kl; i += nt) ABx(i, j) = 0.0;
    if (tid == 0) { sh_ju = 0; sh_info = 0; }
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        if (j + kv < n)
            for (int i = tid; i < kl; i += nt) ABx(i, j + kv) = 0.0;
        const int km = kl < n - 1 - j ? kl : n - 1 - j;
        // IDAMAX over rows kv..kv+km of column j (first maximum wins), wave 0
        if (tid < 64) {
            double best = -1.0;
            int bi = 1 << 30;
            for (int t = tid; t <= km; t += 64) {
                double v = fabs(ABx(kv + t, j));
                if (v > best) { best = v; bi = t; }
            }
            for (int off = 32; off > 0; off >>= 1) {
                double ob = __shfl_down(best, off, 64);
                int oi = __shfl_down(bi, off, 64);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            if (tid == 0) {
                sh_jp = bi;
                sh_piv = ABx(kv + bi, j);
                sh_old = ABx(kv, j);
            }
        }
        __syncthreads();
        const int jp = sh_jp;
        const double piv = sh_piv;
        if (tid == 0) ipiv[j] = jp + j + 1;
        if (piv != 0.0) {
            int jt = j + ku + jp;
            if (jt > n - 1) jt = n - 1;
            const int ju = sh_ju > jt ? sh_ju : jt;
            // column j: swap + DSCAL by the reciprocal pivot; columns j+1..ju: swap
            const double rp = 1.0 / piv;
            if (tid == 0) ABx(kv, j) = piv;
            for (int t = 1 + tid; t <= km; t += nt) {
                double v = (t == jp) ? sh_old : ABx(kv + t, j);
                ABx(kv + t, j) = v * rp;
            }
            if (jp != 0)
                for (int t = 1 + tid; t <= ju - j; t += nt) {
                    double a = ABx(kv + jp - t, j + t);
                    ABx(kv + jp - t, j + t) = ABx(kv - t, j + t);
                    ABx(kv - t, j + t) = a;
                }
            __syncthreads();
            if (tid == 0) sh_ju = ju;
            // DGER rank-1 update of the trailing band
            if (km > 0 && ju > j) {
                const int ncol = ju - j;
                const int tot = ncol * km;
                for (int e = tid; e < tot; e += nt) {
                    int jj = e / km + 1, t = e % km + 1;
                    double y = ABx(kv - jj, j + jj);
                    if (y != 0.0) {
                        double temp = -1.0 * y;
                        ABx(kv + t - jj, j + jj) += ABx(kv + t, j) * temp;
                    }
                }
            }
        } else if (tid == 0 && sh_info == 0) {
            sh_info = j + 1;
        }
        __syncthreads();
    }
    if (tid == 0) *info_out = sh_info;
}

// DGBTF2 with the active part of the band held in LDS.  At step j the
// elimination touches rows j..j+kl of columns j..j+kl+ku; element (r, c)
// of that window lives at W[(c % LU_CS) * LU_RS + (r % LU_RS)], so the
// window slides by one row and one column per step without moving data.
// Each step writes the finished row j (U) and the multipliers of column j
// back to the band, loads the entering row j+kl+1 and zeroes the fill-in
// rows of the entering column j+kl+ku+1.  The arithmetic per element is
// DGBTF2's (swap, scale by the reciprocal pivot, rank-1 update skipping
// zero multiplicands), so the factors equal k_dgbtf2's.
constexpr int LU_CS = 136;  // column slots >= kl+ku+1 (and <= LU_THREADS: one entering element per thread)
constexpr int LU_RS = 128;  // row slots >= kl+2
constexpr int LU_THREADS = 320;  // 4 waves: a lane per row 1..64; wave 4: rows 65..km

__global__ __launch_bounds__(LU_THREADS) void k_dgbtf2w(double* __restrict__ ab, int ldab, int n, int kl, int ku,
                                                       int32_t* __restrict__ ipiv, int32_t* __restrict__ info_out)
{
    extern __shared__ double W[];  // LU_CS * LU_RS
    __shared__ int sh_jp, sh_ju, sh_info;
    __shared__ double sh_piv, sh_old;
    const int kv = ku + kl;
    const int tid = threadIdx.x;
    auto at = [&](int r, int c) -> double& { return W[(c % LU_CS) * LU_RS + (r % LU_RS)]; };
    // band element (r, c) of the stored (fill-in padded) band, 0 outside it
    auto band = [&](int r, int c) -> double {
        const int b = kv + r - c;
        return (b >= 0 && b < ldab && r >= 0 && r < n) ? ab[(size_t)b + (size_t)c * ldab] : 0.0;
    };
    // initial window: columns 0..min(kv, n-1), rows 0..kl; fill-in rows are zero
    const int c_end0 = kv < n - 1 ? kv : n - 1;
    for (int e = tid; e < (c_end0 + 1) * (kl + 1); e += LU_THREADS) {
        const int c = e / (kl + 1), r = e % (kl + 1);
        const int b = kv + r - c;
        at(r, c) = (b >= kl) ? band(r, c) : 0.0;
    }
    if (tid == 0) { sh_ju = 0; sh_info = 0; }
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        const int km = kl < n - 1 - j ? kl : n - 1 - j;
        // the row entering at the end of this step is still unmodified in
        // the band: fetch it now, store it into the window when sliding
        const int rn = j + 1 + kl, cn = j + 1 + kv;
        const int c_hi = cn < n - 1 ? cn : n - 1;
        const int c_in = j + 1 + tid;
        const double v_in = (rn <= n - 1 && c_in <= c_hi) ? band(rn, c_in) : 0.0;
        // IDAMAX over rows j..j+km of column j (first maximum wins), wave 0
        if (tid < 64) {
            double best = -1.0;
            int bi = 1 << 30;
            for (int t = tid; t <= km; t += 64) {
                const double v = fabs(at(j + t, j));
                if (v > best) { best = v; bi = t; }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_down(best, off, 64);
                const int oi = __shfl_down(bi, off, 64);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            if (tid == 0) {
                sh_jp = bi;
                sh_piv = at(j + bi, j);
                sh_old = at(j, j);
                ipiv[j] = bi + j + 1;
            }
        }
        __syncthreads();
        const int jp = sh_jp;
        const double piv = sh_piv;
        int ju = sh_ju;
        if (piv != 0.0) {
            int jt = j + ku + jp;
            if (jt > n - 1) jt = n - 1;
            ju = ju > jt ? ju : jt;
            // column j: swap + scale by the reciprocal pivot; columns j+1..ju: swap
            const double rp = 1.0 / piv;
            for (int t = 1 + tid; t <= km; t += LU_THREADS) {
                const double v = (t == jp) ? sh_old : at(j + t, j);
                at(j + t, j) = v * rp;
            }
            if (jp != 0)
                for (int c = j + 1 + tid; c <= ju; c += LU_THREADS) {
                    const double a = at(j + jp, c);
                    at(j + jp, c) = at(j, c);
                    at(j, c) = a;
                }
            __syncthreads();
            if (tid == 0) { at(j, j) = piv; sh_ju = ju; }
            // rank-1 update of rows j+1..j+km, columns j+1..ju.  Waves 0-3:
            // a lane per row 1..64, the waves striding over the columns
            // (the row's multiplier stays in a register, the column's U
            // element is a broadcast read); wave 4: rows 65..km, a lane per
            // column.
            const int lane = tid & 63, wv = tid >> 6;
            if (wv < 4) {
                const int rr = lane + 1;
                if (rr <= km) {
                    const double L = at(j + rr, j);
                    for (int c = j + 1 + wv; c <= ju; c += 4) {
                        const double y = at(j, c);
                        if (y != 0.0) {
                            const double temp = -1.0 * y;
                            double& a = at(j + rr, c);
                            a += L * temp;
                        }
                    }
                }
            } else {
                for (int rr = 65; rr <= km; ++rr) {
                    const double L = at(j + rr, j);
                    for (int c = j + 1 + lane; c <= ju; c += 64) {
                        const double y = at(j, c);
                        if (y != 0.0) {
                            const double temp = -1.0 * y;
                            double& a = at(j + rr, c);
                            a += L * temp;
                        }
                    }
                }
            }
        } else if (tid == 0 && sh_info == 0) {
            sh_info = j + 1;
        }
        __syncthreads();
        // write back the finished row j of columns j..min(j+kv, n-1) and the
        // multipliers of column j
        const int c_last = j + kv < n - 1 ? j + kv : n - 1;
        for (int c = j + tid; c <= c_last; c += LU_THREADS) ab[(size_t)(kv + j - c) + (size_t)c * ldab] = at(j, c);
        for (int t = 1 + tid; t <= km; t += LU_THREADS) ab[(size_t)(kv + t) + (size_t)j * ldab] = at(j + t, j);
        __syncthreads();
        // slide: the entering column j+1+kv (fill-in rows j+1..j+kl zero) and
        // the entering row j+1+kl of every column of the next window
        if (cn <= n - 1)
            for (int r = j + 1 + tid; r <= j + kl; r += LU_THREADS) at(r, cn) = 0.0;
        if (rn <= n - 1 && c_in <= c_hi) at(rn, c_in) = v_in;
        __syncthreads();
    }
    if (tid == 0) *info_out = sh_info;
}

// Explicit inverse from the DGBTF2 factors: lane c of a wavefront applies
// DGBTRS('N') to the unit vector e_c (the same operations, in the same
// order, as the reference algorithm), writing column c of A^-1 into the
// row-major Ainv (lanes touch consecutive columns: coalesced).  The rows a
// substitution step touches form a window of kl+1 (forward) or kl+ku+1
// (backward) rows that slides by one row per step; it lives in LDS as a
// circular buffer, so each step is LDS traffic plus one coalesced global
// store of the row that leaves the window.
template <int INV_COLS>
__global__ __launch_bounds__(64) void k_band_inverse(const double* __restrict__ ab, int ldab, int n, int kl, int ku,
                                                     const int32_t* __restrict__ ipiv, double* __restrict__ X)
{
    extern __shared__ double win[];  // [W][INV_COLS]
    const int lane = threadIdx.x;
    const int c = blockIdx.x * INV_COLS + lane;
    const bool active = c < n;
    const size_t nn = (size_t)n;
    const int kd = ku + kl;
    // ---- forward: L^-1 P e_c, window rows [j, j+kl] -----------------------
    const int Wf = kl + 1;
    const int cmin = blockIdx.x * INV_COLS;
    const int jstart = cmin - kl > 0 ? cmin - kl : 0;  // rows before this are zero for every lane
    for (int t = 0; t < Wf; ++t) {
        const int row = jstart + t;
        win[((row % Wf) * INV_COLS) + lane] = (row == c) ? 1.0 : 0.0;
    }
    for (int i = 0; i < jstart && active; ++i) X[(size_t)i * nn + c] = 0.0;
    for (int j = jstart; j < n; ++j) {
        if (j < n - 1) {
            const int lm = kl < n - 1 - j ? kl : n - 1 - j;
            const int l = ipiv[j] - 1;
            if (l != j) {
                double& a = win[(l % Wf) * INV_COLS + lane];
                double& b = win[(j % Wf) * INV_COLS + lane];
                const double tmp = a;
                a = b;
                b = tmp;
            }
            const double bj = win[(j % Wf) * INV_COLS + lane];
            if (bj != 0.0) {
                const double temp = -1.0 * bj;
                for (int t = 1; t <= lm; ++t) {
                    double& x = win[((j + t) % Wf) * INV_COLS + lane];
                    x += ab[(size_t)(kd + t) + (size_t)j * ldab] * temp;
                }
            }
        }
        // row j is final for the forward pass; row j+kl+1 enters the window
        const double yj = win[(j % Wf) * INV_COLS + lane];
        if (active) X[(size_t)j * nn + c] = yj;
        const int enter = j + Wf;
        win[(j % Wf) * INV_COLS + lane] = (enter == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    // ---- backward: U^-1 y, window rows [j-k, j] ---------------------------
    const int k = kl + ku;
    const int Wb = k + 1;
    for (int t = 0; t < Wb && t < n; ++t) {
        const int row = n - 1 - t;
        win[(row % Wb) * INV_COLS + lane] = active ? X[(size_t)row * nn + c] : 0.0;
    }
    for (int j = n - 1; j >= 0; --j) {
        double bj = win[(j % Wb) * INV_COLS + lane];
        if (bj != 0.0) {
            bj = bj / ab[(size_t)k + (size_t)j * ldab];
            const int i0 = j - k > 0 ? j - k : 0;
            for (int i = j - 1; i >= i0; --i) {
                double& x = win[(i % Wb) * INV_COLS + lane];
                x -= bj * ab[(size_t)(k + i - j) + (size_t)j * ldab];
            }
        }
        if (active) X[(size_t)j * nn + c] = bj;
        const int enter = j - Wb;
        if (enter >= 0) win[(j % Wb) * INV_COLS + lane] = active ? X[(size_t)enter * nn + c] : 0.0;
    }
}

// ---------------------------------------------------------------------------
// Explicit inverse from the DGBTF2 factors, column-parallel, in two passes.
//
// Forward (k_inv_fwd): lane c of a wavefront computes y = L^-1 P e_c with
// the DGBTRS 'N' elimination order (swap b(j) <-> b(ipiv(j)), then
// b(j+1:j+lm) -= L(:,j) b(j)).  The kl+1 rows the step touches are a
// register window that shifts by one row per step; the multipliers L(:,j)
// are uniform across the wavefront (scalar loads).  Rows above c-kl stay 0
// (a value moves up by at most one swap of reach kl), so each wavefront
// starts at its first column minus kl.  y is stored row-major in X.
//
// Backward (k_inv_bwd): x(j) = (y(j) - sum_t U(j,j+t) x(j+t)) / U(j,j) for
// j = n-1..0, the last kl+ku values of x of each lane held in an LDS ring,
// U read row-wise from a row-major copy of its band (k_u_rows).
// ---------------------------------------------------------------------------
constexpr int INV_CH = 32;     // elimination steps whose coefficients are staged at a time (forward)
constexpr int INV_CHB = 16;    // rows of U staged at a time (backward)
constexpr int INV_RING = 256;  // >= kl+ku+1 rows of x per lane

template <int KLT>
__global__ __launch_bounds__(64) void k_inv_fwd(const double* __restrict__ ab, int ldab, int n, int kl, int ku,
                                               const int32_t* __restrict__ ipiv, double* __restrict__ X)
{
    __shared__ double Ls[INV_CH][KLT];  // Ls[s][t-1] = L(j+s+t, j+s), 0 past lm
    __shared__ int ps[INV_CH];          // pivot offset of step j+s
    const int lane = threadIdx.x;
    const int c = blockIdx.x * 64 + lane;
    const bool active = c < n;
    const int kd = kl + ku;
    const int c0 = blockIdx.x * 64;
    const int jstart = c0 - kl > 0 ? c0 - kl : 0;
    // rows [0, jstart) of y are zero for every lane of this wavefront
    if (active)
        for (int j = 0; j < jstart; ++j) X[(size_t)j * n + c] = 0.0;
    double r[KLT + 1];
#pragma unroll
    for (int t = 0; t <= KLT; ++t) r[t] = (jstart + t == c) ? 1.0 : 0.0;
    for (int j0 = jstart; j0 < n; j0 += INV_CH) {
        const int cnt = n - j0 < INV_CH ? n - j0 : INV_CH;
        for (int e = lane; e < cnt * KLT; e += 64) {
            const int sdx = e / KLT, t = e % KLT + 1, jj = j0 + sdx;
            const int lm = kl < n - 1 - jj ? kl : n - 1 - jj;
            Ls[sdx][t - 1] = (t <= lm) ? ab[(size_t)kd + t + (size_t)jj * ldab] : 0.0;
        }
        for (int sdx = lane; sdx < cnt; sdx += 64) ps[sdx] = (j0 + sdx < n - 1) ? ipiv[j0 + sdx] - 1 - (j0 + sdx) : 0;
        __syncthreads();
        for (int sdx = 0; sdx < cnt; ++sdx) {
            const int j = j0 + sdx;
            const int p = __builtin_amdgcn_readfirstlane(ps[sdx]);  // uniform, 0..kl
#pragma unroll
            for (int t = 1; t <= KLT; ++t) {
                if (t == p) {
                    const double a = r[t];
                    r[t] = r[0];
                    r[0] = a;
                }
            }
            const double b0 = r[0];
            if (active) X[(size_t)j * n + c] = b0;
            const double temp = -1.0 * b0;
#pragma unroll
            for (int t = 1; t <= KLT; ++t) r[t - 1] = r[t] + Ls[sdx][t - 1] * temp;
            r[KLT] = (j + 1 + KLT == c) ? 1.0 : 0.0;
        }
        __syncthreads();
    }
}

// Ur[j*(kd+1) + t] = U(j, j+t), t = 0..kd (0 past the last column).
__global__ __launch_bounds__(256) void k_u_rows(const double* __restrict__ ab, int ldab, int n, int kd,
                                               double* __restrict__ Ur)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t w = (size_t)kd + 1;
    if (e >= (size_t)n * w) return;
    const int j = (int)(e / w), t = (int)(e % w);
    Ur[e] = (j + t < n) ? ab[(size_t)(kd - t) + (size_t)(j + t) * ldab] : 0.0;
}

template <int KDT>
__global__ __launch_bounds__(64) void k_inv_bwd(const double* __restrict__ Ur, int n, int kd, double* __restrict__ X)
{
    __shared__ double ring[INV_RING * 64];
    __shared__ double Us[INV_CHB][KDT + 1];  // Us[s][t] = U(j, j+t) for j = jb - s; 0 for t > kd
    const int lane = threadIdx.x;
    const int c = blockIdx.x * 64 + lane;
    const bool active = c < n;
    const size_t w = (size_t)kd + 1;
    for (int q = 0; q < INV_RING; ++q) ring[q * 64 + lane] = 0.0;
    for (int jb = n - 1; jb >= 0; jb -= INV_CHB) {
        const int cnt = jb + 1 < INV_CHB ? jb + 1 : INV_CHB;
        for (int e = lane; e < cnt * (KDT + 1); e += 64) {
            const int sdx = e / (KDT + 1), t = e % (KDT + 1);
            Us[sdx][t] = (t <= kd) ? Ur[(size_t)(jb - sdx) * w + t] : 0.0;
        }
        __syncthreads();
        for (int sdx = 0; sdx < cnt; ++sdx) {
            const int j = jb - sdx;
            // four partial sums break the dependence chain of the dot product
            double a0 = active ? X[(size_t)j * n + c] : 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
            static_assert(KDT % 4 == 0, "KDT is a multiple of 4");
#pragma unroll
            for (int t = 1; t <= KDT; t += 4) {
                a0 -= Us[sdx][t] * ring[((j + t) & (INV_RING - 1)) * 64 + lane];
                a1 -= Us[sdx][t + 1] * ring[((j + t + 1) & (INV_RING - 1)) * 64 + lane];
                a2 -= Us[sdx][t + 2] * ring[((j + t + 2) & (INV_RING - 1)) * 64 + lane];
                a3 -= Us[sdx][t + 3] * ring[((j + t + 3) & (INV_RING - 1)) * 64 + lane];
            }
            const double x = ((a0 + a1) + (a2 + a3)) / Us[sdx][0];
            ring[(j & (INV_RING - 1)) * 64 + lane] = x;
            if (active) X[(size_t)j * n + c] = x;
        }
        __syncthreads();
    }
}

// Coarse solve: B(p) = RHS(grids,i,k)*R(p) on interior rows (0 on walls),
// x = A^-1 B, error(grids) <- x on the interior with the ghost mirrors of
// enforceBCs_stream (walls stay 0).  One workgroup per interior row.
__global__ __launch_bounds__(256) void k_coarse_apply(const double* __restrict__ Ainv, const double* __restrict__ RowB,
                                                      const double* __restrict__ rhs, double* __restrict__ E, Grid g,
                                                      int n, double zsg)
{
    const int nxc = g.nx, nzc = g.nz, N = nzc + 1;
    const int ip = blockIdx.x + 1, kp = blockIdx.y + 1;  // interior unknown
    const int prow = N * ip + kp;                         // 0-based row of A^-1
    const double* row = Ainv + (size_t)prow * (size_t)n;
    double acc = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int kq = q % N, iq = q / N;
        if (iq > 0 && iq < nxc && kq > 0 && kq < nzc) acc += row[q] * (rhs[g.at(iq, kq)] * RowB[q]);
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) {
        E[g.at(ip, kp)] = acc;
        stream_ghosts(E, g, ip, kp, acc, zsg);
    }
}

// Coarse solve by banded substitution with the DGBTF2 factors (DGBTRS 'N',
// iterative_solver.f90:139-142): one workgroup, the solution vector in LDS,
// one barrier per substitution step.
__global__ __launch_bounds__(256) void k_coarse_lusolve(const double* __restrict__ ab, int ldab, int n, int kl, int ku,
                                                        const int32_t* __restrict__ ipiv,
                                                        const double* __restrict__ RowB, const double* __restrict__ rhs,
                                                        double* __restrict__ E, Grid g, double zsg)
{
    extern __shared__ double b[];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nxc = g.nx, nzc = g.nz, N = nzc + 1;
    for (int q = tid; q < n; q += nt) {
        const int kq = q % N, iq = q / N;
        b[q] = (iq > 0 && iq < nxc && kq > 0 && kq < nzc) ? rhs[g.at(iq, kq)] * RowB[q] : 0.0;
    }
    __syncthreads();
    const int kd = ku + kl;
    for (int j = 0; j < n - 1; ++j) {
        const int lm = kl < n - 1 - j ? kl : n - 1 - j;
        const int l = ipiv[j] - 1;
        double bj;
        if (l != j) {
            const double bl = b[l], bjj = b[j];
            __syncthreads();
            if (tid == 0) { b[l] = bjj; b[j] = bl; }
            bj = bl;
        } else {
            bj = b[j];
        }
        __syncthreads();
        if (bj != 0.0) {
            const double temp = -1.0 * bj;
            for (int t = 1 + tid; t <= lm; t += nt) b[j + t] += ABx(kd + t, j) * temp;
        }
        __syncthreads();
    }
    const int k = kl + ku;
    for (int j = n - 1; j >= 0; --j) {
        double bj = b[j];
        if (bj != 0.0) {
            bj = bj / ABx(k, j);
            const int i0 = j - k > 0 ? j - k : 0;
            for (int i = j - 1 - tid; i >= i0; i -= nt) b[i] -= bj * ABx(k + i - j, j);
        }
        __syncthreads();
        if (tid == 0) b[j] = bj;
        __syncthreads();
    }
    for (int q = tid; q < n; q += nt) {
        const int kq = q % N, iq = q / N;
        if (iq > 0 && iq < nxc && kq > 0 && kq < nzc) {
            E[g.at(iq, kq)] = b[q];
            stream_ghosts(E, g, iq, kq, b[q], zsg);
        }
    }
}

#undef ABx

void launch_coarse_lusolve(hipStream_t st, const double* ab, int ldab, int n, int kl, int ku, const int32_t* ipiv,
                           const double* RowB, const Level& L, double* E, double zsg)
{
    hipLaunchKernelGGL(k_coarse_lusolve, dim3(1), dim3(256), sizeof(double) * (size_t)n, st, ab, ldab, n, kl, ku, ipiv,
                       RowB, L.rhs, E, L.g, zsg);
}

void launch_coarse_assemble(hipStream_t st, double* Mc, double* RowB, const Level& L, const FD& F, const Spacing& h,
                            int ivis, int Vbc, int kl, int ku, int ngc)
{
    hipLaunchKernelGGL(k_coarse_assemble, dim3((ngc + 255) / 256), dim3(256), 0, st, Mc, RowB, L.vis, L.visx, L.visz,
                       L.visxx, L.viszz, L.visxz, L.g, F, h, ivis, Vbc, kl, ku, ngc);
}

void launch_dgbtf2(hipStream_t st, double* ab, int ldab, int n, int kl, int ku, int32_t* ipiv, int32_t* info)
{
    if (kl + ku + 1 <= LU_CS && kl + 2 <= LU_RS) {
        const size_t lds = sizeof(double) * LU_CS * LU_RS;
        hipFuncSetAttribute((const void*)k_dgbtf2w, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k_dgbtf2w, dim3(1), dim3(LU_THREADS), lds, st, ab, ldab, n, kl, ku, ipiv, info);
    } else {
        hipLaunchKernelGGL(k_dgbtf2, dim3(1), dim3(256), 0, st, ab, ldab, n, kl, ku, ipiv, info);
    }
}

void launch_band_inverse(hipStream_t st, const double* ab, int ldab, int n, int kl, int ku, const int32_t* ipiv,
                         double* Ainv, double* Ur)
{
    const int kd = kl + ku;
    const dim3 gw((n + 63) / 64), bw(64);
    bool done = true;
    if (kl <= 16) hipLaunchKernelGGL((k_inv_fwd<16>), gw, bw, 0, st, ab, ldab, n, kl, ku, ipiv, Ainv);
    else if (kl <= 32) hipLaunchKernelGGL((k_inv_fwd<32>), gw, bw, 0, st, ab, ldab, n, kl, ku, ipiv, Ainv);
    else if (kl <= 66) hipLaunchKernelGGL((k_inv_fwd<66>), gw, bw, 0, st, ab, ldab, n, kl, ku, ipiv, Ainv);
    else if (kl <= 96) hipLaunchKernelGGL((k_inv_fwd<96>), gw, bw, 0, st, ab, ldab, n, kl, ku, ipiv, Ainv);
    else done = false;
    if (done && kd <= 136) {
        const size_t tot = (size_t)n * (size_t)(kd + 1);
        hipLaunchKernelGGL(k_u_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, ab, ldab, n, kd, Ur);
        if (kd <= 32) hipLaunchKernelGGL((k_inv_bwd<32>), gw, bw, 0, st, Ur, n, kd, Ainv);
        else if (kd <= 64) hipLaunchKernelGGL((k_inv_bwd<64>), gw, bw, 0, st, Ur, n, kd, Ainv);
        else hipLaunchKernelGGL((k_inv_bwd<136>), gw, bw, 0, st, Ur, n, kd, Ainv);
        return;
    }
    // wide bands: one lane per column with an LDS window of kl+ku+1 rows
    const size_t row_bytes = sizeof(double) * (size_t)(kl + ku + 1);
    if (64 * row_bytes <= 65536)
        hipLaunchKernelGGL((k_band_inverse<64>), dim3((n + 63) / 64), dim3(64), 64 * row_bytes, st, ab, ldab, n, kl, ku,
                           ipiv, Ainv);
    else if (32 * row_bytes <= 65536)
        hipLaunchKernelGGL((k_band_inverse<32>), dim3((n + 31) / 32), dim3(32), 32 * row_bytes, st, ab, ldab, n, kl, ku,
                           ipiv, Ainv);
    else if (16 * row_bytes <= 65536)
        hipLaunchKernelGGL((k_band_inverse<16>), dim3((n + 15) / 16), dim3(16), 16 * row_bytes, st, ab, ldab, n, kl, ku,
                           ipiv, Ainv);
    else
        hipLaunchKernelGGL((k_band_inverse<8>), dim3((n + 7) / 8), dim3(8), 8 * row_bytes, st, ab, ldab, n, kl, ku, ipiv,
                           Ainv);
}

void launch_coarse_apply(hipStream_t st, const double* Ainv, const double* RowB, const Level& L, double* E, int ngc,
                         double zsg)
{
    hipLaunchKernelGGL(k_coarse_apply, dim3(L.g.nx - 1, L.g.nz - 1), dim3(256), 0, st, Ainv, RowB, L.rhs, E, L.g, ngc,
                       zsg);
}

}  // namespace tc
