// Field and tracer kernels: compute_velocities/velocityBCs, the strain-rate
// invariant, the T and C filters (smoother_time_T, smoother_time), the
// tracer-ratio binning (tracers_to_corners), load_conductivity_array,
// compute_derivatives, the RK tracer stages and the T-field RK stages of
// energy_time.  Uniform grids (xr=zs=1, xrr=zss=0).
#include "tc_kernels.h"

namespace tc {

static inline dim3 grid2(int nx, int ny, dim3 b) { return dim3((nx + b.x - 1) / b.x, (ny + b.y - 1) / b.y); }

// compute_velocities (momentum.f90:538-549): u = dSF/ds, w = -dSF/dr on 0..nx x 0..nz.
__global__ __launch_bounds__(256) void k_velocities(const double* __restrict__ SF, double* __restrict__ u,
                                                    double* __restrict__ w, Grid g, FD F, double dr, double ds)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > g.nx || k > g.nz) return;
    const size_t c = g.at(i, k);
    double SFr = dot3(F.D1, SF, c, 1) / dr;
    double SFs = dot3(F.D1, SF, c, g.ld) / ds;
    u[c] = SFs;
    w[c] = -SFr;
}

// velocityBCs (composition.f90:806-836), reference order, one workgroup.
__global__ void k_velocity_bcs(double* __restrict__ u, double* __restrict__ w, Grid g, int Vbc)
{
    const int nx = g.nx, nz = g.nz, tid = threadIdx.x, nt = blockDim.x;
    for (int k = tid; k <= nz; k += nt) { w[g.at(-1, k)] = w[g.at(1, k)]; w[g.at(nx + 1, k)] = w[g.at(nx - 1, k)]; }
    for (int i = tid; i <= nx; i += nt) {
        if (Vbc == 0) { u[g.at(i, -1)] = u[g.at(i, 1)]; u[g.at(i, nz + 1)] = u[g.at(i, nz - 1)]; }
        else { u[g.at(i, -1)] = -u[g.at(i, 1)]; u[g.at(i, nz + 1)] = -u[g.at(i, nz - 1)]; }
    }
    __syncthreads();
    for (int i = -1 + tid; i <= nx + 1; i += nt) { w[g.at(i, -1)] = -w[g.at(i, 1)]; w[g.at(i, nz + 1)] = -w[g.at(i, nz - 1)]; }
    for (int k = -1 + tid; k <= nz + 1; k += nt) { u[g.at(-1, k)] = -u[g.at(1, k)]; u[g.at(nx + 1, k)] = -u[g.at(nx - 1, k)]; }
    __syncthreads();
    for (int k = -1 + tid; k <= nz + 1; k += nt) { u[g.at(0, k)] = 0.0; u[g.at(nx, k)] = 0.0; }
    for (int i = -1 + tid; i <= nx + 1; i += nt) { w[g.at(i, 0)] = 0.0; w[g.at(i, nz)] = 0.0; }
    if (Vbc == 1) {
        __syncthreads();
        for (int i = -1 + tid; i <= nx + 1; i += nt) { u[g.at(i, 0)] = 0.0; u[g.at(i, nz)] = 0.0; }
    }
}

// compute_strain_rate_invariant (momentum.f90:39-52).
__global__ __launch_bounds__(256) void k_strain(const double* __restrict__ u, const double* __restrict__ w,
                                                double* __restrict__ st, Grid g, FD F, double dr, double ds)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > g.nx || k > g.nz) return;
    const size_t c = g.at(i, k);
    double ur = dot3(F.D1, u, c, 1) / dr;
    double us = dot3(F.D1, u, c, g.ld) / ds;
    double wr = dot3(F.D1, w, c, 1) / dr;
    double ws = dot3(F.D1, w, c, g.ld) / ds;
    double a = us + wr;
    st[c] = sqrt(0.5 * (2.0 * (ur * ur + ws * ws) + a * a));
}

// Symmetric ghosts of a node field: sides for k=0..nz, then full rows
// (strain momentum.f90:54-61, conductivity energy.f90:42-49).
__global__ void k_sym_bcs(double* __restrict__ f, Grid g)
{
    const int nx = g.nx, nz = g.nz, tid = threadIdx.x, nt = blockDim.x;
    for (int k = tid; k <= nz; k += nt)
        for (int i = 1; i <= SPAN; ++i) { f[g.at(-i, k)] = f[g.at(i, k)]; f[g.at(nx + i, k)] = f[g.at(nx - i, k)]; }
    __syncthreads();
    for (int i = -SPAN + tid; i <= nx + SPAN; i += nt)
        for (int k = 1; k <= SPAN; ++k) { f[g.at(i, -k)] = f[g.at(i, k)]; f[g.at(i, nz + k)] = f[g.at(i, nz - k)]; }
}

// ---------------------------------------------------------------------------
// Filter iterations of smoother_time_T / smoother_time, one launch each.
// The smoothness measure (smoother.f90:75-95) is the largest jump from a
// node to its five forward neighbours over i=0..nx-1, k=1..nz-1.  Maxima are
// combined by atomicMax on the bit patterns (the measure is >= 0, whose IEEE
// order is the integer order) into FILT_SLOTS slots on separate cache
// lines, block b into slot b % FILT_SLOTS, so that the atomics of different
// blocks rarely meet on one address.  Three slot sets rotate:
//   k_smoothness_max writes the measure of the loop's first iterate into
//     set 0;
//   k_filter_iter (iteration m) reads set m%3 and runs the stop test on it
//     (the whole grid reads the same value, so every block takes the same
//     branch; block (0,0) records it).  Otherwise it does the filter step
//     with the boundary conditions folded in: each cell gets its final value
//     after the step and the reference's BC routine, with the step
//     evaluated at the interior cell that a ghost mirrors.  It also writes
//     the measure of the iterate it produces into set (m+1)%3, from an LDS
//     tile of that iterate with a one-cell halo that the block recomputes,
//     and clears set (m+2)%3 for the iteration after.
// ---------------------------------------------------------------------------
constexpr int FILT_SLOTS = 64, FILT_STRIDE = 16;  // 16 x 8 B = one 128-byte line per slot
constexpr int FILT_SET = FILT_SLOTS * FILT_STRIDE;  // u64 words per slot set
constexpr int FILT_ROWS = 4;                      // rows per thread of both kernels

// the largest jump at one node from its own value t0 to the five forward
// neighbours, merged into the running maximum a.  fmax ignores a NaN jump,
// as the reference's max() (gfortran's MAX returns the non-NaN argument)
// and its "if (a.gt.diff)" do; a starts at 0 and never becomes NaN.  The
// result is one of the jumps (or 0) whatever the order of the merges.
__device__ __forceinline__ double smooth_merge(double a, double t0, double n1, double n2, double n3, double n4,
                                               double n5)
{
    a = fmax(a, fabs(n1 - t0));
    a = fmax(a, fabs(n2 - t0));
    a = fmax(a, fabs(n3 - t0));
    a = fmax(a, fabs(n4 - t0));
    return fmax(a, fabs(n5 - t0));
}

__device__ __forceinline__ void slot_max(unsigned long long* __restrict__ set, double a)
{
    const int s = (int)((blockIdx.y * gridDim.x + blockIdx.x) % FILT_SLOTS);
    atomicMax(set + s * FILT_STRIDE, (unsigned long long)__double_as_longlong(a));
}

__global__ __launch_bounds__(256) void k_smoothness_max(const double* __restrict__ T, Grid g,
                                                        unsigned long long* __restrict__ slots,
                                                        const int32_t* __restrict__ stop)
{
    if (stop[0]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0;
    for (int r = 0; r < FILT_ROWS; ++r) {
        const int k = (blockIdx.y * blockDim.y + threadIdx.y) * FILT_ROWS + r + 1;
        if (i <= g.nx - 1 && k <= g.nz - 1)
            a = smooth_merge(a, T[g.at(i, k)], T[g.at(i, k - 1)], T[g.at(i + 1, k - 1)], T[g.at(i + 1, k)],
                             T[g.at(i, k + 1)], T[g.at(i + 1, k + 1)]);
    }
    a = block_max(a);
    if (threadIdx.x == 0 && threadIdx.y == 0) slot_max(slots, a);
}

// dr^2 and ds^2 of the filter step, with their reciprocals: when both are
// powers of two (the kernels' POW2 instance) the quotient a/dr2 equals
// a*idr2 exactly, and the multiplication replaces the FP64 division sequence
struct FScale {
    double dr2, ds2, idr2, ids2;
    int pow2;
};
template <bool POW2>
__device__ __forceinline__ double fdiv_r(double a, const FScale& sc)
{
    if constexpr (POW2) return a * sc.idr2; else return a / sc.dr2;
}
template <bool POW2>
__device__ __forceinline__ double fdiv_s(double a, const FScale& sc)
{
    if constexpr (POW2) return a * sc.ids2; else return a / sc.ds2;
}

// the step at a node 0..nx x 0..nz (smoother.f90:319-329 with the clamp to
// [0,1] for C, :363-371 for T)
template <bool COMP, bool POW2>
__device__ __forceinline__ double filter_new(const double* __restrict__ T0, const double* __restrict__ dts,
                                             const Grid& g, const FD& F, int i, int k, const FScale& sc,
                                             double cfac, double courant)
{
    const size_t c = g.at(i, k);
    double Trr = fdiv_r<POW2>(dot3(F.D2, T0, c, 1), sc);
    double Tss = fdiv_s<POW2>(dot3(F.D2, T0, c, g.ld), sc);
    double temp = Trr + Tss;
    double d = dts[c] * courant;
    if constexpr (COMP) {
        double v = T0[c] + cfac * d * temp;
        if (v < 0.0) v = 0.0;
        if (v > 1.0) v = 1.0;
        return v;
    } else {
        return T0[c] + d * temp;
    }
}

// the value of cell (i,k), i in -SPAN..nx+SPAN, k in -SPAN..nz+SPAN, after
// one filter step and the boundary routine
template <bool COMP, bool POW2>
__device__ __forceinline__ double filter_cell(const double* __restrict__ T0, const double* __restrict__ dts,
                                              const Grid& g, const FD& F, int i, int k, const FScale& sc,
                                              double cfac, double courant, int Tbot)
{
    const int nx = g.nx, nz = g.nz;
    // column mirror of the side BCs: ghost columns read interior column 2nx-i / -i
    const int im = (i < 0) ? -i : (i > nx ? 2 * nx - i : i);
    auto W = [&](int ii, int kk) { return filter_new<COMP, POW2>(T0, dts, g, F, ii, kk, sc, cfac, courant); };
    if constexpr (COMP) {
        // enforceBCs_comp (composition.f90:792-803) in its interleaved
        // order: column +-1, row +-1, column +-2, row +-2; the first row
        // mirror still reads the old ghost columns +-2 (T0, which the step
        // copies there)
        if (k >= 0 && k <= nz) return W(im, k);
        const int ks = (k < 0) ? -k : 2 * nz - k;  // mirrored row
        const bool row1 = (k == -1 || k == nz + 1);
        if (row1 && (i == -2 || i == nx + 2)) return T0[g.at(i, ks)];
        return W(im, ks);
    } else {
        // enforceBCs (T): side mirrors on rows 0..nz, then the bottom/top
        // ghost rows from the mirrored rows, then the wall values
        if (k == nz) return 0.0;
        if (k == 0 && Tbot == 0) return 1.0;
        if (k >= 0 && k < nz) return W(im, k);
        if (k < 0) {
            const double x = W(im, -k);
            return (Tbot == 0) ? -(x - 1.0) + 1.0 : x;
        }
        return -W(im, 2 * nz - k);
    }
}

// A block covers 64 columns x FI_TILE rows; the LDS tile holds the new
// iterate on those rows plus the row above and below, and one more column.
constexpr int FI_BX = 64, FI_BY = 4, FI_TILE = FI_BY * FILT_ROWS;
constexpr int FI_HALO = (FI_BX + 1) * 2 + FI_TILE;  // cells the block evaluates beyond its own

template <bool COMP, bool POW2>
__global__ __launch_bounds__(256) void k_filter_iter(const double* __restrict__ T0, double* __restrict__ T,
                                                     const double* __restrict__ dts, Grid g, FD F, FScale sc,
                                                     double cfac, double courant, int Tbot,
                                                     unsigned long long* __restrict__ slots, int m, double delta,
                                                     int32_t* __restrict__ ctl)
{
    if (ctl[0]) return;
    __shared__ double s_max;
    __shared__ double tile[FI_TILE + 2][FI_BX + 1];
    const int tid = threadIdx.y * FI_BX + threadIdx.x;
    // the iteration's maximum over its slot set (first wave, one slot per lane)
    if (tid < 64) {
        unsigned long long v = slots[(size_t)(m % 3) * FILT_SET + tid * FILT_STRIDE];
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long w = __shfl_xor(v, o, 64);
            v = (w > v) ? w : v;
        }
        if (tid == 0) s_max = __longlong_as_double((long long)v);
    }
    __syncthreads();
    const bool first = blockIdx.x == 0 && blockIdx.y == 0;
    if (s_max < delta) {
        if (first && tid == 0) {
            ctl[0] = 1;
            ctl[1] = m;
        }
        return;
    }
    if (first && tid < FILT_SLOTS) slots[(size_t)((m + 2) % 3) * FILT_SET + tid * FILT_STRIDE] = 0ull;
    const int nx = g.nx, nz = g.nz;
    // x from i = -ORG_PAD, so that each row segment starts a cache line
    const int i0 = blockIdx.x * FI_BX - ORG_PAD, k0 = blockIdx.y * FI_TILE - SPAN;
    const int i = i0 + threadIdx.x;
    auto V = [&](int ii, int kk) { return filter_cell<COMP, POW2>(T0, dts, g, F, ii, kk, sc, cfac, courant, Tbot); };
    for (int r = 0; r < FILT_ROWS; ++r) {
        const int tk = threadIdx.y + FI_BY * r, k = k0 + tk;
        if (i >= -SPAN && i <= nx + SPAN && k <= nz + SPAN) {
            const double v = V(i, k);
            T[g.at(i, k)] = v;
            tile[tk + 1][threadIdx.x] = v;
        }
    }
    // the halo the measure reads: the rows above and below (65 columns) and
    // the column to the right; only nodes 0..nx x 0..nz are read
    for (int e = tid; e < FI_HALO; e += FI_BX * FI_BY) {
        int ti, tk;
        if (e < FI_BX + 1) { ti = e; tk = 0; }
        else if (e < 2 * (FI_BX + 1)) { ti = e - (FI_BX + 1); tk = FI_TILE + 1; }
        else { ti = FI_BX; tk = e - 2 * (FI_BX + 1) + 1; }
        const int ii = i0 + ti, kk = k0 + tk - 1;
        if (ii >= 0 && ii <= nx && kk >= 0 && kk <= nz) tile[tk][ti] = V(ii, kk);
    }
    __syncthreads();
    double a = 0.0;
    const int tx = threadIdx.x;
    for (int r = 0; r < FILT_ROWS; ++r) {
        const int tk = threadIdx.y + FI_BY * r + 1, k = k0 + tk - 1;
        if (i >= 0 && i <= nx - 1 && k >= 1 && k <= nz - 1)
            a = smooth_merge(a, tile[tk][tx], tile[tk - 1][tx], tile[tk - 1][tx + 1], tile[tk][tx + 1],
                             tile[tk + 1][tx], tile[tk + 1][tx + 1]);
    }
    a = block_max(a);
    if (tid == 0) slot_max(slots + (size_t)((m + 1) % 3) * FILT_SET, a);
}

// ---------------------------------------------------------------------------
// Several iterations of smoother_time's C filter per launch, with the
// iterates held in registers.  Each wavefront owns a tile of FM_TXO x FM_TY
// cells of the ghosted plane and holds it with a halo of FM_H = K+1 cells:
// 64 columns, one per lane, by FM_RY rows in the lane's registers.  An
// iteration updates the held rows from the previous iterate -- the step at
// every interior node with the left/right values from the neighbouring
// lanes (cross-lane shuffles) and the upper/lower ones from the lane's own
// rows -- then applies enforceBCs_comp: ghost columns copy the lane they
// mirror, ghost rows the row they mirror, and the row-1 corner cells
// (-2 or nx+2, -1 or nz+1) the previous iterate's value at row 1 (nz-1), as
// the reference's interleaved mirrors do.  The iterates are therefore
// bit-identical to one launch per iteration (filter_cell).  The valid part
// of the held region shrinks by one cell per iteration; the halo keeps the
// tile valid through K iterations and the last iterate's measure.  A ghost
// takes the step at the cell it mirrors, so a tile starting right of the
// last interior column (or above the last interior row) holds its region
// shifted inward by the shortfall.  dts_array is one value on the uniform
// grids the library accepts.  After each iteration the tile's own nodes give
// that iterate's smoothness maximum (one slot set per iterate).
//
// The stop test needs the whole grid's maximum, which a launch only sees
// for the iterates of the launch before it.  Launch l (iterate m = input)
// therefore first tests iterates m-K+1..m: if iterate j0 of them is smooth,
// the loop stops there.  j0 = m: nothing is computed (the input is the
// result).  j0 < m: the previous launch's input (iterate m-K, kept in a
// third buffer) is advanced j0-(m-K) iterations into the output.  Block
// (0,0) records the stop, the stopping iterate and the buffer holding it.
// Slot sets: 3K, set j mod 3K for iterate j; a launch reads K, writes K and
// clears the K its successor writes.
// ---------------------------------------------------------------------------
// a double from lane l-1 (wave_shr:1) or lane l+1 (wave_shl:1) of the
// wavefront, as DPP moves of its two halves; lane 0 (63) gets 0
__device__ __forceinline__ double dpp_from_left(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, 0x138, 0xF, 0xF, true),
                            __builtin_amdgcn_mov_dpp(lo, 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ double dpp_from_right(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, 0x130, 0xF, 0xF, true),
                            __builtin_amdgcn_mov_dpp(lo, 0x130, 0xF, 0xF, true));
}

constexpr int FM_K = 4;                  // iterations per launch
constexpr int FM_H = FM_K + 1;           // halo: one cell per iteration, one for the last measure
constexpr int FM_TXO = 64 - 2 * FM_H;    // own columns of a wavefront's tile
constexpr int FM_TY = 16;                // own rows
constexpr int FM_RY = FM_TY + 2 * FM_H;  // held rows
constexpr int FM_WAVES = 4;              // tiles (stacked in z) per workgroup
constexpr int FM_SETS = 3 * FM_K;

struct FmFac { double v[2 * FM_K]; };    // factors of the previous launch's iterations, then this one's
struct FmBufs { double* v[3]; };

// FAST: dr^2 and ds^2 are powers of two and D2 is (1, -2, 1) at offsets
// -1, 0, 1 (the second-order table); the products by 1 and -2 are then exact
// and are left out, and the quotients are products by the reciprocals: the
// same bits as the general form.
template <bool FAST>
__global__ __launch_bounds__(256) void k_filter_multi(FmBufs bufs, int b_in, const double* __restrict__ dts, Grid g,
                                                      FD F, FScale sc, double courant, FmFac fac, int m,
                                                      double delta, unsigned long long* __restrict__ slots,
                                                      int32_t* __restrict__ ctl)
{
    // ctl[0] = m+1 marks the stop taken by the launch with input iterate m:
    // the later launches return, the blocks of that launch still do its work
    {
        const int s = ctl[0];
        if (s != 0 && s != m + 1) return;
    }
    __shared__ int s_j0;
    const int lane = threadIdx.x, tid = threadIdx.y * 64 + lane;
    // the stop test on the iterates the previous launch produced (or the
    // input): the first wave reads one slot per lane of each set and takes
    // the maxima by shuffles
    if (tid < 64) {
        int j0 = -1;
        for (int j = (m - FM_K + 1 > 0 ? m - FM_K + 1 : 0); j <= m; ++j) {
            unsigned long long v = slots[(size_t)(j % FM_SETS) * FILT_SET + tid * FILT_STRIDE];
            for (int o = 32; o > 0; o >>= 1) {
                const unsigned long long w = __shfl_xor(v, o, 64);
                v = (w > v) ? w : v;
            }
            if (j0 < 0 && __longlong_as_double((long long)v) < delta) j0 = j;
        }
        if (tid == 0) s_j0 = j0;
    }
    __syncthreads();
    const int j0 = s_j0;
    const bool first = blockIdx.x == 0 && blockIdx.y == 0;
    const int b_prev = (b_in + 2) % 3, b_out = (b_in + 1) % 3;
    if (j0 == m) {
        if (first && tid == 0) { ctl[0] = m + 1; ctl[1] = m; ctl[2] = b_in; }
        return;
    }
    const bool replay = j0 >= 0;
    const int n = replay ? j0 - (m - FM_K) : FM_K;
    const double* __restrict__ src = replay ? bufs.v[b_prev] : bufs.v[b_in];
    double* __restrict__ dst = bufs.v[b_out];
    const int fo = replay ? 0 : FM_K;  // factor offset
    if (!replay && first && tid < FILT_SLOTS)
        for (int j = m + FM_K + 1; j <= m + 2 * FM_K; ++j)
            slots[(size_t)(j % FM_SETS) * FILT_SET + tid * FILT_STRIDE] = 0ull;
    const int nx = g.nx, nz = g.nz;
    const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);  // the wave's tile (uniform)
    const int i0 = blockIdx.x * FM_TXO - SPAN, k0 = (blockIdx.y * FM_WAVES + wy) * FM_TY - SPAN;
    const int sx = i0 - nx + 1 > 0 ? i0 - nx + 1 : 0, sk = k0 - nz + 1 > 0 ? k0 - nz + 1 : 0;
    const int ri0 = i0 - FM_H - sx, rk0 = k0 - FM_H - sk;  // held region origin
    const int i = ri0 + lane;                               // this lane's column
    const bool col_in = i >= -SPAN && i <= nx + SPAN, col_int = i >= 0 && i <= nx;
    const bool corner_col = i == -SPAN || i == nx + SPAN;
    const int im = (i < 0) ? -i : (i > nx ? 2 * nx - i : i);
    int ml = lane + (im - i);                               // lane holding the mirrored column
    ml = ml < 0 ? 0 : (ml > 63 ? 63 : ml);
    const double dtc = dts[g.at(0, 0)] * courant;
    double c[FM_RY];
#pragma unroll
    for (int r = 0; r < FM_RY; ++r) {
        const int k = rk0 + r;
        c[r] = (col_in && k >= -SPAN && k <= nz + SPAN) ? src[g.at(i, k)] : 0.0;
    }
    const bool ghost_col = col_in && !col_int;
    const bool gcols = ri0 <= -1 || ri0 + 63 >= nx + 1;  // the held columns include ghost columns (uniform)
    const bool own_col = i >= i0 && i < i0 + FM_TXO && i >= 0 && i <= nx - 1;  // columns measured
    for (int it = 1; it <= n; ++it) {
        const double cfd = fac.v[fo + it - 1] * dtc;
        // the row predicates below are recomputed from this copy in every
        // iteration rather than held across the loop
        int rk = rk0;
        asm volatile("" : "+s"(rk));
        const bool top = rk <= -1, bot = rk + FM_RY - 1 >= nz + 1;  // the held rows include ghost rows
        // the previous iterate's rows 1 and nz-1 (the row-1 corners read them)
        double old1 = 0.0, oldn1 = 0.0;
        if (top) {
#pragma unroll
            for (int r = 0; r < FM_RY; ++r)
                if (rk + r == 1) old1 = c[r];
        }
        if (bot) {
#pragma unroll
            for (int r = 0; r < FM_RY; ++r)
                if (rk + r == nz - 1) oldn1 = c[r];
        }
        // the step (filter_new's operation order) on every held row but the
        // first and last; ghost cells are overwritten below, cells outside
        // the plane are never read by the valid part of the region
        double om = c[0];
#pragma unroll
        for (int r = 1; r < FM_RY - 1; ++r) {
            const double oc = c[r];
            const double lf = dpp_from_left(oc), rt = dpp_from_right(oc);
            double Trr, Tss;
            if constexpr (FAST) {
                const double m2 = -2.0 * oc;
                Trr = ((lf + m2) + rt) * sc.idr2;
                Tss = ((om + m2) + c[r + 1]) * sc.ids2;
            } else {
                double acc = F.D2[1] * lf;
                acc += F.D2[2] * oc;
                acc += F.D2[3] * rt;
                Trr = acc / sc.dr2;
                acc = F.D2[1] * om;
                acc += F.D2[2] * oc;
                acc += F.D2[3] * c[r + 1];
                Tss = acc / sc.ds2;
            }
            const double temp = Trr + Tss;
            double v = oc + cfd * temp;
            if (v < 0.0) v = 0.0;
            if (v > 1.0) v = 1.0;
            c[r] = v;
            om = oc;
        }
        // ghost columns: the mirrored lane's new value (ghost rows are
        // overwritten next)
        if (gcols) {
#pragma unroll
            for (int r = 1; r < FM_RY - 1; ++r) {
                const double mv = __shfl(c[r], ml, 64);
                c[r] = ghost_col ? mv : c[r];
            }
        }
        // ghost rows: the mirrored row (columns already mirrored), the row-1
        // corners from the previous iterate
        if (top) {
#pragma unroll
            for (int r = 0; r + 4 < FM_RY; ++r) {
                if (rk + r == -1) c[r] = corner_col ? old1 : c[r + 2];
                if (rk + r == -2) c[r] = c[r + 4];
            }
        }
        if (bot) {
#pragma unroll
            for (int r = 4; r < FM_RY; ++r) {
                if (rk + r == nz + 1) c[r] = corner_col ? oldn1 : c[r - 2];
                if (rk + r == nz + 2) c[r] = c[r - 4];
            }
        }
        if (!replay) {
            // the new iterate's smoothness over the tile's own nodes of
            // rows 1..nz-1 (held rows ra..rb): the jumps of smooth_merge,
            // each vertical pair once (node r's jump up is node r-1's jump
            // down, |x-y| = |y-x|); the maximum is that of smooth_merge
            const int ka = k0 > 1 ? k0 : 1, kb = k0 + FM_TY - 1 < nz - 1 ? k0 + FM_TY - 1 : nz - 1;
            const int ra = ka - rk, rb = kb - rk;  // ra >= FM_H, rb <= FM_RY - 1 - FM_H
            double a = 0.0;
            double rm = 0.0, r0 = 0.0;  // the right lane's c[r-1], c[r]
#pragma unroll
            for (int r = 1; r < FM_RY - 1; ++r) {
                // rows outside ra-2..rb are skipped by the whole wave
                if (r + 2 >= ra && r <= rb) {
                    const double rp = dpp_from_right(c[r + 1]);
                    if (r >= ra) {
                        if (r == ra) a = fmax(a, fabs(c[r - 1] - c[r]));
                        a = fmax(a, fabs(rm - c[r]));
                        a = fmax(a, fabs(r0 - c[r]));
                        a = fmax(a, fabs(c[r + 1] - c[r]));
                        a = fmax(a, fabs(rp - c[r]));
                    }
                    rm = r0;
                    r0 = rp;
                }
            }
            a = own_col ? a : 0.0;
            a = block_max(a);
            if (tid == 0) slot_max(slots + (size_t)((m + it) % FM_SETS) * FILT_SET, a);
        }
    }
    // the tile's own cells of the last iterate
    const bool own_w = col_in && i >= i0 && i < i0 + FM_TXO;
#pragma unroll
    for (int r = 0; r < FM_RY; ++r) {
        const int k = rk0 + r;
        if (own_w && k >= k0 && k < k0 + FM_TY && k >= -SPAN && k <= nz + SPAN) dst[g.at(i, k)] = c[r];
    }
    if (replay && first && tid == 0) { ctl[0] = m + 1; ctl[1] = j0; ctl[2] = b_out; }
}

// enforceBCs (energy.f90:455-475), one workgroup.
__global__ void k_bcs_T(double* __restrict__ T, Grid g, int Tbot, const int32_t* __restrict__ stop)
{
    if (stop && stop[0]) return;
    const int nx = g.nx, nz = g.nz, tid = threadIdx.x, nt = blockDim.x;
    for (int k = tid; k <= nz; k += nt)
        for (int i = 1; i <= SPAN; ++i) { T[g.at(-i, k)] = T[g.at(i, k)]; T[g.at(nx + i, k)] = T[g.at(nx - i, k)]; }
    __syncthreads();
    for (int i = -SPAN + tid; i <= nx + SPAN; i += nt)
        for (int k = 1; k <= SPAN; ++k) {
            if (Tbot == 0) T[g.at(i, -k)] = -(T[g.at(i, k)] - 1.0) + 1.0;
            else T[g.at(i, -k)] = T[g.at(i, k)];
            T[g.at(i, nz + k)] = -T[g.at(i, nz - k)];
        }
    __syncthreads();
    for (int i = -SPAN + tid; i <= nx + SPAN; i += nt) {
        if (Tbot == 0) T[g.at(i, 0)] = 1.0;
        T[g.at(i, nz)] = 0.0;
    }
}

// enforceBCs_comp (composition.f90:792-803) in the reference's interleaved
// order, one workgroup.
__global__ void k_bcs_comp(double* __restrict__ C, Grid g, const int32_t* __restrict__ stop)
{
    if (stop && stop[0]) return;
    const int nx = g.nx, nz = g.nz, tid = threadIdx.x, nt = blockDim.x;
    for (int i = 1; i <= SPAN; ++i) {
        for (int k = tid; k <= nz; k += nt) { C[g.at(-i, k)] = C[g.at(i, k)]; C[g.at(nx + i, k)] = C[g.at(nx - i, k)]; }
        __syncthreads();
        for (int ii = -SPAN + tid; ii <= nx + SPAN; ii += nt) { C[g.at(ii, -i)] = C[g.at(ii, i)]; C[g.at(ii, nz + i)] = C[g.at(ii, nz - i)]; }
        __syncthreads();
    }
}

__global__ void k_zero_i32(int32_t* __restrict__ p, size_t n)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = 0;
}

// tracers_to_corners counting loop (composition.f90:704-715): nearest node
// by Fortran NINT (half away from zero) of r/dr and s/ds, domain check.
template <bool POW2>
__global__ __launch_bounds__(256) void k_bin(const double* __restrict__ r, const double* __restrict__ s,
                                             const int32_t* __restrict__ ty, int64_t n, CntArr cnt, Grid g, double dr,
                                             double ds, double idr, double ids, double aspect, int32_t* __restrict__ flag)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // key of the (type, node) counter this tracer adds to; none = ~0
    uint32_t key = ~0u;
    int t = 0;
    size_t node = 0;
    if (id < n) {
        const double rv = r[id], sv = s[id];
        // the reference's test (r<=0 or r>=aspect or s<=0 or s>=1), written so
        // that a NaN position is also rejected before it can form an index
        if (!(rv > 0.0 && rv < aspect && sv > 0.0 && sv < 1.0)) {
            atomicOr(flag, 1);
        } else {
            const int i = (int)round(sdiv<POW2>(rv, dr, idr));
            const int k = (int)round(sdiv<POW2>(sv, ds, ids));
            node = g.at(i, k);
            t = ty[id];
            key = (uint32_t)node * (uint32_t)(TC_MAX_TYPES + 1) + (uint32_t)t;
        }
    }
    // wavefront segmented reduction: consecutive lanes with the same key form
    // a run (tracers are stored node by node, and move less than a cell per
    // step, so runs of ~tpc are the rule); the run's first lane adds the
    // run's length, one atomic per run instead of one per tracer
    const int lane = threadIdx.x & 63;
    const uint32_t prev = __shfl_up(key, 1, 64);
    const bool boundary = lane == 0 || prev != key;
    const unsigned long long starts = __ballot(boundary);
    if (boundary && key != ~0u) {
        const unsigned long long later = (lane == 63) ? 0ull : (starts >> (lane + 1));
        const int len = later ? __builtin_ctzll(later) + 1 : 64 - lane;
        atomicAdd(&cnt.v[t][node], len);
    }
}

// Mirrored ghost counts (composition.f90:745-754), nempty (:756) and the
// tracer ratio (:757-776) in one pass over the ghosted grid.
__global__ __launch_bounds__(256) void k_ratio(CntArr cnt, int32_t* __restrict__ nempty, PtrArr Cnew, Grid g, int ntype,
                                               int comp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x - SPAN;
    const int k = blockIdx.y * blockDim.y + threadIdx.y - SPAN;
    const int nx = g.nx, nz = g.nz;
    if (i > nx + SPAN || k > nz + SPAN) return;
    const bool gi = i < 0 || i > nx, gk = k < 0 || k > nz;
    const size_t c = g.at(i, k);
    int tot = 0;
    if (gi && gk) {
        for (int t = 0; t <= ntype; ++t) cnt.v[t][c] = 0;
    } else if (gi || gk) {
        const int mi = i < 0 ? -i : (i > nx ? 2 * nx - i : i);
        const int mk = k < 0 ? -k : (k > nz ? 2 * nz - k : k);
        const size_t m = g.at(mi, mk);
        for (int t = 0; t <= ntype; ++t) {
            int v = cnt.v[t][m];
            cnt.v[t][c] = v;
            tot += v;
        }
    } else {
        int ct[TC_MAX_TYPES + 1];
        for (int t = 0; t <= ntype; ++t) { ct[t] = cnt.v[t][c]; tot += ct[t]; }
        if (tot > 0 && comp == 1)
            for (int t = 1; t <= ntype; ++t) {
                double v;
                if (ct[t] == tot) v = 1.0;
                else if (ct[t] == 0) v = 0.0;
                else v = (double)ct[t] / (double)tot;
                Cnew.v[t - 1][c] = v;
            }
    }
    nempty[c] = tot;
}

// conductivity + load_conductivity_array (energy.f90:16-58): the ghost
// values of conduct are the mirrored interior ones, so every point of the
// ghosted grid is computed from Cvis at its mirror point.
__global__ __launch_bounds__(256) void k_conduct(PtrArr Cvis, double* __restrict__ K, Grid g, int comp, int ntype,
                                                 RaCArr lcf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x - SPAN;
    const int k = blockIdx.y * blockDim.y + threadIdx.y - SPAN;
    const int nx = g.nx, nz = g.nz;
    if (i > nx + SPAN || k > nz + SPAN) return;
    const int mi = i < 0 ? -i : (i > nx ? 2 * nx - i : i);
    const int mk = k < 0 ? -k : (k > nz ? 2 * nz - k : k);
    const size_t m = g.at(mi, mk);
    double v = 1.0;
    if (comp != 0)
        for (int j = 0; j < ntype; ++j) v = exp(-Cvis.v[j][m] * lcf.v[j]);
    K[g.at(i, k)] = v;
}

__global__ __launch_bounds__(256) void k_conduct_grad(const double* __restrict__ K, double* __restrict__ Kr,
                                                      double* __restrict__ Ks, Grid g, FD F, double dr, double ds)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x - 1;
    const int k = blockIdx.y * blockDim.y + threadIdx.y - 1;
    if (i > g.nx + 1 || k > g.nz + 1) return;
    const size_t c = g.at(i, k);
    Kr[c] = dot3(F.D1, K, c, 1) / dr;
    Ks[c] = dot3(F.D1, K, c, g.ld) / ds;
}

// compute_derivatives (composition.f90:900-919) of u and w, k=-1..nz+1.
__global__ __launch_bounds__(256) void k_derivs(const double* __restrict__ u, const double* __restrict__ w,
                                                double* __restrict__ Du, double* __restrict__ Dw, Grid g, D3Arr Dint,
                                                double dr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y * blockDim.y + threadIdx.y - 1;
    if (i > g.nx || k > g.nz + 1) return;
    const size_t c = g.at(i, k);
    double a = Dint.v[0] * u[c - 1];
    a += Dint.v[1] * u[c];
    a += Dint.v[2] * u[c + 1];
    Du[c] = a / dr;
    double b = Dint.v[0] * w[c - 1];
    b += Dint.v[1] * w[c];
    b += Dint.v[2] * w[c + 1];
    Dw[c] = b / dr;
}

// interpolate (composition.f90:838-898), order=2.  POW2: dr and ds are
// powers of two, so the quotients are the products by the reciprocals
// (idr, ids), bit for bit.
template <bool POW2>
__device__ __forceinline__ void interp_uw(double r, double s, const double* __restrict__ u, const double* __restrict__ w,
                                          const double* __restrict__ Du, const double* __restrict__ Dw, const Grid& g,
                                          const D3Arr& D, double dr, double ds, double idr, double ids, double& ou,
                                          double& ow)
{
    int i = (int)sdiv<POW2>(r, dr, idr);
    int k = (int)sdiv<POW2>(s, ds, ids);
    // in-domain positions give 0<=i<=nx-1, 0<=k<=nz-1; the clamp only keeps
    // the gathers inside the ghost ring for invalid input
    i = i < 0 ? 0 : (i > g.nx - 1 ? g.nx - 1 : i);
    k = k < 0 ? 0 : (k > g.nz - 1 ? g.nz - 1 : k);
    double hr = r - (double)i * dr;
    double fu[4], fw[4];
    int ib;
    double a1;
    if (hr <= 0.5 * dr) { ib = i; a1 = hr; }
    else { hr = dr - hr; ib = i + 1; a1 = -hr; }
#pragma unroll
    for (int jj = -1; jj <= 2; ++jj) {
        const size_t c = g.at(ib, k + jj);
        fu[jj + 1] = u[c] + Du[c] * a1;
        fw[jj + 1] = w[c] + Dw[c] * a1;
    }
    double hs = s - (double)k * ds;
    if (hs <= 0.5 * ds) {
        double au = D.v[0] * fu[0]; au += D.v[1] * fu[1]; au += D.v[2] * fu[2];
        double aw = D.v[0] * fw[0]; aw += D.v[1] * fw[1]; aw += D.v[2] * fw[2];
        ou = fu[1] + sdiv<POW2>(au, ds, ids) * hs;
        ow = fw[1] + sdiv<POW2>(aw, ds, ids) * hs;
    } else {
        hs = ds - hs;
        double au = D.v[0] * fu[1]; au += D.v[1] * fu[2]; au += D.v[2] * fu[3];
        double aw = D.v[0] * fw[1]; aw += D.v[1] * fw[2]; aw += D.v[2] * fw[3];
        ou = fu[2] + sdiv<POW2>(au, ds, ids) * (-hs);
        ow = fw[2] + sdiv<POW2>(aw, ds, ids) * (-hs);
    }
}

// RK stage `stage` of the tracer loops (energy.f90:182-225, 270-312,
// 355-378, 420-440), tracer=0, uniform grids, global dt.
template <bool POW2>
__global__ __launch_bounds__(256) void k_tracer_stage(int stage, int torder, double dt, int64_t n, double* __restrict__ r,
                                                      double* __restrict__ s, double* __restrict__ r0,
                                                      double* __restrict__ s0, double* __restrict__ r1,
                                                      double* __restrict__ s1, double* __restrict__ r2,
                                                      double* __restrict__ s2, const double* __restrict__ u,
                                                      const double* __restrict__ w, const double* __restrict__ Du,
                                                      const double* __restrict__ Dw, Grid g, D3Arr D, double dr, double ds,
                                                      double idr, double ids)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    double tu, tw;
    if (stage == 1) {
        const double x0 = r[id], z0 = s[id];  // rtr0 = rtr (energy.f90:147-148)
        r0[id] = x0; s0[id] = z0;
        interp_uw<POW2>(x0, z0, u, w, Du, Dw, g, D, dr, ds, idr, ids, tu, tw);
        const double dtstage = (torder == 2) ? dt : 0.5 * dt;
        r[id] = tu; s[id] = tw;
        r1[id] = x0 + dtstage * tu;
        s1[id] = z0 + dtstage * tw;
    } else if (stage == 2) {
        interp_uw<POW2>(r1[id], s1[id], u, w, Du, Dw, g, D, dr, ds, idr, ids, tu, tw);
        const double dtstage = 0.5 * dt;
        if (torder == 2) {
            double a = r[id] + tu, b = s[id] + tw;
            r[id] = r0[id] + dtstage * a;
            s[id] = s0[id] + dtstage * b;
        } else {
            r[id] = r[id] + 2.0 * tu;
            s[id] = s[id] + 2.0 * tw;
            r2[id] = r0[id] + dtstage * tu;
            s2[id] = s0[id] + dtstage * tw;
        }
    } else if (stage == 3) {
        interp_uw<POW2>(r2[id], s2[id], u, w, Du, Dw, g, D, dr, ds, idr, ids, tu, tw);
        r[id] = r[id] + 2.0 * tu;
        r1[id] = r0[id] + dt * tu;
        s[id] = s[id] + 2.0 * tw;
        s1[id] = s0[id] + dt * tw;
    } else {
        interp_uw<POW2>(r1[id], s1[id], u, w, Du, Dw, g, D, dr, ds, idr, ids, tu, tw);
        const double dto6 = dt / 6.0;
        r[id] = r0[id] + dto6 * (r[id] + tu);
        s[id] = s0[id] + dto6 * (s[id] + tw);
    }
}

// energy_space (energy.f90:60-80) with heating (energy.f90:1-14).
__device__ __forceinline__ double energy_space(const double* __restrict__ T0, size_t c, ptrdiff_t ld, const FD& F,
                                               double dr, double ds, double dr2, double ds2, double K, double Kr,
                                               double Ks, double uu, double ww, double heat)
{
    double Tr = dot3(F.D1, T0, c, 1) / dr;
    double Ts = dot3(F.D1, T0, c, ld) / ds;
    double Trr = dot3(F.D2, T0, c, 1) / dr2;
    double Tss = dot3(F.D2, T0, c, ld) / ds2;
    return K * (Trr + Tss) + (Kr * Tr + Ks * Ts) - uu * Tr - ww * Ts + heat;
}

__global__ __launch_bounds__(256) void k_energy_stage(int stage, double dt, double* __restrict__ T,
                                                      const double* __restrict__ T0, double* __restrict__ T1,
                                                      double* __restrict__ T2, const double* __restrict__ u,
                                                      const double* __restrict__ w, const double* __restrict__ K,
                                                      const double* __restrict__ Kr, const double* __restrict__ Ks,
                                                      PtrArr Cvis, Grid g, FD F, double dr, double ds, double dr2,
                                                      double ds2, int comp, int ntype, double H, RaCArr Htype)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > g.nx || k > g.nz) return;
    const size_t c = g.at(i, k);
    double heat;
    if (comp == 0) heat = H;
    else {
        double sum = 0.0;
        for (int j = 0; j < ntype; ++j) sum += Cvis.v[j][c];
        heat = H * (1.0 - sum);
        for (int j = 0; j < ntype; ++j) heat = Htype.v[j] * Cvis.v[j][c] + heat;
    }
    const double* src = (stage == 1) ? T0 : (stage == 3 ? T2 : T1);
    double temp = energy_space(src, c, g.ld, F, dr, ds, dr2, ds2, K[c], Kr[c], Ks[c], u[c], w[c], heat);
    if (stage == 1) { T[c] = temp; T1[c] = T0[c] + 0.5 * dt * temp; }
    else if (stage == 2) { T[c] = T[c] + 2.0 * temp; T2[c] = T0[c] + 0.5 * dt * temp; }
    else if (stage == 3) { T[c] = T[c] + 2.0 * temp; T1[c] = T0[c] + dt * temp; }
    else { T[c] = T0[c] + (dt / 6.0) * (T[c] + temp); }
}

// stability_RK4 maxima of eigA / eigD (energy.f90:614-623, 686-706), tracer=0.
__global__ __launch_bounds__(256) void k_stability(const double* __restrict__ u, const double* __restrict__ w,
                                                   const double* __restrict__ K, const double* __restrict__ Kr,
                                                   const double* __restrict__ Ks, Grid g, double Tr_max, double Ts_max,
                                                   double Trr_max, double Tss_max, double* __restrict__ pA,
                                                   double* __restrict__ pD)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y * blockDim.y + threadIdx.y;
    double eA = 0.0, eD = 0.0;
    if (i <= g.nx && k <= g.nz) {
        const size_t c = g.at(i, k);
        const double Kc = K[c];
        eA = fabs(-(Kc * 0.0 / 1.0) - u[c] / 1.0 + Kr[c] / 1.0) * Tr_max +
             fabs(-(Kc * 0.0 / 1.0) - w[c] / 1.0 + Ks[c] / 1.0) * Ts_max;
        eD = Kc * (fabs(Trr_max / 1.0) + fabs(Tss_max / 1.0));
    }
    eA = block_max(eA);
    eD = block_max(eD);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        pA[blockIdx.y * gridDim.x + blockIdx.x] = eA;
        pD[blockIdx.y * gridDim.x + blockIdx.x] = eD;
    }
}

// ---------------------------------------------------------------------------
void launch_velocities(hipStream_t st, const double* SF, double* u, double* w, const Grid& g, const FD& F, double dr,
                       double ds, int Vbc)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_velocities, grid2(g.nx + 1, g.nz + 1, b), b, 0, st, SF, u, w, g, F, dr, ds);
    hipLaunchKernelGGL(k_velocity_bcs, dim3(1), dim3(256), 0, st, u, w, g, Vbc);
}

void launch_strain(hipStream_t st, const double* u, const double* w, double* strain, const Grid& g, const FD& F,
                   double dr, double ds)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_strain, grid2(g.nx + 1, g.nz + 1, b), b, 0, st, u, w, strain, g, F, dr, ds);
    hipLaunchKernelGGL(k_sym_bcs, dim3(1), dim3(256), 0, st, strain, g);
}

// ---------------------------------------------------------------------------
// stats (basics.f90:236-324) on the device: trapezoid integrals
// (integrate2D / integrate1Dx, basics.f90:341-390; uniform grids, metrics 1)
// as weighted block sums, the tracer-count extremes and the surface strain
// maximum as integer / bit-pattern atomics, and the two Nusselt integrals
// along the walls in one workgroup.  The sums run in another order than the
// reference's row-by-row loops, so they agree to rounding.
// ---------------------------------------------------------------------------
// part[3*b + {0,1,2}]: the block's weighted sums of u^2+w^2, T and C;
// ints[0] += nodes with nempty == 0, ints[1] = min nempty, ints[2] = max
// nempty; smax: max strain on the top row (bits of a non-negative double)
__global__ __launch_bounds__(256) void k_diag_part(const double* __restrict__ u, const double* __restrict__ w,
                                                   const double* __restrict__ T, const double* __restrict__ C,
                                                   const int32_t* __restrict__ nempty,
                                                   const double* __restrict__ strain, Grid g, double dr, double ds,
                                                   double* __restrict__ part, int32_t* __restrict__ ints,
                                                   unsigned long long* __restrict__ smax)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y * blockDim.y + threadIdx.y;
    double v2 = 0.0, tt = 0.0, cc = 0.0;
    if (i <= g.nx && k <= g.nz) {
        const size_t c = g.at(i, k);
        const double wgt = ((i == 0 || i == g.nx) ? 0.5 : 1.0) * ((k == 0 || k == g.nz) ? 0.5 : 1.0) * dr * ds;
        v2 = (u[c] * u[c] + w[c] * w[c]) * wgt;
        tt = T[c] * wgt;
        cc = C ? C[c] * wgt : 0.0;
        const int32_t n = nempty[c];
        if (n == 0) atomicAdd(&ints[0], 1);
        atomicMin(&ints[1], n);
        atomicMax(&ints[2], n);
        if (k == g.nz) atomicMax(smax, (unsigned long long)__double_as_longlong(strain[c]));
    }
    v2 = block_sum(v2);
    tt = block_sum(tt);
    cc = block_sum(cc);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double* q = part + 3 * (blockIdx.y * gridDim.x + blockIdx.x);
        q[0] = v2;
        q[1] = tt;
        q[2] = cc;
    }
}

// out[0..2] = the sums of part (fixed order); out[3..5] = the wall integrals
// of -conduct*dT/dz at the top, T at the bottom and -conduct*dT/dz at the
// bottom (integrate1Dx)
__global__ __launch_bounds__(1024) void k_diag_final(const double* __restrict__ part, int nb,
                                                     const double* __restrict__ T,
                                                     const double* __restrict__ conduct, Grid g, FD F, double dr,
                                                     double ds, double* __restrict__ out)
{
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, a = 0.0, b = 0.0, e = 0.0;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) {
        s0 += part[3 * t];
        s1 += part[3 * t + 1];
        s2 += part[3 * t + 2];
    }
    for (int i = threadIdx.x; i <= g.nx; i += blockDim.x) {
        const double wi = ((i == 0 || i == g.nx) ? 0.5 : 1.0) * dr;
        const size_t ct = g.at(i, g.nz), cb = g.at(i, 0);
        const double Tz_surf = dot3(F.D1, T, ct, g.ld) / ds;
        const double Tz_bot = dot3(F.D1, T, cb, g.ld) / ds;
        a += -conduct[ct] * Tz_surf * wi;
        b += T[cb] * wi;
        e += -conduct[cb] * Tz_bot * wi;
    }
    s0 = block_sum(s0);
    s1 = block_sum(s1);
    s2 = block_sum(s2);
    a = block_sum(a);
    b = block_sum(b);
    e = block_sum(e);
    if (threadIdx.x == 0) {
        out[0] = s0; out[1] = s1; out[2] = s2; out[3] = a; out[4] = b; out[5] = e;
    }
}

int diag_blocks(const Grid& g) { return ((g.nx + 1 + 63) / 64) * ((g.nz + 1 + 3) / 4); }

void launch_diagnostics(hipStream_t st, const double* u, const double* w, const double* T, const double* C,
                        const int32_t* nempty, const double* strain, const double* conduct, const Grid& g,
                        const FD& F, double dr, double ds, double* part, int32_t* ints, unsigned long long* smax,
                        double* out)
{
    dim3 b(64, 4);
    dim3 gr = grid2(g.nx + 1, g.nz + 1, b);
    hipLaunchKernelGGL(k_diag_part, gr, b, 0, st, u, w, T, C, nempty, strain, g, dr, ds, part, ints, smax);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(1024), 0, st, part, (int)(gr.x * gr.y), T, conduct, g, F, dr, ds,
                       out);
}

// The rest of the stats.dat row (basics.f90:236-324): per block, the
// weighted sums of RaT/conduct/vis, of each Cnew and of each Cnew above row
// kent (integrate1Dz_ent, basics.f90:393-404: z weights ds/2 at kent and nz,
// ds between, 0 below kent).  part[STATS_NP*b + {0, 1+t, 1+TC_MAX_TYPES+t}].
constexpr int STATS_NP = 1 + 2 * TC_MAX_TYPES;

__global__ __launch_bounds__(256) void k_stats_part(const double* __restrict__ conduct, const double* __restrict__ vis,
                                                    PtrArr Cnew, int nC, int kent, double RaT, Grid g, double dr,
                                                    double ds, double* __restrict__ part)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y * blockDim.y + threadIdx.y;
    double acc[STATS_NP];
    for (int j = 0; j < STATS_NP; ++j) acc[j] = 0.0;
    if (i <= g.nx && k <= g.nz) {
        const size_t c = g.at(i, k);
        const double wx = ((i == 0 || i == g.nx) ? 0.5 : 1.0) * dr;
        const double wz = ((k == 0 || k == g.nz) ? 0.5 : 1.0) * ds;
        const double we = (kent >= g.nz || k < kent) ? 0.0 : ((k == kent || k == g.nz) ? 0.5 : 1.0) * ds;
        acc[0] = RaT / conduct[c] / vis[c] * (wx * wz);
        for (int t = 0; t < nC; ++t) {
            const double v = Cnew.v[t][c];
            acc[1 + t] = v * (wx * wz);
            acc[1 + TC_MAX_TYPES + t] = v * (wx * we);
        }
    }
    for (int j = 0; j < STATS_NP; ++j) acc[j] = block_sum(acc[j]);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        double* q = part + STATS_NP * (blockIdx.y * gridDim.x + blockIdx.x);
        for (int j = 0; j < STATS_NP; ++j) q[j] = acc[j];
    }
}

// out[0..STATS_NP-1] = the sums of part; out[STATS_NP] = integrate1Dx of
// u^2+w^2 on the top row; out[STATS_NP+1..+4] = q1..q4, the heat fluxes at
// the corners (0,nz), (nx,nz), (nx,0), (0,0)
__global__ __launch_bounds__(1024) void k_stats_final(const double* __restrict__ part, int nb,
                                                      const double* __restrict__ u, const double* __restrict__ w,
                                                      const double* __restrict__ T,
                                                      const double* __restrict__ conduct, Grid g, FD F, double dr,
                                                      double ds, double* __restrict__ out)
{
    double acc[STATS_NP + 1];
    for (int j = 0; j <= STATS_NP; ++j) acc[j] = 0.0;
    for (int t = threadIdx.x; t < nb; t += blockDim.x)
        for (int j = 0; j < STATS_NP; ++j) acc[j] += part[STATS_NP * t + j];
    for (int i = threadIdx.x; i <= g.nx; i += blockDim.x) {
        const size_t c = g.at(i, g.nz);
        acc[STATS_NP] += (u[c] * u[c] + w[c] * w[c]) * (((i == 0 || i == g.nx) ? 0.5 : 1.0) * dr);
    }
    for (int j = 0; j <= STATS_NP; ++j) acc[j] = block_sum(acc[j]);
    if (threadIdx.x == 0) {
        for (int j = 0; j <= STATS_NP; ++j) out[j] = acc[j];
        const int ci[4] = {0, g.nx, g.nx, 0}, ck[4] = {g.nz, g.nz, 0, 0};
        for (int j = 0; j < 4; ++j) {
            const size_t c = g.at(ci[j], ck[j]);
            out[STATS_NP + 1 + j] = -conduct[c] * dot3(F.D1, T, c, g.ld) / ds;
        }
    }
}

int stats_np() { return STATS_NP; }

void launch_stats_rest(hipStream_t st, const double* u, const double* w, const double* T, const double* conduct,
                       const double* vis, const PtrArr& Cnew, int nC, int kent, double RaT, const Grid& g,
                       const FD& F, double dr, double ds, double* part, double* out)
{
    dim3 b(64, 4);
    dim3 gr = grid2(g.nx + 1, g.nz + 1, b);
    hipLaunchKernelGGL(k_stats_part, gr, b, 0, st, conduct, vis, Cnew, nC, kent, RaT, g, dr, ds, part);
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(1024), 0, st, part, (int)(gr.x * gr.y), u, w, T, conduct, g, F, dr,
                       ds, out);
}

void launch_sym_bcs(hipStream_t st, double* f, const Grid& g)
{
    hipLaunchKernelGGL(k_sym_bcs, dim3(1), dim3(256), 0, st, f, g);
}

void launch_bcs_T(hipStream_t st, double* T, const Grid& g, int Tbot, const int32_t* stop)
{
    hipLaunchKernelGGL(k_bcs_T, dim3(1), dim3(1024), 0, st, T, g, Tbot, stop);
}

void launch_bcs_comp(hipStream_t st, double* C, const Grid& g, const int32_t* stop)
{
    hipLaunchKernelGGL(k_bcs_comp, dim3(1), dim3(1024), 0, st, C, g, stop);
}

void launch_smoothness_max(hipStream_t st, const double* T0, const Grid& g, unsigned long long* slots,
                           const int32_t* stop)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_smoothness_max, grid2(g.nx, (g.nz - 1 + FILT_ROWS - 1) / FILT_ROWS, b), b, 0, st, T0, g,
                       slots, stop);
}
size_t filter_slot_words() { return (size_t)FM_SETS * FILT_SET; }

// the filter step's divisors; reciprocals used when both are powers of two
static FScale filter_scale(double dr2, double ds2)
{
    int e;
    FScale sc{dr2, ds2, 1.0 / dr2, 1.0 / ds2, 0};
    sc.pow2 = std::frexp(dr2, &e) == 0.5 && std::frexp(ds2, &e) == 0.5;
    return sc;
}
int filter_multi_k() { return FM_K; }

void launch_filter_multi(hipStream_t st, double* const bufs[3], int b_in, const double* dts, const Grid& g,
                         const FD& F, double dr2, double ds2, double courant, const double* fac, int m, double delta,
                         unsigned long long* slots, int32_t* ctl)
{
    FmBufs b{{bufs[0], bufs[1], bufs[2]}};
    FmFac f{};
    for (int j = 0; j < 2 * FM_K; ++j) f.v[j] = fac[j];
    // tiles of FM_TXO x FM_TY cells over the plane's -SPAN..nx+SPAN x
    // -SPAN..nz+SPAN, FM_WAVES of them stacked in z per workgroup
    const int tz = (g.nz + 1 + 2 * SPAN + FM_TY - 1) / FM_TY;
    dim3 blk(64, FM_WAVES);
    dim3 gr((unsigned)((g.nx + 1 + 2 * SPAN + FM_TXO - 1) / FM_TXO), (unsigned)((tz + FM_WAVES - 1) / FM_WAVES));
    const FScale sc = filter_scale(dr2, ds2);
    if (sc.pow2 && F.D2[1] == 1.0 && F.D2[2] == -2.0 && F.D2[3] == 1.0)
        hipLaunchKernelGGL((k_filter_multi<true>), gr, blk, 0, st, b, b_in, dts, g, F, sc, courant, f, m, delta, slots,
                           ctl);
    else
        hipLaunchKernelGGL((k_filter_multi<false>), gr, blk, 0, st, b, b_in, dts, g, F, sc, courant, f, m, delta, slots,
                           ctl);
}

void launch_filter_iter(hipStream_t st, bool comp, const double* T0, double* T, const double* dts, const Grid& g,
                        const FD& F, double dr2, double ds2, double cfac, double courant, int Tbot,
                        unsigned long long* slots, int m, double delta, int32_t* ctl)
{
    dim3 b(FI_BX, FI_BY);
    dim3 gr((unsigned)((g.nx + 1 + SPAN + ORG_PAD + FI_BX - 1) / FI_BX),
            (unsigned)((g.nz + 1 + 2 * SPAN + FI_TILE - 1) / FI_TILE));
    const FScale sc = filter_scale(dr2, ds2);
    if (comp && sc.pow2)
        hipLaunchKernelGGL((k_filter_iter<true, true>), gr, b, 0, st, T0, T, dts, g, F, sc, cfac, courant, Tbot, slots,
                           m, delta, ctl);
    else if (comp)
        hipLaunchKernelGGL((k_filter_iter<true, false>), gr, b, 0, st, T0, T, dts, g, F, sc, cfac, courant, Tbot, slots,
                           m, delta, ctl);
    else if (sc.pow2)
        hipLaunchKernelGGL((k_filter_iter<false, true>), gr, b, 0, st, T0, T, dts, g, F, sc, cfac, courant, Tbot, slots,
                           m, delta, ctl);
    else
        hipLaunchKernelGGL((k_filter_iter<false, false>), gr, b, 0, st, T0, T, dts, g, F, sc, cfac, courant, Tbot,
                           slots, m, delta, ctl);
}

void launch_zero_i32(hipStream_t st, int32_t* p, size_t n)
{
    unsigned nb = (unsigned)((n + 255) / 256);
    if (nb > 4096) nb = 4096;
    if (nb == 0) nb = 1;
    hipLaunchKernelGGL(k_zero_i32, dim3(nb), dim3(256), 0, st, p, n);
}

void launch_bin(hipStream_t st, const double* r, const double* s, const int32_t* ty, int64_t n, const CntArr& cnt,
                const Grid& g, double dr, double ds, double aspect, int32_t* flag)
{
    if (n <= 0) return;
    int e;
    const bool pow2 = std::frexp(dr, &e) == 0.5 && std::frexp(ds, &e) == 0.5;
    if (pow2)
        hipLaunchKernelGGL((k_bin<true>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, r, s, ty, n, cnt, g, dr, ds,
                           1.0 / dr, 1.0 / ds, aspect, flag);
    else
        hipLaunchKernelGGL((k_bin<false>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, r, s, ty, n, cnt, g, dr,
                           ds, 1.0 / dr, 1.0 / ds, aspect, flag);
}

void launch_ratio(hipStream_t st, const CntArr& cnt, int32_t* nempty, const PtrArr& Cnew, const Grid& g, int ntype,
                  int comp)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_ratio, grid2(g.nx + 2 * SPAN + 1, g.nz + 2 * SPAN + 1, b), b, 0, st, cnt, nempty, Cnew, g, ntype,
                       comp);
}

void launch_conduct(hipStream_t st, const PtrArr& Cvis, double* K, double* Kr, double* Ks, const Grid& g, const FD& F,
                    double dr, double ds, int comp, int ntype, const RaCArr& lcf)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_conduct, grid2(g.nx + 2 * SPAN + 1, g.nz + 2 * SPAN + 1, b), b, 0, st, Cvis, K, g, comp, ntype,
                       lcf);
    hipLaunchKernelGGL(k_conduct_grad, grid2(g.nx + 3, g.nz + 3, b), b, 0, st, K, Kr, Ks, g, F, dr, ds);
}

void launch_derivs(hipStream_t st, const double* u, const double* w, double* Du, double* Dw, const Grid& g,
                   const D3Arr& Dint, double dr)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_derivs, grid2(g.nx + 1, g.nz + 3, b), b, 0, st, u, w, Du, Dw, g, Dint, dr);
}

void launch_tracer_stage(hipStream_t st, int stage, int torder, double dt, int64_t n, double* r, double* s, double* r0,
                         double* s0, double* r1, double* s1, double* r2, double* s2, const double* u, const double* w,
                         const double* Du, const double* Dw, const Grid& g, const D3Arr& Dint, double dr, double ds)
{
    if (n <= 0) return;
    int e;
    const bool pow2 = std::frexp(dr, &e) == 0.5 && std::frexp(ds, &e) == 0.5;
    if (pow2)
        hipLaunchKernelGGL((k_tracer_stage<true>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, stage, torder, dt,
                           n, r, s, r0, s0, r1, s1, r2, s2, u, w, Du, Dw, g, Dint, dr, ds, 1.0 / dr, 1.0 / ds);
    else
        hipLaunchKernelGGL((k_tracer_stage<false>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, stage, torder,
                           dt, n, r, s, r0, s0, r1, s1, r2, s2, u, w, Du, Dw, g, Dint, dr, ds, 1.0 / dr, 1.0 / ds);
}

void launch_energy_stage(hipStream_t st, int stage, double dt, double* T, const double* T0, double* T1, double* T2,
                         const double* u, const double* w, const double* K, const double* Kr, const double* Ks,
                         const PtrArr& Cvis, const Grid& g, const FD& F, double dr, double ds, double dr2, double ds2,
                         int comp, int ntype, double H, const RaCArr& Htype)
{
    dim3 b(64, 4);
    hipLaunchKernelGGL(k_energy_stage, grid2(g.nx + 1, g.nz + 1, b), b, 0, st, stage, dt, T, T0, T1, T2, u, w, K, Kr, Ks,
                       Cvis, g, F, dr, ds, dr2, ds2, comp, ntype, H, Htype);
}

int launch_stability(hipStream_t st, const double* u, const double* w, const double* K, const double* Kr,
                     const double* Ks, const Grid& g, double TrA, double TsA, double TrrD, double TssD, double* partialA,
                     double* partialD)
{
    dim3 b(64, 4);
    dim3 gr = grid2(g.nx + 1, g.nz + 1, b);
    hipLaunchKernelGGL(k_stability, gr, b, 0, st, u, w, K, Kr, Ks, g, TrA, TsA, TrrD, TssD, partialA, partialD);
    return (int)(gr.x * gr.y);
}

}  // namespace tc
