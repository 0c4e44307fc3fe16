// equalize_tracers (composition.f90:83-580) on the device.
//
// The reference moves tracers from nodes with more than their type's share
// to the nodes that lack the most, in lottery passes over a shuffled visit
// order, drawing from one mzran stream per OpenMP thread.  Here every draw
// is a counter-based hash of (seed, call, stream, index), so each step is a
// data-parallel kernel whose result does not depend on the launch shape or
// the tracer order:
//   * per-node tracer lists: one radix sort of (node, tracer index) keys;
//   * per-node shuffles and donor choice: one thread per node with excess;
//   * the per-type donor order: a radix sort by a keyed bijection of the
//     tracer index (unique keys, so the order is fully determined);
//   * a lottery pass: the nodes that qualify, in visit order, take the next
//     donors of their type; the visit-order rank of a node among those of its
//     type comes from wavefront ballots and a scan of per-block counts.
// The node index is c = k*(nx+1) + i; the tracer index is the caller's.

#include "tc_kernels.h"

namespace tc {

namespace {

constexpr uint64_t EQ_GOLD = 0x9e3779b97f4a7c15ULL;
constexpr unsigned EQ_S_NODE = 1, EQ_S_PLACE = 2, EQ_S_ROWS = 3, EQ_S_COLS = 4;
constexpr int EQ_BLOCK = 256;

// SplitMix64 finalizer (Steele, Lea and Flood, 2014)
__host__ __device__ __forceinline__ uint64_t eq_mix(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}

// uniform double in [0, 1) with 53 random bits
__host__ __device__ __forceinline__ double eq_u(uint64_t key, unsigned stream, uint64_t a, uint64_t b)
{
    uint64_t x = eq_mix(key ^ ((uint64_t)stream << 48) ^ a);
    x = eq_mix(x + EQ_GOLD * (b + 1));
    return (double)(x >> 11) * 0x1.0p-53;
}

// keyed bijection of [0, 2^40): xor, odd multiplications and right
// xorshifts are each invertible modulo 2^40
__device__ __forceinline__ uint64_t eq_p40(uint64_t key, uint64_t id)
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

// shuffle_serial (composition.f90:22-33): each element swaps with a later
// one drawn uniformly (Sattolo's algorithm)
template <typename T>
__host__ __device__ __forceinline__ void eq_shuffle(T* a, int n, uint64_t key, unsigned stream, uint64_t id)
{
    for (int j = 0; j + 1 < n; ++j) {
        const double u = eq_u(key, stream, id, (uint64_t)j);
        const int r = (int)floor(u * (double)(n - 1 - j)) + j + 1;
        const T t = a[j];
        a[j] = a[r];
        a[r] = t;
    }
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int64_t wave_sum_i64(int64_t v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int wave_min_i32(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}

// workgroup reductions (EQ_BLOCK threads); every thread gets the result
__device__ __forceinline__ long long eq_block_sum(long long v, long long* lds)
{
    v = wave_sum_i64(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    long long t = 0;
    for (int w = 0; w < EQ_BLOCK / 64; ++w) t += lds[w];
    return t;
}
__device__ __forceinline__ int eq_block_min(int v, long long* lds)
{
    v = wave_min_i32(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = 2147483647;
    for (int w = 0; w < EQ_BLOCK / 64; ++w) t = min(t, (int)lds[w]);
    return t;
}

// one workgroup's partial record: sums s[] (0 if unset), minima m[]
struct EqAcc {
    long long s[EQ_NS];
    int m[EQ_NM];
    __device__ void init()
    {
        for (int j = 0; j < EQ_NS; ++j) s[j] = 0;
        for (int j = 0; j < EQ_NM; ++j) m[j] = 2147483647;
    }
    __device__ void store(EqPart* part) const
    {
        if (threadIdx.x != 0) return;
        EqPart& q = part[blockIdx.x];
        for (int j = 0; j < EQ_NS; ++j) q.s[j] = s[j];
        for (int j = 0; j < EQ_NM; ++j) q.m[j] = m[j];
    }
};

// interior node of a flat index over (nx+1)(nz+1)
__device__ __forceinline__ bool eq_node(int64_t t, int nx, int nz, int& i, int& k)
{
    if (t >= (int64_t)(nx + 1) * (nz + 1)) return false;
    k = (int)(t / (nx + 1));
    i = (int)(t - (int64_t)k * (nx + 1));
    return true;
}

// the selection masks of tpc_type and minexcess (composition.f90:120-140,
// 388-423): type 0 = ambient nodes (no enriched type), type p = pure nodes
__device__ __forceinline__ bool eq_mask(const PtrArr& Cnew, size_t c, int p, int ntype, int comp)
{
    if (comp == 0) return true;
    if (p == 0) {
        double acc = 0.0;
        for (int t = 0; t < ntype; ++t) acc += Cnew.v[t][c];
        return acc == 0.0;
    }
    return Cnew.v[p - 1][c] == 1.0;
}

// first type of largest count (maxloc, composition.f90:438)
__device__ __forceinline__ int eq_type_save(const CntArr& cnt, size_t c, int nt)
{
    int best = 0, bv = cnt.v[0][c];
    for (int t = 1; t <= nt; ++t) {
        const int v = cnt.v[t][c];
        if (v > bv) { bv = v; best = t; }
    }
    return best;
}

}  // namespace

// -------------------------------------------------------------------------
__global__ void k_eq_clear(EqRed* red)
{
    const int t = threadIdx.x;
    if (t < EQ_MAXT) { red->dcount[t] = 0; red->qtot[t] = 0; }
    if (t == 0) { red->found = 0; red->bad = 0; }
}

// the final sums and minima of the per-workgroup partials, in fixed order
__global__ __launch_bounds__(1024) void k_eq_final(const EqPart* __restrict__ part, int nblk, EqRed* red)
{
    __shared__ long long lds[16];
    for (int j = 0; j < EQ_NS + EQ_NM; ++j) {
        long long a = j < EQ_NS ? 0 : 2147483647;
        for (int b = threadIdx.x; b < nblk; b += blockDim.x)
            a = j < EQ_NS ? a + part[b].s[j] : min(a, (long long)part[b].m[j - EQ_NS]);
        if (j < EQ_NS) {
            for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
        } else {
            for (int o = 32; o > 0; o >>= 1) a = min(a, __shfl_xor(a, o, 64));
        }
        __syncthreads();
        if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = a;
        __syncthreads();
        if (threadIdx.x == 0) {
            long long t = lds[0];
            for (int w = 1; w < (int)(blockDim.x / 64); ++w) t = j < EQ_NS ? t + lds[w] : min(t, lds[w]);
            if (j < EQ_NS) red->fs[j] = t;
            else red->fm[j - EQ_NS] = (int32_t)t;
        }
    }
}

// tpc_type statistics (composition.f90:114-140): nodes of each mask and
// the sum of that type's counts over them: s[2p], s[2p+1]
__global__ __launch_bounds__(EQ_BLOCK) void k_eq_tpc_stats(CntArr cnt, PtrArr Cnew, Grid g, int ntype, int comp,
                                                           EqPart* part)
{
    __shared__ long long lds[EQ_BLOCK / 64];
    EqAcc acc;
    acc.init();
    int i = 0, k = 0;
    const bool in = eq_node((int64_t)blockIdx.x * EQ_BLOCK + threadIdx.x, g.nx, g.nz, i, k);
    const size_t c = in ? g.at(i, k) : 0;
    for (int p = 0; p <= ntype; ++p) {
        const bool sel = in && eq_mask(Cnew, c, p, ntype, comp);
        acc.s[2 * p] = eq_block_sum(sel ? 1 : 0, lds);
        acc.s[2 * p + 1] = eq_block_sum(sel ? (long long)cnt.v[p][c] : 0, lds);
    }
    acc.store(part);
}

// nexcess (composition.f90:118-161)
__global__ __launch_bounds__(EQ_BLOCK) void k_eq_nexcess(CntArr cnt, const int32_t* __restrict__ nempty, Grid g,
                                                         int nt, int comp, EqTypes tpc, int32_t* __restrict__ ne,
                                                         EqPart* part)
{
    __shared__ long long lds[EQ_BLOCK / 64];
    int i = 0, k = 0;
    const int64_t t = (int64_t)blockIdx.x * EQ_BLOCK + threadIdx.x;
    const bool in = eq_node(t, g.nx, g.nz, i, k);
    int v = 0;
    if (in) {
        const size_t c = g.at(i, k);
        const int n0 = nempty[c];
        if (comp == 0) {
            if (n0 > 0) v = n0 - tpc.v[0];
        } else if (n0 > 0) {
            for (int p = 0; p <= nt; ++p) {
                bool all = true;
                for (int dk = -1; dk <= 1 && all; ++dk)
                    for (int di = -1; di <= 1; ++di) {
                        const size_t m = g.at(i + di, k + dk);
                        if (cnt.v[p][m] != nempty[m]) { all = false; break; }
                    }
                if (all) { v = n0 - tpc.v[p]; break; }
            }
        }
        ne[t] = v;
    }
    EqAcc acc;
    acc.init();
    acc.s[0] = eq_block_sum(v > 0 ? v : 0, lds);   // nfree
    acc.s[1] = eq_block_sum(v != 0 ? 1 : 0, lds);  // nodes with nexcess != 0
    acc.store(part);
}

// (node << 32 | tracer index, slot) pairs; node = tracers_to_corners' NINT
// node (composition.f90:198-203).  A tracer outside the domain (or NaN)
// sets red->bad and gets node 0; the host stops before the keys are used.
__global__ __launch_bounds__(EQ_BLOCK) void k_eq_node_keys(const double* __restrict__ r, const double* __restrict__ s,
                                                           const int64_t* __restrict__ id, int64_t n, int nx,
                                                           double dr, double ds, double aspect,
                                                           uint64_t* __restrict__ key, uint32_t* __restrict__ val,
                                                           EqRed* red)
{
    const int64_t j = (int64_t)blockIdx.x * EQ_BLOCK + threadIdx.x;
    if (j >= n) return;
    const double rv = r[j], sv = s[j];
    uint64_t c = 0;
    if (rv > 0.0 && rv < aspect && sv > 0.0 && sv < 1.0) {
        const int i = (int)round(rv / dr), k = (int)round(sv / ds);
        c = (uint64_t)k * (uint64_t)(nx + 1) + (uint64_t)i;
    } else {
        atomicOr(&red->bad, 1);
    }
    const uint64_t who = id ? (uint64_t)id[j] : (uint64_t)j;
    key[j] = (c << 32) | who;
    val[j] = (uint32_t)j;
}

__global__ __launch_bounds__(EQ_BLOCK) void k_eq_fill2(int32_t* __restrict__ a, int32_t* __restrict__ b, int64_t n)
{
    const int64_t j = (int64_t)blockIdx.x * EQ_BLOCK + threadIdx.x;
    if (j < n) { a[j] = 0; b[j] = 0; }
}

// [start, end) of each node's run in the sorted keys
__global__ __launch_bounds__(EQ_BLOCK) void k_eq_bounds(const uint64_t* __restrict__ key, int64_t n,
                                                        int32_t* __restrict__ start, int32_t* __restrict__ end)
{
    const int64_t j = (int64_t)blockIdx.x * EQ_BLOCK + threadIdx.x;
    if (j >= n) return;
    const uint64_t c = key[j] >> 32;
    if (j == 0 || (key[j - 1] >> 32) != c) start[c] = (int32_t)j;
    if (j == n - 1 || (key[j + 1] >> 32) != c) end[c] = (int32_t)(j + 1);
}

// donors (composition.f90:229-335): the node's list is shuffled in place;
// wall nodes give their tracers within dist of a wall first (segment 0),
// then the others (segment 1); interior nodes their first nexcess tracers.
// Each donor gets the sort key (type*2 + segment) << 40 | p40(index).
__global__ __launch_bounds__(EQ_BLOCK) void k_eq_select(const int32_t* __restrict__ ne,
                                                        const int32_t* __restrict__ start,
                                                        const int32_t* __restrict__ end, uint32_t* __restrict__ list,
                                                        const double* __restrict__ r, const double* __restrict__ s,
                                                        const int32_t* __restrict__ ty, const int64_t* __restrict__ id,
                                                        int nx, int nz, double dist_r, double dist_s, double aspect,
                                                        uint64_t key, uint64_t* __restrict__ dkey,
                                                        uint32_t* __restrict__ dval, EqRed* red)
{
    int i = 0, k = 0;
    const int64_t c = (int64_t)blockIdx.x * EQ_BLOCK + threadIdx.x;
    if (!eq_node(c, nx, nz, i, k)) return;
    int want = ne[c];
    if (want <= 0) return;
    const int b0 = start[c];
    const int n = end[c] - b0;
    uint32_t* L = list + b0;
    eq_shuffle(L, n, key, EQ_S_NODE, (uint64_t)c);
    const int take = min(want, n);
    const unsigned long long base = atomicAdd(&red->found, (unsigned long long)take);
    int got = 0;
    int per[EQ_MAXT] = {0, 0, 0, 0, 0, 0, 0, 0};
    const bool edge = i == 0 || i == nx || k == 0 || k == nz;
    for (int pass = edge ? 0 : 1; pass < 2 && got < take; ++pass)
        for (int q = 0; q < n && got < take; ++q) {
            const uint32_t slot = L[q];
            const double rv = r[slot], sv = s[slot];
            const bool near = sv < dist_s || sv > 1.0 - dist_s || rv < dist_r || rv > aspect - dist_r;
            if (edge && near != (pass == 0)) continue;
            const int t = ty[slot];
            const uint64_t who = id ? (uint64_t)id[slot] : (uint64_t)slot;
            dkey[base + got] = ((uint64_t)(t * 2 + pass) << 40) | eq_p40(key, who);
            dval[base + got] = slot;
            per[t]++;
            got++;
        }
    for (int t = 0; t < EQ_MAXT; ++t)
        if (per[t]) atomicAdd(&red->dcount[t], (unsigned long long)per[t]);
}

// minexcess selections before the type gates (composition.f90:387-423):
// m[p]; per-type totals (tracer_count, :381-385): s[p]; the deficit
// (:566-574): s[EQ_MAXT]
__global__ __launch_bounds__(EQ_BLOCK) void k_eq_reduce(const int32_t* __restrict__ ne, CntArr cnt, PtrArr Cnew,
                                                        Grid g, int nt, int ntype, int comp, EqPart* part)
{
    __shared__ long long lds[EQ_BLOCK / 64];
    EqAcc acc;
    acc.init();
    int i = 0, k = 0;
    const int64_t t = (int64_t)blockIdx.x * EQ_BLOCK + threadIdx.x;
    const bool in = eq_node(t, g.nx, g.nz, i, k);
    const size_t c = in ? g.at(i, k) : 0;
    const int v = in ? ne[t] : 1;
    for (int p = 0; p <= nt; ++p) {
        acc.m[p] = eq_block_min((in && v <= 0 && eq_mask(Cnew, c, p, ntype, comp)) ? v : 2147483647, lds);
        acc.s[p] = eq_block_sum(in ? (long long)cnt.v[p][c] : 0, lds);
    }
    acc.s[EQ_MAXT] = eq_block_sum((in && v <= 0) ? (long long)v : 0, lds);
    acc.store(part);
}

namespace {

// type of the donor a visited node asks for, or -1 when it does not
// qualify (composition.f90:437-451 and :505-508)
__device__ __forceinline__ int eq_flag(int64_t v, const int32_t* rows, const int32_t* cols, const int32_t* ne,
                                       const CntArr& cnt, const Grid& g, int nt, int tg, const EqTypes& mx,
                                       int& i, int& k, int& ts)
{
    const int64_t nzc = g.nz + 1;
    if (v >= (int64_t)(g.nx + 1) * nzc) return -1;
    const int ii = (int)(v / nzc), kk = (int)(v - (int64_t)ii * nzc);
    i = rows[ii];
    k = cols[kk];
    const int64_t node = (int64_t)k * (g.nx + 1) + i;
    const int e = ne[node];
    ts = eq_type_save(cnt, g.at(i, k), nt);
    if (!(e == mx.v[ts] && e < 0)) return -1;
    return tg >= 0 ? tg : ts;
}

// rank of this lane among the lanes of its workgroup with the same flag;
// wtot[w*EQ_MAXT+t] are the per-wave counts
__device__ __forceinline__ int eq_block_rank(int f, int nt, int* wtot)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    int r = 0;
    for (int t = 0; t <= nt; ++t) {
        const unsigned long long m = __ballot(f == t);
        if (lane == 0) wtot[w * EQ_MAXT + t] = __popcll(m);
        if (f == t) r = __popcll(m & lt);
    }
    __syncthreads();
    if (f >= 0)
        for (int w2 = 0; w2 < w; ++w2) r += wtot[w2 * EQ_MAXT + f];
    return r;
}

}  // namespace

// per-block counts of the qualifying nodes of each donor type
__global__ __launch_bounds__(EQ_BLOCK) void k_eq_flags(const int32_t* __restrict__ rows,
                                                       const int32_t* __restrict__ cols,
                                                       const int32_t* __restrict__ ne, CntArr cnt, Grid g, int nt,
                                                       int tg, EqTypes mx, int32_t* __restrict__ bcnt)
{
    __shared__ int wtot[(EQ_BLOCK / 64) * EQ_MAXT];
    int i, k, ts;
    const int64_t v = (int64_t)blockIdx.x * EQ_BLOCK + threadIdx.x;
    const int f = eq_flag(v, rows, cols, ne, cnt, g, nt, tg, mx, i, k, ts);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int t = 0; t <= nt; ++t) {
        const unsigned long long m = __ballot(f == t);
        if (lane == 0) wtot[w * EQ_MAXT + t] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x <= nt) {
        int s = 0;
        for (int w2 = 0; w2 < EQ_BLOCK / 64; ++w2) s += wtot[w2 * EQ_MAXT + threadIdx.x];
        bcnt[(size_t)blockIdx.x * EQ_MAXT + threadIdx.x] = s;
    }
}

// exclusive scan of the block counts over the blocks, per type; totals to
// red->qtot.  One workgroup.
__global__ __launch_bounds__(256) void k_eq_scan(int32_t* __restrict__ bcnt, int nblk, int nt, EqRed* red)
{
    __shared__ int64_t part[256];
    const int per = (nblk + 255) / 256;
    const int b0 = threadIdx.x * per, b1 = min(nblk, b0 + per);
    for (int t = 0; t <= nt; ++t) {
        int64_t s = 0;
        for (int b = b0; b < b1; ++b) s += bcnt[(size_t)b * EQ_MAXT + t];
        part[threadIdx.x] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t acc = 0;
            for (int j = 0; j < 256; ++j) { const int64_t x = part[j]; part[j] = acc; acc += x; }
            red->qtot[t] = (unsigned long long)acc;
        }
        __syncthreads();
        int64_t acc = part[threadIdx.x];
        for (int b = b0; b < b1; ++b) {
            const int x = bcnt[(size_t)b * EQ_MAXT + t];
            bcnt[(size_t)b * EQ_MAXT + t] = (int32_t)acc;
            acc += x;
        }
        __syncthreads();
    }
}

// placement (composition.f90:452-470 and :514-534): the j-th qualifying node
// of type tt in visit order takes donor used[tt] + j while donors last
__global__ __launch_bounds__(EQ_BLOCK) void k_eq_assign(const int32_t* __restrict__ rows,
                                                        const int32_t* __restrict__ cols, int32_t* __restrict__ ne,
                                                        CntArr cnt, Grid g, int nt, int tg, EqTypes mx,
                                                        const int32_t* __restrict__ boff,
                                                        const uint32_t* __restrict__ donors, EqOff first,
                                                        EqOff avail, double* __restrict__ r, double* __restrict__ s,
                                                        int32_t* __restrict__ ty, const int64_t* __restrict__ id,
                                                        uint64_t key, double dr, double ds, double dist_r,
                                                        double dist_s, double aspect, int comp, int update_counts)
{
#pragma clang fp contract(off)
    __shared__ int wtot[(EQ_BLOCK / 64) * EQ_MAXT];
    int i = 0, k = 0, ts = 0;
    const int64_t v = (int64_t)blockIdx.x * EQ_BLOCK + threadIdx.x;
    const int f = eq_flag(v, rows, cols, ne, cnt, g, nt, tg, mx, i, k, ts);
    const int rank = eq_block_rank(f, nt, wtot);
    if (f < 0) return;
    const int64_t j = (int64_t)boff[(size_t)blockIdx.x * EQ_MAXT + f] + rank;
    if (j >= avail.v[f]) return;
    const uint32_t slot = donors[first.v[f] + j];
    const uint64_t who = id ? (uint64_t)id[slot] : (uint64_t)slot;
    const double pr = (eq_u(key, EQ_S_PLACE, who, 0) - 0.5) * dr;
    const double ps = (eq_u(key, EQ_S_PLACE, who, 1) - 0.5) * ds;
    double rv = dr * (double)i;
    rv = rv + pr;
    double sv = ds * (double)k;
    sv = sv + ps;
    if (rv <= 0.0) rv = dist_r;
    if (rv >= aspect) rv = aspect - dist_r;
    if (sv <= 0.0) sv = dist_s;
    if (sv >= 1.0) sv = 1.0 - dist_s;
    r[slot] = rv;
    s[slot] = sv;
    const int old = ty[slot];
    const int nt2 = comp == 1 ? ts : old;
    if (comp == 1) ty[slot] = ts;
    const size_t c = g.at(i, k);
    if (update_counts) {
        cnt.v[f][c] -= 1;
        cnt.v[nt2][c] += 1;
    }
    ne[(int64_t)k * (g.nx + 1) + i] += 1;
}

// -------------------------------------------------------------------------
uint64_t eq_stream_key(uint64_t seed, uint64_t call) { return eq_mix(seed + EQ_GOLD * (call + 1)); }

static inline unsigned eq_blocks(int64_t n) { return (unsigned)((n + EQ_BLOCK - 1) / EQ_BLOCK); }

void launch_eq_clear(hipStream_t st, EqRed* red)
{
    hipLaunchKernelGGL(k_eq_clear, dim3(1), dim3(64), 0, st, red);
}

static void eq_final(hipStream_t st, const Grid& g, EqPart* part, EqRed* red)
{
    const int nblk = (int)eq_blocks((int64_t)(g.nx + 1) * (g.nz + 1));
    hipLaunchKernelGGL(k_eq_final, dim3(1), dim3(1024), 0, st, part, nblk, red);
}

void launch_eq_tpc_stats(hipStream_t st, const CntArr& cnt, const PtrArr& Cnew, const Grid& g, int ntype, int comp,
                         EqPart* part, EqRed* red)
{
    const int64_t V = (int64_t)(g.nx + 1) * (g.nz + 1);
    hipLaunchKernelGGL(k_eq_tpc_stats, dim3(eq_blocks(V)), dim3(EQ_BLOCK), 0, st, cnt, Cnew, g, ntype, comp, part);
    eq_final(st, g, part, red);
}

void launch_eq_nexcess(hipStream_t st, const CntArr& cnt, const int32_t* nempty, const Grid& g, int nt, int comp,
                       const EqTypes& tpc, int32_t* ne, EqPart* part, EqRed* red)
{
    const int64_t V = (int64_t)(g.nx + 1) * (g.nz + 1);
    hipLaunchKernelGGL(k_eq_nexcess, dim3(eq_blocks(V)), dim3(EQ_BLOCK), 0, st, cnt, nempty, g, nt, comp, tpc, ne,
                       part);
    eq_final(st, g, part, red);
}

void launch_eq_node_lists(hipStream_t st, const double* r, const double* s, const int64_t* id, int64_t n,
                          const Grid& g, double dr, double ds, double aspect, uint64_t* key, uint32_t* val,
                          int32_t* start, int32_t* end, EqRed* red)
{
    const int64_t V = (int64_t)(g.nx + 1) * (g.nz + 1);
    hipLaunchKernelGGL(k_eq_fill2, dim3(eq_blocks(V)), dim3(EQ_BLOCK), 0, st, start, end, V);
    hipLaunchKernelGGL(k_eq_node_keys, dim3(eq_blocks(n)), dim3(EQ_BLOCK), 0, st, r, s, id, n, g.nx, dr, ds, aspect, key,
                       val, red);
}

void launch_eq_bounds(hipStream_t st, const uint64_t* key, int64_t n, int32_t* start, int32_t* end)
{
    hipLaunchKernelGGL(k_eq_bounds, dim3(eq_blocks(n)), dim3(EQ_BLOCK), 0, st, key, n, start, end);
}

void launch_eq_select(hipStream_t st, const int32_t* ne, const int32_t* start, const int32_t* end, uint32_t* list,
                      const double* r, const double* s, const int32_t* ty, const int64_t* id, const Grid& g,
                      double dist_r, double dist_s, double aspect, uint64_t key, uint64_t* dkey, uint32_t* dval,
                      EqRed* red)
{
    const int64_t V = (int64_t)(g.nx + 1) * (g.nz + 1);
    hipLaunchKernelGGL(k_eq_select, dim3(eq_blocks(V)), dim3(EQ_BLOCK), 0, st, ne, start, end, list, r, s, ty, id, g.nx,
                       g.nz, dist_r, dist_s, aspect, key, dkey, dval, red);
}

void launch_eq_reduce(hipStream_t st, const int32_t* ne, const CntArr& cnt, const PtrArr& Cnew, const Grid& g, int nt,
                      int ntype, int comp, EqPart* part, EqRed* red)
{
    const int64_t V = (int64_t)(g.nx + 1) * (g.nz + 1);
    hipLaunchKernelGGL(k_eq_reduce, dim3(eq_blocks(V)), dim3(EQ_BLOCK), 0, st, ne, cnt, Cnew, g, nt, ntype, comp, part);
    eq_final(st, g, part, red);
}

// the row and column visit orders of lottery pass `pass`: each pass
// shuffles the previous orders (composition.f90:427-429); on the host, which
// runs the pass loop
void eq_visit_orders(int32_t* rows, int nr, int32_t* cols, int nc, uint64_t key, uint64_t pass)
{
    eq_shuffle(rows, nr, key, EQ_S_ROWS, pass);
    eq_shuffle(cols, nc, key, EQ_S_COLS, pass);
}

int eq_pass_blocks(const Grid& g) { return (int)eq_blocks((int64_t)(g.nx + 1) * (g.nz + 1)); }

void launch_eq_pass(hipStream_t st, const int32_t* rows, const int32_t* cols, int32_t* ne, const CntArr& cnt,
                    const Grid& g, int nt, int tg, const EqTypes& mx, int32_t* bcnt, const uint32_t* donors,
                    const EqOff& first, const EqOff& avail, double* r, double* s, int32_t* ty, const int64_t* id,
                    uint64_t key, double dr, double ds, double dist_r, double dist_s, double aspect, int comp,
                    int update_counts, EqRed* red)
{
    const int nblk = eq_pass_blocks(g);
    hipLaunchKernelGGL(k_eq_flags, dim3(nblk), dim3(EQ_BLOCK), 0, st, rows, cols, ne, cnt, g, nt, tg, mx, bcnt);
    hipLaunchKernelGGL(k_eq_scan, dim3(1), dim3(256), 0, st, bcnt, nblk, nt, red);
    hipLaunchKernelGGL(k_eq_assign, dim3(nblk), dim3(EQ_BLOCK), 0, st, rows, cols, ne, cnt, g, nt, tg, mx, bcnt, donors,
                       first, avail, r, s, ty, id, key, dr, ds, dist_r, dist_s, aspect, comp, update_counts);
}

// LSD radix sort of (64-bit key, slot) pairs on bits [0, bits) with
// ping-pong buffers (radix_sort_pairs); the result is in k[*cur], v[*cur].
// With temp == nullptr only sets *temp_bytes.
int eq_sort(hipStream_t st, void* temp, size_t* temp_bytes, uint64_t* k0, uint64_t* k1, uint32_t* v0, uint32_t* v1,
            int64_t n, int bits, int* cur)
{
    const size_t words = radix_scratch_words(n);
    if (!temp) {
        *temp_bytes = words * sizeof(uint32_t);
        return 0;
    }
    if (*temp_bytes < words * sizeof(uint32_t)) return TC_ERR_ARG;
    int c = 0;
    const int e = radix_sort_pairs<uint64_t>(st, k0, k1, v0, v1, false, n, bits, static_cast<uint32_t*>(temp), words, &c);
    if (cur) *cur = c;
    return e;
}

}  // namespace tc
