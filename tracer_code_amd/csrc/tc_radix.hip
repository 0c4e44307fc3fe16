// Stable LSD radix sort of (key, 32-bit value) pairs, hand-written for
// wave64: 8-bit digits, one pass per digit of the key's low `bits` bits
// (bits at and above `bits` are ignored).  k[0], v[0] hold the input (or
// only k[0], with values 0..n-1 implied); the result is in k[cur], v[cur].
//
// A pass is three launches:
//   k_rs_hist     each workgroup counts the digits of its tile of
//                 RS_TILE items (LDS counters) into hist[digit][tile];
//   k_rs_scan     one workgroup per digit turns its row of hist into the
//                 exclusive prefix over the tiles and writes the row total;
//   k_rs_scatter  each workgroup reloads its tile and moves every item to
//                 (items of smaller digits) + (items of its digit in earlier
//                 tiles) + (its rank among the equal digits of the tile).
// The rank inside a wavefront comes from eight ballots (the lanes whose
// digit equals the lane's own), the rank across the four wavefronts and the
// tile's sub-tiles from per-digit counters in LDS.  Items keep their input
// order among equal digits, so the sort is stable and its permutation is
// unique: the result does not depend on timing or on the launch shape.
#include "tc_kernels.h"

namespace tc {

constexpr int RS_THREADS = 256, RS_IPT = 8, RS_TILE = RS_THREADS * RS_IPT, RS_BINS = 256;

static inline int64_t rs_tiles(int64_t n) { return (n + RS_TILE - 1) / RS_TILE; }

size_t radix_scratch_words(int64_t n) { return (size_t)RS_BINS * (size_t)rs_tiles(n) + RS_BINS; }

template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const K* __restrict__ key, int64_t n, int shift, uint32_t dmask,
                                                        uint32_t* __restrict__ hist, int64_t tiles)
{
    __shared__ uint32_t cnt[RS_BINS];
    const int tid = threadIdx.x;
    cnt[tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
    for (int j = 0; j < RS_IPT; ++j) {
        const int64_t i = base + (int64_t)j * RS_THREADS + tid;
        if (i < n) atomicAdd(&cnt[(uint32_t)(key[i] >> shift) & dmask], 1u);
    }
    __syncthreads();
    hist[(size_t)tid * tiles + blockIdx.x] = cnt[tid];
}

// row `digit` of hist: exclusive prefix over the tiles, in place; the row
// total into total[digit].  1024 threads, each a contiguous chunk.
__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t* __restrict__ hist, int64_t tiles,
                                                  uint32_t* __restrict__ total)
{
    __shared__ uint32_t part[1024];
    uint32_t* row = hist + (size_t)blockIdx.x * tiles;
    const int tid = threadIdx.x;
    const int64_t per = (tiles + 1023) / 1024;
    const int64_t a = (int64_t)tid * per, b = a + per < tiles ? a + per : tiles;
    uint32_t s = 0;
    for (int64_t t = a; t < b; ++t) s += row[t];
    part[tid] = s;
    __syncthreads();
    // inclusive scan of the chunk sums (Hillis-Steele in LDS)
    for (int o = 1; o < 1024; o <<= 1) {
        const uint32_t v = tid >= o ? part[tid - o] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t run = tid > 0 ? part[tid - 1] : 0u;
    for (int64_t t = a; t < b; ++t) {
        const uint32_t c = row[t];
        row[t] = run;
        run += c;
    }
    if (tid == 1023) total[blockIdx.x] = part[1023];
}

template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                           K* __restrict__ kout, uint32_t* __restrict__ vout, int64_t n,
                                                           int shift, uint32_t dmask, const uint32_t* __restrict__ hist,
                                                           const uint32_t* __restrict__ total, int64_t tiles)
{
    __shared__ uint32_t base[RS_BINS];
    __shared__ uint32_t wcnt[RS_THREADS / 64][RS_BINS];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // base[d] = items of digits < d, plus items of digit d in earlier tiles
    base[tid] = total[tid];
    for (int q = 0; q < RS_THREADS / 64; ++q) wcnt[q][tid] = 0;
    __syncthreads();
    for (int o = 1; o < RS_BINS; o <<= 1) {
        const uint32_t v = tid >= o ? base[tid - o] : 0u;
        __syncthreads();
        base[tid] += v;
        __syncthreads();
    }
    const uint32_t excl = (tid > 0 ? base[tid - 1] : 0u) + hist[(size_t)tid * tiles + blockIdx.x];
    __syncthreads();
    base[tid] = excl;
    __syncthreads();
    const unsigned long long lt = (1ull << lane) - 1ull;
    const int64_t tile0 = (int64_t)blockIdx.x * RS_TILE;
    for (int j = 0; j < RS_IPT; ++j) {
        const int64_t i = tile0 + (int64_t)j * RS_THREADS + tid;
        const bool valid = i < n;
        K k = 0;
        uint32_t v = 0, d = 0;
        if (valid) {
            k = kin[i];
            v = vin ? vin[i] : (uint32_t)i;
            d = (uint32_t)(k >> shift) & dmask;
        }
        // the lanes of this wavefront whose digit equals this lane's
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const unsigned long long b = __ballot((d >> bit) & 1u);
            peers &= ((d >> bit) & 1u) ? b : ~b;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + rank;
            for (int q = 0; q < w; ++q) pos += wcnt[q][d];
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int q = 0; q < RS_THREADS / 64; ++q) {
            add += wcnt[q][tid];
            wcnt[q][tid] = 0;
        }
        base[tid] += add;
        __syncthreads();
    }
}

template <typename K>
int radix_sort_pairs(hipStream_t st, K* k0, K* k1, uint32_t* v0, uint32_t* v1, bool identity_values, int64_t n,
                     int bits, uint32_t* scratch, size_t scratch_words, int* cur)
{
    *cur = 0;
    if (n <= 0) return 0;
    if (n >= ((int64_t)1 << 31) || bits < 1 || bits > (int)(8 * sizeof(K))) return TC_ERR_ARG;
    if (scratch_words < radix_scratch_words(n)) return TC_ERR_ARG;
    const int64_t tiles = rs_tiles(n);
    uint32_t* hist = scratch;
    uint32_t* total = scratch + (size_t)RS_BINS * tiles;
    K* kb[2] = {k0, k1};
    uint32_t* vb[2] = {v0, v1};
    int c = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        // the digit's bits that lie below `bits` (higher key bits are ignored)
        const uint32_t dmask = bits - shift >= 8 ? (uint32_t)(RS_BINS - 1) : (1u << (bits - shift)) - 1u;
        hipLaunchKernelGGL((k_rs_hist<K>), dim3((unsigned)tiles), dim3(RS_THREADS), 0, st, kb[c], n, shift, dmask, hist,
                           tiles);
        hipLaunchKernelGGL(k_rs_scan, dim3(RS_BINS), dim3(1024), 0, st, hist, tiles, total);
        const uint32_t* vin = (shift == 0 && identity_values) ? nullptr : vb[c];
        hipLaunchKernelGGL((k_rs_scatter<K>), dim3((unsigned)tiles), dim3(RS_THREADS), 0, st, kb[c], vin, kb[1 - c],
                           vb[1 - c], n, shift, dmask, hist, total, tiles);
        c = 1 - c;
    }
    *cur = c;
    return hipGetLastError() == hipSuccess ? 0 : TC_ERR_HIP;
}

template int radix_sort_pairs<uint32_t>(hipStream_t, uint32_t*, uint32_t*, uint32_t*, uint32_t*, bool, int64_t, int,
                                        uint32_t*, size_t, int*);
template int radix_sort_pairs<uint64_t>(hipStream_t, uint64_t*, uint64_t*, uint32_t*, uint32_t*, bool, int64_t, int,
                                        uint32_t*, size_t, int*);

}  // namespace tc
