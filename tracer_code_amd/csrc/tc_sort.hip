// Tracer state sorted by node.  The tracer arrays are structure-of-arrays;
// sorting them by the node each tracer is binned to (tracers_to_corners'
// nearest node, composition.f90:704-715) keeps the tracers of one node
// together: the RK stages then read the velocity of neighbouring cells with
// neighbouring lanes, and binning finds long runs of one (node, type) key
// per wavefront.  The order carries no meaning in the reference; tr_id maps
// each slot back to the caller's tracer index for the boundary copies.
#include <algorithm>

#include "tc_kernels.h"

namespace tc {

// key = the binning node (row-major k * (nx+1) + i), out-of-domain or NaN
// positions sort last; val = the slot
__global__ __launch_bounds__(256) void k_tracer_keys(const double* __restrict__ r, const double* __restrict__ s,
                                                     int64_t n, int nx, double dr, double ds, double aspect,
                                                     uint32_t bad, uint32_t* __restrict__ key,
                                                     uint32_t* __restrict__ val)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double rv = r[j], sv = s[j];
    uint32_t kk = bad;
    if (rv > 0.0 && rv < aspect && sv > 0.0 && sv < 1.0) {
        const int i = (int)round(rv / dr), k = (int)round(sv / ds);
        kk = (uint32_t)k * (uint32_t)(nx + 1) + (uint32_t)i;
    }
    key[j] = kk;
    val[j] = (uint32_t)j;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gather(const T* __restrict__ in, const uint32_t* __restrict__ idx,
                                                T* __restrict__ out, int64_t n)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[j] = in[idx[j]];
}

// new id of slot j: the id of the slot it came from (identity if none yet)
__global__ __launch_bounds__(256) void k_gather_id(const int64_t* __restrict__ id, const uint32_t* __restrict__ idx,
                                                   int64_t* __restrict__ out, int64_t n)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[j] = id ? id[idx[j]] : (int64_t)idx[j];
}

// caller order: out[id[j]] = in[j]
template <typename T>
__global__ __launch_bounds__(256) void k_scatter(const T* __restrict__ in, const int64_t* __restrict__ id,
                                                 T* __restrict__ out, int64_t n)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[id[j]] = in[j];
}

static inline unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

void launch_tracer_keys(hipStream_t st, const double* r, const double* s, int64_t n, int nx, double dr, double ds,
                        double aspect, uint32_t bad, uint32_t* key, uint32_t* val)
{
    hipLaunchKernelGGL(k_tracer_keys, dim3(blocks(n)), dim3(256), 0, st, r, s, n, nx, dr, ds, aspect, bad, key, val);
}

void launch_gather_d(hipStream_t st, const double* in, const uint32_t* idx, double* out, int64_t n)
{
    hipLaunchKernelGGL(k_gather<double>, dim3(blocks(n)), dim3(256), 0, st, in, idx, out, n);
}

void launch_gather_i32(hipStream_t st, const int32_t* in, const uint32_t* idx, int32_t* out, int64_t n)
{
    hipLaunchKernelGGL(k_gather<int32_t>, dim3(blocks(n)), dim3(256), 0, st, in, idx, out, n);
}

void launch_gather_id(hipStream_t st, const int64_t* id, const uint32_t* idx, int64_t* out, int64_t n)
{
    hipLaunchKernelGGL(k_gather_id, dim3(blocks(n)), dim3(256), 0, st, id, idx, out, n);
}

void launch_scatter_d(hipStream_t st, const double* in, const int64_t* id, double* out, int64_t n)
{
    hipLaunchKernelGGL(k_scatter<double>, dim3(blocks(n)), dim3(256), 0, st, in, id, out, n);
}

void launch_scatter_i32(hipStream_t st, const int32_t* in, const int64_t* id, int32_t* out, int64_t n)
{
    hipLaunchKernelGGL(k_scatter<int32_t>, dim3(blocks(n)), dim3(256), 0, st, in, id, out, n);
}

// One launch copying up to COPY_MAX device buffers into one staging block:
// blockIdx.y picks the buffer, 16-byte words with a 4-byte tail.  A blit
// kernel streams at HBM rate where queued device-to-device copies may go
// through the copy engines.
__global__ __launch_bounds__(256) void k_copy_list(CopyList L, char* __restrict__ dst)
{
    const int b = blockIdx.y;
    const char* src = static_cast<const char*>(L.src[b]);
    char* out = dst + L.dst_off[b];
    const size_t n16 = L.bytes[b] / 16;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n16; j += stride)
        reinterpret_cast<uint4*>(out)[j] = reinterpret_cast<const uint4*>(src)[j];
    const size_t tail = (L.bytes[b] - n16 * 16) / 4;
    if (blockIdx.x == 0 && threadIdx.x < tail)
        reinterpret_cast<uint32_t*>(out + n16 * 16)[threadIdx.x] = reinterpret_cast<const uint32_t*>(src + n16 * 16)[threadIdx.x];
}

void launch_copy_list(hipStream_t st, const CopyList& L, void* dst)
{
    size_t most = 0;
    for (int j = 0; j < L.n; ++j) most = std::max(most, L.bytes[j] / 16);
    const unsigned gx = (unsigned)std::min<size_t>(std::max<size_t>((most + 255) / 256, 1), 1024);
    hipLaunchKernelGGL(k_copy_list, dim3(gx, (unsigned)L.n), dim3(256), 0, st, L, static_cast<char*>(dst));
}

}  // namespace tc
