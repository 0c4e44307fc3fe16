// Binning of tracers to nodes on caller-owned device buffers, without any
// device allocation (tc_bin_tracers_noalloc): for callers that cannot
// allocate (code under HIP graph capture, serving loops).  Its working
// storage -- the per-(type, node) counts when the caller passes no
// count_type, and the out-of-domain flag -- lives in one fixed device array
// per GPU, `scratch`, whose address is taken once per device with
// hipGetSymbolAddress.
#include <hip/hip_runtime.h>

#include <mutex>

#include "tc_internal.h"
#include "tc_kernels.h"

using namespace tc;

namespace {

// int32 words: the count planes, then the flag slot (its own 256-byte line)
constexpr size_t SCRATCH_WORDS = (size_t)TC_BIN_NOALLOC_MAX_COUNTS + 64;

static __device__ __align__(256) int32_t scratch[SCRATCH_WORDS];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
int32_t* g_scratch[MAX_DEVICES] = {};  // device address of `scratch` per device

// The device address of this device's `scratch` (looked up on the first
// call made with that device current).
int scratch_ptr(int32_t** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return TC_ERR_HIP;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TC_ERR_HIP;
        g_scratch[dev] = static_cast<int32_t*>(p);
    }
    *out = g_scratch[dev];
    return 0;
}

}  // namespace

extern "C" int tc_bin_tracers_noalloc(int32_t nx, int32_t nz, double aspect, int32_t ntype, int32_t comp,
                                      const double* r, const double* s, const int32_t* ttype, int64_t ntr,
                                      int32_t* count_type, int32_t* nempty, double* Cnew, void* stream)
{
    if (nx < 2 || nz < 2 || !(aspect > 0.0) || ntype < 0 || ntype > TC_MAX_TYPES || (comp != 0 && comp != 1) ||
        ntr < 0 || !nempty || (comp == 1 && ntype > 0 && !Cnew) || (ntr > 0 && (!r || !s || !ttype)))
        return TC_ERR_ARG;
    // the reference layout f(-2:nx+2,-2:nz+2), i fastest
    Grid g;
    g.nx = nx;
    g.nz = nz;
    g.ld = nx + 2 * SPAN + 1;
    g.org = SPAN * g.ld + SPAN;
    g.elems = (size_t)g.ld * (size_t)(nz + 2 * SPAN + 1);
    const size_t need = count_type ? 0 : (size_t)(ntype + 1) * g.elems;
    if (need > (size_t)TC_BIN_NOALLOC_MAX_COUNTS) return TC_ERR_NOMEM;
    // the bins are 32-bit keys of (node, type) in k_bin
    if (g.elems * (size_t)(TC_MAX_TYPES + 1) >= 0xffffffffull) return TC_ERR_UNSUPPORTED;
    int32_t* base = nullptr;
    if (const int st = scratch_ptr(&base)) return st;
    int32_t* flag = base + TC_BIN_NOALLOC_MAX_COUNTS;
    hipStream_t st = static_cast<hipStream_t>(stream);
    int32_t* counts = count_type ? count_type : base;
    CntArr cnt{};
    for (int t = 0; t <= ntype; ++t) {
        cnt.v[t] = counts + (size_t)t * g.elems;
        launch_zero_i32(st, cnt.v[t], g.elems);
    }
    launch_zero_i32(st, flag, 1);
    const double dr = aspect / nx, ds = 1.0 / nz;
    if (ntr > 0) launch_bin(st, r, s, ttype, ntr, cnt, g, dr, ds, aspect, flag);
    PtrArr cn{};
    for (int t = 0; t < ntype; ++t) cn.v[t] = Cnew + (size_t)t * g.elems;
    launch_ratio(st, cnt, nempty, cn, g, ntype, comp);
    if (comp == 1)
        for (int t = 0; t < ntype; ++t) launch_bcs_comp(st, cn.v[t], g);
    if (hipGetLastError() != hipSuccess) return TC_ERR_HIP;
    int32_t h = 0;
    if (hipMemcpyAsync(&h, flag, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess) return TC_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return TC_ERR_HIP;
    return h ? TC_ERR_TRACER_DOMAIN : 0;
}
