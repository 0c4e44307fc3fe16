// C-ABI of the MI355X hot path (include/tracer_hip.h) and the host-side
// orchestration of the kernels: the multigrid driver with its data-dependent
// control flow, the T/C filter loops, tracers_to_corners and the RK stages
// of energy_time.  One HIP stream carries the hot path; a second stream
// (default priority) factors and inverts the coarsest-grid matrix while the
// first W-cycle descent runs.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <new>
#include <thread>

#include "tc_kernels.h"
#include "tc_nn.h"

using namespace tc;

namespace {

constexpr double PII = 3.1415926535897932;

// slots of the device scalar array
enum { S_RESID1 = 0, S_RESREF1, S_RMIN, S_SMOOTH, S_EIGA, S_EIGD, S_COUNT };

bool is_pow2(double x)
{
    if (!(x > 0.0) || !std::isfinite(x)) return false;
    int e;
    double m = std::frexp(x, &e);
    return m == 0.5;
}

Spacing make_spacing(const Level& L)
{
    Spacing h;
    h.dr = L.dr; h.ds = L.ds; h.dr2 = L.dr2; h.ds2 = L.ds2; h.dr3 = L.dr3; h.ds3 = L.ds3; h.dr4 = L.dr4; h.ds4 = L.ds4;
    h.drds = L.drds; h.drds2 = L.drds2; h.dr2ds = L.dr2ds; h.dr2ds2 = L.dr2ds2;
    h.idr = 1.0 / L.dr; h.ids = 1.0 / L.ds; h.idr2 = 1.0 / L.dr2; h.ids2 = 1.0 / L.ds2; h.idr3 = 1.0 / L.dr3;
    h.ids3 = 1.0 / L.ds3; h.idr4 = 1.0 / L.dr4; h.ids4 = 1.0 / L.ds4; h.idrds = 1.0 / L.drds;
    h.idrds2 = 1.0 / L.drds2; h.idr2ds = 1.0 / L.dr2ds; h.idr2ds2 = 1.0 / L.dr2ds2;
    return h;
}

bool level_pow2(const Level& L)
{
    return is_pow2(L.dr) && is_pow2(L.ds) && is_pow2(L.dr2) && is_pow2(L.ds2) && is_pow2(L.dr3) && is_pow2(L.ds3) &&
           is_pow2(L.dr4) && is_pow2(L.ds4) && is_pow2(L.drds) && is_pow2(L.drds2) && is_pow2(L.dr2ds) &&
           is_pow2(L.dr2ds2);
}

}  // namespace

// Device-side objects that need only host bookkeeping.
struct tc_dev {
    FD F;
    Spacing h[TC_MAX_LEVELS + 1];
    bool pow2[TC_MAX_LEVELS + 1];
    double* d_scal = nullptr;
    double* h_scal = nullptr;
    double* partialA = nullptr;
    int learn_slot = -1;       // recorded sample of the solve in progress
    double* partialB = nullptr;
    int max_partial = 0;
    void* diag = nullptr;           // tc_diagnostics: atomics and results
    double* diag_part = nullptr;    // tc_diagnostics: 3 partial sums per block
    double* stats_part = nullptr;   // tc_stats_record: partial sums per block
    double* stats_out = nullptr;    // tc_stats_record: results
    bool coarse_ready = false;
    bool trace = false;
    bool coarse_lu = false;
    bool tb = true;            // temporally blocked sweeps on levels >= 2
    bool fuse_rr = true;       // residual+restriction in the last blocked launch of a visit (TC_FUSED_RR=0: apart)
    // defect-correction cycle: level-1 residual updated by increments (k_absorb_inc) with an exact
    // double-double residual before the stop test accepts (TC_INC_RESID=0: exact after every correction)
    bool inc_res = true;
    bool res_exact = true;     // L1.res is the double-double residual of psi (not an incremental one)
    int cus = 256;             // compute units of the device
    bool inv_pending = false;  // the main stream has not yet waited for the coarse inverse
    int32_t* ctl = nullptr;    // device loop control of the filter loops: stop flag, stop index
    unsigned long long* fslots = nullptr;  // the filter iterations' smoothness maxima (rotating slot sets)
    double* fbuf = nullptr;    // third iterate buffer of the blocked C filter loop
    int32_t* h_ctl = nullptr;  // pinned mirror
    // device loop control of smoother_vis (vis_decide in k_vis_sweep_bcs):
    // state, factor and ratio, their pinned mirrors, and the pinned initial state
    int32_t* vctl = nullptr;   // two loop-state buffers of smoother_vis (VC_COUNT each)
    double* vdctl = nullptr;   // two (VD_COUNT each)
    unsigned long long* vslots = nullptr;  // two sets of VIS_SLOTS minimum-ratio slots
    int32_t* h_vctl = nullptr;
    double* h_vdctl = nullptr;
    int32_t* h_vinit = nullptr;
    int vis_hint[TC_MAX_LEVELS + 1][2] = {};  // iterations the last call of (level, interpolate) took
    std::vector<double> hbuf;  // host staging plane
};

static tc_dev& DEV(tc_ctx* ctx) { return *static_cast<tc_dev*>(ctx->impl); }

namespace {

// Allocations are zeroed on the context stream, which orders them before
// every later use (the streams are non-blocking, so the null stream would not).
int dalloc(tc_ctx* ctx, double** p, size_t n)
{
    TC_HIP(hipMalloc((void**)p, n * sizeof(double)));
    TC_HIP(hipMemsetAsync(*p, 0, n * sizeof(double), ctx->stream));
    return 0;
}
int falloc(tc_ctx* ctx, float** p, size_t n);
int dalloc_i(tc_ctx* ctx, int32_t** p, size_t n)
{
    TC_HIP(hipMalloc((void**)p, n * sizeof(int32_t)));
    TC_HIP(hipMemsetAsync(*p, 0, n * sizeof(int32_t), ctx->stream));
    return 0;
}

EigScalars eig_scalars(const tc_params& p, const Level& L)
{
    // (iterative_solver.f90:674-685)
    EigScalars e;
    e.Ts_max = p.Tr_base / L.ds;
    e.Tss_max = -p.Trr_base / L.ds2;
    e.Tsss_max = p.Trrr_base / L.ds3;
    e.Tssss_max = p.Trrrr_base / L.ds4;
    e.Tr_max = p.Tr_base / L.dr;
    e.Trr_max = -p.Trr_base / L.dr2;
    e.Trrr_max = p.Trrr_base / L.dr3;
    e.Trrrr_max = p.Trrrr_base / L.dr4;
    e.Trs_max = -p.Trs_base / L.drds;
    e.Trrss_max = p.Trrss_base / L.dr2ds2;
    e.Trss_max = p.Trrs_base / L.drds2;
    e.Trrs_max = p.Trrs_base / L.dr2ds;
    return e;
}

size_t plane_bytes(const Grid& g) { return g.elems * sizeof(double); }

// Host <-> device copies of rows of a pitched device array through two pinned
// chunk buffers: the caller's arrays are pageable (a Fortran driver's module
// arrays), and a pitched copy from pageable memory runs at ~2 GB/s.  Chunk c
// goes by DMA between the device and buffer c%2 while the host copies the
// previous chunk between the other buffer and the caller's array.  Both
// return with the host array read or written (the boundary's contract).
constexpr size_t XFER_CHUNK = (size_t)2 << 20;  // bytes per pinned buffer
static int xfer_init(tc_ctx* ctx)
{
    for (int b = 0; b < 2; ++b) {
        if (!ctx->xfer_buf[b]) TC_HIP(hipHostMalloc(&ctx->xfer_buf[b], XFER_CHUNK, hipHostMallocDefault));
        if (!ctx->xfer_ev[b]) TC_HIP(hipEventCreateWithFlags(&ctx->xfer_ev[b], hipEventDisableTiming));
    }
    return 0;
}
static int copy_rows_d2h(tc_ctx* ctx, void* host, const void* dev, size_t dpitch, size_t width, size_t rows)
{
    TC_TRY(xfer_init(ctx));
    const size_t per = std::max<size_t>(1, XFER_CHUNK / width);
    if (per * width > XFER_CHUNK) {  // one row wider than a buffer: a plain copy
        TC_HIP(hipMemcpy2DAsync(host, width, dev, dpitch, width, rows, hipMemcpyDeviceToHost, ctx->stream));
        TC_HIP(hipStreamSynchronize(ctx->stream));
        return 0;
    }
    const size_t nch = (rows + per - 1) / per;
    auto issue = [&](size_t c) -> int {
        const size_t r0 = c * per, nr = std::min(per, rows - r0);
        TC_HIP(hipMemcpy2DAsync(ctx->xfer_buf[c % 2], width, static_cast<const char*>(dev) + r0 * dpitch, dpitch, width,
                                nr, hipMemcpyDeviceToHost, ctx->stream));
        TC_HIP(hipEventRecord(ctx->xfer_ev[c % 2], ctx->stream));
        return 0;
    };
    for (size_t c = 0; c < std::min<size_t>(2, nch); ++c) TC_TRY(issue(c));
    for (size_t c = 0; c < nch; ++c) {
        const size_t r0 = c * per, nr = std::min(per, rows - r0);
        TC_HIP(hipEventSynchronize(ctx->xfer_ev[c % 2]));
        std::memcpy(static_cast<char*>(host) + r0 * width, ctx->xfer_buf[c % 2], nr * width);
        if (c + 2 < nch) TC_TRY(issue(c + 2));
    }
    return 0;
}
static int copy_rows_h2d(tc_ctx* ctx, void* dev, size_t dpitch, const void* host, size_t width, size_t rows)
{
    TC_TRY(xfer_init(ctx));
    const size_t per = std::max<size_t>(1, XFER_CHUNK / width);
    if (per * width > XFER_CHUNK) {
        TC_HIP(hipMemcpy2DAsync(dev, dpitch, host, width, width, rows, hipMemcpyHostToDevice, ctx->stream));
        TC_HIP(hipStreamSynchronize(ctx->stream));
        return 0;
    }
    const size_t nch = (rows + per - 1) / per;
    for (size_t c = 0; c < nch; ++c) {
        const size_t r0 = c * per, nr = std::min(per, rows - r0);
        if (c >= 2) TC_HIP(hipEventSynchronize(ctx->xfer_ev[c % 2]));  // its buffer's last DMA is done
        std::memcpy(ctx->xfer_buf[c % 2], static_cast<const char*>(host) + r0 * width, nr * width);
        TC_HIP(hipMemcpy2DAsync(static_cast<char*>(dev) + r0 * dpitch, dpitch, ctx->xfer_buf[c % 2], width, width, nr,
                                hipMemcpyHostToDevice, ctx->stream));
        TC_HIP(hipEventRecord(ctx->xfer_ev[c % 2], ctx->stream));
    }
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

// host <-> device copies of a ghosted Fortran-layout plane (ld = nx+5),
// complete on return
int put_plane(tc_ctx* ctx, double* dev, const Grid& g, const double* host)
{
    const size_t w = (size_t)(g.nx + 2 * SPAN + 1);
    return copy_rows_h2d(ctx, dev + g.at(-SPAN, -SPAN), g.ld * sizeof(double), host, w * sizeof(double),
                         (size_t)(g.nz + 2 * SPAN + 1));
}
int get_plane(tc_ctx* ctx, const double* dev, const Grid& g, double* host)
{
    const size_t w = (size_t)(g.nx + 2 * SPAN + 1);
    return copy_rows_d2h(ctx, host, dev + g.at(-SPAN, -SPAN), g.ld * sizeof(double), w * sizeof(double),
                         (size_t)(g.nz + 2 * SPAN + 1));
}
// the same as put_plane without the pinned buffers (one-time set-up copies)
int put_plane_direct(tc_ctx* ctx, double* dev, const Grid& g, const double* host)
{
    const size_t w = (size_t)(g.nx + 2 * SPAN + 1);
    TC_HIP(hipMemcpy2DAsync(dev + g.at(-SPAN, -SPAN), g.ld * sizeof(double), host, w * sizeof(double),
                            w * sizeof(double), g.nz + 2 * SPAN + 1, hipMemcpyHostToDevice, ctx->stream));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}
int put_iplane(tc_ctx* ctx, int32_t* dev, const Grid& g, const int32_t* host)
{
    const size_t w = (size_t)(g.nx + 2 * SPAN + 1);
    TC_HIP(hipMemcpy2DAsync(dev + g.at(-SPAN, -SPAN), g.ld * sizeof(int32_t), host, w * sizeof(int32_t),
                            w * sizeof(int32_t), g.nz + 2 * SPAN + 1, hipMemcpyHostToDevice, ctx->stream));
    return 0;
}
int get_iplane(tc_ctx* ctx, const int32_t* dev, const Grid& g, int32_t* host)
{
    const size_t w = (size_t)(g.nx + 2 * SPAN + 1);
    TC_HIP(hipMemcpy2DAsync(host, w * sizeof(int32_t), dev + g.at(-SPAN, -SPAN), g.ld * sizeof(int32_t),
                            w * sizeof(int32_t), g.nz + 2 * SPAN + 1, hipMemcpyDeviceToHost, ctx->stream));
    return 0;
}

int copy_plane(tc_ctx* ctx, double* dst, const double* src, const Grid& g)
{
    TC_HIP(hipMemcpyAsync(dst, src, plane_bytes(g), hipMemcpyDeviceToDevice, ctx->stream));
    return 0;
}

// kernel timing: event pairs around timed launches, folded into totals
constexpr int PROF_PAIRS = 4096;
int prof_flush(tc_ctx* ctx)
{
    if (ctx->prof_used == 0) return 0;
    TC_HIP(hipEventSynchronize(ctx->prof_ev[2 * ctx->prof_used - 1]));
    for (int q = 0; q < ctx->prof_used; ++q) {
        float ms = 0.f;
        TC_HIP(hipEventElapsedTime(&ms, ctx->prof_ev[2 * q], ctx->prof_ev[2 * q + 1]));
        ctx->prof_ms[TC_K_SWEEP_L1] += ms;
        ctx->prof_n[TC_K_SWEEP_L1] += ctx->prof_cnt[q];
    }
    ctx->prof_used = 0;
    return 0;
}
int prof_begin(tc_ctx* ctx)
{
    if (ctx->prof_used == PROF_PAIRS) TC_TRY(prof_flush(ctx));
    TC_HIP(hipEventRecord(ctx->prof_ev[2 * ctx->prof_used], ctx->stream));
    return 0;
}
// the pair brackets `launches` back-to-back launches of the timed kernel
int prof_end(tc_ctx* ctx, int launches)
{
    TC_HIP(hipEventRecord(ctx->prof_ev[2 * ctx->prof_used + 1], ctx->stream));
    ctx->prof_cnt[ctx->prof_used] = launches;
    ctx->prof_used++;
    return 0;
}

// read device scalars [0, n) into the pinned mirror (synchronises the stream)
int fetch_scalars(tc_ctx* ctx, tc_dev& d, int n)
{
    TC_HIP(hipMemcpyAsync(d.h_scal, d.d_scal, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// viscosity hierarchy
// ---------------------------------------------------------------------------
int grad_eig_min(tc_ctx* ctx, tc_dev& d, int level, double* rmin)
{
    Level& L = ctx->lv[level];
    EigScalars e = eig_scalars(ctx->p, L);
    int nb = launch_vis_grad_eig(ctx->stream, d.pow2[level], L, d.F, d.h[level], ctx->p.ivis, e, 1, d.partialA);
    ctx->ratio_level = (ctx->p.ivis == 0) ? -1 : level;
    launch_reduce_final(ctx->stream, d.partialA, nb, 1, d.d_scal, S_RMIN);
    TC_TRY(fetch_scalars(ctx, d, S_COUNT));
    *rmin = (ctx->p.ivis == 0) ? 1e99 : d.h_scal[S_RMIN];
    ctx->stats.vis_iters++;
    return 0;
}

// smoother_vis (smoother.f90:165-263), host-controlled: one read-back of
// the ratio minimum per iteration (passes longer than VIS_CMAX)
int smoother_vis_host(tc_ctx* ctx, tc_dev& d, int level, int interpolate)
{
    Level& L = ctx->lv[level];
    const tc_params& p = ctx->p;
    int option, Niter;
    double lb, la;
    if (level == 1 && interpolate == 0) { option = 1; Niter = 6; lb = -5.0; la = -2.0; }
    else { option = 0; Niter = p.vis_iter; lb = -2.0; la = -1.0; }
    if (interpolate == 0 || level == 1) { ctx->Nvis_store = 0; ctx->nn_store = 1; }
    int kmin, kmax;
    if (p.RaT == 0.0) { kmin = 0; kmax = L.g.nz; } else { kmin = 1; kmax = L.g.nz - 1; }
    int n = 0, trigger = 0;
    const Grid* gf = level > 1 ? &ctx->lv[level - 1].g : &L.g;
    const double* Vf = level > 1 ? ctx->lv[level - 1].vis : nullptr;
    for (;;) {
        for (int nn = ctx->nn_store; nn <= Niter; ++nn) {
            double rmin;
            TC_TRY(grad_eig_min(ctx, d, level, &rmin));
            if (option == 1) {
                if ((ctx->Nvis_store >= 1 && rmin > p.eig_ratio_min) || p.ivis == 0 || n > 20) {
                    ctx->stats.eig_fine = rmin;
                    ctx->stats.n_eta = n;
                    trigger = 1;
                    break;
                }
            } else if (interpolate == 0) {
                if (rmin > p.eig_ratio || p.ivis == 0) { trigger = 1; ctx->nn_store = nn; break; }
            } else {
                if (ctx->Nvis_store >= 1 || p.ivis == 0) { trigger = 1; break; }
            }
            const int ntemp = Niter - nn + 1;
            const double c = 2.0 / (-lb - la + (lb - la) * std::cos((double)(2 * ntemp - 1) * PII / (double)(2 * Niter)));
            TC_TRY(copy_plane(ctx, L.vtmp, L.vis, L.g));
            launch_vis_sweep(ctx->stream, d.pow2[level], L.vtmp, L.vis, L.dtvis, L.g, d.F, d.h[level], kmin, kmax, c);
            launch_visc_bcs(ctx->stream, L.vis, L.g, Vf, *gf, level, 0, p.RaT, p.visT * p.visT);
        }
        if (trigger) break;
        ctx->Nvis_store++;
        ctx->nn_store = 1;
        if (option == 1) { ctx->nn_store = 1; Niter = 6; lb = -2.0; la = -1.0; }
        n++;
        // the reference loops until the eigenvalue ratio test passes; a NaN
        // ratio (or a failed launch) would never pass it
        if (n > 100000) {
            ctx->err = "smoother_vis: eigenvalue-ratio test not met after 100000 passes";
            return TC_ERR_HIP;
        }
    }
    return 0;
}

// smoother_vis (smoother.f90:165-263) with the loop control on the device:
// the host queues iterations -- two launches each: the eigenvalue ratios'
// minimum and the copy; the exit test, the Chebyshev sweep and the boundary
// values -- in batches sized by the iterations the previous call on this
// level took, and reads the loop state once per batch; queued iterations
// after the exit do nothing.  The Chebyshev factors are the host table of
// the same formula, so the sweeps see the same values.
int smoother_vis(tc_ctx* ctx, tc_dev& d, int level, int interpolate)
{
    Level& L = ctx->lv[level];
    const tc_params& p = ctx->p;
    int option, Niter;
    double lb, la;
    if (level == 1 && interpolate == 0) { option = 1; Niter = 6; lb = -5.0; la = -2.0; }
    else { option = 0; Niter = p.vis_iter; lb = -2.0; la = -1.0; }
    if (Niter > VIS_CMAX || p.vis_iter > VIS_CMAX) return smoother_vis_host(ctx, d, level, interpolate);
    if (interpolate == 0 || level == 1) { ctx->Nvis_store = 0; ctx->nn_store = 1; }
    int kmin, kmax;
    if (p.RaT == 0.0) { kmin = 0; kmax = L.g.nz; } else { kmin = 1; kmax = L.g.nz - 1; }
    const Grid* gf = level > 1 ? &ctx->lv[level - 1].g : &L.g;
    const double* Vf = level > 1 ? ctx->lv[level - 1].vis : nullptr;
    auto cfac = [](int N, int nn, double lbv, double lav) {
        const int ntemp = N - nn + 1;
        return 2.0 / (-lbv - lav + (lbv - lav) * std::cos((double)(2 * ntemp - 1) * PII / (double)(2 * N)));
    };
    VisDecide P{};
    P.option = option;
    P.interpolate = interpolate;
    P.ivis = p.ivis;
    P.eig_ratio_min = p.eig_ratio_min;
    P.eig_ratio = p.eig_ratio;
    const int N1 = (option == 1) ? 6 : Niter;  // the passes after the first (option 1) or every pass
    for (int nn = 1; nn <= Niter; ++nn) P.c0[nn - 1] = cfac(Niter, nn, lb, la);
    for (int nn = 1; nn <= N1; ++nn) P.c1[nn - 1] = cfac(N1, nn, -2.0, -1.0);
    // the state at entry; a first pass that starts beyond its end is only
    // the pass bookkeeping (the reference's inner loop does not run)
    int nn0 = ctx->nn_store, n0 = 0, Nvs = ctx->Nvis_store, tab = (option == 1) ? 0 : 1, Ncur = Niter;
    while (nn0 > Ncur) {
        ++Nvs;
        nn0 = 1;
        if (option == 1) { Ncur = 6; tab = 1; }
        ++n0;
    }
    int32_t* h = d.h_vinit;
    for (int j = 0; j < VC_COUNT; ++j) h[j] = 0;
    h[VC_NN] = nn0; h[VC_N] = n0; h[VC_NITER] = Ncur; h[VC_NVIS] = Nvs; h[VC_NNSTORE] = (nn0 != ctx->nn_store) ? 1 : ctx->nn_store;
    h[VC_TAB] = tab;
    // the loop state of iteration j is buffer j % 2 (k_vis_sweep_bcs writes the
    // next one); the slot sets of the first two iterations start at the
    // largest bit pattern (atomicMin)
    TC_HIP(hipMemcpyAsync(d.vctl, h, VC_COUNT * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    TC_HIP(hipMemsetAsync(d.vdctl, 0, VD_COUNT * sizeof(double), ctx->stream));
    TC_HIP(hipMemsetAsync(d.vslots, 0xff, 2 * VIS_SLOTS * VIS_SLOT_STRIDE * sizeof(unsigned long long), ctx->stream));
    ctx->ratio_level = (p.ivis == 0) ? -1 : level;
    const EigScalars e = eig_scalars(p, L);
    int batch = std::max(1, d.vis_hint[level][interpolate ? 1 : 0]);
    const size_t slot_set = (size_t)VIS_SLOTS * VIS_SLOT_STRIDE;
    for (int queued = 0;;) {
        for (int b = 0; b < batch; ++b) {
            const int j = (queued + b) & 1, jn = 1 - j;
            VisLoopArgs lp{d.vctl + j * VC_COUNT, d.vctl + jn * VC_COUNT, d.vdctl + j * VD_COUNT,
                           d.vdctl + jn * VD_COUNT, d.vslots + j * slot_set, d.vslots + jn * slot_set, P};
            // the ratio's block minima into the slots and the copy that the sweep reads
            launch_vis_grad_eig(ctx->stream, d.pow2[level], L, d.F, d.h[level], p.ivis, e, 1, nullptr,
                                lp.ctl_in + VC_DONE, L.vtmp, false, d.vslots + j * slot_set);
            // the decision, the sweep and viscosityBCs in one launch
            launch_vis_sweep_bcs(ctx->stream, d.pow2[level], L.vtmp, L.vis, L.dtvis, L.g, d.F, d.h[level], kmin, kmax,
                                 lp, Vf, *gf, level, p.RaT, p.visT * p.visT);
        }
        queued += batch;
        const int jf = queued & 1;  // the state after the last queued iteration
        TC_HIP(hipGetLastError());
        TC_HIP(hipMemcpyAsync(d.h_vctl, d.vctl + jf * VC_COUNT, VC_COUNT * sizeof(int32_t), hipMemcpyDeviceToHost,
                              ctx->stream));
        TC_HIP(hipMemcpyAsync(d.h_vdctl, d.vdctl + jf * VD_COUNT, VD_COUNT * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
        TC_HIP(hipStreamSynchronize(ctx->stream));
        if (d.h_vctl[VC_ERR]) {
            ctx->err = "smoother_vis: eigenvalue-ratio test not met after 100000 passes";
            return TC_ERR_HIP;
        }
        if (d.h_vctl[VC_DONE]) break;
        // the reference loops until the test passes; a NaN ratio never passes it
        if (queued > 100000 * (VIS_CMAX + 1)) {
            ctx->err = "smoother_vis: eigenvalue-ratio test not met";
            return TC_ERR_HIP;
        }
        batch = std::max(2, batch / 2);
    }
    // the gradient and ratio planes of the final viscosity (what the loop's
    // last iteration computed before its stop)
    launch_vis_grad_eig(ctx->stream, d.pow2[level], L, d.F, d.h[level], p.ivis, e, 1, nullptr);
    const int iters = d.h_vctl[VC_ITERS];
    d.vis_hint[level][interpolate ? 1 : 0] = iters;
    ctx->stats.vis_iters += iters;
    ctx->Nvis_store = d.h_vctl[VC_NVIS];
    ctx->nn_store = d.h_vctl[VC_NNSTORE];
    if (option == 1) {
        ctx->stats.eig_fine = d.h_vdctl[VD_EIG];
        ctx->stats.n_eta = d.h_vctl[VC_NETA];
    }
    return 0;
}

int save_restore_level(tc_ctx* ctx, Level& L, bool save)
{
    double* a[6] = {L.vis, L.visx, L.visz, L.visxx, L.viszz, L.visxz};
    double* b[6] = {L.visf, L.visxf, L.viszf, L.visxxf, L.viszzf, L.visxzf};
    for (int t = 0; t < 6; ++t) TC_TRY(save ? copy_plane(ctx, b[t], a[t], L.g) : copy_plane(ctx, a[t], b[t], L.g));
    return 0;
}

// compute_viscosity (iterative_solver.f90:331-377)
int compute_viscosity(tc_ctx* ctx, tc_dev& d, int level)
{
    const tc_params& p = ctx->p;
    if (level == 1) {
        RaCArr lc{};
        PtrArr cv{};
        for (int t = 0; t < p.ntype; ++t) { lc.v[t] = (p.comp == 1 && p.ivis == 2) ? std::log(p.visC[t]) : 0.0; cv.v[t] = ctx->Cvis[t]; }
        launch_visf(ctx->stream, ctx->lv[1].vis, ctx->Tvis, cv, ctx->zg, ctx->lv[1].g, p.ivis, p.comp, p.ntype, p.RaT,
                    std::log(p.visT), std::log(p.visP), lc);
        return smoother_vis(ctx, d, 1, 0);
    }
    Level& Lf = ctx->lv[level - 1];
    Level& L = ctx->lv[level];
    TC_TRY(save_restore_level(ctx, Lf, true));
    TC_TRY(smoother_vis(ctx, d, level - 1, 1));
    launch_restrict(ctx->stream, Lf.vis, Lf.g, L.vis, L.g);
    launch_visc_bcs(ctx->stream, L.vis, L.g, Lf.vis, Lf.g, level, 1, p.RaT, p.visT * p.visT);
    if (level < ctx->grids) {
        TC_TRY(smoother_vis(ctx, d, level, 0));
    } else {
        EigScalars e = eig_scalars(p, L);
        launch_vis_grad_eig(ctx->stream, d.pow2[level], L, d.F, d.h[level], p.ivis, e, 0, nullptr);
    }
    TC_TRY(save_restore_level(ctx, Lf, false));
    return 0;
}

// ---------------------------------------------------------------------------
// multigrid (iterative_solver.f90:79-173), Newtonian branch
// ---------------------------------------------------------------------------
// Chebyshev factors of sweep n of a visit (iterative_solver.f90:873-899):
// temp for the per-point ratio form, cconst for the scalar form (c=1 on
// the upward leg, lb=-2 / la=-1 otherwise).
void cheb_factors(int n, int Niter, int dir, double& temp, double& cconst)
{
    const double ntemp = (double)(Niter - n + 1);
    temp = std::cos((2.0 * ntemp - 1.0) * PII / (double)(2 * Niter));
    cconst = 1.0;
    if (dir != 1) {
        const double lb = -2.0, la = -1.0;
        cconst = 2.0 / (-lb - la + (lb - la) * temp);
    }
}

// The Niter sweeps of one visit of `level` on its error buffers (FP64
// err[] or FP32 err32[]).  view.rhs is the FP64 right-hand side, rhs32 the
// FP32 one.  In the defect-correction modes the levels >= 2 run several
// sweeps per launch (k_sweep_tb, the 13-tap form of the operator) unless
// TC_TB=0; the reference-form path and level 1 run one sweep per launch.
// rr (FP32 blocked sweeps): the visit's last launch also forms the next
// coarser level's right-hand side; *rr_done says whether it did.
int run_sweeps(tc_ctx* ctx, tc_dev& d, int level, int Niter, int dir, bool f32, const Level& view, const float* rhs32,
               bool* first_sweep, bool zero_start = false, const RRTarget* rr = nullptr, bool* rr_done = nullptr)
{
    if (rr_done) *rr_done = false;
    Level& L = ctx->lv[level];
    const tc_params& P = ctx->p;
    const double zsg = (P.Vbc == 0) ? -1.0 : 1.0;
    const bool use_ratio = (dir == 0) && (level == ctx->ratio_level);
    if (f32 && level == 1 && ctx->gs_sweeps > 0 && ctx->mode == TC_MODE_MIXED) {
        // multicolour Gauss-Seidel in place of the visit's Jacobi-Richardson sweeps
        for (int n = 0; n < ctx->gs_sweeps;) {
            const int nk = std::min(ctx->gs_per_launch, ctx->gs_sweeps - n);
            launch_gs32(ctx->stream, L, rhs32, L.err32[L.cur32], L.err32[1 - L.cur32], nk, ctx->gs_omega, zsg);
            L.cur32 = 1 - L.cur32;
            n += nk;
        }
        ctx->stats.sweeps_exec[0] += ctx->gs_sweeps;
        return 0;
    }
    const bool tb = d.tb && level >= 2 && ctx->mode != TC_MODE_FP64;
    // sweeps per launch: 4, or 8 on FP32 levels of at most 16 tiles
    const int K = (tb && f32 && L.tb_tile != 32 && L.tb_tile != 48) ? tb_sweeps_per_launch(L.g) : 4;
    // level-1 timing (tc_set_profiling): one event pair around the visit's
    // back-to-back launches, or around each launch when the first sweep is
    // followed by the wall reset of the reference form
    const bool timed = ctx->prof && level == 1 && !tb;
    const bool timed_visit = timed && !(first_sweep && *first_sweep);
    if (timed_visit) TC_TRY(prof_begin(ctx));
    int n = 1;
    while (n <= Niter) {
        if (tb) {
            const int nk = std::min(tb_max_sweeps(K), Niter - n + 1);
            double temp[8], cc[8];
            for (int s = 0; s < nk; ++s) cheb_factors(n + s, Niter, dir, temp[s], cc[s]);
            if (f32) {
                // the first launch of a zero-started visit stages zeros instead of reading them;
                // the first launch after an ascent adds the prolonged coarse correction
                const float* in = (zero_start && n == 1) ? nullptr : L.err32[L.cur32];
                const bool pro = n == 1 && ctx->pend_src && ctx->pend_level == level;
                const bool last = n + nk - 1 == Niter;
                const bool fused =
                    launch_sweep_tb32(ctx->stream, use_ratio, K, L, rhs32, in, L.err32[1 - L.cur32], d.F, d.h[level],
                                      temp, cc, nk, zsg, pro ? ctx->pend_src : nullptr,
                                      pro ? &ctx->lv[ctx->pend_from].g : nullptr, last ? rr : nullptr);
                if (fused && rr_done) *rr_done = true;
                if (pro) ctx->pend_src = nullptr;
                L.cur32 = 1 - L.cur32;
            } else {
                launch_sweep_tb64(ctx->stream, use_ratio, K, L, view.rhs, L.err[L.cur], L.err[1 - L.cur], d.F,
                                  d.h[level], temp, cc, nk, zsg);
                L.cur = 1 - L.cur;
            }
            n += nk;
            continue;
        }
        double temp, cconst;
        cheb_factors(n, Niter, dir, temp, cconst);
        // the FP32 cycle's level-1 interval (tc_set_dc_cheb_interval), at a
        // tuned count only: the Nmulti retry of a solve that diverged or
        // stalled runs the reference's factors (a wide interval's steps reach
        // 4x the local time step, and Nmulti of them amplify the FP32
        // rounding of the high modes until the solve blows up)
        if (f32 && level == 1 && dir != 1 && ctx->dc_sweeps > 0 && ctx->dc_sweeps != P.Nmulti &&
            ctx->dc_sweeps_active == ctx->dc_sweeps)
            cconst = 2.0 / (-ctx->cheb1_lb - ctx->cheb1_la + (ctx->cheb1_lb - ctx->cheb1_la) * temp);
        if (timed && !timed_visit) TC_TRY(prof_begin(ctx));
        if (f32) {
            launch_sweep32(ctx->stream, use_ratio, L, L.err32[L.cur32], L.err32[1 - L.cur32], rhs32, d.F, d.h[level],
                           temp, cconst, zsg);
            L.cur32 = 1 - L.cur32;
        } else {
            launch_sweep(ctx->stream, d.pow2[level], use_ratio, view, L.err[L.cur], L.err[1 - L.cur], d.F, d.h[level],
                         P.ivis, temp, cconst, zsg);
            L.cur = 1 - L.cur;
        }
        if (timed && !timed_visit) TC_TRY(prof_end(ctx, 1));
        if (first_sweep && *first_sweep) {
            // error(1) started as a copy of SF; from the second sweep on the
            // reference reads walls zeroed by enforceBCs_stream
            launch_bc_stream(ctx->stream, L.err[1 - L.cur], L.g, P.Vbc);
            *first_sweep = false;
        }
        ++n;
    }
    if (timed_visit) TC_TRY(prof_end(ctx, Niter));
    ctx->stats.sweeps_exec[level - 1] += Niter;
    return 0;
}

int iterate_stream(tc_ctx* ctx, tc_dev& d, int level, int Niter, int dir, bool& first_sweep)
{
    Level& L = ctx->lv[level];
    const tc_params& P = ctx->p;
    TC_TRY(run_sweeps(ctx, d, level, Niter, dir, false, L, nullptr, &first_sweep));
    // compute_residual
    int nb = launch_residual(ctx->stream, d.pow2[level], L, L.err[L.cur], L.res, d.F, d.h[level], P.ivis,
                             level == 1 ? d.partialA : nullptr);
    ctx->stats.residuals_exec[level - 1]++;
    if (level == 1) launch_reduce_final(ctx->stream, d.partialA, nb, 0, d.d_scal, S_RESID1);
    return 0;
}

// Assembly, DGBEQU/DGBTF2 and the explicit inverse of the coarsest
// operator on the side stream.  lagged (defect-correction modes): the result
// goes to the *_next buffers for the next refresh, and this solve uses the
// previous refresh's (swapped in here); the first solve waits for its own.
// Between refreshes (every coarse_refresh-th solve) the inverse in use stays.
int coarse_setup(tc_ctx* ctx, tc_dev& d, bool lagged)
{
    Level& L = ctx->lv[ctx->grids];
    const tc_params& P = ctx->p;
    bool to_next = false;
    if (lagged && d.coarse_ready) {
        if (++ctx->coarse_since < ctx->coarse_refresh) return 0;
        ctx->coarse_since = 0;
        if (ctx->next_pending) {
            std::swap(ctx->Ainv, ctx->Ainv_next);
            std::swap(ctx->Ainv32, ctx->Ainv32_next);
            std::swap(ctx->RowB, ctx->RowB_next);
            std::swap(ctx->d_info, ctx->d_info_next);
            std::swap(ctx->ev_inv_ready, ctx->ev_inv_next);
        }
        if (!ctx->Ainv_next) {
            TC_TRY(dalloc(ctx, &ctx->Ainv_next, (size_t)ctx->ngc * ctx->ngc));
            TC_TRY(falloc(ctx, &ctx->Ainv32_next, (size_t)ctx->ngc * ctx->ngc));
            TC_TRY(dalloc(ctx, &ctx->RowB_next, ctx->ngc));
            TC_HIP(hipMalloc((void**)&ctx->d_info_next, sizeof(int32_t)));
            TC_HIP(hipMemsetAsync(ctx->d_info_next, 0, sizeof(int32_t), ctx->stream));
        }
        to_next = true;
    }
    double* Ainv = to_next ? ctx->Ainv_next : ctx->Ainv;
    double* RowB = to_next ? ctx->RowB_next : ctx->RowB;
    int32_t* info = to_next ? ctx->d_info_next : ctx->d_info;
    // the side stream starts after this solve's viscosity and after every
    // earlier use of the buffers it overwrites (stream order of the main stream)
    TC_HIP(hipEventRecord(ctx->ev_visc_ready, ctx->stream));
    TC_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_visc_ready, 0));
    TC_HIP(hipMemsetAsync(ctx->Mc, 0, sizeof(double) * (size_t)ctx->ldab * ctx->ngc, ctx->side));
    launch_coarse_assemble(ctx->side, ctx->Mc, RowB, L, d.F, d.h[ctx->grids], P.ivis, P.Vbc, ctx->kl, ctx->ku,
                           ctx->ngc);
    launch_dgbtf2(ctx->side, ctx->Mc, ctx->ldab, ctx->ngc, ctx->kl, ctx->ku, ctx->ipiv, info);
    launch_band_inverse(ctx->side, ctx->Mc, ctx->ldab, ctx->ngc, ctx->kl, ctx->ku, ctx->ipiv, Ainv, ctx->Ur);
    launch_to_f32(ctx->side, Ainv, to_next ? ctx->Ainv32_next : ctx->Ainv32, (size_t)ctx->ngc * ctx->ngc);
    TC_HIP(hipEventRecord(to_next ? ctx->ev_inv_next : ctx->ev_inv_ready, ctx->side));
    ctx->next_pending = to_next;
    d.coarse_ready = true;
    return 0;
}

// compute_RHS(1), compute_reference_residual(1), the viscosity hierarchy,
// the coarse matrix factorisation/inverse (side stream) and the local time
// steps: the set-up part of multigrid (iterative_solver.f90:79-100).
int solve_setup(tc_ctx* ctx, tc_dev& d)
{
    const tc_params& P = ctx->p;
    const int G = ctx->grids;
    Level& L1 = ctx->lv[1];
    {
        RaCArr rc{};
        PtrArr cb{};
        for (int t = 0; t < P.ntype; ++t) { rc.v[t] = P.RaC[t]; cb.v[t] = ctx->Cbuoy[t]; }
        int nb = launch_rhs1(ctx->stream, ctx->Tbuoy, cb, L1.rhs, L1.g, d.F, L1.dr, P.RaT, P.comp, P.ntype, rc, d.partialA);
        launch_reduce_final(ctx->stream, d.partialA, nb, 0, d.d_scal, S_RESREF1);
    }
    TC_TRY(fetch_scalars(ctx, d, S_COUNT));
    ctx->res_ref[1] = d.h_scal[S_RESREF1] / (double)(L1.g.nx - 2) / (double)(L1.g.nz - 2);
    for (int L = 1; L <= G; ++L) TC_TRY(compute_viscosity(ctx, d, L));
    if (ctx->time == 0.0 || P.ivis > 0 || !d.coarse_ready) TC_TRY(coarse_setup(ctx, d, ctx->mode != TC_MODE_FP64));
    for (int L = 1; L <= G; ++L) {
        EigScalars e = eig_scalars(P, ctx->lv[L]);
        launch_local_dt(ctx->stream, ctx->lv[L], e, P.courant_stream);
    }
    d.inv_pending = true;
    return 0;
}

// The coarse factorisation runs on the side stream while the first descent
// of the cycle smooths the finer levels; the first coarse solve waits for it.
int wait_coarse(tc_ctx* ctx, tc_dev& d)
{
    if (d.inv_pending) {
        TC_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_inv_ready, 0));
        d.inv_pending = false;
    }
    return 0;
}

// Stop test of the reference after a level-1 visit (iterative_solver.f90:
// 130-145): reads the device sum of |residual(1)|.
int stop_test(tc_ctx* ctx, tc_dev& d, int q, bool& done)
{
    const Level& L1 = ctx->lv[1];
    TC_TRY(fetch_scalars(ctx, d, S_COUNT));
    ctx->residual_mag[1] = d.h_scal[S_RESID1] / (double)(L1.g.nx - 1) / (double)(L1.g.nz - 1);
    // the reference's per-cycle line (iterative_solver.f90:135), TC_TRACE=1
    if (d.trace) fprintf(stderr, "%5d %20.8g %20.8g\n", q, ctx->residual_mag[1], ctx->res_ref[1]);
    if (ctx->residual_mag[1] > 1e10 * ctx->res_ref[1]) {
        ctx->err = "Multigrid Convergence Problem: Abnormally High Residual Detected -- stopping";
        return TC_ERR_RESID_BLOWUP;
    }
    done = ctx->residual_mag[1] / ctx->res_ref[1] <= ctx->p.tolerance;
    return 0;
}

int guess_pre_solve(tc_ctx* ctx, bool dc, int* slot, bool* nn_guess);
int guess_post_solve(tc_ctx* ctx, int slot, const double* psi);

int finish_solve(tc_ctx* ctx, int status, int q, const double* psi, int32_t* q_out)
{
    TC_TRY(wait_coarse(ctx, DEV(ctx)));
    // LU singularity check (DGBTRF INFO)
    TC_HIP(hipMemcpyAsync(ctx->h_flag, ctx->d_info, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    const int32_t info = *ctx->h_flag;
    if (info != 0 && status == 0) {
        ctx->err = "DGBTRF: singular coarse-grid matrix";
        status = TC_ERR_SINGULAR;
    }
    if (status == 0) {
        TC_TRY(copy_plane(ctx, ctx->SF, psi, ctx->lv[1].g));
        TC_TRY(guess_post_solve(ctx, DEV(ctx).learn_slot, ctx->SF));
    }
    DEV(ctx).learn_slot = -1;
    ctx->stats.q_last = q;
    if (ctx->solve_stage >= 0 && ctx->solve_stage < 4) ctx->stats.q_step[ctx->solve_stage] = q;
    ctx->stats.residual_last = ctx->residual_mag[1];
    ctx->stats.res_ref_last = ctx->res_ref[1];
    if (q_out) *q_out = q;
    return status;
}

int run_multigrid_dc(tc_ctx* ctx, int32_t* q_out);

int run_multigrid(tc_ctx* ctx, int32_t* q_out)
{
    if (ctx->mode != TC_MODE_FP64) return run_multigrid_dc(ctx, q_out);
    tc_dev& d = DEV(ctx);
    const tc_params& P = ctx->p;
    const int G = ctx->grids;
    Level& L1 = ctx->lv[1];
    const double zsg = (P.Vbc == 0) ? -1.0 : 1.0;
    // error(1,:,:) = SF
    L1.cur = 0;
    TC_TRY(copy_plane(ctx, L1.err[0], ctx->SF, L1.g));
    TC_TRY(solve_setup(ctx, d));
    {
        bool hg = false;
        TC_TRY(guess_pre_solve(ctx, false, &d.learn_slot, &hg));
    }
    bool first_sweep = true;
    const int* multi = P.multi;
    int q = 1;
    int status = 0;
    for (;;) {
        bool done = false;
        for (int p = 1; p <= P.nsteps; ++p) {
            const int level = multi[p];
            const int dir = (multi[p + 1] > multi[p]) ? 0 : 1;
            if (multi[p - 1] < multi[p]) {
                // compute_RHS(level) = restrict(residual of level-1)
                Level& Lf = ctx->lv[level - 1];
                launch_restrict(ctx->stream, Lf.res, Lf.g, ctx->lv[level].rhs, ctx->lv[level].g);
                if (P.fixes & TC_FIX_ZERO_COARSE) {
                    Level& Lc = ctx->lv[level];
                    TC_HIP(hipMemsetAsync(Lc.err[Lc.cur], 0, plane_bytes(Lc.g), ctx->stream));
                }
            }
            if (level < G) {
                TC_TRY(iterate_stream(ctx, d, level, P.Nmulti, dir, first_sweep));
            } else {
                Level& Lc = ctx->lv[G];
                TC_TRY(wait_coarse(ctx, d));
                if (d.coarse_lu)
                    launch_coarse_lusolve(ctx->stream, ctx->Mc, ctx->ldab, ctx->ngc, ctx->kl, ctx->ku, ctx->ipiv,
                                          ctx->RowB, Lc, Lc.err[Lc.cur], zsg);
                else
                    launch_coarse_apply(ctx->stream, ctx->Ainv, ctx->RowB, Lc, Lc.err[Lc.cur], ctx->ngc, zsg);
                ctx->stats.coarse_solves++;
            }
            if (level == 1) {
                status = stop_test(ctx, d, q, done);
                if (status || done) break;
            }
            if (multi[p + 1] < multi[p]) {
                Level& Lc = ctx->lv[level];
                Level& Lf = ctx->lv[level - 1];
                launch_prolong(ctx->stream, Lf.err[Lf.cur], Lf.g, Lc.err[Lc.cur], Lc.g, zsg);
            }
        }
        ctx->stats.mg_cycles++;
        if (status) break;
        if (done) break;
        q++;
        if (q > P.max_cycles) {
            ctx->err = "Multigrid Convergence Problem: Too Many Multigrid Cycles Needed";
            status = TC_ERR_MAX_CYCLES;
            break;
        }
    }
    return finish_solve(ctx, status, q, L1.err[L1.cur], q_out);
}

// ---------------------------------------------------------------------------
// learned initial guess: inputs before a solve, targets and history after
// ---------------------------------------------------------------------------
// f (level-1 plane) -> level m by the cycle's restriction (full weighting);
// the intermediate levels use the nn_ra/nn_rb scratch planes, and each
// result's boundary nodes are zero.  m = 1 copies.
int restrict_to(tc_ctx* ctx, const double* f, int m, double* dst)
{
    if (m == 1) return copy_plane(ctx, dst, f, ctx->lv[1].g);
    double* bufs[2] = {ctx->nn_ra, ctx->nn_rb};
    const double* in = f;
    for (int L = 2; L <= m; ++L) {
        double* out = (L == m) ? dst : bufs[L & 1];
        TC_HIP(hipMemsetAsync(out, 0, plane_bytes(ctx->lv[L].g), ctx->stream));
        launch_restrict(ctx->stream, in, ctx->lv[L - 1].g, out, ctx->lv[L].g);
        in = out;
    }
    return 0;
}

// fine += P_2 ... P_m e (e on level m, zero boundary), by the cycle's
// prolongation with the streamfunction ghost mirrors at every level.
int prolong_add(tc_ctx* ctx, const double* e, int m, double* fine, double zsg)
{
    double* bufs[2] = {ctx->nn_ra, ctx->nn_rb};
    const double* in = e;
    for (int L = m - 1; L >= 2; --L) {
        double* out = bufs[L & 1];
        TC_HIP(hipMemsetAsync(out, 0, plane_bytes(ctx->lv[L].g), ctx->stream));
        launch_prolong(ctx->stream, out, ctx->lv[L].g, in, ctx->lv[L + 1].g, zsg);
        in = out;
    }
    launch_prolong(ctx->stream, fine, ctx->lv[1].g, in, ctx->lv[2].g, zsg);
    return 0;
}

// Candidate initial guesses of a solve inside the time step, built before
// it (stage s = solve_stage, with solutions of stage s in the last two
// steps):
//   inc_psi   = SF + (psi_1 - psi'_1): the plain start (the previous RK
//               stage's solution) plus the increment from the previous stage
//               to this one in the last step (psi'_1: the previous stage's
//               solution in that step);
//   ext_psi   = psi_1 + c (psi_1 - psi_2), c = (tau - tau_1) / (tau_1 - tau_2):
//               the stage's linear extrapolation in time, the network's base;
//   guess_psi = ext_psi + the network's predicted correction of it.
// The network works on multigrid level m = nn_level (its H x W are that
// level's nz x nx): the inputs are restricted there, and its correction is
// prolonged to the fine grid by the cycle's own transfer operators, so it
// arrives as smooth as a coarse-grid correction.  Inputs: T, C, log10 eta,
// psi_1/s, (psi_1 - psi_2)/sd, (T - T_1)/dT (psi_i, T_1 of stage s);
// output: (psi - ext)/e, e = max |psi - ext| of the stage's last solve.
// Also records the network's training inputs (the target follows the solve).
// A candidate that lost its last GUESS_BACKOFF judgements is built and
// judged only every GUESS_BACKOFF-th solve until it wins again: a start that
// does not help then costs little (its first visit, the network's forward
// pass), and one that starts helping is picked up within that many solves.
constexpr int GUESS_BACKOFF = 8;
bool guess_due(const tc_ctx* ctx, int c)
{
    return ctx->guess_losses[c] < GUESS_BACKOFF || ctx->guess_solves % GUESS_BACKOFF == 0;
}

int guess_pre_solve(tc_ctx* ctx, bool dc, int* slot, bool* nn_guess)
{
    *slot = -1;
    *nn_guess = false;
    ctx->have_ext = false;
    ctx->have_inc = false;
    const int s = ctx->solve_stage;
    if (s < 0 || !ctx->sh1[s] || ctx->scount[s] < 2) return 0;
    const bool want_ext = dc && ctx->guess >= TC_GUESS_EXTRAP;
    if (!want_ext && !ctx->nn) return 0;
    const Grid& g = ctx->lv[1].g;
    // the stage increment: the composition is rebinned from the tracers before
    // every stage, so psi is not smooth in time, but the change from one
    // stage to the next is much alike from step to step
    const int sp = (s + 3) % 4;
    if (want_ext && ctx->sh1[sp] && ctx->scount[sp] >= 2) {
        launch_add_diff(ctx->stream, ctx->SF, ctx->sh1[s], ctx->sh2[sp], ctx->inc_psi, g.elems);
        ctx->have_inc = true;
    }
    const double den = ctx->tau1[s] - ctx->tau2[s];
    const double c = (den != 0.0) ? (ctx->solve_tau - ctx->tau1[s]) / den : 1.0;
    launch_extrap(ctx->stream, ctx->sh1[s], ctx->sh2[s], c, ctx->ext_psi, g.elems);
    ctx->have_ext = true;
    if (!ctx->nn) return 0;
    UNet& u = ctx->nn->net;
    const int m = ctx->nn_level;
    const Grid& gc = ctx->lv[m].g;
    const size_t np = (size_t)u.npx[0];
    TC_TRY(restrict_to(ctx, ctx->ext_psi, m, ctx->nn_ext));
    const bool record = ctx->rec_n < ctx->rec_cap && ctx->ecount[s] >= 1;
    const bool predict = dc && ctx->guess == TC_GUESS_LEARNED && ctx->ecount[s] >= 1 && guess_due(ctx, 2);
    if (!record && !predict) return 0;
    const double* C = nullptr;
    TC_TRY(restrict_to(ctx, ctx->Tbuoy, m, ctx->nn_T));
    if (ctx->p.comp == 1 && ctx->p.ntype > 0) {
        TC_TRY(restrict_to(ctx, ctx->Cbuoy[0], m, ctx->nn_C));
        C = ctx->nn_C;
    }
    TC_TRY(restrict_to(ctx, ctx->sh1[s], m, ctx->nn_p1));
    TC_TRY(restrict_to(ctx, ctx->sh2[s], m, ctx->nn_p2));
    launch_nn_scales(ctx->stream, ctx->nn_p1, ctx->nn_p2, ctx->nn_T, ctx->Tc1[s], gc, ctx->nn_part, ctx->nn_scales);
    // output scale: the stage's last extrapolation error
    TC_HIP(hipMemcpyAsync(ctx->nn_scales + 1, ctx->stage_err + NN_NSCALE * s + 3, sizeof(double),
                          hipMemcpyDeviceToDevice, ctx->stream));
    if (record) {
        *slot = ctx->rec_n;
        launch_nn_input(ctx->stream, ctx->nn_T, C, ctx->lv[m].vis, ctx->nn_p1, ctx->nn_p2, ctx->Tc1[s], gc, u.W, u.H,
                        ctx->nn_scales, ctx->rec_x + (size_t)*slot * np * 8);
        TC_HIP(hipMemcpyAsync(ctx->rec_scales + NN_NSCALE * *slot, ctx->nn_scales, NN_NSCALE * sizeof(double),
                              hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (predict) {
        launch_nn_input(ctx->stream, ctx->nn_T, C, ctx->lv[m].vis, ctx->nn_p1, ctx->nn_p2, ctx->Tc1[s], gc, u.W, u.H,
                        ctx->nn_scales, u.x0);
        u.forward(ctx->stream);
        const double zsg = (ctx->p.Vbc == 0) ? -1.0 : 1.0;
        if (m == 1) {
            launch_nn_apply(ctx->stream, u.out, ctx->ext_psi, ctx->guess_psi, g, u.W, ctx->nn_scales, zsg);
        } else {
            TC_HIP(hipMemsetAsync(ctx->nn_inc, 0, plane_bytes(gc), ctx->stream));
            launch_nn_apply(ctx->stream, u.out, nullptr, ctx->nn_inc, gc, u.W, ctx->nn_scales, zsg);
            TC_TRY(copy_plane(ctx, ctx->guess_psi, ctx->ext_psi, g));
            TC_TRY(prolong_add(ctx, ctx->nn_inc, m, ctx->guess_psi, zsg));
        }
        *nn_guess = true;
    }
    return 0;
}

// After a solve inside the time step: the training target, the stage's
// extrapolation error (the next output scale) and the history.
int guess_post_solve(tc_ctx* ctx, int slot, const double* psi)
{
    const int s = ctx->solve_stage;
    if (s < 0 || !ctx->sh1[s]) return 0;
    if (ctx->nn) {
        UNet& u = ctx->nn->net;
        const int m = ctx->nn_level;
        const Grid& gc = ctx->lv[m].g;
        if (ctx->have_ext) {
            TC_TRY(restrict_to(ctx, psi, m, ctx->nn_psi));
            if (slot >= 0) {
                const size_t np = (size_t)u.npx[0];
                TC_HIP(hipMemcpyAsync(ctx->nn_scales, ctx->rec_scales + NN_NSCALE * slot, NN_NSCALE * sizeof(double),
                                      hipMemcpyDeviceToDevice, ctx->stream));
                launch_nn_target(ctx->stream, ctx->nn_psi, ctx->nn_ext, gc, u.W, u.H, ctx->nn_scales,
                                 ctx->rec_t + (size_t)slot * np, ctx->rec_mask);
                ctx->rec_n++;
            }
            // [3] = max |R psi - R ext| (the T arguments only fill [2])
            launch_nn_scales(ctx->stream, ctx->nn_psi, ctx->nn_ext, ctx->nn_psi, ctx->nn_psi, gc, ctx->nn_part,
                             ctx->stage_err + NN_NSCALE * s);
            ctx->ecount[s]++;
        }
        TC_TRY(restrict_to(ctx, ctx->Tbuoy, m, ctx->Tc1[s]));
    }
    std::swap(ctx->sh1[s], ctx->sh2[s]);
    TC_TRY(copy_plane(ctx, ctx->sh1[s], psi, ctx->lv[1].g));
    ctx->tau2[s] = ctx->tau1[s];
    ctx->tau1[s] = ctx->solve_tau;
    if (ctx->scount[s] < 2) ctx->scount[s]++;
    return 0;
}

// fine-grid history and candidate planes (TC_GUESS_EXTRAP and up, or a network)
int guess_alloc(tc_ctx* ctx)
{
    if (ctx->ext_psi) return 0;
    const Grid& g = ctx->lv[1].g;
    for (int s = 0; s < 4; ++s) {
        TC_TRY(dalloc(ctx, &ctx->sh1[s], g.elems));
        TC_TRY(dalloc(ctx, &ctx->sh2[s], g.elems));
        ctx->scount[s] = 0;
    }
    TC_TRY(dalloc(ctx, &ctx->ext_psi, g.elems));
    TC_TRY(dalloc(ctx, &ctx->inc_psi, g.elems));
    TC_TRY(dalloc(ctx, &ctx->guess_psi, g.elems));
    TC_TRY(dalloc(ctx, &ctx->guess_lo, g.elems));
    TC_TRY(dalloc(ctx, &ctx->stage_err, 4 * NN_NSCALE));
    return 0;
}

// ---------------------------------------------------------------------------
// defect-correction multigrid (TC_MODE_MIXED, TC_MODE_DC64)
// ---------------------------------------------------------------------------
int falloc(tc_ctx* ctx, float** p, size_t n)
{
    TC_HIP(hipMalloc((void**)p, n * sizeof(float)));
    TC_HIP(hipMemsetAsync(*p, 0, n * sizeof(float), ctx->stream));
    return 0;
}

int dc_alloc(tc_ctx* ctx)
{
    if (ctx->psi_hi) return 0;
    for (int j = 1; j <= ctx->grids; ++j) {
        Level& L = ctx->lv[j];
        float** a[] = {&L.err32[0], &L.err32[1], &L.rhs32, &L.res32, &L.vis32, &L.visx32, &L.visz32, &L.a132,
                       &L.visxz32, &L.dts32, &L.lb32};
        for (float** q : a) TC_TRY(falloc(ctx, q, L.g.elems));
    }
    TC_TRY(dalloc(ctx, &ctx->psi_hi, ctx->g.elems));
    TC_TRY(dalloc(ctx, &ctx->psi_lo, ctx->g.elems));
    TC_TRY(dalloc(ctx, &ctx->coarse_b, (size_t)std::max(ctx->ngc, 1)));
    return 0;
}

// Jacobi-Richardson sweeps of one visit on a correction level (FP32 or
// FP64), followed (level > 1) by compute_residual of the level.
// blocked FP32 sweeps on this level: its visits can start from a zero
// correction without a fill (run_sweeps' zero_start)
bool tb_zero_start(const tc_ctx* ctx, const tc_dev& d, int level, bool f32)
{
    return f32 && d.tb && level >= 2 && level < ctx->grids && ctx->mode != TC_MODE_FP64;
}

int dc_iterate(tc_ctx* ctx, tc_dev& d, int level, int dir, bool f32, bool zero_start = false,
               const RRTarget* rr = nullptr, bool* rr_done = nullptr)
{
    Level& L = ctx->lv[level];
    const tc_params& P = ctx->p;
    // level 1 smooths the correction against the defect of psi
    Level view = L;
    if (level == 1) view.rhs = L.res;
    const float* rhs32 = (level == 1) ? L.res32 : L.rhs32;
    // a per-level count applies to the tuned cycle, not to the Nmulti retry
    int n = ctx->dc_sweeps_active;
    if (ctx->dc_level_sweeps[level] > 0 && ctx->dc_sweeps > 0 && n == ctx->dc_sweeps) n = ctx->dc_level_sweeps[level];
    TC_TRY(run_sweeps(ctx, d, level, n, dir, f32, view, rhs32, nullptr, zero_start, rr, rr_done));
    // FP64: compute_residual now; FP32: fused into the restriction of the
    // next descent (k_residual_restrict32), the only reader of the residual
    if (level > 1 && !f32) {
        launch_residual(ctx->stream, d.pow2[level], L, L.err[L.cur], L.res, d.F, d.h[level], P.ivis, nullptr);
        ctx->stats.residuals_exec[level - 1]++;
    }
    return 0;
}

// psi += e(1); e(1) = 0; residual(1) = RHS - L(psi) in double-double with
// the |residual| block sums of the stop test.
int dc_absorb_residual(tc_ctx* ctx, tc_dev& d, bool f32)
{
    Level& L1 = ctx->lv[1];
    const double zsg = (ctx->p.Vbc == 0) ? -1.0 : 1.0;
    if (d.inc_res) {
        const int nb = launch_absorb_inc(ctx->stream, ctx->psi_hi, ctx->psi_lo, f32 ? L1.err32[L1.cur32] : nullptr,
                                         f32 ? nullptr : L1.err[L1.cur], nullptr, L1, L1.res,
                                         f32 ? L1.res32 : nullptr, d.F, d.h[1], ctx->p.ivis, d.pow2[1], zsg,
                                         d.partialA);
        // e(1) back to zero for the next visit (a fill kernel: capturable)
        if (f32) launch_zero_i32(ctx->stream, reinterpret_cast<int32_t*>(L1.err32[L1.cur32]), L1.g.elems);
        else launch_zero_i32(ctx->stream, reinterpret_cast<int32_t*>(L1.err[L1.cur]), 2 * L1.g.elems);
        launch_reduce_final(ctx->stream, d.partialA, nb, 0, d.d_scal, S_RESID1);
        ctx->stats.residuals_exec[0]++;
        d.res_exact = false;
        return 0;
    }
    launch_dd_absorb(ctx->stream, ctx->psi_hi, ctx->psi_lo, f32 ? L1.err32[L1.cur32] : nullptr,
                     f32 ? nullptr : L1.err[L1.cur], L1.g, zsg);
    int nb = launch_residual_dd(ctx->stream, ctx->psi_hi, ctx->psi_lo, L1, L1.rhs, L1.res, f32 ? L1.res32 : nullptr,
                                d.F, d.h[1], ctx->p.ivis, d.partialA);
    launch_reduce_final(ctx->stream, d.partialA, nb, 0, d.d_scal, S_RESID1);
    ctx->stats.residuals_exec[0]++;
    d.res_exact = true;
    return 0;
}

// GCR over the cycles.  A cycle maps psi_s (residual r_s) to psi'
// (residual r'): it is a preconditioner, and z = psi' - psi_s its
// direction, with A z = r_s - r' from the two residuals.  The new direction
// is orthogonalised (in the operator's image) against the kept ones, and
// psi = psi_s + alpha z minimises |r|_2 over them; when the set is full it
// starts again empty.  The residual of the new psi is r_s - alpha q: the
// directions and their images are FP64 differences of double-double psi and
// of level-1 defects (exact, or kept by FP64 increments -L(e)), so it is
// RHS - L(psi) to FP64 rounding of q and of the increments (an FP32 copy of
// z would carry its rounding times the fourth-order operator's spread of
// eigenvalues).  The next level-1 visit smooths against it; a stop test that
// passes is confirmed on the exact residual, so the tolerance is the
// cycles'.  With no kept direction and alpha = 1 this is the plain cycle.
static_assert(GCR_MAX == 16, "tc_ctx keeps 16 GCR slots");

int gcr_alloc(tc_ctx* ctx)
{
    const Grid& g = ctx->lv[1].g;
    if (!ctx->gcr_hs) {
        TC_TRY(dalloc(ctx, &ctx->gcr_hs, g.elems));
        TC_TRY(dalloc(ctx, &ctx->gcr_ls, g.elems));
        TC_TRY(dalloc(ctx, &ctx->gcr_rs, g.elems));
        TC_TRY(dalloc(ctx, &ctx->gcr_part, (size_t)gcr_blocks(g) * GCR_NDOT));
        TC_TRY(dalloc(ctx, &ctx->gcr_out, GCR_NDOT));
        TC_TRY(dalloc(ctx, &ctx->gcr_qq, GCR_MAX));
        TC_TRY(dalloc(ctx, &ctx->gcr_coef, (sizeof(GcrCoef) + sizeof(double) - 1) / sizeof(double)));
        TC_TRY(dalloc_i(ctx, &ctx->gcr_flag, 1));
        TC_HIP(hipHostMalloc((void**)&ctx->gcr_hflag, sizeof(int32_t), hipHostMallocDefault));
        TC_HIP(hipEventCreateWithFlags(&ctx->gcr_ev, hipEventDisableTiming));
    }
    for (int j = 0; j < ctx->gcr_m; ++j) {
        if (!ctx->gcr_Z[j]) TC_TRY(dalloc(ctx, &ctx->gcr_Z[j], g.elems));
        if (!ctx->gcr_Q[j]) TC_TRY(dalloc(ctx, &ctx->gcr_Q[j], g.elems));
    }
    return 0;
}

// The outcome of the last step (its flag, copied after it): the direction
// kept, kept alone, or no direction kept (the cycle's result became the
// state).  The stop test after that step has synchronised the stream, so
// the event wait normally returns at once.
int gcr_settle(tc_ctx* ctx)
{
    if (!ctx->gcr_pending) return 0;
    TC_HIP(hipEventSynchronize(ctx->gcr_ev));
    ctx->gcr_pending = false;
    const int f = *ctx->gcr_hflag;
    if (f == 0) {
        ctx->gcr_order[ctx->gcr_nb++] = ctx->gcr_new;
        ctx->gcr_updates++;
    } else if (f == 1) {
        ctx->gcr_order[0] = ctx->gcr_new;
        ctx->gcr_nb = 1;
        ctx->gcr_updates++;
        ctx->gcr_restarts++;
    } else {
        ctx->gcr_nb = 0;
    }
    return 0;
}

// One step after a cycle: the dots, the coefficients and the update are
// queued with no host synchronisation.
int gcr_step(tc_ctx* ctx, bool f32)
{
    Level& L1 = ctx->lv[1];
    const Grid& g = L1.g;
    TC_TRY(gcr_settle(ctx));
    if (!ctx->gcr_have) {
        // psi_s, r_s = psi, r: the state the next cycle starts from
        const size_t bytes = plane_bytes(g);
        TC_HIP(hipMemcpyAsync(ctx->gcr_hs, ctx->psi_hi, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        TC_HIP(hipMemcpyAsync(ctx->gcr_ls, ctx->psi_lo, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        TC_HIP(hipMemcpyAsync(ctx->gcr_rs, L1.res, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        ctx->gcr_have = true;
        ctx->gcr_nb = 0;
        return 0;
    }
    if (ctx->gcr_nb >= ctx->gcr_m) ctx->gcr_nb = 0;  // a full set starts again
    const int nb = ctx->gcr_nb;
    int sn = 0;  // a slot no kept direction uses (nb < m of them)
    for (;; ++sn) {
        bool used = false;
        for (int j = 0; j < nb; ++j) used = used || ctx->gcr_order[j] == sn;
        if (!used) break;
    }
    GcrIdx idx{};
    GcrSet Qk{};
    GcrSetW Z{}, Q{};
    for (int j = 0; j < nb; ++j) {
        idx.v[j] = ctx->gcr_order[j];
        Qk.v[j] = ctx->gcr_Q[idx.v[j]];
        Z.v[j] = ctx->gcr_Z[idx.v[j]];
        Q.v[j] = ctx->gcr_Q[idx.v[j]];
    }
    idx.v[nb] = sn;
    Z.v[nb] = ctx->gcr_Z[sn];
    Q.v[nb] = ctx->gcr_Q[sn];
    launch_gcr_dots(ctx->stream, nb, ctx->gcr_rs, L1.res, Qk, g, ctx->gcr_part, ctx->gcr_out);
    GcrCoef* cf = reinterpret_cast<GcrCoef*>(ctx->gcr_coef);
    launch_gcr_coef(ctx->stream, nb, ctx->gcr_out, idx, ctx->gcr_qq, cf, ctx->gcr_flag);
    const double zsg = (ctx->p.Vbc == 0) ? -1.0 : 1.0;
    launch_gcr_update(ctx->stream, nb, Z, Q, cf, ctx->psi_hi, ctx->psi_lo, ctx->gcr_hs, ctx->gcr_ls, ctx->gcr_rs,
                      L1.res, f32 ? L1.res32 : nullptr, g, zsg);
    if (DEV(ctx).inc_res) DEV(ctx).res_exact = false;
    TC_HIP(hipMemcpyAsync(ctx->gcr_hflag, ctx->gcr_flag, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    TC_HIP(hipEventRecord(ctx->gcr_ev, ctx->stream));
    ctx->gcr_new = sn;
    ctx->gcr_pending = true;
    return 0;
}

// The reference's cycle schedule applied to corrections: psi is a
// double-double; level 1 smooths a zero-started correction e(1) against the
// exact defect of psi and folds it into psi after each visit; the level-2
// correction is prolonged straight into psi (and the defect recomputed)
// before the level-1 post-smoothing; coarser levels start every descent
// from a zero correction; the coarsest grid is solved in FP64.  In exact
// arithmetic one cycle is the reference's cycle with zeroed coarse
// corrections; the stop test reads the double-double residual of psi.
int dc_solve(tc_ctx* ctx, int32_t* q_out, int max_cycles)
{
    tc_dev& d = DEV(ctx);
    const tc_params& P = ctx->p;
    const int G = ctx->grids;
    const bool f32 = ctx->mode == TC_MODE_MIXED;
    Level& L1 = ctx->lv[1];
    const double zsg = (P.Vbc == 0) ? -1.0 : 1.0;
    TC_TRY(dc_alloc(ctx));
    TC_TRY(copy_plane(ctx, ctx->psi_hi, ctx->SF, L1.g));
    launch_bc_stream(ctx->stream, ctx->psi_hi, L1.g, P.Vbc);
    TC_HIP(hipMemsetAsync(ctx->psi_lo, 0, plane_bytes(L1.g), ctx->stream));
    TC_TRY(solve_setup(ctx, d));
    if (f32) {
        for (int L = 1; L <= G; ++L) launch_coeffs_to_f32(ctx->stream, ctx->lv[L], L == ctx->ratio_level, P.ivis);
        if (d.tb)
            for (int L = 2; L < G; ++L)
                if (tb_uses_w13(ctx->lv[L])) launch_tb_weights32(ctx->stream, ctx->lv[L], d.F, d.h[L]);
        if (ctx->gs_sweeps > 0) {
            launch_tb_weights32(ctx->stream, L1, d.F, d.h[1]);
            if (!L1.w13) {
                ctx->err = "out of device memory for the level-1 Gauss-Seidel weights";
                return TC_ERR_NOMEM;
            }
        }
    }
    L1.cur = 0;
    L1.cur32 = 0;
    if (f32) TC_HIP(hipMemsetAsync(L1.err32[0], 0, L1.g.elems * sizeof(float), ctx->stream));
    else TC_HIP(hipMemsetAsync(L1.err[0], 0, plane_bytes(L1.g), ctx->stream));
    // initial defect.  With more than one candidate start (the plain warm
    // start, the extrapolation, the extrapolation plus the network's
    // correction), each takes the cycle's first level-1 smoothing and the
    // one with the smallest residual after it continues: the smoother
    // removes the rough part of any start's error at once, what remains is
    // what the cycles must do.
    bool nn_guess = false;
    TC_TRY(guess_pre_solve(ctx, true, &d.learn_slot, &nn_guess));
    auto defect = [&](const double* hi, double* mag) -> int {
        const int nb = launch_residual_dd(ctx->stream, hi, ctx->psi_lo, L1, L1.rhs, L1.res, f32 ? L1.res32 : nullptr,
                                          d.F, d.h[1], P.ivis, d.partialA);
        d.res_exact = true;
        if (mag) {
            launch_reduce_final(ctx->stream, d.partialA, nb, 0, d.d_scal, S_RESID1);
            TC_TRY(fetch_scalars(ctx, d, S_COUNT));
            *mag = d.h_scal[S_RESID1];
        }
        return 0;
    };
    ctx->stats.guess_used = 0;
    for (int c = 0; c < 3; ++c) ctx->stats.guess_res[c] = 0.0;
    const int* multi = P.multi;
    TC_TRY(defect(ctx->psi_hi, nullptr));
    // TC_GUESS_EXTRAP judges the extrapolation, TC_GUESS_LEARNED the
    // extrapolation plus the network's correction (the network's base)
    double** cand[3] = {nullptr,
                        (ctx->have_inc && guess_due(ctx, 1)) ? &ctx->inc_psi : nullptr,
                        nn_guess ? &ctx->guess_psi : nullptr};
    const bool judge = cand[1] || cand[2];
    if (judge) {
        const int dir1 = (multi[2] > multi[1]) ? 0 : 1;
        const double norm = (double)(L1.g.nx - 1) * (double)(L1.g.nz - 1);
        auto first_visit = [&](double* mag) -> int {
            TC_TRY(dc_iterate(ctx, d, 1, dir1, f32));
            TC_TRY(dc_absorb_residual(ctx, d, f32));
            TC_TRY(fetch_scalars(ctx, d, S_COUNT));
            *mag = d.h_scal[S_RESID1];
            return 0;
        };
        double best = 0.0;
        TC_TRY(first_visit(&best));
        ctx->stats.guess_res[0] = best / norm;
        bool stale = false;  // the residual planes belong to a start that lost
        for (int c = 1; c < 3; ++c) {
            if (!cand[c]) continue;
            std::swap(ctx->psi_hi, *cand[c]);
            std::swap(ctx->psi_lo, ctx->guess_lo);
            TC_HIP(hipMemsetAsync(ctx->psi_lo, 0, plane_bytes(L1.g), ctx->stream));
            TC_TRY(defect(ctx->psi_hi, nullptr));
            double mag = 0.0;
            TC_TRY(first_visit(&mag));
            ctx->stats.guess_res[c] = mag / norm;
            if (mag < best) {
                best = mag;
                ctx->stats.guess_used = c;
                stale = false;
            } else {
                std::swap(ctx->psi_hi, *cand[c]);
                std::swap(ctx->psi_lo, ctx->guess_lo);
                stale = true;
            }
        }
        if (stale) {
            double mag = 0.0;
            TC_TRY(defect(ctx->psi_hi, &mag));
        }
        for (int c = 1; c < 3; ++c)
            if (cand[c]) ctx->guess_losses[c] = (ctx->stats.guess_used == c) ? 0 : ctx->guess_losses[c] + 1;
    }
    if (ctx->solve_stage >= 0) {
        ctx->stats.guesses_used[ctx->stats.guess_used]++;
        if (ctx->guess != TC_GUESS_PLAIN) ctx->guess_solves++;
    }
    const bool first_visit_done = judge;
    int q = 1;
    int status = 0;
    ctx->pend_src = nullptr;
    TC_TRY(gcr_settle(ctx));
    ctx->gcr_have = false;
    ctx->gcr_nb = 0;
    // the Nmulti retry of a tuned solve runs the plain cycles
    const bool gcr = ctx->gcr_m > 0 && !ctx->gcr_suspended &&
                     !(ctx->dc_sweeps > 0 && ctx->dc_sweeps_active != ctx->dc_sweeps);
    if (gcr) TC_TRY(gcr_alloc(ctx));
    // the stop test on an incremental residual: a pass is confirmed on the
    // double-double residual of psi (which then replaces it)
    auto stop_test_dc = [&](int qq, bool& done) -> int {
        int st = stop_test(ctx, d, qq, done);
        if (st == 0 && done && !d.res_exact) {
            double mag = 0.0;
            TC_TRY(defect(ctx->psi_hi, &mag));
            d.res_exact = true;
            ctx->stats.residuals_exec[0]++;
            st = stop_test(ctx, d, qq, done);
        }
        return st;
    };
    bool coarse_b_ready = false;  // the coarse right-hand side was formed by the descent
    int rr_level = 0;             // the level whose right-hand side the last visit formed
    // one step p of the cycle schedule: the descent into the level, its
    // visit, and (level 1) the stop test or the ascent out of it
    auto step = [&](int p, bool& done, bool& brk) -> int {
        const int level = multi[p];
        const int dir = (multi[p + 1] > multi[p]) ? 0 : 1;
        Level& L = ctx->lv[level];
        const bool zero_start = multi[p - 1] < multi[p] && tb_zero_start(ctx, d, level, f32);
        if (multi[p - 1] < multi[p]) {
            Level& Lf = ctx->lv[level - 1];
            if (f32) {
                if (level - 1 == 1) {
                    launch_restrict32(ctx->stream, Lf.res32, Lf.g, L.rhs32, L.g);
                } else if (rr_level == level) {
                    // formed by the last launch of the visit above
                    rr_level = 0;
                    if (level == G) coarse_b_ready = true;
                } else if (level == G) {
                    // the coarse solve's right-hand side is formed here, from
                    // the equilibration of the inverse in use (ready after the wait)
                    TC_TRY(wait_coarse(ctx, d));
                    launch_residual_restrict32(ctx->stream, Lf, Lf.err32[Lf.cur32], Lf.rhs32, d.F, d.h[level - 1],
                                               L.rhs32, L.g, ctx->RowB, ctx->coarse_b);
                    ctx->stats.residuals_exec[level - 2]++;
                    coarse_b_ready = true;
                } else {
                    launch_residual_restrict32(ctx->stream, Lf, Lf.err32[Lf.cur32], Lf.rhs32, d.F, d.h[level - 1],
                                               L.rhs32, L.g);
                    ctx->stats.residuals_exec[level - 2]++;
                }
                // a fill kernel, not a memset: this can be recorded in a graph.  The
                // coarse solve writes every interior cell and ghost of its level
                // (the walls stay 0), so it needs none.
                if (!zero_start && level < G)
                    launch_zero_i32(ctx->stream, reinterpret_cast<int32_t*>(L.err32[L.cur32]), L.g.elems);
            } else {
                launch_restrict(ctx->stream, Lf.res, Lf.g, L.rhs, L.g);
                TC_HIP(hipMemsetAsync(L.err[L.cur], 0, plane_bytes(L.g), ctx->stream));
            }
        }
        if (level < G) {
            // a descent follows: the visit's last blocked launch forms the
            // next level's right-hand side (for the coarsest level with the
            // equilibration of the inverse in use, ready after the wait)
            RRTarget rrt;
            const bool want_rr = f32 && d.fuse_rr && level >= 2 && multi[p + 1] == level + 1;
            if (want_rr) {
                Level& Lc = ctx->lv[level + 1];
                rrt.out = Lc.rhs32;
                rrt.go = Lc.g;
                rrt.cus = d.cus;
                if (level + 1 == G) {
                    TC_TRY(wait_coarse(ctx, d));
                    rrt.RowB = ctx->RowB;
                    rrt.bout = ctx->coarse_b;
                }
            }
            bool fused = false;
            TC_TRY(dc_iterate(ctx, d, level, dir, f32, zero_start, want_rr ? &rrt : nullptr, &fused));
            if (fused) {
                rr_level = level + 1;
                ctx->stats.residuals_exec[level - 1]++;
            }
        } else {
            TC_TRY(wait_coarse(ctx, d));
            if (f32) {
                launch_coarse_apply32(ctx->stream, ctx->Ainv32, ctx->RowB, L, L.rhs32, L.err32[L.cur32], ctx->ngc, zsg,
                                      ctx->coarse_b, coarse_b_ready);
                coarse_b_ready = false;
            }
            else launch_coarse_apply(ctx->stream, ctx->Ainv, ctx->RowB, L, L.err[L.cur], ctx->ngc, zsg);
            ctx->stats.coarse_solves++;
        }
        if (level == 1) {
            TC_TRY(dc_absorb_residual(ctx, d, f32));
            status = stop_test_dc(q, done);
            if (status || done) {
                brk = true;
                return 0;
            }
        }
        if (multi[p + 1] < multi[p]) {
            Level& Lf = ctx->lv[level - 1];
            if (level == 2) {
                // fold the coarse-grid correction into psi; level 1 then
                // post-smooths a zero correction against the new defect
                if (d.inc_res) {
                    launch_absorb_inc(ctx->stream, ctx->psi_hi, ctx->psi_lo, f32 ? L.err32[L.cur32] : nullptr,
                                      f32 ? nullptr : L.err[L.cur], &L.g, Lf, Lf.res, f32 ? Lf.res32 : nullptr, d.F,
                                      d.h[1], P.ivis, d.pow2[1], zsg, nullptr);
                    d.res_exact = false;
                } else {
                    launch_prolong_dd(ctx->stream, ctx->psi_hi, ctx->psi_lo, Lf.g, f32 ? L.err32[L.cur32] : nullptr,
                                      f32 ? nullptr : L.err[L.cur], L.g, zsg);
                    launch_residual_dd(ctx->stream, ctx->psi_hi, ctx->psi_lo, Lf, Lf.rhs, Lf.res,
                                       f32 ? Lf.res32 : nullptr, d.F, d.h[1], P.ivis, nullptr);
                }
                ctx->stats.residuals_exec[0]++;
            } else if (f32 && tb_zero_start(ctx, d, level - 1, f32) && multi[p + 1] == level - 1) {
                // folded into the first blocked launch of level-1's next visit
                ctx->pend_src = L.err32[L.cur32];
                ctx->pend_level = level - 1;
                ctx->pend_from = level;
            } else if (f32) {
                launch_prolong32(ctx->stream, Lf.err32[Lf.cur32], Lf.g, L.err32[L.cur32], L.g, zsg);
            } else {
                launch_prolong(ctx->stream, Lf.err[Lf.cur], Lf.g, L.err[L.cur], L.g, zsg);
            }
        }
        return 0;
    };
    // the coarse segment p = 2..nsteps: no level-1 visit and no host
    // synchronisation, the same launches in every cycle of the solve
    bool seg_graph = ctx->graphs && f32 && multi[1] == 1;
    for (int p = 2; p <= P.nsteps && seg_graph; ++p)
        if (multi[p] < 2) seg_graph = false;
    for (auto& sg : ctx->seg) sg.valid = false;
    // (eager, a schedule with a level-1 visit inside it, e.g. the F-cycle,
    // stops at that visit's stop test as the whole cycle would)
    auto segment = [&](bool& done, bool& brk) -> int {
        for (int p = 2; p <= P.nsteps && !brk; ++p) TC_TRY(step(p, done, brk));
        return 0;
    };
    auto run_segment = [&](bool& done, bool& brk) -> int {
        if (!seg_graph) return segment(done, brk);
        int par[TC_MAX_LEVELS + 1] = {};
        for (int l = 2; l <= G; ++l) par[l] = ctx->lv[l].cur32;
        for (auto& sg : ctx->seg) {
            if (!sg.valid || std::memcmp(sg.par0, par, sizeof par) != 0) continue;
            TC_HIP(hipGraphLaunch(sg.exec, ctx->stream));
            for (int l = 2; l <= G; ++l) ctx->lv[l].cur32 = sg.par1[l];
            for (int l = 0; l < TC_MAX_LEVELS; ++l) {
                ctx->stats.sweeps_exec[l] += sg.dsweeps[l];
                ctx->stats.residuals_exec[l] += sg.dres[l];
            }
            ctx->stats.coarse_solves += sg.dcoarse;
            ctx->seg_replays++;
            return 0;
        }
        // capture this parity's segment (the coarse inverse must be ready:
        // a wait on the side stream's event cannot be recorded)
        tc_ctx::SegGraph* sg = !ctx->seg[0].valid ? &ctx->seg[0] : (!ctx->seg[1].valid ? &ctx->seg[1] : nullptr);
        if (!sg) return segment(done, brk);
        TC_TRY(wait_coarse(ctx, d));
        const tc_stats s0 = ctx->stats;
        TC_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int st = segment(done, brk);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
        if (st) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        TC_HIP(ec);
        bool updated = false;
        if (sg->exec) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult res;
            updated = hipGraphExecUpdate(sg->exec, graph, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
            if (!updated) {
                (void)hipGetLastError();
                TC_HIP(hipGraphExecDestroy(sg->exec));
                sg->exec = nullptr;
            }
        }
        if (!updated) {
            const hipError_t ei = hipGraphInstantiate(&sg->exec, graph, nullptr, nullptr, 0);
            if (ei != hipSuccess) {
                (void)hipGraphDestroy(graph);
                sg->exec = nullptr;
                ctx->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(ei);
                return TC_ERR_HIP;
            }
        }
        TC_HIP(hipGraphDestroy(graph));
        TC_HIP(hipGraphLaunch(sg->exec, ctx->stream));
        std::memcpy(sg->par0, par, sizeof par);
        for (int l = 2; l <= G; ++l) sg->par1[l] = ctx->lv[l].cur32;
        for (int l = 0; l < TC_MAX_LEVELS; ++l) {
            sg->dsweeps[l] = ctx->stats.sweeps_exec[l] - s0.sweeps_exec[l];
            sg->dres[l] = ctx->stats.residuals_exec[l] - s0.residuals_exec[l];
        }
        sg->dcoarse = ctx->stats.coarse_solves - s0.coarse_solves;
        sg->valid = true;
        ctx->seg_captures++;
        return 0;
    };
    for (;;) {
        bool done = false, brk = false;
        if (first_visit_done && q == 1) {
            // level 1's first visit ran while judging the guess
            status = stop_test_dc(q, done);
            brk = status || done;
        } else {
            TC_TRY(step(1, done, brk));
        }
        if (!brk) TC_TRY(run_segment(done, brk));
        if (!brk && gcr) TC_TRY(gcr_step(ctx, f32));
        ctx->stats.mg_cycles++;
        if (status || done) break;
        q++;
        if (q > max_cycles) {
            ctx->err = "Multigrid Convergence Problem: Too Many Multigrid Cycles Needed";
            status = TC_ERR_MAX_CYCLES;
            break;
        }
    }
    TC_TRY(gcr_settle(ctx));
    return finish_solve(ctx, status, q, ctx->psi_hi, q_out);
}

// A defect-correction solve with the tuned sweep count, redone with the
// parameters' Nmulti if it diverges or stalls (SF is only written by a
// solve that converged, so the retry starts from the same state).  With
// the parameters' own count and GCR on, the retry runs the plain cycles:
// GCR's recursive residual r_s is never recomputed from psi_s, so its
// rounding drift could stall a solve that the plain cycles finish.
int run_multigrid_dc(tc_ctx* ctx, int32_t* q_out)
{
    const tc_params& P = ctx->p;
    const int fast = ctx->dc_sweeps;
    if (fast <= 0 || fast == P.Nmulti) {
        ctx->dc_sweeps_active = P.Nmulti;
        if (ctx->gcr_m <= 0) return dc_solve(ctx, q_out, P.max_cycles);
        const int st = dc_solve(ctx, q_out, std::min(P.max_cycles, TC_DC_FAST_MAX_CYCLES));
        if (st != TC_ERR_RESID_BLOWUP && st != TC_ERR_MAX_CYCLES) return st;
        ctx->stats.dc_retries++;
        ctx->gcr_suspended = true;
        const int st2 = dc_solve(ctx, q_out, P.max_cycles);
        ctx->gcr_suspended = false;
        return st2;
    }
    ctx->dc_sweeps_active = fast;
    const int st = dc_solve(ctx, q_out, std::min(P.max_cycles, TC_DC_FAST_MAX_CYCLES));
    if (st != TC_ERR_RESID_BLOWUP && st != TC_ERR_MAX_CYCLES) return st;
    ctx->stats.dc_retries++;
    ctx->dc_sweeps_active = P.Nmulti;
    return dc_solve(ctx, q_out, P.max_cycles);
}

// ---------------------------------------------------------------------------
// filters, binning, stages
// ---------------------------------------------------------------------------
double* field_ptr(tc_ctx* ctx, int which)
{
    switch (which) {
    case TC_F_T: return ctx->T; case TC_F_T0: return ctx->T0; case TC_F_T1: return ctx->T1; case TC_F_T2: return ctx->T2;
    default: return nullptr;
    }
}

// Data-dependent filter loop of smoother_time_T / smoother_time on the
// device: iteration m checks the smoothness of the current iterate (stop
// when below delta), else smooths it into the other buffer with factor
// cfac(m) and applies the boundary routine.  Iterations are enqueued in
// batches; the stop flag turns the rest of a batch into no-ops, and the
// host reads the flag (and the stopping index) once per batch.  Returns the
// number of smoothing iterations; the result ends in A.
// the filter loops stop on the smoothness test (smoother.f90); a bound far
// above any count seen (hundreds) turns a loop that cannot stop into an error
constexpr int FILTER_MAX_ITERS = 1 << 20;

template <typename CFac>
int filter_loop(tc_ctx* ctx, tc_dev& d, double* A, double* B, bool comp, double delta, double courant, CFac cfac,
                int* iters)
{
    const tc_params& P = ctx->p;
    // ctl: [0] stop flag, [1] stop index; fslots: the iterations' maxima
    // (three rotating sets; the first iterate's measure goes into set 0)
    launch_zero_i32(ctx->stream, d.ctl, 2);
    launch_zero_i32(ctx->stream, reinterpret_cast<int32_t*>(d.fslots), 2 * filter_slot_words());
    unsigned long long* slots = d.fslots;
    launch_smoothness_max(ctx->stream, A, ctx->g, slots, d.ctl);
    int m = 0, batch = 8;
    for (;;) {
        for (int b = 0; b < batch; ++b, ++m) {
            double* cur = (m % 2 == 0) ? A : B;
            double* nxt = (m % 2 == 0) ? B : A;
            launch_filter_iter(ctx->stream, comp, cur, nxt, ctx->dts, ctx->g, d.F, ctx->lv[1].dr2, ctx->lv[1].ds2,
                               cfac(m), courant, P.Tbot, slots, m, delta, d.ctl);
        }
        // a launch that failed would leave the stop flag unset for ever
        TC_HIP(hipGetLastError());
        TC_HIP(hipMemcpyAsync(d.h_ctl, d.ctl, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        TC_HIP(hipStreamSynchronize(ctx->stream));
        if (d.h_ctl[0]) break;
        if (m >= FILTER_MAX_ITERS) {
            ctx->err = "filter loop: no stop after " + std::to_string(m) + " iterations";
            return TC_ERR_HIP;
        }
        if (batch < 64) batch *= 2;
    }
    const int mt = d.h_ctl[1];
    if (mt % 2 == 1) TC_TRY(copy_plane(ctx, A, B, ctx->g));
    *iters = mt;
    return 0;
}

// The C filter loop with K iterations per launch (launch_filter_multi):
// the same iterates, stop test and stopping index as filter_loop, with the
// buffers A (the first iterate and the result), B and C rotating.
template <typename CFac>
int filter_loop_multi(tc_ctx* ctx, tc_dev& d, double* A, double* B, double* C, double delta, double courant,
                      CFac cfac, int* iters)
{
    const int K = filter_multi_k();
    launch_zero_i32(ctx->stream, d.ctl, 3);
    launch_zero_i32(ctx->stream, reinterpret_cast<int32_t*>(d.fslots), 2 * filter_slot_words());
    launch_smoothness_max(ctx->stream, A, ctx->g, d.fslots, d.ctl);
    double* bufs[3] = {A, B, C};
    std::vector<double> fac(2 * K);
    int l = 0, batch = 4;
    for (;;) {
        for (int b = 0; b < batch; ++b, ++l) {
            const int m = l * K;
            for (int j = 0; j < K; ++j) {
                fac[j] = (m - K + j >= 0) ? cfac(m - K + j) : 0.0;
                fac[K + j] = cfac(m + j);
            }
            launch_filter_multi(ctx->stream, bufs, l % 3, ctx->dts, ctx->g, d.F, ctx->lv[1].dr2, ctx->lv[1].ds2, courant,
                                fac.data(), m, delta, d.fslots, d.ctl);
        }
        TC_HIP(hipGetLastError());
        TC_HIP(hipMemcpyAsync(d.h_ctl, d.ctl, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        TC_HIP(hipStreamSynchronize(ctx->stream));
        if (d.h_ctl[0]) break;
        if (l * K >= FILTER_MAX_ITERS) {
            ctx->err = "filter loop: no stop after " + std::to_string(l * K) + " iterations";
            return TC_ERR_HIP;
        }
        if (batch < 16) batch *= 2;
    }
    const int where = d.h_ctl[2];
    if (where < 0 || where > 2) {
        ctx->err = "filter loop: bad result buffer";
        return TC_ERR_HIP;
    }
    if (where != 0) TC_TRY(copy_plane(ctx, A, bufs[where], ctx->g));
    *iters = d.h_ctl[1];
    return 0;
}

// smoother_time_T (smoother.f90:342-376)
int smoother_time_T(tc_ctx* ctx, tc_dev& d, const double* Tin, int32_t* n_out)
{
    const tc_params& P = ctx->p;
    if (Tin != ctx->Tvis) TC_TRY(copy_plane(ctx, ctx->Tvis, Tin, ctx->g));
    TC_TRY(copy_plane(ctx, ctx->Tbuoy, ctx->Tvis, ctx->g));
    int n = 0;
    if (P.RaT != 0.0) {
        TC_TRY(filter_loop(ctx, d, ctx->Tvis, ctx->tmp2, false, P.delta_T, P.courantT, [](int) { return 1.0; }, &n));
        ctx->stats.t_iters += n;
        ctx->stats.n_temp = n;
    }
    if (n_out) *n_out = n;
    return 0;
}

// smoother_time (smoother.f90:286-340): a Chebyshev pass of Niter
// iterations (lb=-5, la=-1), then single iterations with factor 1 until
// smooth; nn counts the passes after the first.  Niter is 10 on entry and is
// set to 1 (smoother.f90:335) once a type's first pass has run to its end
// without meeting the stop test; it is not reset for the later types.
int smoother_time(tc_ctx* ctx, tc_dev& d, int32_t* n_comp)
{
    const tc_params& P = ctx->p;
    int Niter = 10;
    const double lb = -5.0, la = -1.0;
    for (int t = 0; t < P.ntype; ++t) {
        TC_TRY(copy_plane(ctx, ctx->Cvis[t], ctx->Cnew[t], ctx->g));
        TC_TRY(copy_plane(ctx, ctx->Cbuoy[t], ctx->Cnew[t], ctx->g));
    }
    for (int t = 0; t < P.ntype; ++t) {
        const int N = Niter;
        auto cfac = [N, lb, la](int m) {
            if (m >= N) return 1.0;
            const int ntemp = N - (m + 1) + 1;
            return 2.0 / (-lb - la + (lb - la) * std::cos((double)(2 * ntemp - 1) * PII / (double)(2 * N)));
        };
        int iters = 0;
        if (ctx->filter_blocking)
            TC_TRY(filter_loop_multi(ctx, d, ctx->Cvis[t], ctx->tmp2, d.fbuf, P.delta_C[t], P.courantC, cfac, &iters));
        else
            TC_TRY(filter_loop(ctx, d, ctx->Cvis[t], ctx->tmp2, true, P.delta_C[t], P.courantC, cfac, &iters));
        const int nn = iters < N ? 0 : iters - N + 1;
        if (iters >= N) Niter = 1;
        ctx->stats.c_iters += iters;
        ctx->stats.n_comp[t] = nn;
        if (n_comp) n_comp[t] = nn;
    }
    return 0;
}

int load_conductivity(tc_ctx* ctx, tc_dev& d)
{
    const tc_params& P = ctx->p;
    RaCArr lcf{};
    PtrArr cv{};
    for (int t = 0; t < P.ntype; ++t) {
        lcf.v[t] = (P.comp != 0) ? std::log(1.0 / P.conduct_factor[t]) : 0.0;
        cv.v[t] = ctx->Cvis[t];
    }
    launch_conduct(ctx->stream, cv, ctx->conduct, ctx->conduct_r, ctx->conduct_s, ctx->g, d.F, ctx->lv[1].dr,
                   ctx->lv[1].ds, P.comp, P.ntype, lcf);
    return 0;
}

void tracer_set(tc_ctx* ctx, int which, double** r, double** s)
{
    switch (which) {
    case TC_TR_X0: *r = ctx->rtr0; *s = ctx->str0; break;
    case TC_TR_X1: *r = ctx->rtr1; *s = ctx->str1; break;
    case TC_TR_X2: *r = ctx->rtr2; *s = ctx->str2; break;
    default: *r = ctx->rtr; *s = ctx->str; break;
    }
}

// tracers_to_corners (composition.f90:669-790)
int tracers_to_corners(tc_ctx* ctx, tc_dev& d, int which, int smooth, int32_t* n_comp)
{
    const tc_params& P = ctx->p;
    double *r, *s;
    tracer_set(ctx, which, &r, &s);
    CntArr cnt{};
    for (int t = 0; t <= P.ntype; ++t) {
        cnt.v[t] = ctx->count[t];
        launch_zero_i32(ctx->stream, ctx->count[t], ctx->g.elems);
    }
    launch_zero_i32(ctx->stream, ctx->d_flag, 1);
    launch_bin(ctx->stream, r, s, ctx->ttype, ctx->ntr, cnt, ctx->g, ctx->lv[1].dr, ctx->lv[1].ds, P.aspect, ctx->d_flag);
    PtrArr cn{};
    for (int t = 0; t < P.ntype; ++t) cn.v[t] = ctx->Cnew[t];
    launch_ratio(ctx->stream, cnt, ctx->nempty, cn, ctx->g, P.ntype, P.comp);
    TC_HIP(hipMemcpyAsync(ctx->h_flag, ctx->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    if (*ctx->h_flag) {
        ctx->err = "Tracer advected outside of domain - stopping";
        return TC_ERR_TRACER_DOMAIN;
    }
    if (P.comp == 1) {
        for (int t = 0; t < P.ntype; ++t) launch_bcs_comp(ctx->stream, ctx->Cnew[t], ctx->g);
        if (smooth) TC_TRY(smoother_time(ctx, d, n_comp));
    }
    return load_conductivity(ctx, d);
}

int tracer_stage(tc_ctx* ctx, tc_dev& d, int stage, double dt)
{
    D3Arr Dint{{ctx->p.Dint[0][0], ctx->p.Dint[0][1], ctx->p.Dint[0][2]}};
    launch_derivs(ctx->stream, ctx->u, ctx->w, ctx->DERr_u, ctx->DERr_w, ctx->g, Dint, ctx->lv[1].dr);
    launch_tracer_stage(ctx->stream, stage, ctx->p.torder, dt, ctx->ntr, ctx->rtr, ctx->str, ctx->rtr0, ctx->str0,
                        ctx->rtr1, ctx->str1, ctx->rtr2, ctx->str2, ctx->u, ctx->w, ctx->DERr_u, ctx->DERr_w, ctx->g,
                        Dint, ctx->lv[1].dr, ctx->lv[1].ds);
    (void)d;
    return 0;
}

int energy_stage(tc_ctx* ctx, tc_dev& d, int stage, double dt)
{
    const tc_params& P = ctx->p;
    if (P.RaT == 0.0) return 0;
    RaCArr ht{};
    PtrArr cv{};
    for (int t = 0; t < P.ntype; ++t) { ht.v[t] = P.Htype[t]; cv.v[t] = ctx->Cvis[t]; }
    launch_energy_stage(ctx->stream, stage, dt, ctx->T, ctx->T0, ctx->T1, ctx->T2, ctx->u, ctx->w, ctx->conduct,
                        ctx->conduct_r, ctx->conduct_s, cv, ctx->g, d.F, ctx->lv[1].dr, ctx->lv[1].ds, ctx->lv[1].dr2,
                        ctx->lv[1].ds2, P.comp, P.ntype, P.H, ht);
    double* dst = (stage == 1 || stage == 3) ? ctx->T1 : (stage == 2 ? ctx->T2 : ctx->T);
    launch_bcs_T(ctx->stream, dst, ctx->g, P.Tbot);
    return 0;
}

int stability_RK4(tc_ctx* ctx, tc_dev& d, double* dt_out)
{
    const tc_params& P = ctx->p;
    const Level& L = ctx->lv[1];
    double dt;
    if (P.RaT != 0.0) {
        const double Ts_max = P.Tr_base / L.ds, Tss_max = -P.Trr_base / L.ds2;
        const double Tr_max = P.Tr_base / L.dr, Trr_max = -P.Trr_base / L.dr2;
        int nb = launch_stability(ctx->stream, ctx->u, ctx->w, ctx->conduct, ctx->conduct_r, ctx->conduct_s, ctx->g,
                                  Tr_max, Ts_max, Trr_max, Tss_max, d.partialA, d.partialB);
        launch_reduce_final(ctx->stream, d.partialA, nb, 2, d.d_scal, S_EIGA);
        launch_reduce_final(ctx->stream, d.partialB, nb, 2, d.d_scal, S_EIGD);
        TC_TRY(fetch_scalars(ctx, d, S_COUNT));
        const double eA = d.h_scal[S_EIGA], eD = d.h_scal[S_EIGD];
        double dtA, dtD;
        if (P.torder == 2) { dtA = 1.0 / eA; dtD = 2.0 / eD; }
        else { dtA = 2.8 / eA; dtD = 2.8 / eD; }
        if (eA == 0.0) dt = dtD;
        else dt = 1.0 / (1.0 / dtA + 1.0 / dtD);
    } else {
        dt = 0.5;
    }
    *dt_out = P.courant * dt;
    return 0;
}

int velocities(tc_ctx* ctx, tc_dev& d)
{
    launch_velocities(ctx->stream, ctx->SF, ctx->u, ctx->w, ctx->g, d.F, ctx->lv[1].dr, ctx->lv[1].ds, ctx->p.Vbc);
    return 0;
}

int copy_C(tc_ctx* ctx, double* const* dst, double* const* src)
{
    for (int t = 0; t < ctx->p.ntype; ++t) TC_TRY(copy_plane(ctx, dst[t], src[t], ctx->g));
    return 0;
}

// buffers of equalize (tracer- and node-sized), allocated at the first call
int eq_alloc(tc_ctx* ctx)
{
    if (ctx->eq_red) return 0;
    const Grid& g = ctx->g;
    const int64_t n = std::max<int64_t>(ctx->ntr, 1);
    const size_t V = (size_t)(g.nx + 1) * (size_t)(g.nz + 1);
    TC_HIP(hipMalloc((void**)&ctx->eq_ne, V * sizeof(int32_t)));
    TC_HIP(hipMalloc((void**)&ctx->eq_start, V * sizeof(int32_t)));
    TC_HIP(hipMalloc((void**)&ctx->eq_end, V * sizeof(int32_t)));
    TC_HIP(hipMalloc((void**)&ctx->eq_rows, (size_t)(g.nx + 1) * sizeof(int32_t)));
    TC_HIP(hipMalloc((void**)&ctx->eq_cols, (size_t)(g.nz + 1) * sizeof(int32_t)));
    TC_HIP(hipMalloc((void**)&ctx->eq_bcnt, (size_t)eq_pass_blocks(g) * EQ_MAXT * sizeof(int32_t)));
    TC_HIP(hipMalloc((void**)&ctx->eq_part, (size_t)eq_pass_blocks(g) * sizeof(EqPart)));
    TC_HIP(hipHostMalloc((void**)&ctx->eq_hord, (size_t)(g.nx + g.nz + 2) * sizeof(int32_t), hipHostMallocDefault));
    for (int b = 0; b < 2; ++b) {
        TC_HIP(hipMalloc((void**)&ctx->eq_k[b], (size_t)n * sizeof(uint64_t)));
        TC_HIP(hipMalloc((void**)&ctx->eq_v[b], (size_t)n * sizeof(uint32_t)));
    }
    size_t need = 0;
    TC_TRY(eq_sort(ctx->stream, nullptr, &need, ctx->eq_k[0], ctx->eq_k[1], ctx->eq_v[0], ctx->eq_v[1], n, 64,
                   nullptr));
    TC_HIP(hipMalloc(&ctx->eq_tmp, std::max<size_t>(need, 1)));
    ctx->eq_tmp_bytes = std::max<size_t>(need, 1);
    TC_HIP(hipHostMalloc((void**)&ctx->eq_hred, sizeof(EqRed), hipHostMallocDefault));
    TC_HIP(hipMalloc((void**)&ctx->eq_red, sizeof(EqRed)));
    return 0;
}

// sort scratch of n items on `bits` key bits, grown when too small
int eq_temp(tc_ctx* ctx, int64_t n, int bits)
{
    size_t need = 0;
    TC_TRY(eq_sort(ctx->stream, nullptr, &need, ctx->eq_k[0], ctx->eq_k[1], ctx->eq_v[0], ctx->eq_v[1], n, bits,
                   nullptr));
    if (need > ctx->eq_tmp_bytes) {
        TC_HIP(hipStreamSynchronize(ctx->stream));
        TC_HIP(hipFree(ctx->eq_tmp));
        ctx->eq_tmp = nullptr;
        TC_HIP(hipMalloc(&ctx->eq_tmp, need));
        ctx->eq_tmp_bytes = need;
    }
    return 0;
}

int eq_fetch(tc_ctx* ctx)
{
    TC_HIP(hipGetLastError());
    TC_HIP(hipMemcpyAsync(ctx->eq_hred, ctx->eq_red, sizeof(EqRed), hipMemcpyDeviceToHost, ctx->stream));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

// equalize_tracers (composition.f90:83-580) in its counter-based form: the
// host runs the reference's loop control on per-pass reductions, the
// kernels of tc_equalize.hip do the per-node and per-tracer work
int equalize(tc_ctx* ctx, tc_dev& d, tc_equalize_info* out)
{
    const tc_params& P = ctx->p;
    const Grid& g = ctx->g;
    const hipStream_t st = ctx->stream;
    const int nt = P.comp == 1 ? P.ntype : 0;
    const int64_t n = ctx->ntr;
    tc_equalize_info info;
    std::memset(&info, 0, sizeof info);
    const uint64_t key = eq_stream_key(ctx->eq_seed, ctx->eq_calls++);
    if (n >= ((int64_t)1 << 31)) {
        ctx->err = "equalize supports fewer than 2^31 tracers";
        return TC_ERR_UNSUPPORTED;
    }
    if (std::max(g.nx, g.nz) + 1 > 16384) {
        ctx->err = "equalize supports at most 16384 nodes per row or column";
        return TC_ERR_UNSUPPORTED;
    }
    TC_TRY(eq_alloc(ctx));
    EqRed& h = *ctx->eq_hred;
    CntArr cnt{};
    for (int t = 0; t <= P.ntype; ++t) cnt.v[t] = ctx->count[t];
    PtrArr cn{};
    for (int t = 0; t < P.ntype; ++t) cn.v[t] = ctx->Cnew[t];
    const double dr = ctx->lv[1].dr, ds = ctx->lv[1].ds;
    const double dist_r = 1e-4 * dr, dist_s = 1e-4 * ds;  // dist_factor*dr (main.f90:58, basics.f90:159-160)
    // Nideal (composition.f90:103-111)
    int64_t Nideal[EQ_MAXT] = {};
    if (P.comp == 1) {
        int64_t acc = 0;
        for (int t = 1; t <= nt; ++t) {
            Nideal[t] = std::llround(ctx->eq_Cinit[t - 1] * (double)n);
            acc += Nideal[t];
        }
        Nideal[0] = n - acc;
    } else {
        Nideal[0] = n;
    }
    // tpc_type (composition.f90:114-140) and nexcess (:118-161)
    EqTypes tpc{};
    launch_eq_clear(st, ctx->eq_red);
    if (P.comp == 0) {
        tpc.v[0] = ctx->eq_tpc;
    } else {
        launch_eq_tpc_stats(st, cnt, cn, g, P.ntype, P.comp, ctx->eq_part, ctx->eq_red);
        TC_TRY(eq_fetch(ctx));
        for (int t = 0; t <= nt; ++t) {
            const long long nsel = h.fs[2 * t], sum = h.fs[2 * t + 1];
            tpc.v[t] = nsel ? (int32_t)std::llround((double)sum / (double)nsel) : 0;
        }
    }
    for (int t = 0; t <= nt; ++t) info.tpc_type[t] = tpc.v[t];
    launch_eq_nexcess(st, cnt, ctx->nempty, g, nt, P.comp, tpc, ctx->eq_ne, ctx->eq_part, ctx->eq_red);
    TC_TRY(eq_fetch(ctx));
    const int64_t nfree = h.fs[0];
    info.nfree = nfree;
    if (h.fs[1] == 0) {   // every nexcess is 0 (composition.f90:166-171)
        if (out) *out = info;
        return 0;
    }
    info.did_work = 1;
    // tracers of each node in index order, then the donors
    launch_eq_node_lists(st, ctx->rtr, ctx->str, ctx->tr_id, n, g, dr, ds, P.aspect, ctx->eq_k[0], ctx->eq_v[0],
                         ctx->eq_start, ctx->eq_end, ctx->eq_red);
    TC_TRY(eq_fetch(ctx));
    if (h.bad) {
        ctx->err = "Tracer advected outside of domain - stopping";
        return TC_ERR_TRACER_DOMAIN;
    }
    const uint64_t V = (uint64_t)(g.nx + 1) * (uint64_t)(g.nz + 1);
    int cbits = 1;
    while (cbits < 32 && ((uint64_t)1 << cbits) < V) ++cbits;
    int cur = 0;
    TC_TRY(eq_temp(ctx, n, 32 + cbits));
    size_t bytes = ctx->eq_tmp_bytes;
    TC_TRY(eq_sort(st, ctx->eq_tmp, &bytes, ctx->eq_k[0], ctx->eq_k[1], ctx->eq_v[0], ctx->eq_v[1], n, 32 + cbits,
                   &cur));
    launch_eq_bounds(st, ctx->eq_k[cur], n, ctx->eq_start, ctx->eq_end);
    const int o = 1 - cur;
    launch_eq_select(st, ctx->eq_ne, ctx->eq_start, ctx->eq_end, ctx->eq_v[cur], ctx->rtr, ctx->str, ctx->ttype,
                     ctx->tr_id, g, dist_r, dist_s, P.aspect, key, ctx->eq_k[o], ctx->eq_v[o], ctx->eq_red);
    TC_TRY(eq_fetch(ctx));
    if ((int64_t)h.found != nfree) {   // composition.f90:344-347
        ctx->err = "equalize: donor tracers found inconsistent with nfree (counts out of date?)";
        return TC_ERR_ARG;
    }
    int dsel = 0;
    TC_TRY(eq_temp(ctx, nfree, 44));
    bytes = ctx->eq_tmp_bytes;
    TC_TRY(eq_sort(st, ctx->eq_tmp, &bytes, ctx->eq_k[o], ctx->eq_k[cur], ctx->eq_v[o], ctx->eq_v[cur], nfree, 44,
                   &dsel));
    const uint32_t* donors = dsel == 0 ? ctx->eq_v[o] : ctx->eq_v[cur];
    int64_t nfree_t[EQ_MAXT] = {}, off[EQ_MAXT] = {}, used[EQ_MAXT] = {};
    int64_t acc = 0;
    for (int t = 0; t <= nt; ++t) {
        nfree_t[t] = (int64_t)h.dcount[t];
        off[t] = acc;
        acc += nfree_t[t];
    }
    // lottery (composition.f90:362-463), then the second loop (:466-541)
    std::vector<int32_t> rows(g.nx + 1), cols(g.nz + 1);
    for (int i = 0; i <= g.nx; ++i) rows[i] = i;
    for (int k = 0; k <= g.nz; ++k) cols[k] = k;
    launch_eq_reduce(st, ctx->eq_ne, cnt, cn, g, nt, P.ntype, P.comp, ctx->eq_part, ctx->eq_red);
    TC_TRY(eq_fetch(ctx));
    auto balanced = [&]() {
        for (int t = 0; t <= nt; ++t)
            if (h.fs[t] != Nideal[t]) return false;
        return true;
    };
    bool escape = balanced();
    int loop = 0;
    uint64_t pass = 0;
    for (;;) {
        if (loop == 0 && !escape) escape = balanced();
        EqTypes mx{};
        bool allz = true;
        int64_t su = 0;
        for (int p = 0; p <= nt; ++p) {
            const bool gate = loop == 0 ? tpc.v[p] > 0 : tpc.v[0] > 0;
            mx.v[p] = gate ? h.fm[p] : 0;
            if (mx.v[p] != 0) allz = false;
            su += used[p];
        }
        if (allz || su >= nfree) {
            if (loop == 0) { loop = 1; continue; }
            break;
        }
        // this pass's visit orders (the previous host copy from eq_hord is done: eq_fetch synchronised)
        eq_visit_orders(rows.data(), g.nx + 1, cols.data(), g.nz + 1, key, pass);
        std::memcpy(ctx->eq_hord, rows.data(), rows.size() * sizeof(int32_t));
        std::memcpy(ctx->eq_hord + rows.size(), cols.data(), cols.size() * sizeof(int32_t));
        TC_HIP(hipMemcpyAsync(ctx->eq_rows, ctx->eq_hord, rows.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
        TC_HIP(hipMemcpyAsync(ctx->eq_cols, ctx->eq_hord + rows.size(), cols.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, st));
        pass++;
        int tg = -1;
        if (loop == 0 && !escape)
            for (int p = 0; p <= nt; ++p)
                if (h.fs[p] > Nideal[p]) { tg = p; break; }
        EqOff first{}, avail{};
        for (int t = 0; t <= nt; ++t) {
            first.v[t] = off[t] + used[t];
            avail.v[t] = nfree_t[t] - used[t];
        }
        launch_eq_pass(st, ctx->eq_rows, ctx->eq_cols, ctx->eq_ne, cnt, g, nt, tg, mx, ctx->eq_bcnt, donors, first,
                       avail, ctx->rtr, ctx->str, ctx->ttype, ctx->tr_id, key, dr, ds, dist_r, dist_s, P.aspect, P.comp,
                       loop == 0 ? 1 : 0, ctx->eq_red);
        launch_eq_reduce(st, ctx->eq_ne, cnt, cn, g, nt, P.ntype, P.comp, ctx->eq_part, ctx->eq_red);
        TC_TRY(eq_fetch(ctx));
        int64_t moved = 0;
        for (int t = 0; t <= nt; ++t) {
            const int64_t q = std::min<int64_t>((int64_t)h.qtot[t], avail.v[t]);
            used[t] += q;
            moved += q;
        }
        if (!moved) {
            if (loop == 0) { loop = 1; continue; }
            break;
        }
    }
    info.passes = (int32_t)pass;
    for (int t = 0; t <= nt; ++t) {
        info.minexcess[t] = h.fm[t];
        info.count[t] = h.fs[t];
        info.moved += used[t];
    }
    info.deficit = h.fs[EQ_MAXT];
    if (out) *out = info;
    // tracers_to_corners without the C filter (composition.f90:576-578)
    return tracers_to_corners(ctx, d, TC_TR_X, 0, nullptr);
}

// energy_time (energy.f90:117-453) for tracer=0, comp=1, local_time=0
int energy_time(tc_ctx* ctx, tc_dev& d)
{
    const tc_params& P = ctx->p;
    double dt;
    TC_TRY(stability_RK4(ctx, d, &dt));
    ctx->dt = dt;
    if (ctx->eq_in_step) TC_TRY(equalize(ctx, d, nullptr));
    TC_TRY(copy_C(ctx, ctx->C0, ctx->Cnew));
    TC_TRY(copy_plane(ctx, ctx->T0, ctx->T, ctx->g));
    TC_TRY(energy_stage(ctx, d, 1, dt));
    TC_TRY(tracer_stage(ctx, d, 1, dt));
    if (P.torder >= 2) {
        if (P.RK_update == 1) {
            TC_TRY(tracers_to_corners(ctx, d, TC_TR_X1, 1, nullptr));
            TC_TRY(smoother_time_T(ctx, d, ctx->T1, nullptr));
            ctx->solve_stage = 1;
            ctx->solve_tau = ctx->time + 0.5 * dt;
            TC_TRY(run_multigrid(ctx, nullptr));
            TC_TRY(velocities(ctx, d));
        }
        TC_TRY(energy_stage(ctx, d, 2, dt));
        TC_TRY(tracer_stage(ctx, d, 2, dt));
    }
    if (P.torder >= 3) {
        if (P.RK_update == 1) {
            TC_TRY(copy_C(ctx, ctx->Cnew, ctx->C0));
            TC_TRY(tracers_to_corners(ctx, d, TC_TR_X2, 1, nullptr));
            TC_TRY(smoother_time_T(ctx, d, ctx->T2, nullptr));
            ctx->solve_stage = 2;
            ctx->solve_tau = ctx->time + 0.5 * dt;
            TC_TRY(run_multigrid(ctx, nullptr));
            TC_TRY(velocities(ctx, d));
        }
        TC_TRY(energy_stage(ctx, d, 3, dt));
        TC_TRY(tracer_stage(ctx, d, 3, dt));
    }
    if (P.torder >= 4) {
        if (P.RK_update == 1) {
            TC_TRY(copy_C(ctx, ctx->Cnew, ctx->C0));
            TC_TRY(tracers_to_corners(ctx, d, TC_TR_X1, 1, nullptr));
            TC_TRY(smoother_time_T(ctx, d, ctx->T1, nullptr));
            ctx->solve_stage = 3;
            ctx->solve_tau = ctx->time + dt;
            TC_TRY(run_multigrid(ctx, nullptr));
            TC_TRY(velocities(ctx, d));
        }
        TC_TRY(energy_stage(ctx, d, 4, dt));
        TC_TRY(tracer_stage(ctx, d, 4, dt));
    }
    ctx->solve_stage = -1;
    TC_TRY(copy_C(ctx, ctx->Cnew, ctx->C0));
    TC_TRY(tracers_to_corners(ctx, d, TC_TR_X, 1, nullptr));
    TC_TRY(smoother_time_T(ctx, d, ctx->T, nullptr));
    return 0;
}

// host copies of the grid-dependent time-step tables
void host_dt_vis(const tc_params& p, const Level& L, std::vector<double>& plane)
{
    // stability_vis_smoother (smoother.f90:465-486), uniform grid
    const Grid& g = L.g;
    const int w = g.nx + 2 * SPAN + 1;
    plane.assign((size_t)w * (g.nz + 2 * SPAN + 1), 0.0);
    const double Ts_max = p.Tr_base / L.ds, Tss_max = -p.Trr_base / L.ds2;
    const double Tr_max = p.Tr_base / L.dr, Trr_max = -p.Trr_base / L.dr2;
    for (int i = 0; i <= g.nx; ++i)
        for (int k = 1; k <= g.nz - 1; ++k) {
            double eR = std::fabs(Trr_max / (1.0 * 1.0)) + std::fabs(Tss_max / (1.0 * 1.0));
            double eI = std::fabs(0.0 / std::pow(1.0, 3.0)) * Tr_max + std::fabs(0.0 / std::pow(1.0, 3.0)) * Ts_max;
            plane[(size_t)(k + SPAN) * w + (i + SPAN)] = 2.0 * eR / (eR * eR + eI * eI);
        }
}

void host_dts(const tc_params& p, const Level& L, std::vector<double>& plane)
{
    // stability_RK4_smoother (smoother.f90:423-444), uniform grid
    const Grid& g = L.g;
    const int w = g.nx + 2 * SPAN + 1;
    plane.assign((size_t)w * (g.nz + 2 * SPAN + 1), 0.0);
    const double Ts_max = p.Tr_base / L.ds, Tss_max = -p.Trr_base / L.ds2;
    const double Tr_max = p.Tr_base / L.dr, Trr_max = -p.Trr_base / L.dr2;
    for (int i = 0; i <= g.nx; ++i)
        for (int k = 0; k <= g.nz; ++k) {
            double eR = std::fabs(Trr_max / (1.0 * 1.0)) + std::fabs(Tss_max / (1.0 * 1.0));
            double eI = std::fabs(0.0 / std::pow(1.0, 3.0)) * Tr_max + std::fabs(0.0 / std::pow(1.0, 3.0)) * Ts_max;
            plane[(size_t)(k + SPAN) * w + (i + SPAN)] = 2.0 * eR / (eR * eR + eI * eI);
        }
}

int validate(const tc_params* p, std::string& why)
{
    if (p->order != 2) { why = "order must be 2"; return TC_ERR_UNSUPPORTED; }
    if (p->iterate_order != 1) { why = "iterate_order must be 1"; return TC_ERR_UNSUPPORTED; }
    if (p->grids < 2 || p->grids > TC_MAX_LEVELS) { why = "grids must be 2..6"; return TC_ERR_ARG; }
    if (p->ivis < 0 || p->ivis > 2) { why = "ivis must be 0, 1 or 2 (Newtonian)"; return TC_ERR_UNSUPPORTED; }
    if (p->ntype < 0 || p->ntype > TC_MAX_TYPES || (p->comp == 1 && p->ntype < 1)) { why = "bad ntype"; return TC_ERR_ARG; }
    if (p->torder != 2 && p->torder != 4) { why = "torder must be 2 or 4"; return TC_ERR_ARG; }
    if (p->nsteps < 2 || p->nsteps > TC_MAX_MULTI - 2) { why = "bad nsteps"; return TC_ERR_ARG; }
    if (p->nx < 4 || p->nz < 4) { why = "grid too small"; return TC_ERR_ARG; }
    const int f = 1 << (p->grids - 1);
    if (p->nx % f || p->nz % f || p->nx / f < 2 || p->nz / f < 2) { why = "nx, nz must be divisible by 2^(grids-1)"; return TC_ERR_ARG; }
    for (int q = 1; q <= p->nsteps; ++q)
        if (p->multi[q] < 1 || p->multi[q] > p->grids) { why = "multi() out of range"; return TC_ERR_ARG; }
    if (p->multi[0] != p->multi[p->nsteps] || p->multi[p->nsteps + 1] != p->multi[1]) { why = "multi() must wrap"; return TC_ERR_ARG; }
    if (p->Nmulti < 1 || p->max_cycles < 1) { why = "bad Nmulti/max_cycles"; return TC_ERR_ARG; }
    return 0;
}

void free_ctx(tc_ctx* ctx)
{
    auto F = [](void* p) { if (p) (void)hipFree(p); };
    for (int L = 1; L <= TC_MAX_LEVELS; ++L) {
        Level& l = ctx->lv[L];
        double* a[] = {l.err[0], l.err[1], l.rhs, l.vis, l.visx, l.visz, l.visxx, l.viszz, l.visxz, l.dtstream, l.ratio,
                       l.dtvis, l.visf, l.visxf, l.viszf, l.visxxf, l.viszzf, l.visxzf, l.vtmp, l.res};
        for (double* p : a) F(p);
        float* a32[] = {l.err32[0], l.err32[1], l.rhs32, l.res32, l.vis32, l.visx32, l.visz32, l.a132, l.visxz32,
                        l.dts32, l.lb32, l.w13};
        for (float* p : a32) F(p);
    }
    double* b[] = {ctx->T, ctx->SF, ctx->u, ctx->w, ctx->strain, ctx->Tvis, ctx->Tbuoy, ctx->conduct, ctx->conduct_r,
                   ctx->conduct_s, ctx->dts, ctx->DERr_u, ctx->DERr_w, ctx->T0, ctx->T1, ctx->T2, ctx->tmp, ctx->tmp2,
                   ctx->zg, ctx->Mc, ctx->RowB, ctx->Ainv, ctx->RowB_next, ctx->Ainv_next, ctx->Bc, ctx->Ur, ctx->rtr, ctx->str, ctx->rtr0, ctx->str0, ctx->rtr1,
                   ctx->str1, ctx->rtr2, ctx->str2, ctx->psi_hi, ctx->psi_lo, ctx->ext_psi, ctx->inc_psi, ctx->stage_err,
                   ctx->nn_scales, ctx->rec_scales, ctx->nn_part, ctx->guess_psi, ctx->guess_lo, ctx->nn_T,
                   ctx->nn_C, ctx->nn_p1, ctx->nn_p2, ctx->nn_ext, ctx->nn_psi, ctx->nn_inc, ctx->nn_ra, ctx->nn_rb};
    for (double* p : b) F(p);
    F(ctx->Ainv32); F(ctx->Ainv32_next);
    for (int t = 0; t < TC_MAX_TYPES; ++t) { F(ctx->Cnew[t]); F(ctx->Cvis[t]); F(ctx->Cbuoy[t]); F(ctx->C0[t]); }
    for (int t = 0; t < 4; ++t) { F(ctx->sh1[t]); F(ctx->sh2[t]); F(ctx->Tc1[t]); }
    for (int t = 0; t <= TC_MAX_TYPES; ++t) F(ctx->count[t]);
    F(ctx->nempty); F(ctx->ipiv); F(ctx->d_info); F(ctx->ttype); F(ctx->d_flag); F(ctx->d_info_next);
    F(ctx->tr_id); F(ctx->tr_id_spare); F(ctx->spare_d); F(ctx->spare_i); F(ctx->sort_tmp);
    for (int b = 0; b < 2; ++b) { F(ctx->sort_k[b]); F(ctx->sort_v[b]); }
    F(ctx->rec_x); F(ctx->rec_t); F(ctx->rec_mask);
    F(ctx->eq_ne); F(ctx->eq_start); F(ctx->eq_end); F(ctx->eq_rows); F(ctx->eq_cols); F(ctx->eq_bcnt);
    F(ctx->eq_tmp); F(ctx->eq_red); F(ctx->eq_part); F(ctx->coarse_b);
    F(ctx->gcr_hs); F(ctx->gcr_ls); F(ctx->gcr_rs); F(ctx->gcr_part); F(ctx->gcr_out); F(ctx->gcr_qq);
    F(ctx->gcr_coef); F(ctx->gcr_flag);
    for (int j = 0; j < GCR_MAX; ++j) { F(ctx->gcr_Z[j]); F(ctx->gcr_Q[j]); }
    if (ctx->gcr_hflag) (void)hipHostFree(ctx->gcr_hflag);
    if (ctx->gcr_ev) (void)hipEventDestroy(ctx->gcr_ev);
    if (ctx->eq_hord) (void)hipHostFree(ctx->eq_hord);
    for (int b = 0; b < 2; ++b) { F(ctx->eq_k[b]); F(ctx->eq_v[b]); }
    if (ctx->eq_hred) (void)hipHostFree(ctx->eq_hred);
    if (ctx->snap_stream) (void)hipStreamSynchronize(ctx->snap_stream);
    F(ctx->snap_dev);
    if (ctx->snap_host) (void)hipHostFree(ctx->snap_host);
    if (ctx->ev_snap_d) (void)hipEventDestroy(ctx->ev_snap_d);
    if (ctx->ev_snap_h) (void)hipEventDestroy(ctx->ev_snap_h);
    if (ctx->snap_stream) (void)hipStreamDestroy(ctx->snap_stream);
    if (ctx->h_flag) (void)hipHostFree(ctx->h_flag);
    for (int b = 0; b < 2; ++b) {
        if (ctx->xfer_buf[b]) (void)hipHostFree(ctx->xfer_buf[b]);
        if (ctx->xfer_ev[b]) (void)hipEventDestroy(ctx->xfer_ev[b]);
    }
    tc_dev* d = static_cast<tc_dev*>(ctx->impl);
    if (d) {
        F(d->d_scal); F(d->partialA); F(d->partialB); F(d->ctl); F(d->fslots); F(d->fbuf); F(d->diag); F(d->diag_part); F(d->stats_part); F(d->stats_out);
        F(d->vctl); F(d->vdctl); F(d->vslots);
        if (d->h_scal) (void)hipHostFree(d->h_scal);
        if (d->h_ctl) (void)hipHostFree(d->h_ctl);
        if (d->h_vctl) (void)hipHostFree(d->h_vctl);
        if (d->h_vdctl) (void)hipHostFree(d->h_vdctl);
        if (d->h_vinit) (void)hipHostFree(d->h_vinit);
        delete d;
    }
    for (hipEvent_t e : ctx->prof_ev) (void)hipEventDestroy(e);
    if (ctx->ev_inv_ready) (void)hipEventDestroy(ctx->ev_inv_ready);
    if (ctx->ev_inv_next) (void)hipEventDestroy(ctx->ev_inv_next);
    if (ctx->ev_visc_ready) (void)hipEventDestroy(ctx->ev_visc_ready);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
}

int init_ctx(tc_ctx* ctx, const tc_params* p, int64_t ntr)
{
    ctx->p = *p;
    ctx->ntr = ntr;
    ctx->grids = p->grids;
    tc_dev* d = new (std::nothrow) tc_dev();
    if (!d) return TC_ERR_NOMEM;
    ctx->impl = d;
    {
        const char* e = getenv("TC_TRACE");
        d->trace = e && e[0] == '1';
        const char* c = getenv("TC_COARSE_LU");
        d->coarse_lu = c && c[0] == '1';
        const char* t = getenv("TC_TB");
        d->tb = !(t && t[0] == '0');
        const char* r = getenv("TC_FUSED_RR");
        d->fuse_rr = !(r && r[0] == '0');
        const char* ir = getenv("TC_INC_RESID");
        d->inc_res = !(ir && ir[0] == '0');
    }
    {
        int dev = 0, cus = 0;
        TC_HIP(hipGetDevice(&dev));
        TC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        if (cus > 0) d->cus = cus;
    }
    TC_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    TC_HIP(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    TC_HIP(hipEventCreateWithFlags(&ctx->ev_inv_ready, hipEventDisableTiming));
    TC_HIP(hipEventCreateWithFlags(&ctx->ev_inv_next, hipEventDisableTiming));
    TC_HIP(hipEventCreateWithFlags(&ctx->ev_visc_ready, hipEventDisableTiming));
    // FD tables (order 2): Drs(i,k)=D1(k)*D1(i) ... (main.f90:448-459)
    for (int j = 0; j < 5; ++j) { d->F.D1[j] = p->D1[j]; d->F.D2[j] = p->D2[j]; d->F.D3[j] = p->D3[j]; d->F.D4[j] = p->D4[j]; }
    for (int dk = -1; dk <= 1; ++dk)
        for (int di = -1; di <= 1; ++di) {
            const int t = (dk + 1) * 3 + (di + 1);
            d->F.Drs[t] = p->D1[dk + 2] * p->D1[di + 2];
            d->F.Drss[t] = p->D2[dk + 2] * p->D1[di + 2];
            d->F.Drrs[t] = p->D1[dk + 2] * p->D2[di + 2];
            d->F.Drrss[t] = p->D2[dk + 2] * p->D2[di + 2];
        }
    // level hierarchy (basics.f90:167-182)
    for (int j = 1; j <= ctx->grids; ++j) {
        Level& L = ctx->lv[j];
        const int nxl = p->nx / (1 << (j - 1)), nzl = p->nz / (1 << (j - 1));
        L.g = make_grid(nxl, nzl);
        L.dr = p->aspect / (double)nxl;
        L.ds = 1.0 / (double)nzl;
        L.dr2 = L.dr * L.dr; L.ds2 = L.ds * L.ds;
        L.dr3 = std::pow(L.dr, 3.0); L.ds3 = std::pow(L.ds, 3.0);
        L.dr4 = std::pow(L.dr, 4.0); L.ds4 = std::pow(L.ds, 4.0);
        L.drds = L.dr * L.ds; L.drds2 = L.dr * L.ds2; L.dr2ds = L.dr2 * L.ds; L.dr2ds2 = L.dr2 * L.ds2;
        d->h[j] = make_spacing(L);
        d->pow2[j] = level_pow2(L);
        const size_t n = L.g.elems;
        double** a[] = {&L.err[0], &L.err[1], &L.rhs, &L.vis, &L.visx, &L.visz, &L.visxx, &L.viszz, &L.visxz,
                        &L.dtstream, &L.ratio, &L.dtvis, &L.visf, &L.visxf, &L.viszf, &L.visxxf, &L.viszzf, &L.visxzf,
                        &L.vtmp, &L.res};
        for (double** q : a) TC_TRY(dalloc(ctx, q, n));
    }
    ctx->g = ctx->lv[1].g;
    const size_t n1 = ctx->g.elems;
    double** b[] = {&ctx->T, &ctx->SF, &ctx->u, &ctx->w, &ctx->strain, &ctx->Tvis, &ctx->Tbuoy, &ctx->conduct,
                    &ctx->conduct_r, &ctx->conduct_s, &ctx->dts, &ctx->DERr_u, &ctx->DERr_w, &ctx->T0, &ctx->T1,
                    &ctx->T2, &ctx->tmp, &ctx->tmp2};
    for (double** q : b) TC_TRY(dalloc(ctx, q, n1));
    for (int t = 0; t < p->ntype; ++t) {
        TC_TRY(dalloc(ctx, &ctx->Cnew[t], n1)); TC_TRY(dalloc(ctx, &ctx->Cvis[t], n1));
        TC_TRY(dalloc(ctx, &ctx->Cbuoy[t], n1)); TC_TRY(dalloc(ctx, &ctx->C0[t], n1));
    }
    for (int t = 0; t <= p->ntype; ++t) TC_TRY(dalloc_i(ctx, &ctx->count[t], n1));
    TC_TRY(dalloc_i(ctx, &ctx->nempty, n1));
    // zg(-3:nz+3), uniform grid with the mirrored ghosts of build_grid (grid.f90:97-108)
    {
        std::vector<double> zg(p->nz + 7);
        const double ds = 1.0 / (double)p->nz;
        for (int k = 0; k <= p->nz; ++k) zg[k + 3] = ds * (double)k;
        for (int k = 1; k <= 3; ++k) { zg[-k + 3] = -zg[k + 3]; zg[p->nz + k + 3] = 1.0 - zg[p->nz - k + 3] + 1.0; }
        zg[3] = 0.0;
        zg[p->nz + 3] = 1.0;
        TC_TRY(dalloc(ctx, &ctx->zg, zg.size()));
        TC_HIP(hipMemcpyAsync(ctx->zg, zg.data(), zg.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        TC_HIP(hipStreamSynchronize(ctx->stream));
    }
    // coarse matrix bandwidth (Compute_Matrix_coarse analysis=1)
    {
        const int nxc = ctx->lv[ctx->grids].g.nx, nzc = ctx->lv[ctx->grids].g.nz, N = nzc + 1;
        ctx->ngc = (nxc + 1) * (nzc + 1);
        const int limit[5] = {0, 1, 2, 1, 0};
        int ku = 0, kl = 0;
        for (int pp = 1; pp <= ctx->ngc; ++pp) {
            int kp = pp - N * ((pp - 1) / N) - 1, ip = (pp - 1) / N;
            for (int j = -2; j <= 2; ++j) {
                int iq = j + ip;
                if (iq > 0 && iq < nxc)
                    for (int kq = -limit[j + 2] + kp; kq <= kp + limit[j + 2]; ++kq)
                        if (kq > 0 && kq < nzc) {
                            int t = (N * iq + kq + 1) - pp;
                            ku = std::max(ku, t);
                            kl = std::max(kl, -t);
                        }
            }
        }
        ctx->kl = kl; ctx->ku = ku; ctx->ldab = 2 * kl + ku + 1;
        if (8 * (size_t)(kl + ku + 1) * sizeof(double) > 65536 || (size_t)ctx->ngc * ctx->ngc > ((size_t)1 << 31)) {
            ctx->err = "coarsest grid too large for the explicit coarse inverse (use more levels)";
            return TC_ERR_UNSUPPORTED;
        }
        TC_TRY(dalloc(ctx, &ctx->Mc, (size_t)ctx->ldab * ctx->ngc));
        TC_TRY(dalloc(ctx, &ctx->RowB, ctx->ngc));
        TC_TRY(dalloc(ctx, &ctx->Ainv, (size_t)ctx->ngc * ctx->ngc));
        TC_TRY(falloc(ctx, &ctx->Ainv32, (size_t)ctx->ngc * ctx->ngc));
        TC_TRY(dalloc(ctx, &ctx->Bc, ctx->ngc));
        TC_TRY(dalloc(ctx, &ctx->Ur, (size_t)ctx->ngc * (size_t)(ctx->kl + ctx->ku + 1)));
        TC_TRY(dalloc_i(ctx, &ctx->ipiv, ctx->ngc));
        TC_TRY(dalloc_i(ctx, &ctx->d_info, 1));
    }
    // tracers
    if (ntr > 0) {
        double** c[] = {&ctx->rtr, &ctx->str, &ctx->rtr0, &ctx->str0, &ctx->rtr1, &ctx->str1, &ctx->rtr2, &ctx->str2};
        for (double** q : c) TC_TRY(dalloc(ctx, q, (size_t)ntr));
        TC_TRY(dalloc_i(ctx, &ctx->ttype, (size_t)ntr));
    }
    TC_TRY(dalloc_i(ctx, &ctx->d_flag, 1));
    TC_HIP(hipHostMalloc((void**)&ctx->h_flag, sizeof(int32_t), hipHostMallocDefault));
    // reduction buffers
    {
        const Grid& g = ctx->g;
        int nb = ((g.nx + 2 * SPAN + 1 + 63) / 64) * ((g.nz + 2 * SPAN + 1 + 3) / 4) + 16;
        d->max_partial = nb;
        TC_TRY(dalloc(ctx, &d->partialA, nb));
        TC_TRY(dalloc(ctx, &d->partialB, nb));
        TC_TRY(dalloc(ctx, &d->d_scal, 16));
        TC_TRY(dalloc_i(ctx, &d->ctl, 4));
        TC_HIP(hipMalloc((void**)&d->fslots, filter_slot_words() * sizeof(unsigned long long)));
        TC_TRY(dalloc(ctx, &d->fbuf, g.elems));
        TC_HIP(hipHostMalloc((void**)&d->h_ctl, 4 * sizeof(int32_t), hipHostMallocDefault));
        TC_TRY(dalloc_i(ctx, &d->vctl, 2 * VC_COUNT));
        TC_TRY(dalloc(ctx, &d->vdctl, 2 * VD_COUNT));
        TC_HIP(hipMalloc((void**)&d->vslots, 2 * VIS_SLOTS * VIS_SLOT_STRIDE * sizeof(unsigned long long)));
        TC_HIP(hipHostMalloc((void**)&d->h_vctl, VC_COUNT * sizeof(int32_t), hipHostMallocDefault));
        TC_HIP(hipHostMalloc((void**)&d->h_vdctl, VD_COUNT * sizeof(double), hipHostMallocDefault));
        TC_HIP(hipHostMalloc((void**)&d->h_vinit, VC_COUNT * sizeof(int32_t), hipHostMallocDefault));
        TC_HIP(hipHostMalloc((void**)&d->h_scal, 16 * sizeof(double), hipHostMallocDefault));
    }
    // grid-dependent tables: dt_vis per level, dts_array
    {
        std::vector<double> plane;
        for (int j = 1; j <= ctx->grids; ++j) {
            host_dt_vis(*p, ctx->lv[j], plane);
            TC_TRY(put_plane_direct(ctx, ctx->lv[j].dtvis, ctx->lv[j].g, plane.data()));
            TC_HIP(hipStreamSynchronize(ctx->stream));
        }
        host_dts(*p, ctx->lv[1], plane);
        TC_TRY(put_plane_direct(ctx, ctx->dts, ctx->g, plane.data()));
        TC_HIP(hipStreamSynchronize(ctx->stream));
    }
    TC_HIP(hipDeviceSynchronize());
    return 0;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

int tc_init(const tc_params* p, int64_t ntr, tc_ctx** out)
{
    if (!p || !out || ntr < 0) return TC_ERR_ARG;
    *out = nullptr;
    std::string why;
    int v = validate(p, why);
    if (v) {
        fprintf(stderr, "tc_init: %s\n", why.c_str());
        return v;
    }
    tc_ctx* ctx = new (std::nothrow) tc_ctx();
    if (!ctx) return TC_ERR_NOMEM;
    int st = init_ctx(ctx, p, ntr);
    if (st) {
        fprintf(stderr, "tc_init: %s\n", ctx->err.c_str());
        free_ctx(ctx);
        delete ctx;
        return st;
    }
    *out = ctx;
    return 0;
}

int tc_finalize(tc_ctx* ctx)
{
    if (!ctx) return TC_ERR_ARG;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->side);
    for (auto& sg : ctx->seg)
        if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
    free_ctx(ctx);
    delete ctx;
    return 0;
}

const char* tc_last_error(const tc_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int tc_set_mode(tc_ctx* ctx, int mode)
{
    if (!ctx) return TC_ERR_ARG;
    if (mode != TC_MODE_FP64 && mode != TC_MODE_MIXED && mode != TC_MODE_DC64) return TC_ERR_ARG;
    ctx->mode = mode;
    return 0;
}

int tc_set_graphs(tc_ctx* ctx, int32_t enable)
{
    if (!ctx) return TC_ERR_ARG;
    ctx->graphs = enable != 0;
    return 0;
}

int tc_set_gcr(tc_ctx* ctx, int32_t directions)
{
    if (!ctx) return TC_ERR_ARG;
    if (directions < 0 || directions > GCR_MAX) {
        ctx->err = "tc_set_gcr: directions must be 0.." + std::to_string(GCR_MAX);
        return TC_ERR_ARG;
    }
    ctx->gcr_m = directions;
    return 0;
}

int tc_set_coarse_refresh(tc_ctx* ctx, int32_t solves)
{
    if (!ctx) return TC_ERR_ARG;
    if (solves < 1) {
        ctx->err = "tc_set_coarse_refresh: solves must be >= 1";
        return TC_ERR_ARG;
    }
    ctx->coarse_refresh = solves;
    ctx->coarse_since = 0;
    return 0;
}

int tc_set_dc_cheb_interval(tc_ctx* ctx, double lb, double la)
{
    if (!ctx) return TC_ERR_ARG;
    if (!(lb < 0.0 && la < 0.0 && lb <= la && std::isfinite(lb))) {
        ctx->err = "tc_set_dc_cheb_interval: need lb <= la < 0";
        return TC_ERR_ARG;
    }
    ctx->cheb1_lb = lb;
    ctx->cheb1_la = la;
    return 0;
}

int tc_set_gs_smoother(tc_ctx* ctx, int32_t sweeps, int32_t per_launch, double omega)
{
    if (!ctx) return TC_ERR_ARG;
    if (sweeps < 0 || per_launch < 1 || per_launch > 8 || !(omega > 0.0 && omega < 2.0)) {
        ctx->err = "tc_set_gs_smoother: sweeps >= 0, per_launch in 1..8, 0 < omega < 2";
        return TC_ERR_ARG;
    }
    ctx->gs_sweeps = sweeps;
    ctx->gs_per_launch = per_launch;
    ctx->gs_omega = omega;
    return 0;
}

int tc_set_filter_blocking(tc_ctx* ctx, int32_t enable)
{
    if (!ctx) return TC_ERR_ARG;
    ctx->filter_blocking = enable != 0;
    return 0;
}

int tc_gcr_counts(const tc_ctx* ctx, int64_t* updates, int64_t* restarts)
{
    if (!ctx || !updates || !restarts) return TC_ERR_ARG;
    *updates = ctx->gcr_updates;
    *restarts = ctx->gcr_restarts;
    return 0;
}

int tc_graph_counts(const tc_ctx* ctx, int64_t* captures, int64_t* replays)
{
    if (!ctx || !captures || !replays) return TC_ERR_ARG;
    *captures = ctx->seg_captures;
    *replays = ctx->seg_replays;
    return 0;
}

int tc_set_dc_level_sweeps(tc_ctx* ctx, int32_t level, int32_t sweeps)
{
    if (!ctx || level < 1 || level > ctx->p.grids || sweeps < 0) return TC_ERR_ARG;
    ctx->dc_level_sweeps[level] = sweeps;
    return 0;
}

int tc_set_dc_sweeps(tc_ctx* ctx, int32_t sweeps)
{
    if (!ctx || sweeps < 0 || sweeps > 10000) return TC_ERR_ARG;
    ctx->dc_sweeps = sweeps;
    return 0;
}

int tc_set_tb_tile(tc_ctx* ctx, int32_t level, int32_t tile)
{
    if (!ctx || level < 2 || level >= ctx->grids || (tile != 0 && tile != 16 && tile != 32 && tile != 48))
        return TC_ERR_ARG;
    ctx->lv[level].tb_tile = tile;
    return 0;
}

int tc_autotune_tb_tiles(tc_ctx* ctx, int32_t steps_each, int32_t* tiles)
{
    if (!ctx || steps_each < 1) return TC_ERR_ARG;
    auto step_ms = [&](double* ms) -> int {
        TC_HIP(hipStreamSynchronize(ctx->stream));
        const auto t0 = std::chrono::steady_clock::now();
        TC_TRY(tc_time_step(ctx, steps_each));
        *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps_each;
        return 0;
    };
    for (int L = 2; L < ctx->grids; ++L) {
        Level& lv = ctx->lv[L];
        const int keep = lv.tb_tile;
        // 16, 32 and 48 (where it still gives 128 workgroups); a candidate
        // replaces the kept tile only when it is 2% faster
        const int cand[3] = {16, 32, 48};
        double best_ms = 0.0;
        int best = keep;
        for (int c = 0; c < 3; ++c) {
            const dim3 t = dim3((lv.g.nx - 1 + cand[c] - 1) / cand[c], (lv.g.nz - 1 + cand[c] - 1) / cand[c]);
            if (cand[c] == 48 && t.x * t.y < 128) continue;
            double ms = 0.0;
            lv.tb_tile = cand[c];
            TC_TRY(step_ms(&ms));
            if (c == 0 || ms < 0.98 * best_ms) { best_ms = ms; best = cand[c]; }
        }
        lv.tb_tile = best;
    }
    if (tiles)
        for (int L = 0; L < TC_MAX_LEVELS; ++L)
            tiles[L] = (L + 1 >= 2 && L + 1 < ctx->grids) ? ctx->lv[L + 1].tb_tile : 0;
    return 0;
}

int tc_autotune_dc_sweeps(tc_ctx* ctx, const int32_t* candidates, int32_t n, int32_t steps_each, int32_t* best,
                          double* ms_per_step)
{
    if (!ctx || !candidates || n < 1 || steps_each < 1 || !best) return TC_ERR_ARG;
    for (int j = 0; j < n; ++j)
        if (candidates[j] < 1 || candidates[j] > 10000) return TC_ERR_ARG;
    std::vector<double> ms(n);
    std::vector<bool> ok(n);
    for (int j = 0; j < n; ++j) {
        ctx->dc_sweeps = candidates[j];
        const int64_t r0 = ctx->stats.dc_retries;
        TC_HIP(hipStreamSynchronize(ctx->stream));
        const auto t0 = std::chrono::steady_clock::now();
        TC_TRY(tc_time_step(ctx, steps_each));
        ms[j] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps_each;
        ok[j] = ctx->stats.dc_retries == r0;
        if (ms_per_step) ms_per_step[j] = ms[j];
    }
    // the largest count within 3% of the fastest: a margin from the counts
    // at which the Chebyshev-weighted sweeps stop converging
    double fastest = 0.0;
    for (int j = 0; j < n; ++j)
        if (ok[j] && (fastest == 0.0 || ms[j] < fastest)) fastest = ms[j];
    int pick = -1;
    for (int j = 0; j < n; ++j)
        if (ok[j] && ms[j] <= 1.03 * fastest && (pick < 0 || candidates[j] > candidates[pick])) pick = j;
    ctx->dc_sweeps = pick >= 0 ? candidates[pick] : 0;
    *best = ctx->dc_sweeps;
    return 0;
}

int tc_autotune_dc_level_sweeps(tc_ctx* ctx, int32_t first_level, int32_t last_level, const int32_t* candidates,
                                int32_t n, int32_t steps_each, int32_t* best, double* ms_per_step)
{
    if (!ctx || !candidates || n < 1 || steps_each < 1 || !best || first_level < 1 || last_level < first_level ||
        last_level >= ctx->p.grids)
        return TC_ERR_ARG;
    for (int j = 0; j < n; ++j)
        if (candidates[j] < 0 || candidates[j] > 10000) return TC_ERR_ARG;
    auto apply = [&](int c) {
        for (int L = first_level; L <= last_level; ++L) ctx->dc_level_sweeps[L] = c;
    };
    std::vector<double> ms(n);
    std::vector<bool> ok(n);
    for (int j = 0; j < n; ++j) {
        apply(candidates[j]);
        const int64_t r0 = ctx->stats.dc_retries;
        TC_HIP(hipStreamSynchronize(ctx->stream));
        const auto t0 = std::chrono::steady_clock::now();
        TC_TRY(tc_time_step(ctx, steps_each));
        ms[j] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps_each;
        ok[j] = ctx->stats.dc_retries == r0;
        if (ms_per_step) ms_per_step[j] = ms[j];
    }
    // as for the common count: the largest count within 3% of the fastest
    // (0 stands for the common count)
    auto eff = [&](int j) { return candidates[j] > 0 ? candidates[j] : ctx->dc_sweeps; };
    double fastest = 0.0;
    for (int j = 0; j < n; ++j)
        if (ok[j] && (fastest == 0.0 || ms[j] < fastest)) fastest = ms[j];
    int pick = -1;
    for (int j = 0; j < n; ++j)
        if (ok[j] && ms[j] <= 1.03 * fastest && (pick < 0 || eff(j) > eff(pick))) pick = j;
    *best = pick >= 0 ? candidates[pick] : 0;
    apply(*best);
    return 0;
}

int tc_get_stats(const tc_ctx* ctx, tc_stats* out)
{
    if (!ctx || !out) return TC_ERR_ARG;
    *out = ctx->stats;
    out->dt_last = ctx->dt;
    out->time = ctx->time;
    return 0;
}

int tc_sync(tc_ctx* ctx)
{
    if (!ctx) return TC_ERR_ARG;
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_put_field(tc_ctx* ctx, int field, const void* host)
{
    if (!ctx || !host) return TC_ERR_ARG;
    const int nt = ctx->p.ntype;
    const Grid& g = ctx->g;
    const size_t w = (size_t)(g.nx + 2 * SPAN + 1), h = (size_t)(g.nz + 2 * SPAN + 1);
    double* dptr = nullptr;
    switch (field) {
    case TC_F_T: dptr = ctx->T; break; case TC_F_SF: dptr = ctx->SF; break; case TC_F_TVIS: dptr = ctx->Tvis; break;
    case TC_F_TBUOY: dptr = ctx->Tbuoy; break; case TC_F_U: dptr = ctx->u; break; case TC_F_W: dptr = ctx->w; break;
    case TC_F_STRAIN: dptr = ctx->strain; break; case TC_F_VIS: dptr = ctx->lv[1].vis; break;
    case TC_F_CONDUCT: dptr = ctx->conduct; break; case TC_F_T0: dptr = ctx->T0; break; case TC_F_T1: dptr = ctx->T1; break;
    case TC_F_T2: dptr = ctx->T2; break;
    default: break;
    }
    if (dptr) {
        TC_TRY(put_plane(ctx, dptr, g, (const double*)host));
        TC_HIP(hipStreamSynchronize(ctx->stream));
        return 0;
    }
    if (field == TC_F_CNEW || field == TC_F_CVIS || field == TC_F_CBUOY) {
        if (nt < 1) return TC_ERR_ARG;
        std::vector<double> tmp(w * h);
        const double* src = (const double*)host;
        for (int t = 0; t < nt; ++t) {
            for (size_t e = 0; e < w * h; ++e) tmp[e] = src[e * nt + t];
            double* dst = field == TC_F_CNEW ? ctx->Cnew[t] : (field == TC_F_CVIS ? ctx->Cvis[t] : ctx->Cbuoy[t]);
            TC_TRY(put_plane(ctx, dst, g, tmp.data()));
            TC_HIP(hipStreamSynchronize(ctx->stream));
        }
        return 0;
    }
    if (field == TC_F_NEMPTY) {
        TC_TRY(put_iplane(ctx, ctx->nempty, g, (const int32_t*)host));
        TC_HIP(hipStreamSynchronize(ctx->stream));
        return 0;
    }
    if (field == TC_F_COUNT_TYPE) {
        std::vector<int32_t> tmp(w * h);
        const int32_t* src = (const int32_t*)host;
        for (int t = 0; t <= nt; ++t) {
            for (size_t e = 0; e < w * h; ++e) tmp[e] = src[e * (nt + 1) + t];
            TC_TRY(put_iplane(ctx, ctx->count[t], g, tmp.data()));
            TC_HIP(hipStreamSynchronize(ctx->stream));
        }
        return 0;
    }
    return TC_ERR_ARG;
}

int tc_get_field(tc_ctx* ctx, int field, void* host)
{
    if (!ctx || !host) return TC_ERR_ARG;
    const int nt = ctx->p.ntype;
    const Grid& g = ctx->g;
    const size_t w = (size_t)(g.nx + 2 * SPAN + 1), h = (size_t)(g.nz + 2 * SPAN + 1);
    const double* dptr = nullptr;
    switch (field) {
    case TC_F_T: dptr = ctx->T; break; case TC_F_SF: dptr = ctx->SF; break; case TC_F_TVIS: dptr = ctx->Tvis; break;
    case TC_F_TBUOY: dptr = ctx->Tbuoy; break; case TC_F_U: dptr = ctx->u; break; case TC_F_W: dptr = ctx->w; break;
    case TC_F_STRAIN: dptr = ctx->strain; break; case TC_F_VIS: dptr = ctx->lv[1].vis; break;
    case TC_F_CONDUCT: dptr = ctx->conduct; break; case TC_F_T0: dptr = ctx->T0; break; case TC_F_T1: dptr = ctx->T1; break;
    case TC_F_T2: dptr = ctx->T2; break;
    case TC_F_RHS: dptr = ctx->lv[1].rhs; break;
    case TC_F_SF_LO:
        if (!ctx->psi_lo) { ctx->err = "no defect-correction solve has run"; return TC_ERR_ARG; }
        dptr = ctx->psi_lo;
        break;
    default: break;
    }
    if (dptr) {
        TC_TRY(get_plane(ctx, dptr, g, (double*)host));
        TC_HIP(hipStreamSynchronize(ctx->stream));
        return 0;
    }
    if (field == TC_F_CNEW || field == TC_F_CVIS || field == TC_F_CBUOY) {
        if (nt < 1) return TC_ERR_ARG;
        std::vector<double> tmp(w * h);
        double* dst = (double*)host;
        for (int t = 0; t < nt; ++t) {
            const double* src = field == TC_F_CNEW ? ctx->Cnew[t] : (field == TC_F_CVIS ? ctx->Cvis[t] : ctx->Cbuoy[t]);
            TC_TRY(get_plane(ctx, src, g, tmp.data()));
            TC_HIP(hipStreamSynchronize(ctx->stream));
            for (size_t e = 0; e < w * h; ++e) dst[e * nt + t] = tmp[e];
        }
        return 0;
    }
    if (field == TC_F_NEMPTY) {
        TC_TRY(get_iplane(ctx, ctx->nempty, g, (int32_t*)host));
        TC_HIP(hipStreamSynchronize(ctx->stream));
        return 0;
    }
    if (field == TC_F_COUNT_TYPE) {
        std::vector<int32_t> tmp(w * h);
        int32_t* dst = (int32_t*)host;
        for (int t = 0; t <= nt; ++t) {
            TC_TRY(get_iplane(ctx, ctx->count[t], g, tmp.data()));
            TC_HIP(hipStreamSynchronize(ctx->stream));
            for (size_t e = 0; e < w * h; ++e) dst[e * (nt + 1) + t] = tmp[e];
        }
        return 0;
    }
    return TC_ERR_ARG;
}

int sort_tracers(tc_ctx* ctx)
{
    const int64_t n = ctx->ntr;
    if (n <= 0) return 0;
    if (n >= ((int64_t)1 << 31)) {
        ctx->err = "tracer sort supports fewer than 2^31 tracers";
        return TC_ERR_UNSUPPORTED;
    }
    const tc_params& P = ctx->p;
    if (!ctx->spare_d) {
        for (int b = 0; b < 2; ++b) {
            TC_HIP(hipMalloc((void**)&ctx->sort_k[b], (size_t)n * sizeof(uint32_t)));
            TC_HIP(hipMalloc((void**)&ctx->sort_v[b], (size_t)n * sizeof(uint32_t)));
        }
        TC_HIP(hipMalloc((void**)&ctx->spare_d, (size_t)n * sizeof(double)));
        TC_HIP(hipMalloc((void**)&ctx->spare_i, (size_t)n * sizeof(int32_t)));
        TC_HIP(hipMalloc((void**)&ctx->tr_id_spare, (size_t)n * sizeof(int64_t)));
    }
    const Grid& g = ctx->g;
    const uint32_t bad = (uint32_t)(g.nx + 1) * (uint32_t)(g.nz + 1);
    int bits = 1;
    while (bits < 32 && ((uint64_t)1 << bits) <= bad) ++bits;
    const size_t need = radix_scratch_words(n) * sizeof(uint32_t);
    if (need > ctx->sort_tmp_bytes) {
        if (ctx->sort_tmp) TC_HIP(hipFree(ctx->sort_tmp));
        TC_HIP(hipMalloc(&ctx->sort_tmp, need));
        ctx->sort_tmp_bytes = need;
    }
    launch_tracer_keys(ctx->stream, ctx->rtr, ctx->str, n, g.nx, ctx->lv[1].dr, ctx->lv[1].ds, P.aspect, bad,
                       ctx->sort_k[0], ctx->sort_v[0]);
    int cur = 0;
    const int e = radix_sort_pairs<uint32_t>(ctx->stream, ctx->sort_k[0], ctx->sort_k[1], ctx->sort_v[0], ctx->sort_v[1],
                                             true, n, bits, static_cast<uint32_t*>(ctx->sort_tmp),
                                             ctx->sort_tmp_bytes / sizeof(uint32_t), &cur);
    if (e) {
        ctx->err = "tracer sort: radix sort launch failed";
        return e;
    }
    const uint32_t* perm = ctx->sort_v[cur];
    double** arrays[] = {&ctx->rtr, &ctx->str, &ctx->rtr0, &ctx->str0, &ctx->rtr1, &ctx->str1, &ctx->rtr2, &ctx->str2};
    for (double** a : arrays) {
        launch_gather_d(ctx->stream, *a, perm, ctx->spare_d, n);
        std::swap(*a, ctx->spare_d);
    }
    launch_gather_i32(ctx->stream, ctx->ttype, perm, ctx->spare_i, n);
    std::swap(ctx->ttype, ctx->spare_i);
    launch_gather_id(ctx->stream, ctx->tr_id, perm, ctx->tr_id_spare, n);
    if (!ctx->tr_id) TC_HIP(hipMalloc((void**)&ctx->tr_id, (size_t)n * sizeof(int64_t)));
    std::swap(ctx->tr_id, ctx->tr_id_spare);
    ctx->sorts++;
    return 0;
}

int tc_sort_tracers(tc_ctx* ctx)
{
    if (!ctx) return TC_ERR_ARG;
    TC_TRY(sort_tracers(ctx));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_tracer_slots(tc_ctx* ctx, int64_t* caller_index)
{
    if (!ctx || (!caller_index && ctx->ntr > 0)) return TC_ERR_ARG;
    if (ctx->ntr == 0) return 0;
    if (ctx->tr_id) {
        TC_HIP(hipMemcpyAsync(caller_index, ctx->tr_id, (size_t)ctx->ntr * sizeof(int64_t), hipMemcpyDeviceToHost,
                              ctx->stream));
        TC_HIP(hipStreamSynchronize(ctx->stream));
    } else {
        for (int64_t j = 0; j < ctx->ntr; ++j) caller_index[j] = j;
    }
    return 0;
}

int tc_set_tracer_sort(tc_ctx* ctx, int32_t interval)
{
    if (!ctx || interval < 0) return TC_ERR_ARG;
    ctx->sort_interval = interval;
    return 0;
}

int tc_put_tracers(tc_ctx* ctx, const double* rtr, const double* str, const int32_t* ttype)
{
    if (!ctx || !rtr || !str || !ttype) return TC_ERR_ARG;
    if (ctx->ntr == 0) return 0;
    for (int64_t i = 0; i < ctx->ntr; ++i)
        if (ttype[i] < 0 || ttype[i] > ctx->p.ntype) {
            ctx->err = "tracer type outside 0..ntype";
            return TC_ERR_ARG;
        }
    const size_t nb = (size_t)ctx->ntr * sizeof(double);
    TC_HIP(hipMemcpyAsync(ctx->rtr, rtr, nb, hipMemcpyHostToDevice, ctx->stream));
    TC_HIP(hipMemcpyAsync(ctx->str, str, nb, hipMemcpyHostToDevice, ctx->stream));
    TC_HIP(hipMemcpyAsync(ctx->ttype, ttype, (size_t)ctx->ntr * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    // the caller's order again
    if (ctx->tr_id) {
        TC_HIP(hipFree(ctx->tr_id));
        ctx->tr_id = nullptr;
    }
    return 0;
}

int tc_get_tracers(tc_ctx* ctx, int which, double* r, double* s)
{
    if (!ctx || !r || !s) return TC_ERR_ARG;
    if (ctx->ntr == 0) return 0;
    double *dr, *ds;
    tracer_set(ctx, which, &dr, &ds);
    const size_t nb = (size_t)ctx->ntr * sizeof(double);
    if (ctx->tr_id) {
        // node-sorted on the device: back to the caller's order
        launch_scatter_d(ctx->stream, dr, ctx->tr_id, ctx->spare_d, ctx->ntr);
        TC_HIP(hipMemcpyAsync(r, ctx->spare_d, nb, hipMemcpyDeviceToHost, ctx->stream));
        launch_scatter_d(ctx->stream, ds, ctx->tr_id, ctx->spare_d, ctx->ntr);
        TC_HIP(hipMemcpyAsync(s, ctx->spare_d, nb, hipMemcpyDeviceToHost, ctx->stream));
    } else {
        TC_HIP(hipMemcpyAsync(r, dr, nb, hipMemcpyDeviceToHost, ctx->stream));
        TC_HIP(hipMemcpyAsync(s, ds, nb, hipMemcpyDeviceToHost, ctx->stream));
    }
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_get_tracer_types(tc_ctx* ctx, int32_t* ttype)
{
    if (!ctx || !ttype) return TC_ERR_ARG;
    if (ctx->ntr == 0) return 0;
    const size_t nb = (size_t)ctx->ntr * sizeof(int32_t);
    if (ctx->tr_id) {
        launch_scatter_i32(ctx->stream, ctx->ttype, ctx->tr_id, ctx->spare_i, ctx->ntr);
        TC_HIP(hipMemcpyAsync(ttype, ctx->spare_i, nb, hipMemcpyDeviceToHost, ctx->stream));
    } else {
        TC_HIP(hipMemcpyAsync(ttype, ctx->ttype, nb, hipMemcpyDeviceToHost, ctx->stream));
    }
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_set_time(tc_ctx* ctx, double time, double dt)
{
    if (!ctx) return TC_ERR_ARG;
    ctx->time = time;
    ctx->dt = dt;
    return 0;
}

int tc_multigrid(tc_ctx* ctx, int32_t* q_out, double* resid_out, double* resref_out)
{
    if (!ctx) return TC_ERR_ARG;
    int st = run_multigrid(ctx, q_out);
    if (resid_out) *resid_out = ctx->residual_mag[1];
    if (resref_out) *resref_out = ctx->res_ref[1];
    if (st == 0) TC_HIP(hipStreamSynchronize(ctx->stream));
    return st;
}

int tc_compute_velocities(tc_ctx* ctx)
{
    if (!ctx) return TC_ERR_ARG;
    TC_TRY(velocities(ctx, DEV(ctx)));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_diagnostics(tc_ctx* ctx, double out[9])
{
    if (!ctx || !out) return TC_ERR_ARG;
    tc_dev& d = DEV(ctx);
    const tc_params& P = ctx->p;
    const Grid& g = ctx->g;
    // scratch: 3 partial sums per block, the atomics and the 6 results
    const int nb = diag_blocks(g);
    if (!d.diag_part) TC_HIP(hipMalloc((void**)&d.diag_part, (size_t)3 * nb * sizeof(double)));
    double* part = d.diag_part;
    struct Mem {
        int32_t ints[4];
        unsigned long long smax;
        double res[6];
    };
    static_assert(sizeof(Mem) <= 128, "diagnostics scratch");
    if (!d.diag) TC_HIP(hipMalloc(&d.diag, sizeof(Mem)));
    Mem init{};
    init.ints[1] = 0x7fffffff;
    init.ints[2] = -1;
    TC_HIP(hipMemcpyAsync(d.diag, &init, sizeof(Mem), hipMemcpyHostToDevice, ctx->stream));
    Mem* m = static_cast<Mem*>(d.diag);
    launch_strain(ctx->stream, ctx->u, ctx->w, ctx->strain, g, d.F, ctx->lv[1].dr, ctx->lv[1].ds);
    const double* C = (P.comp == 1 && P.ntype > 0) ? ctx->Cnew[0] : nullptr;
    launch_diagnostics(ctx->stream, ctx->u, ctx->w, ctx->T, C, ctx->nempty, ctx->strain, ctx->conduct, g, d.F,
                       ctx->lv[1].dr, ctx->lv[1].ds, part, m->ints, &m->smax, m->res);
    Mem h{};
    TC_HIP(hipMemcpyAsync(&h, d.diag, sizeof(Mem), hipMemcpyDeviceToHost, ctx->stream));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    const double a = P.aspect;
    out[0] = std::sqrt(h.res[0] / a);                         // vrms
    out[1] = h.res[3] / h.res[4];                             // Nu
    out[2] = h.res[1] / a;                                    // Tavg
    out[3] = C ? h.res[2] / a : 0.0;                          // Cavg
    out[4] = h.ints[0];                                       // empty nodes
    out[5] = h.ints[1];                                       // min tracers per node
    out[6] = h.ints[2];                                       // max tracers per node
    out[7] = h.res[5] / a;                                    // Nu at the bottom
    double smax;
    std::memcpy(&smax, &h.smax, sizeof(double));
    out[8] = smax;                                            // max strain at the surface
    return 0;
}

int tc_stats_record(tc_ctx* ctx, int32_t kent, const double* mass, tc_stats_row* out)
{
    if (!ctx || !out) return TC_ERR_ARG;
    const tc_params& P = ctx->p;
    const Grid& g = ctx->g;
    if (kent < 0 || kent > g.nz) {
        ctx->err = "tc_stats_record: kent outside 0..nz";
        return TC_ERR_ARG;
    }
    double d9[9];
    TC_TRY(tc_diagnostics(ctx, d9));
    tc_dev& d = DEV(ctx);
    const int np = stats_np();
    const int nb = diag_blocks(g);
    if (!d.stats_part) TC_HIP(hipMalloc((void**)&d.stats_part, (size_t)np * nb * sizeof(double)));
    if (!d.stats_out) TC_HIP(hipMalloc((void**)&d.stats_out, (size_t)(np + 5) * sizeof(double)));
    PtrArr cn{};
    const int nC = P.comp == 1 ? P.ntype : 0;
    for (int t = 0; t < nC; ++t) cn.v[t] = ctx->Cnew[t];
    launch_stats_rest(ctx->stream, ctx->u, ctx->w, ctx->T, ctx->conduct, ctx->lv[1].vis, cn, nC, kent, P.RaT, g, d.F,
                      ctx->lv[1].dr, ctx->lv[1].ds, d.stats_part, d.stats_out);
    std::vector<double> h((size_t)np + 5);
    TC_HIP(hipMemcpyAsync(h.data(), d.stats_out, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    std::memset(out, 0, sizeof(*out));
    const double a = P.aspect;
    out->time = ctx->time;
    out->vrms = d9[0]; out->Nu = d9[1]; out->Tavg = d9[2]; out->Nu_bot = d9[7]; out->strain_max = d9[8];
    out->empty_cells = (int32_t)d9[4]; out->tracer_min = (int32_t)d9[5]; out->tracer_max = (int32_t)d9[6];
    out->kent = kent;
    out->RaT_avg = h[0] / a;
    out->mobility = std::sqrt(h[np] / a) / d9[0];
    for (int j = 0; j < 4; ++j) out->q[j] = h[np + 1 + j];
    for (int t = 0; t < nC; ++t) {
        out->Cavg[t] = h[1 + t] / a;
        const double m = mass ? mass[t] : out->Cavg[t];
        out->ent[t] = h[1 + TC_MAX_TYPES + t] / a / m;
    }
    return 0;
}

// snapshot staging slots: T, SF, u, w, strain, vis, Tvis, nempty (int32 in a
// double-sized slot), Cnew(1..ntype), Cvis(1..ntype)
enum { SNAP_T = 0, SNAP_SF, SNAP_U, SNAP_W, SNAP_STRAIN, SNAP_VIS, SNAP_TVIS, SNAP_NEMPTY, SNAP_C };

int tc_snapshot_stage(tc_ctx* ctx)
{
    if (!ctx) return TC_ERR_ARG;
    const Grid& g = ctx->g;
    const int nt = ctx->p.comp == 1 ? ctx->p.ntype : 0;
    const size_t slots = SNAP_C + 2 * (size_t)nt, pb = plane_bytes(g);
    if (!ctx->snap_dev) {
        TC_HIP(hipMalloc((void**)&ctx->snap_dev, slots * pb));
        TC_HIP(hipHostMalloc((void**)&ctx->snap_host, slots * pb, hipHostMallocDefault));
        TC_HIP(hipStreamCreateWithFlags(&ctx->snap_stream, hipStreamNonBlocking));
        TC_HIP(hipEventCreateWithFlags(&ctx->ev_snap_d, hipEventDisableTiming));
        TC_HIP(hipEventCreateWithFlags(&ctx->ev_snap_h, hipEventDisableTiming));
    }
    // the previous stage's host copy must be done before the staging planes are reused
    if (ctx->snap_staged) TC_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_snap_h, 0));
    tc_dev& d = DEV(ctx);
    // strain of the snapshot: compute_strain_rate_invariant of stats (basics.f90:245)
    launch_strain(ctx->stream, ctx->u, ctx->w, ctx->strain, g, d.F, ctx->lv[1].dr, ctx->lv[1].ds);
    const void* src[SNAP_C] = {ctx->T, ctx->SF, ctx->u, ctx->w, ctx->strain, ctx->lv[1].vis, ctx->Tvis, ctx->nempty};
    CopyList L{};
    for (int j = 0; j < SNAP_C; ++j) {
        L.src[L.n] = src[j];
        L.bytes[L.n] = j == SNAP_NEMPTY ? g.elems * sizeof(int32_t) : pb;
        L.dst_off[L.n] = (size_t)j * pb;
        L.n++;
    }
    for (int t = 0; t < nt; ++t) {
        L.src[L.n] = ctx->Cnew[t];
        L.bytes[L.n] = pb;
        L.dst_off[L.n] = (SNAP_C + (size_t)t) * pb;
        L.n++;
        L.src[L.n] = ctx->Cvis[t];
        L.bytes[L.n] = pb;
        L.dst_off[L.n] = (SNAP_C + (size_t)nt + t) * pb;
        L.n++;
    }
    static_assert(SNAP_C + 2 * TC_MAX_TYPES <= COPY_MAX, "snapshot planes");
    launch_copy_list(ctx->stream, L, ctx->snap_dev);
    TC_HIP(hipEventRecord(ctx->ev_snap_d, ctx->stream));
    TC_HIP(hipStreamWaitEvent(ctx->snap_stream, ctx->ev_snap_d, 0));
    TC_HIP(hipMemcpyAsync(ctx->snap_host, ctx->snap_dev, slots * pb, hipMemcpyDeviceToHost, ctx->snap_stream));
    TC_HIP(hipEventRecord(ctx->ev_snap_h, ctx->snap_stream));
    ctx->snap_staged = true;
    return 0;
}

int tc_snapshot_get(tc_ctx* ctx, int field, void* host)
{
    if (!ctx || !host) return TC_ERR_ARG;
    if (!ctx->snap_staged) {
        ctx->err = "tc_snapshot_get: nothing staged (call tc_snapshot_stage)";
        return TC_ERR_ARG;
    }
    const Grid& g = ctx->g;
    const int nt = ctx->p.comp == 1 ? ctx->p.ntype : 0;
    int slot = -1;
    switch (field) {
    case TC_F_T: slot = SNAP_T; break; case TC_F_SF: slot = SNAP_SF; break; case TC_F_U: slot = SNAP_U; break;
    case TC_F_W: slot = SNAP_W; break; case TC_F_STRAIN: slot = SNAP_STRAIN; break;
    case TC_F_VIS: slot = SNAP_VIS; break; case TC_F_TVIS: slot = SNAP_TVIS; break;
    case TC_F_NEMPTY: slot = SNAP_NEMPTY; break;
    case TC_F_CNEW: case TC_F_CVIS: slot = nt > 0 ? SNAP_C : -1; break;
    default: break;
    }
    if (slot < 0) {
        ctx->err = "tc_snapshot_get: field not in a snapshot";
        return TC_ERR_ARG;
    }
    TC_HIP(hipEventSynchronize(ctx->ev_snap_h));
    const size_t w = (size_t)(g.nx + 2 * SPAN + 1), h = (size_t)(g.nz + 2 * SPAN + 1);
    const size_t first = g.at(-SPAN, -SPAN);
    if (field == TC_F_NEMPTY) {
        const int32_t* src = reinterpret_cast<const int32_t*>(ctx->snap_host + (size_t)slot * g.elems);
        int32_t* dst = (int32_t*)host;
        for (size_t k = 0; k < h; ++k) std::memcpy(dst + k * w, src + first + k * g.ld, w * sizeof(int32_t));
        return 0;
    }
    if (field == TC_F_CNEW || field == TC_F_CVIS) {
        double* dst = (double*)host;
        const int base = SNAP_C + (field == TC_F_CVIS ? nt : 0);
        for (int t = 0; t < nt; ++t) {
            const double* src = ctx->snap_host + (size_t)(base + t) * g.elems;
            for (size_t k = 0; k < h; ++k)
                for (size_t i = 0; i < w; ++i) dst[(k * w + i) * nt + t] = src[first + k * g.ld + i];
        }
        return 0;
    }
    const double* src = ctx->snap_host + (size_t)slot * g.elems;
    double* dst = (double*)host;
    for (size_t k = 0; k < h; ++k) std::memcpy(dst + k * w, src + first + k * g.ld, w * sizeof(double));
    return 0;
}

int tc_set_equalize(tc_ctx* ctx, int32_t tpc, const double* Cinit, uint64_t seed, int32_t in_time_step)
{
    if (!ctx) return TC_ERR_ARG;
    if (ctx->p.comp == 1 && !Cinit) {
        ctx->err = "tc_set_equalize: Cinit is required with comp=1";
        return TC_ERR_ARG;
    }
    if (ctx->p.comp == 0 && tpc <= 0) {
        ctx->err = "tc_set_equalize: tpc must be positive with comp=0";
        return TC_ERR_ARG;
    }
    ctx->eq_tpc = tpc;
    for (int t = 0; t < TC_MAX_TYPES; ++t) ctx->eq_Cinit[t] = (Cinit && t < ctx->p.ntype) ? Cinit[t] : 0.0;
    ctx->eq_seed = seed;
    ctx->eq_calls = 0;
    ctx->eq_on = true;
    ctx->eq_in_step = in_time_step != 0;
    return 0;
}

int tc_equalize_tracers(tc_ctx* ctx, tc_equalize_info* info)
{
    if (!ctx) return TC_ERR_ARG;
    if (!ctx->eq_on) {
        ctx->err = "tc_equalize_tracers: call tc_set_equalize first";
        return TC_ERR_ARG;
    }
    return equalize(ctx, DEV(ctx), info);
}

int tc_compute_strain_rate_invariant(tc_ctx* ctx)
{
    if (!ctx) return TC_ERR_ARG;
    tc_dev& d = DEV(ctx);
    launch_strain(ctx->stream, ctx->u, ctx->w, ctx->strain, ctx->g, d.F, ctx->lv[1].dr, ctx->lv[1].ds);
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_smoother_time_T(tc_ctx* ctx, int src_field, int32_t* n_temp)
{
    if (!ctx) return TC_ERR_ARG;
    double* src = field_ptr(ctx, src_field);
    if (!src) return TC_ERR_ARG;
    TC_TRY(smoother_time_T(ctx, DEV(ctx), src, n_temp));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_tracers_to_corners(tc_ctx* ctx, int which, int smooth, int32_t* n_comp)
{
    if (!ctx) return TC_ERR_ARG;
    if (which < TC_TR_X || which > TC_TR_X2) return TC_ERR_ARG;
    TC_TRY(tracers_to_corners(ctx, DEV(ctx), which, smooth, n_comp));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_smoother_time(tc_ctx* ctx, int32_t* n_comp)
{
    if (!ctx) return TC_ERR_ARG;
    TC_TRY(smoother_time(ctx, DEV(ctx), n_comp));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_load_conductivity(tc_ctx* ctx)
{
    if (!ctx) return TC_ERR_ARG;
    TC_TRY(load_conductivity(ctx, DEV(ctx)));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_tracer_stage(tc_ctx* ctx, int stage, double dt)
{
    if (!ctx || stage < 1 || stage > 4) return TC_ERR_ARG;
    TC_TRY(tracer_stage(ctx, DEV(ctx), stage, dt));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_stability_RK4(tc_ctx* ctx, double* dt_out)
{
    if (!ctx || !dt_out) return TC_ERR_ARG;
    return stability_RK4(ctx, DEV(ctx), dt_out);
}

int tc_energy_stage(tc_ctx* ctx, int stage, double dt)
{
    if (!ctx || stage < 1 || stage > 4) return TC_ERR_ARG;
    TC_TRY(energy_stage(ctx, DEV(ctx), stage, dt));
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_time_step(tc_ctx* ctx, int nsteps)
{
    if (!ctx || nsteps < 0) return TC_ERR_ARG;
    tc_dev& d = DEV(ctx);
    for (int s = 0; s < nsteps; ++s) {
        if (ctx->sort_interval > 0 && ctx->tstep % ctx->sort_interval == 0) TC_TRY(sort_tracers(ctx));
        for (int j = 0; j < 4; ++j) ctx->stats.q_step[j] = 0;
        ctx->solve_stage = 0;
        ctx->solve_tau = ctx->time;
        TC_TRY(run_multigrid(ctx, nullptr));
        TC_TRY(velocities(ctx, d));
        launch_strain(ctx->stream, ctx->u, ctx->w, ctx->strain, ctx->g, d.F, ctx->lv[1].dr, ctx->lv[1].ds);
        TC_TRY(energy_time(ctx, d));
        ctx->time += ctx->dt;
        ctx->tstep++;
        ctx->solve_stage = -1;
    }
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_ensemble_step(tc_ctx* const* ctxs, int32_t n, int32_t nsteps, int32_t threads, int32_t* failed)
{
    if (!ctxs || n < 0 || nsteps < 0 || threads < 0) return TC_ERR_ARG;
    for (int j = 0; j < n; ++j) {
        if (!ctxs[j]) return TC_ERR_ARG;
        for (int q = 0; q < j; ++q)
            if (ctxs[q] == ctxs[j]) return TC_ERR_ARG;  // a context is driven by one thread at a time
    }
    if (failed) *failed = -1;
    if (n == 0) return 0;
    const int nt = std::min<int>(threads > 0 ? threads : n, n);
    std::atomic<int> next{0};
    std::vector<int> status(n, 0);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return TC_ERR_HIP;  // the workers use the caller's device
    auto worker = [&]() {
        (void)hipSetDevice(dev);
        for (;;) {
            const int j = next.fetch_add(1);
            if (j >= n) break;
            status[j] = tc_time_step(ctxs[j], nsteps);
        }
    };
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    for (int j = 0; j < n; ++j)
        if (status[j] != 0) {
            if (failed) *failed = j;
            return status[j];
        }
    return 0;
}

int tc_set_profiling(tc_ctx* ctx, int enable)
{
    if (!ctx) return TC_ERR_ARG;
    if (enable) {
        if (ctx->prof_ev.empty()) {
            ctx->prof_ev.resize(2 * PROF_PAIRS);
            ctx->prof_cnt.assign(PROF_PAIRS, 0);
            for (auto& e : ctx->prof_ev) TC_HIP(hipEventCreate(&e));
        }
        for (int k = 0; k < TC_K_COUNT; ++k) { ctx->prof_ms[k] = 0.0; ctx->prof_n[k] = 0; }
        ctx->prof_used = 0;
        ctx->prof = true;
    } else {
        TC_TRY(prof_flush(ctx));
        ctx->prof = false;
    }
    return 0;
}

int tc_kernel_timing(tc_ctx* ctx, int which, double* avg_us, int64_t* count)
{
    if (!ctx || which < 0 || which >= TC_K_COUNT || !avg_us || !count) return TC_ERR_ARG;
    TC_TRY(prof_flush(ctx));
    *count = ctx->prof_n[which];
    *avg_us = ctx->prof_n[which] ? 1000.0 * ctx->prof_ms[which] / (double)ctx->prof_n[which] : 0.0;
    return 0;
}

int tc_guess_enable(tc_ctx* ctx, int mode)
{
    if (!ctx || mode < TC_GUESS_PLAIN || mode > TC_GUESS_LEARNED) return TC_ERR_ARG;
    if (mode == TC_GUESS_LEARNED && !ctx->nn) {
        ctx->err = "no network attached (tc_learn_attach)";
        return TC_ERR_ARG;
    }
    if (mode != TC_GUESS_PLAIN) TC_TRY(guess_alloc(ctx));
    ctx->guess = mode;
    for (int c = 0; c < 3; ++c) ctx->guess_losses[c] = 0;
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_learn_attach(tc_ctx* ctx, tc_nn* nn)
{
    if (!ctx) return TC_ERR_ARG;
    int m = 0;
    if (nn)
        for (int L = 1; L < ctx->grids; ++L)
            if (nn->net.H == ctx->lv[L].g.nz && nn->net.W == ctx->lv[L].g.nx) { m = L; break; }
    if (nn && !m) {
        ctx->err = "network grid must be H = nz, W = nx of a multigrid level below the coarsest";
        return TC_ERR_ARG;
    }
    if (nn && ctx->nn_level && m != ctx->nn_level) {
        ctx->err = "a network of another level was attached before";
        return TC_ERR_ARG;
    }
    ctx->nn = nn;
    for (int s = 0; s < 4; ++s) ctx->ecount[s] = 0;
    if (!nn) {
        if (ctx->guess == TC_GUESS_LEARNED) ctx->guess = TC_GUESS_EXTRAP;
        return 0;
    }
    ctx->nn_level = m;
    const Grid& gc = ctx->lv[m].g;
    TC_TRY(guess_alloc(ctx));
    if (!ctx->nn_T) {
        double** coarse[] = {&ctx->nn_T, &ctx->nn_C, &ctx->nn_p1, &ctx->nn_p2, &ctx->nn_ext, &ctx->nn_psi,
                             &ctx->nn_inc, &ctx->Tc1[0], &ctx->Tc1[1], &ctx->Tc1[2], &ctx->Tc1[3]};
        for (double** q : coarse) TC_TRY(dalloc(ctx, q, gc.elems));
        TC_TRY(dalloc(ctx, &ctx->nn_ra, ctx->lv[2].g.elems));
        TC_TRY(dalloc(ctx, &ctx->nn_rb, ctx->lv[2].g.elems));
        TC_TRY(dalloc(ctx, &ctx->nn_scales, NN_NSCALE));
        TC_TRY(dalloc(ctx, &ctx->nn_part, 3 * (size_t)nn_scale_blocks(gc)));
    }
    TC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tc_learn_record(tc_ctx* ctx, int32_t capacity)
{
    if (!ctx || capacity < 0) return TC_ERR_ARG;
    if (!ctx->nn) {
        ctx->err = "no network attached (tc_learn_attach)";
        return TC_ERR_ARG;
    }
    const size_t np = (size_t)ctx->p.nx * (size_t)ctx->p.nz;
    auto F = [](void* q) { if (q) (void)hipFree(q); };
    F(ctx->rec_x); F(ctx->rec_t); F(ctx->rec_mask); F(ctx->rec_scales);
    ctx->rec_x = nullptr; ctx->rec_t = nullptr; ctx->rec_mask = nullptr; ctx->rec_scales = nullptr;
    ctx->rec_cap = 0;
    ctx->rec_n = 0;
    if (capacity == 0) return 0;
    TC_HIP(hipMalloc((void**)&ctx->rec_x, (size_t)capacity * np * 8 * sizeof(__bf16)));
    TC_HIP(hipMalloc((void**)&ctx->rec_t, (size_t)capacity * np * sizeof(float)));
    TC_HIP(hipMalloc((void**)&ctx->rec_mask, np * sizeof(float)));
    TC_HIP(hipMalloc((void**)&ctx->rec_scales, (size_t)capacity * NN_NSCALE * sizeof(double)));
    ctx->rec_cap = capacity;
    return 0;
}

int tc_learn_samples(const tc_ctx* ctx, int32_t* count)
{
    if (!ctx || !count) return TC_ERR_ARG;
    *count = ctx->rec_n;
    return 0;
}

// Adam on the recorded samples: each epoch visits them in a seeded random
// order, one forward + backward per sample, an update every `batch` samples.
int tc_learn_train(tc_ctx* ctx, int32_t epochs, float lr, int32_t batch, uint64_t seed, double* mean_loss)
{
    if (!ctx || !ctx->nn || !ctx->nn->net.trainable || epochs < 1 || batch < 1) return TC_ERR_ARG;
    if (ctx->rec_n < 1) {
        ctx->err = "no recorded samples";
        return TC_ERR_ARG;
    }
    UNet& u = ctx->nn->net;
    const size_t np = (size_t)u.npx[0];
    // mask: interior nodes, the same for every sample
    {
        std::vector<float> m(np);
        for (int k = 0; k < u.H; ++k)
            for (int i = 0; i < u.W; ++i) m[(size_t)k * u.W + i] = (i >= 1 && k >= 1) ? 1.0f : 0.0f;
        TC_HIP(hipMemcpyAsync(u.mask, m.data(), np * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        TC_HIP(hipStreamSynchronize(ctx->stream));
        u.inv_n = (float)(1.0 / ((double)(u.W - 1) * (double)(u.H - 1)));
    }
    std::vector<int> order(ctx->rec_n);
    uint64_t st = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    double sum = 0.0;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int q = 0; q < ctx->rec_n; ++q) order[q] = q;
        for (int q = ctx->rec_n - 1; q > 0; --q) {
            st = st * 6364136223846793005ULL + 1442695040888963407ULL;
            const int r = (int)((st >> 33) % (uint64_t)(q + 1));
            std::swap(order[q], order[r]);
        }
        TC_HIP(hipMemsetAsync(u.loss_d, 0, sizeof(double), ctx->stream));
        int inb = 0;
        for (int q = 0; q < ctx->rec_n; ++q) {
            const int sidx = order[q];
            TC_HIP(hipMemcpyAsync(u.x0, ctx->rec_x + (size_t)sidx * np * 8, np * 8 * sizeof(__bf16),
                                  hipMemcpyDeviceToDevice, ctx->stream));
            TC_HIP(hipMemcpyAsync(u.target, ctx->rec_t + (size_t)sidx * np, np * sizeof(float), hipMemcpyDeviceToDevice,
                                  ctx->stream));
            u.forward(ctx->stream);
            u.backward(ctx->stream);
            if (++inb == batch || q == ctx->rec_n - 1) {
                u.adam(ctx->stream, lr, 1.0f / (float)inb);
                inb = 0;
            }
        }
        double l = 0.0;
        TC_HIP(hipMemcpyAsync(&l, u.loss_d, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        TC_HIP(hipStreamSynchronize(ctx->stream));
        sum = l / (double)ctx->rec_n;
    }
    if (mean_loss) *mean_loss = sum;
    return 0;
}

/* derivative_eigenvalues (energy.f90:501-598): Fourier-symbol maxima of the
 * FD operators, with T_test's cmplx() products rounded to single precision
 * as the reference does.  base[] = Tr, Trr, Trrr, Trrrr, Trs, Trrs, Trrss. */
int tc_derivative_eigenvalues(const double D1[5], const double D2[5], const double D3[5], const double D4[5],
                              double base[7])
{
    if (!D1 || !D2 || !D3 || !D4 || !base) return TC_ERR_ARG;
    const int n_angle = (1 << 11) + 1;
    const double d_angle = 2.0 * PII / (double)(n_angle - 1);
    auto ttest = [](int i, int k, double kr, double ks, double& re, double& im) {
        volatile float a = (float)std::cos(kr * (double)i), b = (float)std::sin(kr * (double)i);
        volatile float c = (float)std::cos(ks * (double)k), dd = (float)std::sin(ks * (double)k);
        volatile float ac = a * c, bd = b * dd, ad = a * dd, bc = b * c;
        volatile float x = ac - bd, y = ad + bc;
        re = (double)x;
        im = (double)y;
    };
    double m[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int angle = 1; angle <= n_angle; ++angle) {
        const double kr = (double)(angle - 1) * d_angle;
        double re, im, tr_im = 0, trr_re = 0, t3_im = 0, t4_re = 0;
        for (int j = -1; j <= 1; ++j) {
            ttest(j, 0, kr, 0.0, re, im);
            tr_im += D1[2 + j] * im;
            trr_re += D2[2 + j] * re;
        }
        for (int j = -2; j <= 2; ++j) {
            ttest(j, 0, kr, 0.0, re, im);
            t3_im += D3[2 + j] * im;
            t4_re += D4[2 + j] * re;
        }
        m[0] = std::max(m[0], std::fabs(tr_im));
        m[1] = std::max(m[1], std::fabs(trr_re));
        m[2] = std::max(m[2], std::fabs(t3_im));
        m[3] = std::max(m[3], std::fabs(t4_re));
        for (int angle2 = 1; angle2 <= n_angle; ++angle2) {
            const double ks = (double)(angle2 - 1) * d_angle;
            double trs = 0, trrs = 0, trrss = 0;
            for (int j = -1; j <= 1; ++j)
                for (int jj = -1; jj <= 1; ++jj) {
                    ttest(j, jj, kr, ks, re, im);
                    trs += (D1[2 + jj] * D1[2 + j]) * re;
                    trrs += (D1[2 + jj] * D2[2 + j]) * im;
                    trrss += (D2[2 + jj] * D2[2 + j]) * re;
                }
            m[4] = std::max(m[4], std::fabs(trs));
            m[5] = std::max(m[5], std::fabs(trrs));
            m[6] = std::max(m[6], std::fabs(trrss));
        }
    }
    for (int t = 0; t < 7; ++t) base[t] = m[t];
    return 0;
}

}  // extern "C"
