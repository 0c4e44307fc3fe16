// Internal declarations of the MI355X (gfx950) implementation of the
// tracer_code time-step hot path.  Host orchestration lives in tc_api.hip,
// kernels in tc_mg.hip (multigrid + viscosity hierarchy), tc_coarse.hip
// (coarsest-grid banded LU / inverse) and tc_fields.hip (velocities, T/C
// filters, energy equation, tracers).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/tracer_hip.h"

namespace tc {

constexpr int SPAN = 2;       // ghost width of every grid field (order/2+1)
// elements before column i=0 of each row: with the row pitch a multiple of
// 32 elements, the first interior column i=1 starts a 128-byte cache line
// (FP32) or 256 bytes (FP64), so a 64-wide row segment of a launch touches
// the fewest lines and no line is shared by two workgroups of one row
constexpr int ORG_PAD = 31;

// A ghosted 2-D plane f(-2:nx+2,-2:nz+2) in device memory.  Element (i,k)
// sits at data[org + k*ld + i]; ld is a multiple of 32 doubles.
struct Grid {
    int nx = 0, nz = 0;   // cells of this level
    int ld = 0;           // row pitch in elements
    int org = 0;          // offset of (0,0)
    size_t elems = 0;     // allocation size in elements
    __host__ __device__ inline size_t at(int i, int k) const { return (size_t)org + (size_t)((ptrdiff_t)k * ld + i); }
};

#ifdef __HIPCC__
// Workgroups are dispatched round-robin over the chip's 8 XCDs, each with
// its own L2.  xcd_tile() maps a 2-D launch's linear workgroup id so that
// XCD x takes the x-th contiguous band of the tile grid (row-major): tiles
// that share stencil halos then mostly share an L2.  Identity when 8 does
// not divide the number of workgroups.  Returns the tile (x, y).
__device__ inline int2 xcd_tile()
{
    const int gx = (int)gridDim.x, n = (int)(gridDim.x * gridDim.y);
    int w = (int)(blockIdx.y * gridDim.x + blockIdx.x);
    if ((n & 7) == 0) w = (w & 7) * (n >> 3) + (w >> 3);
    return make_int2(w % gx, w / gx);
}
#endif

inline Grid make_grid(int nx, int nz)
{
    Grid g;
    g.nx = nx;
    g.nz = nz;
    int need = nx + 2 * SPAN + 1 + ORG_PAD;
    g.ld = (need + 31) / 32 * 32;
    g.org = ORG_PAD + SPAN * g.ld;
    g.elems = (size_t)g.ld * (size_t)(nz + 2 * SPAN + 1) + 2 * ORG_PAD;
    return g;
}

// Per-level device state of the multigrid hierarchy (the reference keeps
// these as level-fastest fine-size arrays, basics.f90:87-98; here each level
// owns compact planes).
struct Level {
    Grid g;
    double *err[2] = {nullptr, nullptr};  // error(L): ping-pong for Jacobi sweeps
    int cur = 0;                          // which err buffer holds error(L)
    double *rhs = nullptr, *vis = nullptr, *visx = nullptr, *visz = nullptr, *visxx = nullptr;
    double *viszz = nullptr, *visxz = nullptr, *dtstream = nullptr, *ratio = nullptr, *dtvis = nullptr;
    double *visf = nullptr, *visxf = nullptr, *viszf = nullptr, *visxxf = nullptr, *viszzf = nullptr, *visxzf = nullptr;
    double *vtmp = nullptr;               // viscosity smoother copy (T0 of smoother_vis)
    double *res = nullptr;                // compute_residual output of this level
    // spacings (basics.f90:167-182)
    double dr, ds, dr2, ds2, dr3, ds3, dr4, ds4, drds, drds2, dr2ds, dr2ds2;
    // FP32 copies for the mixed-precision path
    float *err32[2] = {nullptr, nullptr};
    float *rhs32 = nullptr, *vis32 = nullptr, *visx32 = nullptr, *visz32 = nullptr, *a132 = nullptr, *visxz32 = nullptr;
    float *dts32 = nullptr, *lb32 = nullptr, *res32 = nullptr;
    float* w13 = nullptr;  // 13 planes: dt-scaled stencil weights of the FP32 operator, per solve
    int tb_tile = 0;       // tile of the FP32 temporally blocked sweeps: 16, 32, or 0 (by level size)
    int cur32 = 0;
};

// equalize_tracers: per-call reductions read back by the host
constexpr int EQ_MAXT = 8;                 // type slots (>= TC_MAX_TYPES + 1)
struct EqTypes { int32_t v[EQ_MAXT]; };
struct EqOff { int64_t v[EQ_MAXT]; };
constexpr int EQ_NS = 16, EQ_NM = 8;      // sums and minima of one reduction
struct EqPart {                            // one workgroup's share of a reduction
    long long s[EQ_NS];
    int32_t m[EQ_NM];
};
struct EqRed {
    long long fs[EQ_NS];                   // the last reduction: sums
    int32_t fm[EQ_NM];                     //                     minima (2147483647: none)
    int32_t bad, pad;                      // a tracer outside the domain
    unsigned long long found;              // donors found
    unsigned long long dcount[EQ_MAXT];    // donors of each type
    unsigned long long qtot[EQ_MAXT];      // qualifying nodes of each donor type in the last pass
};

}  // namespace tc

struct tc_ctx {
    tc_params p;
    int64_t ntr = 0;
    int mode = TC_MODE_FP64;
    int guess = 0;                // initial guess of the defect-correction modes (TC_GUESS_*)
    // a coarse correction whose prolongation the next visit's first blocked
    // launch folds into its staging (level it goes to, source, source level)
    int pend_level = 0, pend_from = 0;
    const float* pend_src = nullptr;
    // GCR over the defect-correction cycles (tc_set_gcr): directions kept
    // (0: off), directions held now (their slots in gcr_order, oldest
    // first), and whether psi_s / r_s hold the state after the last cycle;
    // Z, Q per slot: a direction and the operator applied to it.  On the
    // device: the dots, the step's coefficients, the squared norm per slot
    // and the step's flag, whose pinned copy is read at the next step
    // (gcr_pending, after the stop test has synchronised the stream)
    int gcr_m = 0, gcr_nb = 0;
    int gcr_order[16] = {};
    bool gcr_have = false, gcr_pending = false;
    bool gcr_suspended = false;  // the plain-cycle retry of a solve that failed under GCR
    double *gcr_hs = nullptr, *gcr_ls = nullptr, *gcr_rs = nullptr;
    double* gcr_Z[16] = {};
    double* gcr_Q[16] = {};
    double *gcr_part = nullptr, *gcr_out = nullptr, *gcr_qq = nullptr, *gcr_coef = nullptr;
    int32_t *gcr_flag = nullptr, *gcr_hflag = nullptr;
    int gcr_new = 0;  // slot of the pending step's direction
    hipEvent_t gcr_ev = nullptr;
    int64_t gcr_updates = 0, gcr_restarts = 0;
    // graph replay of the coarse segment of each FP32 defect-correction cycle
    // (tc_set_graphs): up to two executable graphs per solve, keyed by the
    // ping-pong parity of every level's correction at the segment's start
    bool graphs = false;
    struct SegGraph {
        hipGraphExec_t exec = nullptr;
        bool valid = false;                          // captured in the running solve
        int par0[TC_MAX_LEVELS + 1] = {}, par1[TC_MAX_LEVELS + 1] = {};
        int64_t dsweeps[TC_MAX_LEVELS] = {}, dres[TC_MAX_LEVELS] = {}, dcoarse = 0;
    } seg[2];
    int64_t seg_replays = 0, seg_captures = 0;
    int dc_sweeps = 0;            // sweeps per visit of the defect-correction cycle (0: p.Nmulti)
    int dc_sweeps_active = 0;     // the count the running defect-correction solve uses
    int dc_level_sweeps[TC_MAX_LEVELS + 1] = {};  // per-level counts of the tuned cycle (0: dc_sweeps)
    int solve_stage = -1;         // RK stage of the running solve within the time step (0..3), -1 outside one
    double solve_tau = 0.0;       // model time the stage's buoyancy field stands for
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;   // coarse LU / inverse overlap stream
    hipEvent_t ev_inv_ready = nullptr, ev_visc_ready = nullptr;
    hipEvent_t ev_inv_next = nullptr;  // the lagged factorisation in progress (defect-correction modes)
    std::string err;
    // hierarchy
    int grids = 0;
    tc::Level lv[TC_MAX_LEVELS + 1];
    int ratio_level = -1;           // level whose ratio array is valid (others read as 1e99)
    int Nvis_store = 0, nn_store = 1;
    double res_ref[TC_MAX_LEVELS + 1] = {0};
    double residual_mag[TC_MAX_MULTI] = {0};
    // fine-grid fields (level-1 geometry)
    tc::Grid g;
    double *T = nullptr, *SF = nullptr, *u = nullptr, *w = nullptr, *strain = nullptr, *Tvis = nullptr, *Tbuoy = nullptr;
    double *conduct = nullptr, *conduct_r = nullptr, *conduct_s = nullptr, *dts = nullptr;
    double *DERr_u = nullptr, *DERr_w = nullptr, *T0 = nullptr, *T1 = nullptr, *T2 = nullptr;
    double *residual = nullptr, *tmp = nullptr, *tmp2 = nullptr;
    double *psi_hi = nullptr, *psi_lo = nullptr;  // double-double streamfunction of the defect-correction modes
    double *Cnew[TC_MAX_TYPES] = {}, *Cvis[TC_MAX_TYPES] = {}, *Cbuoy[TC_MAX_TYPES] = {}, *C0[TC_MAX_TYPES] = {};
    int32_t *count[TC_MAX_TYPES + 1] = {}, *nempty = nullptr;
    double* zg = nullptr;  // zg(-3:nz+3)
    // coarsest grid direct solve
    int ngc = 0, kl = 0, ku = 0, ldab = 0;
    double *Mc = nullptr, *RowB = nullptr, *Ainv = nullptr, *Bc = nullptr;
    double* Ur = nullptr;  // row-major copy of the U band (explicit inverse)
    int32_t* ipiv = nullptr;
    int32_t* d_info = nullptr;
    // defect-correction modes: the factorisation of this solve's coarse
    // operator is computed into these while the solve uses the previous one
    double *RowB_next = nullptr, *Ainv_next = nullptr;
    // the explicit inverses rounded to FP32, for the FP32 correction cycle
    float *Ainv32 = nullptr, *Ainv32_next = nullptr;
    double* coarse_b = nullptr;   // equilibrated right-hand side of the FP32-cycle coarse solve
    int32_t* d_info_next = nullptr;
    bool next_pending = false;
    // refactorise the coarse operator every coarse_refresh-th solve of the
    // defect-correction modes (tc_set_coarse_refresh); solves since the last
    int coarse_refresh = 1, coarse_since = 0;
    // the C filter loop of smoother_time with several iterations per launch
    // (tc_set_filter_blocking); off: one launch per iteration
    bool filter_blocking = true;
    // multicolour Gauss-Seidel on level 1 of the FP32 correction cycle
    // (tc_set_gs_smoother): sweeps per visit (0: Jacobi-Richardson), per launch
    int gs_sweeps = 0, gs_per_launch = 2;
    double gs_omega = 1.0;
    // Chebyshev interval of the level-1 visits of the FP32 correction cycle
    // (tc_set_dc_cheb_interval; the reference's -2, -1 by default)
    double cheb1_lb = -2.0, cheb1_la = -1.0;
    // tracers (SoA)
    double *rtr = nullptr, *str = nullptr, *rtr0 = nullptr, *str0 = nullptr;
    double *rtr1 = nullptr, *str1 = nullptr, *rtr2 = nullptr, *str2 = nullptr;
    int32_t* ttype = nullptr;
    // node-sorted tracer order (tc_sort_tracers): caller index of each slot
    // (null: the caller's order), sort scratch, and the automatic interval
    int64_t *tr_id = nullptr, *tr_id_spare = nullptr;
    uint32_t *sort_k[2] = {nullptr, nullptr}, *sort_v[2] = {nullptr, nullptr};
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    double* spare_d = nullptr;
    int32_t* spare_i = nullptr;
    int sort_interval = 0;
    int64_t sorts = 0;
    int32_t* d_flag = nullptr;   // tracer-out-of-domain flag
    int32_t* h_flag = nullptr;   // pinned
    // model time
    double time = 0.0, dt = 0.0;
    int tstep = 0;
    tc_stats stats{};
    // initial-guess history: per RK stage, the solutions of that stage in
    // the last two time steps (fine grid) and the model times they stand for
    double* sh1[4] = {nullptr, nullptr, nullptr, nullptr};
    double* sh2[4] = {nullptr, nullptr, nullptr, nullptr};
    double tau1[4] = {0, 0, 0, 0}, tau2[4] = {0, 0, 0, 0};
    int scount[4] = {0, 0, 0, 0};
    int guess_losses[3] = {0, 0, 0};  // consecutive lost judgements of each candidate
    int64_t guess_solves = 0;         // solves inside time steps with candidates enabled
    double* ext_psi = nullptr;    // the extrapolated guess of the running solve (the network's base)
    bool have_ext = false;
    double* inc_psi = nullptr;    // the stage-increment guess of the running solve
    bool have_inc = false;
    int ecount[4] = {0, 0, 0, 0}; // per stage: extrapolation errors measured with the network attached
    double* stage_err = nullptr;  // 4 x [NN_NSCALE]; [3] = the stage's last max |R psi - R ext|
    // learned correction (tc_nn.hip): attached network and recorded samples
    struct tc_nn* nn = nullptr;
    int nn_level = 0;             // multigrid level the network works on (its H x W = that level's nz x nx)
    double* Tc1[4] = {nullptr, nullptr, nullptr, nullptr};  // per stage: buoyancy T of its last solve (level nn_level)
    double *nn_T = nullptr, *nn_C = nullptr;       // this solve's T and C restricted to level nn_level
    double *nn_p1 = nullptr, *nn_p2 = nullptr;     // the stage's last two solutions restricted
    double *nn_ext = nullptr, *nn_psi = nullptr;   // restricted extrapolation and solution
    double* nn_inc = nullptr;                      // predicted correction (level nn_level)
    double *nn_ra = nullptr, *nn_rb = nullptr;     // restriction/prolongation scratch (level-2 sized)
    int rec_cap = 0, rec_n = 0;
    __bf16* rec_x = nullptr;      // rec_cap x [H*W*8] inputs (the network's grid)
    float* rec_t = nullptr;       // rec_cap x [H*W] targets
    float* rec_mask = nullptr;    // [H*W]
    double* nn_scales = nullptr;  // [4] scales of the current solve (tc_nn.h NN_NSCALE)
    double* rec_scales = nullptr; // rec_cap x [4] scales of the recorded samples
    double* nn_part = nullptr;    // block partials of the scale reduction
    double* guess_psi = nullptr;  // candidate start being judged (high part; after a swap, the other start)
    double* guess_lo = nullptr;   // its low part
    // implementation-private host bookkeeping (tc_api.hip)
    void* impl = nullptr;
    // equalize_tracers (tc_set_equalize, tc_equalize_tracers)
    bool eq_on = false, eq_in_step = false;
    int32_t eq_tpc = 0;
    double eq_Cinit[TC_MAX_TYPES] = {};
    uint64_t eq_seed = 0, eq_calls = 0;
    int32_t *eq_ne = nullptr, *eq_start = nullptr, *eq_end = nullptr, *eq_rows = nullptr, *eq_cols = nullptr;
    int32_t* eq_bcnt = nullptr;
    tc::EqPart* eq_part = nullptr;   // per-workgroup partials of the reductions
    int32_t* eq_hord = nullptr;      // pinned visit orders (rows, then columns)
    uint64_t* eq_k[2] = {nullptr, nullptr};
    uint32_t* eq_v[2] = {nullptr, nullptr};
    void* eq_tmp = nullptr;
    size_t eq_tmp_bytes = 0;
    tc::EqRed *eq_red = nullptr, *eq_hred = nullptr;  // device, pinned host mirror
    // snapshot staging (tc_snapshot_stage / tc_snapshot_get)
    double* snap_dev = nullptr;      // device staging planes
    double* snap_host = nullptr;     // pinned host copy
    hipStream_t snap_stream = nullptr;
    hipEvent_t ev_snap_d = nullptr, ev_snap_h = nullptr;
    bool snap_staged = false;
    // host <-> device plane copies through two pinned chunk buffers
    // (tc_get_field / tc_put_field): allocated on first use
    void* xfer_buf[2] = {nullptr, nullptr};
    hipEvent_t xfer_ev[2] = {nullptr, nullptr};
    // kernel timing (tc_set_profiling)
    bool prof = false;
    std::vector<hipEvent_t> prof_ev;
    std::vector<int> prof_cnt;    // launches bracketed by each event pair
    int prof_used = 0;
    double prof_ms[TC_K_COUNT] = {0};
    int64_t prof_n[TC_K_COUNT] = {0};
};

#define TC_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) {                                                        \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(_e);              \
            return TC_ERR_HIP;                                                         \
        }                                                                              \
    } while (0)

#define TC_TRY(expr)                  \
    do {                              \
        int _s = (expr);              \
        if (_s != 0) return _s;       \
    } while (0)
