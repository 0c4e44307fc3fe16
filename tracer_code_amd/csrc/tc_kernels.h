// Kernel-side helpers and the host launchers of every kernel file.
#pragma once
#include "tc_internal.h"
#include "tc_stencil.h"

namespace tc {

struct RaCArr { double v[TC_MAX_TYPES]; };
struct PtrArr { double* v[TC_MAX_TYPES]; };
struct CntArr { int32_t* v[TC_MAX_TYPES + 1]; };
struct D3Arr { double v[3]; };

// Scaled Fourier-symbol maxima of test_eigenvalues / iterate_local_time_steps
// (iterative_solver.f90:674-685, 709-720).
struct EigScalars {
    double Tr_max, Trr_max, Trrr_max, Trrrr_max, Ts_max, Tss_max, Tsss_max, Tssss_max;
    double Trs_max, Trss_max, Trrs_max, Trrss_max;
};

// Block-wide reductions (any block shape up to 1024 threads, wave64).
// The combination order is fixed, so results are run-to-run deterministic.
__device__ __forceinline__ int flat_tid() { return threadIdx.z * blockDim.x * blockDim.y + threadIdx.y * blockDim.x + threadIdx.x; }
__device__ __forceinline__ int flat_nthreads() { return blockDim.x * blockDim.y * blockDim.z; }

template <int OP>
__device__ __forceinline__ double red_op(double a, double b)
{
    if constexpr (OP == 0) return a + b;
    else if constexpr (OP == 1) return b < a ? b : a;
    else return b > a ? b : a;
}

template <int OP>
__device__ __forceinline__ double block_reduce(double a)
{
    __shared__ double sh[16];
    const int tid = flat_tid(), nt = flat_nthreads();
    for (int off = 32; off > 0; off >>= 1) a = red_op<OP>(a, __shfl_down(a, off, 64));
    const int wid = tid >> 6, nw = (nt + 63) >> 6;
    __syncthreads();
    if ((tid & 63) == 0) sh[wid] = a;
    __syncthreads();
    if (tid == 0) {
        double r = sh[0];
        for (int w = 1; w < nw; ++w) r = red_op<OP>(r, sh[w]);
        a = r;
    }
    return a;
}
// prolong (iterative_solver.f90:175-233) of a coarse FP32 correction at the
// fine interior point (i,k): the value added there.  Explicitly rounded, in
// the reference's order, so that every kernel adding it gets the same bits.
__device__ __forceinline__ float prolong_add32(const float* __restrict__ Ec, const Grid& gc, int i, int k)
{
    const bool ie = (i % 2) == 0, ke = (k % 2) == 0;
    if (ie && ke) return Ec[gc.at(i / 2, k / 2)];
    if (ie) return __fmul_rn(__fadd_rn(Ec[gc.at(i / 2, (k - 1) / 2)], Ec[gc.at(i / 2, (k + 1) / 2)]), 0.5f);
    if (ke) return __fmul_rn(__fadd_rn(Ec[gc.at((i - 1) / 2, k / 2)], Ec[gc.at((i + 1) / 2, k / 2)]), 0.5f);
    const float s = __fadd_rn(__fadd_rn(__fadd_rn(Ec[gc.at((i - 1) / 2, (k - 1) / 2)], Ec[gc.at((i + 1) / 2, (k - 1) / 2)]),
                                        Ec[gc.at((i + 1) / 2, (k + 1) / 2)]),
                              Ec[gc.at((i - 1) / 2, (k + 1) / 2)]);
    return __fmul_rn(s, 0.25f);
}

__device__ __forceinline__ double block_sum(double a) { return block_reduce<0>(a); }
__device__ __forceinline__ double block_min(double a) { return block_reduce<1>(a); }
__device__ __forceinline__ double block_max(double a) { return block_reduce<2>(a); }

// --- tc_mg.hip ---------------------------------------------------------------
void launch_sweep(hipStream_t st, bool pow2, bool use_ratio, const Level& L, const double* S0, double* E, const FD& F,
                  const Spacing& h, int ivis, double temp, double cconst, double zsg);
void launch_bc_stream(hipStream_t st, double* S, const Grid& g, int Vbc);
int launch_residual(hipStream_t st, bool pow2, const Level& L, const double* S, double* R, const FD& F,
                    const Spacing& h, int ivis, double* partial);
void launch_restrict(hipStream_t st, const double* in, const Grid& gi, double* out, const Grid& go);
void launch_prolong(hipStream_t st, double* Ef, const Grid& gf, const double* Ec, const Grid& gc, double zsg);
int launch_rhs1(hipStream_t st, const double* Tb, const PtrArr& Cb, double* rhs, const Grid& g, const FD& F,
                double dr, double RaT, int comp, int ntype, const RaCArr& RaC, double* partial);
int launch_abs_partial(hipStream_t st, const double* f, const Grid& g, double* partial);
void launch_reduce_final(hipStream_t st, const double* partial, int n, int op, double* out, int slot);
void launch_visf(hipStream_t st, double* vis, const double* Tvis, const PtrArr& Cvis, const double* zg,
                 const Grid& g, int ivis, int comp, int ntype, double RaT, double lvisT, double lvisP, const RaCArr& lvisC);
// copy_dst: also copy the plane L.vis there (the source of smoother_vis's
// sweep); store = false: only the block minima of the ratio (no planes);
// slots: the minima also into the loop's slots (VisLoopArgs)
int launch_vis_grad_eig(hipStream_t st, bool pow2, Level& L, const FD& F, const Spacing& h, int ivis,
                        const EigScalars& e, int want_ratio, double* partial, const int32_t* gate = nullptr,
                        double* copy_dst = nullptr, bool store = true, unsigned long long* slots = nullptr);
void launch_local_dt(hipStream_t st, Level& L, const EigScalars& e, double courant_stream);
// cptr / gate (device-controlled smoother_vis): the factor read from the
// device, and the launch skipped while gate[0] is set
void launch_vis_sweep(hipStream_t st, bool pow2, const double* T0, double* V, const double* dtv, const Grid& g,
                      const FD& F, const Spacing& h, int kmin, int kmax, double c, const double* cptr = nullptr,
                      const int32_t* gate = nullptr);
struct VisLoopArgs;
// one smoother_vis iteration: its decision, then (unless it stops the loop)
// the sweep (T0 -> V) and the viscosityBCs after it, in one launch
void launch_vis_sweep_bcs(hipStream_t st, bool pow2, const double* T0, double* V, const double* dtv, const Grid& g,
                          const FD& F, const Spacing& h, int kmin, int kmax, const VisLoopArgs& lp,
                          const double* Vf, const Grid& gf, int level, double RaT, double visT2);
void launch_visc_bcs(hipStream_t st, double* V, const Grid& g, const double* Vf, const Grid& gf, int level, int init,
                     double RaT, double visT2, const int32_t* gate = nullptr);
// smoother_vis's loop state on the device (vis_decide)
enum { VC_DONE = 0, VC_ERR, VC_NOSWEEP, VC_NN, VC_N, VC_NITER, VC_NVIS, VC_NNSTORE, VC_TAB, VC_NETA, VC_ITERS, VC_COUNT };
enum { VD_C = 0, VD_EIG, VD_COUNT };
constexpr int VIS_CMAX = 32;  // longest pass (Niter) the device control takes
struct VisDecide {
    int option, interpolate, ivis;
    double eig_ratio_min, eig_ratio;
    double c0[VIS_CMAX], c1[VIS_CMAX];  // factors of the first pass (option 1) and of the later / only passes
};
// The device-side loop of smoother_vis: k_vis_grad_eig puts each block's
// minimum ratio into VIS_SLOTS slots (atomicMin, VIS_SLOT_STRIDE words apart)
// and k_vis_sweep_bcs makes the iteration's decision from them.  The loop
// state alternates between two buffers with the iteration's parity.
constexpr int VIS_SLOTS = 64, VIS_SLOT_STRIDE = 16;
struct VisLoopArgs {
    const int32_t* ctl_in;
    int32_t* ctl_out;
    const double* dctl_in;
    double* dctl_out;
    const unsigned long long* slots_in;
    unsigned long long* slots_reset;
    VisDecide P;
};

// --- tc_coarse.hip -------------------------------------------------------------
void launch_coarse_assemble(hipStream_t st, double* Mc, double* RowB, const Level& L, const FD& F, const Spacing& h,
                            int ivis, int Vbc, int kl, int ku, int ngc);
void launch_dgbtf2(hipStream_t st, double* ab, int ldab, int n, int kl, int ku, int32_t* ipiv, int32_t* info);
void launch_band_inverse(hipStream_t st, const double* ab, int ldab, int n, int kl, int ku, const int32_t* ipiv,
                         double* Ainv, double* Ur);
void launch_coarse_lusolve(hipStream_t st, const double* ab, int ldab, int n, int kl, int ku, const int32_t* ipiv,
                           const double* RowB, const Level& L, double* E, double zsg);
void launch_coarse_apply(hipStream_t st, const double* Ainv, const double* RowB, const Level& L, double* E, int ngc,
                         double zsg);

// --- tc_dc.hip ------------------------------------------------------------------
void launch_add_diff(hipStream_t st, const double* a, const double* b, const double* c, double* out, size_t n);
// GCR over the defect-correction cycles: at most GCR_MAX directions held
// (FP64 planes; the new one at position nb <= GCR_MAX - 1), GCR_NDOT dot
// values per block
constexpr int GCR_MAX = 16, GCR_NDOT = 2 * GCR_MAX + 2;
struct GcrSet { const double* v[GCR_MAX]; };
struct GcrSetW { double* v[GCR_MAX]; };
struct GcrCoef { double beta[GCR_MAX]; double alpha; };
struct GcrIdx { int v[GCR_MAX]; };  // slot of each kept direction, then of the new one
int gcr_blocks(const Grid& g);
// with q = rs - r: out[2j] = (q, Q_j), out[2j+1] = (rs, Q_j) for j < nb;
// out[2nb] = (q, q), out[2nb+1] = (rs, q)
void launch_gcr_dots(hipStream_t st, int nb, const double* rs, const double* r, const GcrSet& Q, const Grid& g,
                     double* part, double* out);
// the coefficients of the step (device-resident cf; flag: 0, or 1 = the new
// direction used alone, 2 = the cycle's result kept; qq: squared norm per slot)
void launch_gcr_coef(hipStream_t st, int nb, const double* out, const GcrIdx& idx, double* qq, GcrCoef* cf,
                     int32_t* flag);
void launch_gcr_update(hipStream_t st, int nb, const GcrSetW& Z, const GcrSetW& Q, const GcrCoef* cf, double* hi,
                       double* lo, double* hs, double* ls, double* rs, double* r, float* r32, const Grid& g,
                       double zsg);
void launch_extrap(hipStream_t st, const double* a, const double* b, double c, double* out, size_t n);
int launch_residual_dd(hipStream_t st, const double* hi, const double* lo, const Level& L, const double* rhs,
                       double* r64, float* r32, const FD& F, const Spacing& h, int ivis, double* partial);
void launch_dd_absorb(hipStream_t st, double* hi, double* lo, float* e32, double* e64, const Grid& g, double zsg);
// psi(hi,lo) += e and r64 -= L(e) (r32 its FP32 rounding), e the level-1
// correction (gc == nullptr) or the prolongation of the level-2 correction on
// grid *gc; with partial, the block sums of |r| (returns the block count)
int launch_absorb_inc(hipStream_t st, double* hi, double* lo, const float* e32, const double* e64, const Grid* gc,
                      const Level& L, double* r64, float* r32, const FD& F, const Spacing& h, int ivis, bool pow2,
                      double zsg, double* partial);
void launch_prolong_dd(hipStream_t st, double* hi, double* lo, const Grid& gf, const float* e32, const double* e64,
                       const Grid& gc, double zsg);
void launch_sweep32(hipStream_t st, bool use_ratio, const Level& L, const float* S0, float* E, const float* rhs,
                    const FD& F, const Spacing& h, double temp, double cconst, double zsg);
void launch_residual32(hipStream_t st, const Level& L, const float* S, float* R, const float* rhs, const FD& F,
                       const Spacing& h);
void launch_restrict32(hipStream_t st, const float* in, const Grid& gi, float* out, const Grid& go);
// with RowB and bout (descent into the coarsest level): also the coarse
// solve's equilibrated right-hand side, as launch_coarse_apply32 forms it
void launch_residual_restrict32(hipStream_t st, const Level& Lf, const float* S, const float* rhs, const FD& F,
                                const Spacing& h, float* out, const Grid& go, const double* RowB = nullptr,
                                double* bout = nullptr);
void launch_prolong32(hipStream_t st, float* Ef, const Grid& gf, const float* Ec, const Grid& gc, double zsg);
// b_ready: b already holds the right-hand side (launch_residual_restrict32)
void launch_coarse_apply32(hipStream_t st, const float* Ainv, const double* RowB, const Level& L, const float* rhs,
                           float* E, int ngc, double zsg, double* b, bool b_ready);
void launch_to_f32(hipStream_t st, const double* in, float* out, size_t n);
void launch_coeffs_to_f32(hipStream_t st, Level& L, bool with_lb, int ivis);

// --- tc_tb.hip ------------------------------------------------------------------
int tb_max_sweeps(int K);
// levels whose temporally blocked sweeps read precomputed weights (Level::w13)
bool tb_uses_w13(const Level& L);
int tb_sweeps_per_launch(const Grid& g);
void launch_tb_weights32(hipStream_t st, Level& L, const FD& F, const Spacing& h);
void launch_sweep_tb64(hipStream_t st, bool use_ratio, int K, const Level& L, const double* rhs, const double* S0,
                       double* E, const FD& F, const Spacing& h, const double* temp, const double* cconst, int nk,
                       double zsg);
// The next coarser level's right-hand side, formed at the end of a visit's
// last blocked launch (what launch_residual_restrict32 would form from the
// visit's result); RowB/bout as there.
struct RRTarget {
    float* out = nullptr;
    Grid go;
    const double* RowB = nullptr;
    double* bout = nullptr;
    int cus = 256;  // compute units: the fused launch serves levels of at most one tile per CU
};
// Pc (optional): a coarse correction on grid *gc to prolong into S0 while
// staging (the prolongation folded into the visit's first launch).
// rr (optional): also form rr's right-hand side; returns whether it did
// (16x16 tiles on levels of at most rr->cus tiles only).
bool launch_sweep_tb32(hipStream_t st, bool use_ratio, int K, const Level& L, const float* rhs, const float* S0,
                       float* E, const FD& F, const Spacing& h, const double* temp, const double* cconst, int nk,
                       double zsg, const float* Pc = nullptr, const Grid* gc = nullptr,
                       const RRTarget* rr = nullptr);

// --- tc_fields.hip -------------------------------------------------------------
void launch_velocities(hipStream_t st, const double* SF, double* u, double* w, const Grid& g, const FD& F, double dr,
                       double ds, int Vbc);
void launch_strain(hipStream_t st, const double* u, const double* w, double* strain, const Grid& g, const FD& F,
                   double dr, double ds);
void launch_bcs_T(hipStream_t st, double* T, const Grid& g, int Tbot, const int32_t* stop = nullptr);
void launch_bcs_comp(hipStream_t st, double* C, const Grid& g, const int32_t* stop = nullptr);
void launch_zero_i32(hipStream_t st, int32_t* p, size_t n);
int diag_blocks(const Grid& g);
int stats_np();  // partial sums per block of launch_stats_rest
void launch_stats_rest(hipStream_t st, const double* u, const double* w, const double* T, const double* conduct,
                       const double* vis, const PtrArr& Cnew, int nC, int kent, double RaT, const Grid& g,
                       const FD& F, double dr, double ds, double* part, double* out);
void launch_diagnostics(hipStream_t st, const double* u, const double* w, const double* T, const double* C,
                        const int32_t* nempty, const double* strain, const double* conduct, const Grid& g,
                        const FD& F, double dr, double ds, double* part, int32_t* ints, unsigned long long* smax,
                        double* out);
void launch_smoothness_max(hipStream_t st, const double* T0, const Grid& g, unsigned long long* slots,
                           const int32_t* stop);
size_t filter_slot_words();  // u64 words of the slot sets of the filter loops
int filter_multi_k();        // C-filter iterations per launch of launch_filter_multi
// K iterations of the C filter per launch from bufs[b_in] (iterate m) into
// bufs[(b_in+1)%3], with the stop test of the previous launch's iterates and
// the replay from bufs[(b_in+2)%3]; fac: the factors of iterations m-K..m-1
// then m..m+K-1.  ctl: stop flag, stopping iterate, buffer holding it.
void launch_filter_multi(hipStream_t st, double* const bufs[3], int b_in, const double* dts, const Grid& g,
                         const FD& F, double dr2, double ds2, double courant, const double* fac, int m, double delta,
                         unsigned long long* slots, int32_t* ctl);
void launch_filter_iter(hipStream_t st, bool comp, const double* T0, double* T, const double* dts, const Grid& g,
                        const FD& F, double dr2, double ds2, double cfac, double courant, int Tbot,
                        unsigned long long* slots, int m, double delta, int32_t* ctl);
// --- tc_sort.hip ----------------------------------------------------------------
void launch_tracer_keys(hipStream_t st, const double* r, const double* s, int64_t n, int nx, double dr, double ds,
                        double aspect, uint32_t bad, uint32_t* key, uint32_t* val);
void launch_gather_d(hipStream_t st, const double* in, const uint32_t* idx, double* out, int64_t n);
void launch_gather_i32(hipStream_t st, const int32_t* in, const uint32_t* idx, int32_t* out, int64_t n);
void launch_gather_id(hipStream_t st, const int64_t* id, const uint32_t* idx, int64_t* out, int64_t n);
void launch_scatter_d(hipStream_t st, const double* in, const int64_t* id, double* out, int64_t n);
constexpr int COPY_MAX = 16;
struct CopyList {
    const void* src[COPY_MAX];
    size_t bytes[COPY_MAX];     // multiples of 4
    size_t dst_off[COPY_MAX];   // multiples of 16
    int n;
};
void launch_copy_list(hipStream_t st, const CopyList& L, void* dst);
void launch_scatter_i32(hipStream_t st, const int32_t* in, const int64_t* id, int32_t* out, int64_t n);
// multicolour Gauss-Seidel sweeps of a level's FP32 correction (tc_gs.hip):
// nk sweeps from Ein into Eout, with the level's 13 dt-scaled tap planes w13
void launch_gs32(hipStream_t st, const Level& L, const float* rhs, const float* Ein, float* Eout, int nk, double omega,
                 double zsg);

// Stable LSD radix sort of (key, value) pairs on the key's bits [0, bits)
// (tc_radix.hip).  k0/v0 hold the input (identity_values: v0 is not read,
// the values are the indices 0..n-1); the result is in k0/v0 (*cur = 0) or
// k1/v1 (*cur = 1).  scratch: radix_scratch_words(n) words.
size_t radix_scratch_words(int64_t n);
template <typename K>
int radix_sort_pairs(hipStream_t st, K* k0, K* k1, uint32_t* v0, uint32_t* v1, bool identity_values, int64_t n,
                     int bits, uint32_t* scratch, size_t scratch_words, int* cur);

// equalize_tracers (tc_equalize.hip)
uint64_t eq_stream_key(uint64_t seed, uint64_t call);
void launch_eq_clear(hipStream_t st, EqRed* red);
void launch_eq_tpc_stats(hipStream_t st, const CntArr& cnt, const PtrArr& Cnew, const Grid& g, int ntype, int comp,
                         EqPart* part, EqRed* red);
void launch_eq_nexcess(hipStream_t st, const CntArr& cnt, const int32_t* nempty, const Grid& g, int nt, int comp,
                       const EqTypes& tpc, int32_t* ne, EqPart* part, EqRed* red);
void launch_eq_node_lists(hipStream_t st, const double* r, const double* s, const int64_t* id, int64_t n,
                          const Grid& g, double dr, double ds, double aspect, uint64_t* key, uint32_t* val,
                          int32_t* start, int32_t* end, EqRed* red);
void launch_eq_bounds(hipStream_t st, const uint64_t* key, int64_t n, int32_t* start, int32_t* end);
void launch_eq_select(hipStream_t st, const int32_t* ne, const int32_t* start, const int32_t* end, uint32_t* list,
                      const double* r, const double* s, const int32_t* ty, const int64_t* id, const Grid& g,
                      double dist_r, double dist_s, double aspect, uint64_t key, uint64_t* dkey, uint32_t* dval,
                      EqRed* red);
void launch_eq_reduce(hipStream_t st, const int32_t* ne, const CntArr& cnt, const PtrArr& Cnew, const Grid& g, int nt,
                      int ntype, int comp, EqPart* part, EqRed* red);
void eq_visit_orders(int32_t* rows, int nr, int32_t* cols, int nc, uint64_t key, uint64_t pass);
int eq_pass_blocks(const Grid& g);
void launch_eq_pass(hipStream_t st, const int32_t* rows, const int32_t* cols, int32_t* ne, const CntArr& cnt,
                    const Grid& g, int nt, int tg, const EqTypes& mx, int32_t* bcnt, const uint32_t* donors,
                    const EqOff& first, const EqOff& avail, double* r, double* s, int32_t* ty, const int64_t* id,
                    uint64_t key, double dr, double ds, double dist_r, double dist_s, double aspect, int comp,
                    int update_counts, EqRed* red);
int eq_sort(hipStream_t st, void* temp, size_t* temp_bytes, uint64_t* k0, uint64_t* k1, uint32_t* v0, uint32_t* v1,
            int64_t n, int bits, int* cur);

void launch_bin(hipStream_t st, const double* r, const double* s, const int32_t* ty, int64_t n, const CntArr& cnt,
                const Grid& g, double dr, double ds, double aspect, int32_t* flag);
void launch_ratio(hipStream_t st, const CntArr& cnt, int32_t* nempty, const PtrArr& Cnew, const Grid& g, int ntype,
                  int comp);
void launch_conduct(hipStream_t st, const PtrArr& Cvis, double* K, double* Kr, double* Ks, const Grid& g, const FD& F,
                    double dr, double ds, int comp, int ntype, const RaCArr& lcf);
void launch_derivs(hipStream_t st, const double* u, const double* w, double* Du, double* Dw, const Grid& g,
                   const D3Arr& Dint, double dr);
void launch_tracer_stage(hipStream_t st, int stage, int torder, double dt, int64_t n, double* r, double* s, double* r0,
                         double* s0, double* r1, double* s1, double* r2, double* s2, const double* u, const double* w,
                         const double* Du, const double* Dw, const Grid& g, const D3Arr& Dint, double dr, double ds);
void launch_energy_stage(hipStream_t st, int stage, double dt, double* T, const double* T0, double* T1, double* T2,
                         const double* u, const double* w, const double* K, const double* Kr, const double* Ks,
                         const PtrArr& Cvis, const Grid& g, const FD& F, double dr, double ds, double dr2, double ds2,
                         int comp, int ntype, double H, const RaCArr& Htype);
void launch_sym_bcs(hipStream_t st, double* f, const Grid& g);
int launch_stability(hipStream_t st, const double* u, const double* w, const double* K, const double* Kr,
                     const double* Ks, const Grid& g, double TrA, double TsA, double TrrD, double TssD, double* partialA,
                     double* partialD);

}  // namespace tc
