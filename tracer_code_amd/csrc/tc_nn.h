// U-Net of the learned multigrid initial guess (tc_nn.hip).
#pragma once
#include <vector>

#include "tc_internal.h"

#define TC_NN_TRY(expr)              \
    do {                             \
        int _s = (expr);             \
        if (_s != 0) return _s;      \
    } while (0)

namespace tc {

constexpr int NN_IN = 6;          // input channels (stored padded to 8 per pixel)
constexpr int NN_NSCALE = 4;      // per-solve scales: max|psi_prev|, s*dT, dT, max|psi_prev - psi_prev2|

struct Layer {
    int cin, cout, lev;           // input/output channels, pyramid level (0 = full resolution)
    int cinp = 0, kp = 0;         // padded input channels (x8), padded GEMM depth 9*cinp (x32)
    int kpt = 0, cinp16 = 0;      // backward-data GEMM depth 9*cout (x32), cin padded to 16
    size_t woff = 0, boff = 0;    // offsets of W[cout][9][cin] and b[cout] in the parameter vector
    __bf16* wk = nullptr;         // packed forward weights [cout][kp]
    __bf16* wt = nullptr;         // packed backward-data weights [cinp16][kpt] (flipped taps)
    __bf16* z = nullptr;          // pre-activation [npix][cout]
    int nsplit = 1;               // pixel splits of the weight gradient
};

struct UNet {
    int H = 0, W = 0;
    int npx[4] = {0, 0, 0, 0};
    bool trainable = false;
    std::vector<Layer> L;
    size_t nparam = 0, head_off = 0;
    int step = 0;
    float inv_n = 0.0f;           // 1 / number of masked pixels of the loss
    float *w32 = nullptr, *g32 = nullptr, *m32 = nullptr, *v32 = nullptr;
    __bf16 *x0 = nullptr, *a1 = nullptr, *s1 = nullptr, *p1 = nullptr, *a2 = nullptr, *s2 = nullptr, *p2 = nullptr;
    __bf16 *a3 = nullptr, *s3 = nullptr, *p3 = nullptr, *a4 = nullptr, *b4 = nullptr, *d3a = nullptr, *d3b = nullptr;
    __bf16 *d2a = nullptr, *d2b = nullptr, *d1a = nullptr, *d1b = nullptr;
    float *ga1 = nullptr, *gs1 = nullptr, *gp1 = nullptr, *ga2 = nullptr, *gs2 = nullptr, *gp2 = nullptr;
    float *ga3 = nullptr, *gs3 = nullptr, *gp3 = nullptr, *ga4 = nullptr, *gb4 = nullptr, *gd3a = nullptr;
    float *gd3b = nullptr, *gd2a = nullptr, *gd2b = nullptr, *gd1a = nullptr, *gd1b = nullptr, *gx0 = nullptr;
    __bf16* dz = nullptr;
    float *wpart = nullptr, *gw32 = nullptr;
    double *hpart = nullptr, *loss_d = nullptr;
    float *out = nullptr, *target = nullptr, *mask = nullptr;
    std::vector<void*> owned;

    void layer_defs();
    size_t param_count() const;
    int alloc(int H, int W, bool train);
    void release();
    void pack(hipStream_t st);
    void conv(hipStream_t st, const Layer& l, const __bf16* in, int cs_in, int off_in, __bf16* o, int cs_out,
              int off_out);
    void forward(hipStream_t st);
    void conv_back(hipStream_t st, const Layer& l, float* gy, int cs_y, int off_y, const __bf16* in, int cs_in,
                   int off_in, float* gin, int cs_gin, int off_gin);
    double backward(hipStream_t st);
    void adam(hipStream_t st, float lr, float scale);
};

// kernels shared with the solver integration (tc_api.hip)
void launch_nn_input(hipStream_t st, const double* T, const double* C, const double* vis, const double* p1,
                     const double* p2, const double* Tp, const Grid& g, int W, int H, const double* scales, __bf16* x);
void launch_nn_target(hipStream_t st, const double* psi, const double* p1, const Grid& g, int W, int H,
                      const double* scales, float* t, float* mask);
void launch_nn_apply(hipStream_t st, const float* out, const double* p1, double* psi, const Grid& g, int W,
                     const double* scales, double zsg);
int nn_scale_blocks(const Grid& g);  // partials of launch_nn_scales: 3 doubles per block
void launch_nn_scales(hipStream_t st, const double* p1, const double* p2, const double* T, const double* Tp,
                      const Grid& g, double* partial, double* scales);

}  // namespace tc

// the opaque handle of the C-ABI
struct tc_nn {
    tc::UNet net;
    hipStream_t stream = nullptr;
    std::vector<__bf16> hx;
    std::string err;
};
