// Learned warm start: a 3D convolutional encoder-decoder that maps (filtered
// densities, load, clamped cells) on the cell grid to a nodal displacement
// guess, trained on the device with the BF16 MFMA kernels of k_conv.hip.
// No reference counterpart (TopOpt.jl starts CG from zero or, opt-in, from
// the previous solution: src/FEA/solvers_api.jl:193-195).
//
// Network (channels 16/32/64/128, 3x3x3 convolutions, SiLU):
//   enc L0: 8->16->16   pool  L1: 16->32->32   pool  L2: 32->64->64
//   pool L3: 64->128->128
//   dec L2: up(128) | skip 64 -> 64 -> 64      dec L1: up(64) | 32 -> 32 -> 32
//   dec L0: up(32) | 16 -> 16 -> 16            out: 2x2x2 conv, pad 1, cells -> nodes
// Skip tensors are written straight into the channel slice of the decoder's
// concatenation buffer, so no copies are made.  Targets are u / rms(u); in
// the solver the prediction is combined with the previous displacement by a
// 2 x 2 minimal-residual projection.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/topopt_hip.h"
#include "ctx_internal.h"
#include "k_common.h"
#include "nn_internal.h"

using namespace tohip;

namespace {

constexpr int kFeat = 8;   // input channels (6 used)
constexpr int kOutC = 8;   // output layer channels (3 used)
constexpr int kNL = 15;    // conv layers

struct NNLayer {
    int cin, cout, k, pad, T, cin32, cout32;
    size_t woff, boff;
    uint16_t* wf = nullptr;  // [cout][T][cin32]
    uint16_t* wt = nullptr;  // [cin][T][cout32] (tap flipped)
    bool act;
};

inline int up32(int c) { return (c + 31) / 32 * 32; }

}  // namespace

struct tohip_nn {
    hipStream_t stream = nullptr;
    std::string err;
    // solver contexts this network is attached to (tohip_nn_attach); their
    // ctx->nn is cleared when the network is destroyed
    std::vector<tohip_ctx*> attached;
    int nx = 0, ny = 0, nz = 0, maxB = 0, maxS = 0;
    int64_t P[4] = {0, 0, 0, 0};  // cells per sample at each level
    int64_t PN = 0;               // nodes per sample
    NNLayer L[kNL];
    size_t nparam = 0;
    float *params = nullptr, *grads = nullptr, *adm = nullptr, *adv = nullptr;
    long step = 0;
    float* part = nullptr;
    size_t part_cap = 0;
    uint8_t* fixmask = nullptr;  // per node
    // activations (BF16 NDHWC)
    uint16_t *feat = nullptr, *a00 = nullptr, *cat0 = nullptr, *p1 = nullptr, *a10 = nullptr,
             *cat1 = nullptr, *p2 = nullptr, *a20 = nullptr, *cat2 = nullptr, *p3 = nullptr,
             *a30 = nullptr, *e3 = nullptr, *ad20 = nullptr, *d2 = nullptr, *ad10 = nullptr,
             *d1 = nullptr, *ad00 = nullptr, *d0 = nullptr;
    uint16_t* Z[kNL] = {};
    float* out = nullptr;  // [B][nodes][3]
    // gradients
    uint16_t *dout = nullptr, *g1 = nullptr, *g2 = nullptr, *dcat0 = nullptr, *dcat1 = nullptr,
             *dcat2 = nullptr;
    // dataset
    int nsamples = 0;
    uint16_t* ds_feat = nullptr;  // [S][cells][kFeat]
    float* ds_u = nullptr;        // [S][nodes][3] normalised
    float* tgt = nullptr;         // [B][nodes][3]
    double *tmp_x = nullptr, *tmp_f = nullptr, *tmp_u = nullptr;
    double *lpart = nullptr, *ldev = nullptr, *h_loss = nullptr;
    size_t dev_bytes = 0;  // device memory held by this network
};

#define NCK(call)                                                                       \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            nn->err = std::string(#call) + ": " + hipGetErrorString(e_);                \
            return TOHIP_ERR_HIP;                                                       \
        }                                                                               \
    } while (0)

namespace tohip {

// ---------------------------------------------------------------- kernels --
__device__ __forceinline__ uint16_t f2bf_rne(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Input features on the cell grid (kFeat channels):
//   0: filtered density, 1..3: sum of the load over the cell's nodes,
//   4: 1 if a node of the cell is clamped, 5: 1 (constant), 6..7: 0.
__global__ void k_nn_features(int nx, int ny, int nz, const double* __restrict__ xt,
                              const double* __restrict__ f, const uint8_t* __restrict__ fixmask,
                              uint16_t* __restrict__ feat) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nc = (int64_t)nx * ny * nz;
    if (e >= nc) return;
    const int i = e % nx, j = (e / nx) % ny, k = e / ((int64_t)nx * ny);
    const int NX = nx + 1, NY = ny + 1;
    float fs[3] = {0.f, 0.f, 0.f};
    int clamp = 0;
    for (int c = 0; c < 8; ++c) {
        const int64_t n = (i + (c & 1)) + (int64_t)NX * ((j + ((c >> 1) & 1)) +
                                                         (int64_t)NY * (k + ((c >> 2) & 1)));
        for (int d = 0; d < 3; ++d) fs[d] += (float)f[3 * n + d];
        clamp |= fixmask[n] != 0;
    }
    uint16_t* o = feat + e * kFeat;
    o[0] = f2bf_rne((float)xt[e]);
    o[1] = f2bf_rne(fs[0]);
    o[2] = f2bf_rne(fs[1]);
    o[3] = f2bf_rne(fs[2]);
    o[4] = f2bf_rne(clamp ? 1.f : 0.f);
    o[5] = f2bf_rne(1.f);
    o[6] = 0;
    o[7] = 0;
}

// Normalised target u / rms(u) (FP32); rms from *ss = sum u^2.
__global__ void k_nn_target(int64_t n, const double* __restrict__ u, const double* __restrict__ ss,
                            float* __restrict__ t) {
    const double inv = *ss > 0 ? 1.0 / sqrt(*ss / (double)n) : 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        t[i] = (float)(u[i] * inv);
}

__global__ __launch_bounds__(256) void k_sumsq(int64_t n, const double* __restrict__ x,
                                               double* __restrict__ partials) {
    __shared__ double red[kWaves];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        s += x[i] * x[i];
    const double t = block_sum_all(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// MSE loss over 3 of kOutC output channels: dout (BF16, kOutC channels) =
// 2 (out - tgt) / count, per-block partial sums of the squared error.
__global__ __launch_bounds__(256) void k_nn_loss(int64_t npos, const float* __restrict__ out,
                                                 const float* __restrict__ tgt,
                                                 uint16_t* __restrict__ dout, double inv_count,
                                                 double* __restrict__ partials) {
    __shared__ double red[kWaves];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < npos;
         i += (int64_t)gridDim.x * blockDim.x) {
        for (int d = 0; d < kOutC; ++d) {
            float g = 0.f;
            if (d < 3) {
                const float r = out[i * 3 + d] - tgt[i * 3 + d];
                s += (double)r * r;
                g = (float)(2.0 * inv_count) * r;
            }
            dout[i * kOutC + d] = f2bf_rne(g);
        }
    }
    const double t = block_sum_all(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// five dot products of the warm-start combination:
// [v1.k1, v1.k2, v2.k2, v1.f, v2.f]
// The five dot products of the 2 x 2 combination and f.f (the residual ratio).
__global__ __launch_bounds__(256) void k_dots6(int64_t n, const double* __restrict__ v1,
                                               const double* __restrict__ k1,
                                               const double* __restrict__ v2,
                                               const double* __restrict__ k2,
                                               const double* __restrict__ f,
                                               double* __restrict__ partials) {
    __shared__ double red[kWaves];
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        s[0] += v1[i] * k1[i];
        s[1] += v1[i] * k2[i];
        s[2] += v2[i] * k2[i];
        s[3] += v1[i] * f[i];
        s[4] += v2[i] * f[i];
        s[5] += f[i] * f[i];
    }
    for (int q = 0; q < 6; ++q) {
        const double t = block_sum_all(s[q], red);
        if (threadIdx.x == 0) partials[(size_t)blockIdx.x * 6 + q] = t;
    }
}

// network output (FP32, nodes x 3 = internal DOF order) -> FP64 with the
// prescribed DOFs zeroed
__global__ void k_pred_to_f64(int64_t nnodes, const float* __restrict__ o,
                              const uint8_t* __restrict__ fixmask, double* __restrict__ v) {
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < nnodes;
         n += (int64_t)gridDim.x * blockDim.x)
        for (int d = 0; d < 3; ++d)
            v[3 * n + d] = ((fixmask[n] >> d) & 1) ? 0.0 : (double)o[3 * n + d];
}

// Coefficients of the warm start u0 = c1 v1 + c2 v2 from the dots
// d = (a11, a12, a22, b1, b2) of the 2x2 normal equations (one thread):
// the best point of span{v1, v2}, else the better single vector.
// coef[0], coef[1]: the combination; coef[2] (minimal-residual combination
// with a previous displacement): ||f - K u0|| of the combination over that of
// the best multiple of the previous displacement alone (1: the prediction
// adds nothing); 0 otherwise.
__global__ void k_ws_coeffs(int have_prev, int min_res, const double* __restrict__ d,
                            double* __restrict__ coef) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double a11 = d[0], a12 = d[1], a22 = d[2], b1 = d[3], b2 = d[4], ff = d[5];
    double c1 = 0, c2 = 0;
    const double det = a11 * a22 - a12 * a12;
    if (have_prev && det > 1e-12 * a11 * a22 && a11 > 0 && a22 > 0) {
        c1 = (b1 * a22 - b2 * a12) / det;
        c2 = (a11 * b2 - a12 * b1) / det;
    } else if (a22 > 0) {
        c2 = b2 / a22;  // best multiple of the prediction
        if (have_prev && a11 > 0 && b1 * b1 / a11 > b2 * b2 / a22) {
            c2 = 0;  // the previous u alone is better
            c1 = b1 / a11;
        }
    }
    coef[0] = c1;
    coef[1] = c2;
    double ratio = 0.0;
    if (min_res && have_prev && a11 > 0) {
        const double r_span = fmax(ff - (c1 * b1 + c2 * b2), 0.0);
        const double r_prev = fmax(ff - b1 * b1 / a11, 0.0);
        ratio = r_prev > 0 ? sqrt(r_span / r_prev) : 1.0;
    }
    coef[2] = ratio;
}

// out = coef[0] v1 + coef[1] v2 (out may alias v1 or v2: element-wise)
__global__ void k_combine(int64_t n, const double* __restrict__ coef, const double* v1,
                          const double* v2, double* out) {
    const double a = coef[0], b = coef[1];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = a * v1[i] + b * v2[i];
}

}  // namespace tohip

namespace {

unsigned nblk(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 4096));
}

ConvDesc conv_desc(const tohip_nn* nn, const NNLayer& l, int B, int level, bool node_out) {
    ConvDesc c{};
    c.B = B;
    c.D = nn->nz >> level;
    c.H = nn->ny >> level;
    c.W = nn->nx >> level;
    c.kd = c.kh = c.kw = l.k;
    c.pd = c.ph = c.pw = l.pad;
    c.OD = c.D + 2 * l.pad - l.k + 1;
    c.OH = c.H + 2 * l.pad - l.k + 1;
    c.OW = c.W + 2 * l.pad - l.k + 1;
    c.Cin = l.cin;
    c.Cin32 = l.cin32;
    c.Cout = l.cout;
    c.out_cstride = l.cout;
    c.out_coff = 0;
    c.out_cvalid = node_out ? 3 : l.cout;
    c.act = l.act ? 1 : 0;
    return c;
}

// Backward-data conv of layer l: input dY (output grid of l, cout channels),
// output dX (input grid of l, cin channels).
ConvDesc bwd_desc(const tohip_nn* nn, const NNLayer& l, int B, int level) {
    ConvDesc f = conv_desc(nn, l, B, level, false);
    ConvDesc c{};
    c.B = B;
    c.D = f.OD;
    c.H = f.OH;
    c.W = f.OW;
    c.kd = c.kh = c.kw = l.k;
    c.pd = c.ph = c.pw = l.k - 1 - l.pad;
    c.OD = f.D;
    c.OH = f.H;
    c.OW = f.W;
    c.Cin = l.cout;
    c.Cin32 = l.cout32;
    c.Cout = l.cin;
    c.out_cstride = l.cin;
    c.out_coff = 0;
    c.out_cvalid = l.cin;
    c.act = 0;
    return c;
}

int nn_upload_weights(tohip_nn* nn) {
    for (int i = 0; i < kNL; ++i) {
        NNLayer& l = nn->L[i];
        NCK(launch_weights_to_bf16(l.cout, l.T, l.cin, l.cin32, l.cout32, nn->params + l.woff,
                                   l.wf, l.wt, nn->stream));
    }
    return 0;
}

template <typename T>
int nalloc(tohip_nn* nn, T** p, size_t n) {
    NCK(hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)));
    nn->dev_bytes += std::max<size_t>(n, 1) * sizeof(T);
    return 0;
}

// Forward pass on B samples already in nn->feat.  train: keep pre-activations.
int nn_forward(tohip_nn* nn, int B, bool train) {
    hipStream_t s = nn->stream;
    auto conv = [&](int li, int level, const uint16_t* X, uint16_t* Y, int ocs, int oco) -> int {
        NNLayer& l = nn->L[li];
        ConvDesc c = conv_desc(nn, l, B, level, false);
        c.out_cstride = ocs;
        c.out_coff = oco;
        NCK(launch_conv_fwd(c, X, l.wf, nn->params + l.boff, Y, train ? nn->Z[li] : nullptr,
                            nullptr, s));
        return 0;
    };
    const int D0 = nn->nz, H0 = nn->ny, W0 = nn->nx;
    int rc;
    if ((rc = conv(0, 0, nn->feat, nn->a00, 16, 0))) return rc;
    if ((rc = conv(1, 0, nn->a00, nn->cat0, 48, 32))) return rc;
    NCK(launch_avgpool_fwd(B, D0, H0, W0, 16, nn->cat0, 48, 32, nn->p1, s));
    if ((rc = conv(2, 1, nn->p1, nn->a10, 32, 0))) return rc;
    if ((rc = conv(3, 1, nn->a10, nn->cat1, 96, 64))) return rc;
    NCK(launch_avgpool_fwd(B, D0 >> 1, H0 >> 1, W0 >> 1, 32, nn->cat1, 96, 64, nn->p2, s));
    if ((rc = conv(4, 2, nn->p2, nn->a20, 64, 0))) return rc;
    if ((rc = conv(5, 2, nn->a20, nn->cat2, 192, 128))) return rc;
    NCK(launch_avgpool_fwd(B, D0 >> 2, H0 >> 2, W0 >> 2, 64, nn->cat2, 192, 128, nn->p3, s));
    if ((rc = conv(6, 3, nn->p3, nn->a30, 128, 0))) return rc;
    if ((rc = conv(7, 3, nn->a30, nn->e3, 128, 0))) return rc;
    NCK(launch_upsample_fwd(B, D0 >> 3, H0 >> 3, W0 >> 3, 128, nn->e3, nn->cat2, 192, 0, s));
    if ((rc = conv(8, 2, nn->cat2, nn->ad20, 64, 0))) return rc;
    if ((rc = conv(9, 2, nn->ad20, nn->d2, 64, 0))) return rc;
    NCK(launch_upsample_fwd(B, D0 >> 2, H0 >> 2, W0 >> 2, 64, nn->d2, nn->cat1, 96, 0, s));
    if ((rc = conv(10, 1, nn->cat1, nn->ad10, 32, 0))) return rc;
    if ((rc = conv(11, 1, nn->ad10, nn->d1, 32, 0))) return rc;
    NCK(launch_upsample_fwd(B, D0 >> 1, H0 >> 1, W0 >> 1, 32, nn->d1, nn->cat0, 48, 0, s));
    if ((rc = conv(12, 0, nn->cat0, nn->ad00, 16, 0))) return rc;
    if ((rc = conv(13, 0, nn->ad00, nn->d0, 16, 0))) return rc;
    {
        NNLayer& l = nn->L[14];
        ConvDesc c = conv_desc(nn, l, B, 0, true);
        NCK(launch_conv_fwd(c, nn->d0, l.wf, nn->params + l.boff, nullptr, nullptr, nn->out, s));
    }
    return 0;
}

int split_for(int64_t P) { return (int)std::max<int64_t>(1, std::min<int64_t>(64, P / 2048)); }

// weight/bias gradient of layer li (input X, output gradient dZ)
int nn_wgrad(tohip_nn* nn, int li, int level, int B, const uint16_t* X, const uint16_t* dZ) {
    NNLayer& l = nn->L[li];
    ConvDesc c = conv_desc(nn, l, B, level, false);
    const int64_t P = (int64_t)B * c.OD * c.OH * c.OW;
    const int ns = split_for(P);
    if ((size_t)ns * l.cout * l.T * l.cin32 > nn->part_cap) {
        nn->err = "weight-gradient scratch too small";
        return TOHIP_ERR_STATE;
    }
    // dW into a [cout][T][cin32] scratch, then compact into grads [cout][T][cin]
    float* dw32 = nn->part + (size_t)ns * l.cout * l.T * l.cin32;
    NCK(launch_conv_bwd_w(c, X, dZ, nn->part, ns, dw32, nn->grads + l.boff, nn->stream));
    NCK(hipMemcpy2DAsync(nn->grads + l.woff, l.cin * sizeof(float), dw32, l.cin32 * sizeof(float),
                         l.cin * sizeof(float), (size_t)l.cout * l.T, hipMemcpyDeviceToDevice,
                         nn->stream));
    return 0;
}

// dX of layer li from its output gradient dZ
int nn_dgrad(tohip_nn* nn, int li, int level, int B, const uint16_t* dZ, uint16_t* dX) {
    NNLayer& l = nn->L[li];
    ConvDesc c = bwd_desc(nn, l, B, level);
    NCK(launch_conv_fwd(c, dZ, l.wt, nullptr, dX, nullptr, nullptr, nn->stream));
    return 0;
}

// Backward pass for B samples (after nn_forward(train) and the loss).
int nn_backward(tohip_nn* nn, int B) {
    hipStream_t s = nn->stream;
    const int D0 = nn->nz, H0 = nn->ny, W0 = nn->nx;
    const int64_t P0 = B * nn->P[0], P1 = B * nn->P[1], P2 = B * nn->P[2], P3 = B * nn->P[3];
    int rc;
    uint16_t *g1 = nn->g1, *g2 = nn->g2;
    // output layer (identity activation)
    {
        NNLayer& l = nn->L[14];
        ConvDesc c = conv_desc(nn, l, B, 0, true);
        const int64_t P = (int64_t)B * c.OD * c.OH * c.OW;
        const int ns = split_for(P);
        float* dw32 = nn->part + (size_t)ns * l.cout * l.T * l.cin32;
        NCK(launch_conv_bwd_w(c, nn->d0, nn->dout, nn->part, ns, dw32, nn->grads + l.boff, s));
        NCK(hipMemcpy2DAsync(nn->grads + l.woff, l.cin * sizeof(float), dw32,
                             l.cin32 * sizeof(float), l.cin * sizeof(float), (size_t)l.cout * l.T,
                             hipMemcpyDeviceToDevice, s));
        if ((rc = nn_dgrad(nn, 14, 0, B, nn->dout, g1))) return rc;  // d d0
    }
    // a (silu) layer: dZ = silu'(Z) dA into gZ, wgrad, dgrad into gX
    auto layer = [&](int li, int level, int64_t P, int cout, const uint16_t* dA, uint16_t* gZ,
                     const uint16_t* X, uint16_t* gX) -> int {
        NCK(launch_silu_bwd(P * cout, dA, nn->Z[li], gZ, s));
        int r;
        if ((r = nn_wgrad(nn, li, level, B, X, gZ))) return r;
        if (gX)
            if ((r = nn_dgrad(nn, li, level, B, gZ, gX))) return r;
        return 0;
    };
    // decoder L0
    if ((rc = layer(13, 0, P0, 16, g1, g2, nn->ad00, g1))) return rc;
    if ((rc = layer(12, 0, P0, 16, g1, g2, nn->cat0, nn->dcat0))) return rc;
    NCK(launch_upsample_bwd(B, D0 >> 1, H0 >> 1, W0 >> 1, 32, nn->dcat0, 48, 0, g1, s));
    // decoder L1
    if ((rc = layer(11, 1, P1, 32, g1, g2, nn->ad10, g1))) return rc;
    if ((rc = layer(10, 1, P1, 32, g1, g2, nn->cat1, nn->dcat1))) return rc;
    NCK(launch_upsample_bwd(B, D0 >> 2, H0 >> 2, W0 >> 2, 64, nn->dcat1, 96, 0, g1, s));
    // decoder L2
    if ((rc = layer(9, 2, P2, 64, g1, g2, nn->ad20, g1))) return rc;
    if ((rc = layer(8, 2, P2, 64, g1, g2, nn->cat2, nn->dcat2))) return rc;
    NCK(launch_upsample_bwd(B, D0 >> 3, H0 >> 3, W0 >> 3, 128, nn->dcat2, 192, 0, g1, s));
    // encoder L3
    if ((rc = layer(7, 3, P3, 128, g1, g2, nn->a30, g1))) return rc;
    if ((rc = layer(6, 3, P3, 128, g1, g2, nn->p3, g1))) return rc;  // g1 = d p3 (64 ch)
    NCK(launch_avgpool_bwd(B, D0 >> 2, H0 >> 2, W0 >> 2, 64, g1, nn->dcat2, 192, 128, 1, s));
    NCK(launch_slice_copy(P2, 64, nn->dcat2, 192, 128, g1, 0, s));  // d e2
    // encoder L2
    if ((rc = layer(5, 2, P2, 64, g1, g2, nn->a20, g1))) return rc;
    if ((rc = layer(4, 2, P2, 64, g1, g2, nn->p2, g1))) return rc;  // d p2 (32 ch)
    NCK(launch_avgpool_bwd(B, D0 >> 1, H0 >> 1, W0 >> 1, 32, g1, nn->dcat1, 96, 64, 1, s));
    NCK(launch_slice_copy(P1, 32, nn->dcat1, 96, 64, g1, 0, s));  // d e1
    // encoder L1
    if ((rc = layer(3, 1, P1, 32, g1, g2, nn->a10, g1))) return rc;
    if ((rc = layer(2, 1, P1, 32, g1, g2, nn->p1, g1))) return rc;  // d p1 (16 ch)
    NCK(launch_avgpool_bwd(B, D0, H0, W0, 16, g1, nn->dcat0, 48, 32, 1, s));
    NCK(launch_slice_copy(P0, 16, nn->dcat0, 48, 32, g1, 0, s));  // d e0
    // encoder L0
    if ((rc = layer(1, 0, P0, 16, g1, g2, nn->a00, g1))) return rc;
    if ((rc = layer(0, 0, P0, 16, g1, g2, nn->feat, nullptr))) return rc;
    return 0;
}

void nn_free(tohip_nn* nn) {
    if (!nn) return;
    for (tohip_ctx* c : nn->attached) c->nn = nullptr;  // no dangling warm-start pointer
    std::vector<void*> bufs = {nn->params, nn->grads, nn->adm, nn->adv, nn->part, nn->fixmask,
                               nn->feat, nn->a00, nn->cat0, nn->p1, nn->a10, nn->cat1, nn->p2,
                               nn->a20, nn->cat2, nn->p3, nn->a30, nn->e3, nn->ad20, nn->d2,
                               nn->ad10, nn->d1, nn->ad00, nn->d0, nn->out, nn->dout, nn->g1,
                               nn->g2, nn->dcat0, nn->dcat1, nn->dcat2, nn->ds_feat, nn->ds_u,
                               nn->tgt, nn->tmp_x, nn->tmp_f, nn->tmp_u, nn->lpart, nn->ldev};
    for (int i = 0; i < kNL; ++i) {
        bufs.push_back(nn->L[i].wf);
        bufs.push_back(nn->L[i].wt);
        bufs.push_back(nn->Z[i]);
    }
    for (void* b : bufs)
        if (b) hipFree(b);
    if (nn->h_loss) hipHostFree(nn->h_loss);
    if (nn->stream) hipStreamDestroy(nn->stream);
    delete nn;
}

int nn_create_impl(tohip_nn* nn, const int64_t* nels, int maxB, int maxS, uint64_t seed,
                   const uint8_t* fixed) {
    if (!nels || maxB < 1 || maxS < 0 || !fixed) {
        nn->err = "tohip_nn_create: bad arguments";
        return TOHIP_ERR_ARG;
    }
    for (int d = 0; d < 3; ++d)
        if (nels[d] < 8 || nels[d] % 8 != 0 || nels[d] > 4096) {
            nn->err = "tohip_nn_create: cell counts must be multiples of 8 (three 2x poolings)";
            return TOHIP_ERR_ARG;
        }
    nn->nx = (int)nels[0];
    nn->ny = (int)nels[1];
    nn->nz = (int)nels[2];
    nn->maxB = maxB;
    nn->maxS = maxS;
    for (int l = 0; l < 4; ++l)
        nn->P[l] = (int64_t)(nn->nx >> l) * (nn->ny >> l) * (nn->nz >> l);
    nn->PN = (int64_t)(nn->nx + 1) * (nn->ny + 1) * (nn->nz + 1);
    const int spec[kNL][4] = {  // cin, cout, k, pad
        {kFeat, 16, 3, 1}, {16, 16, 3, 1}, {16, 32, 3, 1},  {32, 32, 3, 1},  {32, 64, 3, 1},
        {64, 64, 3, 1},    {64, 128, 3, 1}, {128, 128, 3, 1}, {192, 64, 3, 1}, {64, 64, 3, 1},
        {96, 32, 3, 1},    {32, 32, 3, 1},  {48, 16, 3, 1},  {16, 16, 3, 1},  {16, kOutC, 2, 1}};
    size_t off = 0;
    for (int i = 0; i < kNL; ++i) {
        NNLayer& l = nn->L[i];
        l.cin = spec[i][0];
        l.cout = spec[i][1];
        l.k = spec[i][2];
        l.pad = spec[i][3];
        l.T = l.k * l.k * l.k;
        l.cin32 = up32(l.cin);
        l.cout32 = up32(l.cout);
        l.act = i < kNL - 1;
        l.woff = off;
        off += (size_t)l.cout * l.T * l.cin;
        l.boff = off;
        off += l.cout;
    }
    nn->nparam = off;
    NCK(hipStreamCreateWithFlags(&nn->stream, hipStreamNonBlocking));
    int rc;
    const size_t np = nn->nparam;
    if ((rc = nalloc(nn, &nn->params, np)) || (rc = nalloc(nn, &nn->grads, np)) ||
        (rc = nalloc(nn, &nn->adm, np)) || (rc = nalloc(nn, &nn->adv, np)))
        return rc;
    size_t pc = 0;
    for (int i = 0; i < kNL; ++i) {
        NNLayer& l = nn->L[i];
        if ((rc = nalloc(nn, &l.wf, (size_t)l.cout * l.T * l.cin32)) ||
            (rc = nalloc(nn, &l.wt, (size_t)l.cin * l.T * l.cout32)))
            return rc;
        pc = std::max(pc, (size_t)65 * l.cout * l.T * l.cin32);
    }
    nn->part_cap = pc;
    if ((rc = nalloc(nn, &nn->part, pc + 1))) return rc;
    const int64_t B = maxB, P0 = B * nn->P[0], P1 = B * nn->P[1], P2 = B * nn->P[2],
                  P3 = B * nn->P[3], PN = B * nn->PN;
    if ((rc = nalloc(nn, &nn->fixmask, nn->PN)) || (rc = nalloc(nn, &nn->feat, P0 * kFeat)) ||
        (rc = nalloc(nn, &nn->a00, P0 * 16)) || (rc = nalloc(nn, &nn->cat0, P0 * 48)) ||
        (rc = nalloc(nn, &nn->p1, P1 * 16)) || (rc = nalloc(nn, &nn->a10, P1 * 32)) ||
        (rc = nalloc(nn, &nn->cat1, P1 * 96)) || (rc = nalloc(nn, &nn->p2, P2 * 32)) ||
        (rc = nalloc(nn, &nn->a20, P2 * 64)) || (rc = nalloc(nn, &nn->cat2, P2 * 192)) ||
        (rc = nalloc(nn, &nn->p3, P3 * 64)) || (rc = nalloc(nn, &nn->a30, P3 * 128)) ||
        (rc = nalloc(nn, &nn->e3, P3 * 128)) || (rc = nalloc(nn, &nn->ad20, P2 * 64)) ||
        (rc = nalloc(nn, &nn->d2, P2 * 64)) || (rc = nalloc(nn, &nn->ad10, P1 * 32)) ||
        (rc = nalloc(nn, &nn->d1, P1 * 32)) || (rc = nalloc(nn, &nn->ad00, P0 * 16)) ||
        (rc = nalloc(nn, &nn->d0, P0 * 16)) || (rc = nalloc(nn, &nn->out, PN * 3)) ||
        (rc = nalloc(nn, &nn->dout, PN * kOutC)) || (rc = nalloc(nn, &nn->dcat0, P0 * 48)) ||
        (rc = nalloc(nn, &nn->dcat1, P1 * 96)) || (rc = nalloc(nn, &nn->dcat2, P2 * 192)) ||
        (rc = nalloc(nn, &nn->tgt, PN * 3)) || (rc = nalloc(nn, &nn->lpart, 4096 * 5)) ||
        (rc = nalloc(nn, &nn->ldev, 8)) || (rc = nalloc(nn, &nn->tmp_x, nn->P[0])) ||
        (rc = nalloc(nn, &nn->tmp_f, nn->PN * 3)) || (rc = nalloc(nn, &nn->tmp_u, nn->PN * 3)))
        return rc;
    const int64_t gs = std::max<int64_t>(P0 * 16, PN * kOutC);
    if ((rc = nalloc(nn, &nn->g1, gs)) || (rc = nalloc(nn, &nn->g2, gs))) return rc;
    const int zlev[kNL] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0};
    for (int i = 0; i < kNL - 1; ++i)
        if ((rc = nalloc(nn, &nn->Z[i], B * nn->P[zlev[i]] * nn->L[i].cout))) return rc;
    if (maxS > 0) {
        if ((rc = nalloc(nn, &nn->ds_feat, (size_t)maxS * nn->P[0] * kFeat)) ||
            (rc = nalloc(nn, &nn->ds_u, (size_t)maxS * nn->PN * 3)))
            return rc;
    }
    NCK(hipHostMalloc((void**)&nn->h_loss, 8 * sizeof(double), hipHostMallocDefault));
    NCK(hipMemcpyAsync(nn->fixmask, fixed, nn->PN, hipMemcpyHostToDevice, nn->stream));
    // He initialisation (host RNG, seeded), biases 0; last layer small and
    // zero on the unused output channels.
    std::vector<float> h(np, 0.f);
    std::mt19937_64 rng(seed);
    for (int i = 0; i < kNL; ++i) {
        NNLayer& l = nn->L[i];
        const double sd = std::sqrt(2.0 / (double)(l.T * l.cin)) * (i == kNL - 1 ? 0.1 : 1.0);
        std::normal_distribution<double> nd(0.0, sd);
        for (int co = 0; co < l.cout; ++co)
            for (int t = 0; t < l.T; ++t)
                for (int ci = 0; ci < l.cin; ++ci) {
                    float v = (float)nd(rng);
                    if (i == kNL - 1 && co >= 3) v = 0.f;
                    h[l.woff + ((size_t)co * l.T + t) * l.cin + ci] = v;
                }
    }
    NCK(hipMemcpyAsync(nn->params, h.data(), np * sizeof(float), hipMemcpyHostToDevice,
                       nn->stream));
    NCK(hipMemsetAsync(nn->adm, 0, np * sizeof(float), nn->stream));
    NCK(hipMemsetAsync(nn->adv, 0, np * sizeof(float), nn->stream));
    NCK(hipMemsetAsync(nn->grads, 0, np * sizeof(float), nn->stream));
    if ((rc = nn_upload_weights(nn))) return rc;
    NCK(hipStreamSynchronize(nn->stream));
    return 0;
}

// features for B samples: xt (cells), f (nodes x 3) given as device pointers
int nn_build_features(tohip_nn* nn, int b, const double* xt, const double* f,
                      uint16_t* dst_feat) {
    NCK(hipGetLastError());
    hipLaunchKernelGGL(k_nn_features, dim3(nblk(nn->P[0])), dim3(256), 0, nn->stream, nn->nx,
                       nn->ny, nn->nz, xt, f, nn->fixmask, dst_feat + (size_t)b * nn->P[0] * kFeat);
    NCK(hipGetLastError());
    return 0;
}

thread_local std::string g_nn_error;

}  // namespace

// ------------------------------------------------ warm start in the solver --
namespace tohip {

size_t nn_device_bytes(const tohip_nn* nn) { return nn ? nn->dev_bytes : 0; }

void nn_detach_ctx(tohip_nn* nn, tohip_ctx* ctx) {
    if (!nn) return;
    auto& a = nn->attached;
    a.erase(std::remove(a.begin(), a.end(), ctx), a.end());
}

int nn_warm_start(tohip_ctx* ctx, bool have_prev) {
    tohip_nn* nn = ctx->nn;
    if (!nn) {
        ctx->err = "warm_start = 2 needs a network (tohip_nn_attach)";
        return TOHIP_ERR_STATE;
    }
    hipStream_t s = ctx->stream;
    const int64_t n = ctx->g.ndof;
    auto ck = [&](hipError_t e) {
        if (e != hipSuccess) ctx->err = std::string("nn warm start: ") + hipGetErrorString(e);
        return e != hipSuccess;
    };
    // The network runs on its own stream; events order it after the solver's
    // work and the solver after it, with no host synchronisation.
    for (hipEvent_t& e : ctx->ws_ev)
        if (!e && ck(hipEventCreateWithFlags(&e, hipEventDisableTiming))) return TOHIP_ERR_HIP;
    if (ck(hipEventRecord(ctx->ws_ev[0], s)) ||
        ck(hipStreamWaitEvent(nn->stream, ctx->ws_ev[0], 0)))
        return TOHIP_ERR_HIP;
    if (nn_build_features(nn, 0, ctx->xt, ctx->f, nn->feat)) {
        ctx->err = nn->err;
        return TOHIP_ERR_HIP;
    }
    if (nn_forward(nn, 1, false)) {
        ctx->err = nn->err;
        return TOHIP_ERR_HIP;
    }
    if (ck(hipEventRecord(ctx->ws_ev[1], nn->stream)) ||
        ck(hipStreamWaitEvent(s, ctx->ws_ev[1], 0)))
        return TOHIP_ERR_HIP;
    // v2 = prediction (r64), K v2 (q64); v1 = previous u, K v1 (t64)
    hipLaunchKernelGGL(k_pred_to_f64, dim3(nblk(ctx->g.nnodes)), dim3(256), 0, s, ctx->g.nnodes,
                       nn->out, ctx->fixmask, ctx->r64);
    if (ck(launch_matvec<double>(ctx->g, ctx->op64, ctx->r64, ctx->fixmask, ctx->E64, ctx->q64,
                                 ctx->meandiag, s)))
        return TOHIP_ERR_HIP;
    const double* v1 = have_prev ? ctx->u : ctx->r64;
    if (have_prev &&
        ck(launch_matvec<double>(ctx->g, ctx->op64, ctx->u, ctx->fixmask, ctx->E64, ctx->t64,
                                 ctx->meandiag, s)))
        return TOHIP_ERR_HIP;
    // Galerkin (energy-optimal) combination: dots of v_i with K v_j; minimal
    // residual combination: dots of K v_i with K v_j (smallest ||f - K u0||
    // on the span, which is what the residual-based stopping rule sees).
    // TOHIP_WS_COMBINE=galerkin selects the former (measurement).
    static const bool galerkin = [] {
        const char* e = getenv("TOHIP_WS_COMBINE");
        return e && std::string(e) == "galerkin";
    }();
    const double* kv1 = have_prev ? ctx->t64 : ctx->q64;
    hipLaunchKernelGGL(k_dots6, dim3(kVecBlocks), dim3(256), 0, s, n, galerkin ? v1 : kv1, kv1,
                       galerkin ? ctx->r64 : ctx->q64, ctx->q64, ctx->b64, ctx->partials);
    if (ck(launch_reduce(ctx->partials, kVecBlocks, 6, ctx->dscal + 3, s))) return TOHIP_ERR_HIP;
    hipLaunchKernelGGL(k_ws_coeffs, dim3(1), dim3(64), 0, s, have_prev ? 1 : 0, galerkin ? 0 : 1,
                       ctx->dscal + 3, ctx->dscal + kWsCoef);
    // u0 written in place over u (without a previous u, v1 is the prediction)
    hipLaunchKernelGGL(k_combine, dim3(kVecBlocks), dim3(256), 0, s, n, ctx->dscal + kWsCoef, v1,
                       ctx->r64, ctx->u);
    if (ck(hipGetLastError())) return TOHIP_ERR_HIP;
    return 0;
}

}  // namespace tohip

extern "C" {

const char* tohip_nn_last_error(const tohip_nn* nn) {
    return nn ? nn->err.c_str() : g_nn_error.c_str();
}

int tohip_nn_create(tohip_nn** out, const int64_t* nels, int max_batch, int max_samples,
                    uint64_t seed, const uint8_t* fixed_mask) {
    if (!out) return TOHIP_ERR_ARG;
    *out = nullptr;
    tohip_nn* nn = new (std::nothrow) tohip_nn();
    if (!nn) return TOHIP_ERR_HIP;
    // fixed_mask: ndofs bytes (node-major), reduced to a per-node bit mask
    std::vector<uint8_t> nodemask;
    if (nels && fixed_mask) {
        const int64_t nn_nodes = (nels[0] + 1) * (nels[1] + 1) * (nels[2] + 1);
        nodemask.assign(nn_nodes, 0);
        for (int64_t i = 0; i < nn_nodes; ++i)
            for (int d = 0; d < 3; ++d)
                if (fixed_mask[3 * i + d]) nodemask[i] |= (uint8_t)(1u << d);
    }
    int rc = nn_create_impl(nn, nels, max_batch, max_samples, seed,
                            nodemask.empty() ? nullptr : nodemask.data());
    if (rc) {
        g_nn_error = nn->err;
        nn_free(nn);
        return rc;
    }
    *out = nn;
    return 0;
}

void tohip_nn_destroy(tohip_nn* nn) { nn_free(nn); }

int64_t tohip_nn_num_params(const tohip_nn* nn) { return nn ? (int64_t)nn->nparam : -1; }

int tohip_nn_get_params(tohip_nn* nn, float* out) {
    if (!nn || !out) return TOHIP_ERR_ARG;
    NCK(hipMemcpyAsync(out, nn->params, nn->nparam * sizeof(float), hipMemcpyDeviceToHost,
                       nn->stream));
    NCK(hipStreamSynchronize(nn->stream));
    return 0;
}

int tohip_nn_set_params(tohip_nn* nn, const float* in) {
    if (!nn || !in) return TOHIP_ERR_ARG;
    NCK(hipMemcpyAsync(nn->params, in, nn->nparam * sizeof(float), hipMemcpyHostToDevice,
                       nn->stream));
    int rc = nn_upload_weights(nn);
    if (rc) return rc;
    NCK(hipStreamSynchronize(nn->stream));
    return 0;
}

int tohip_nn_clear_samples(tohip_nn* nn) {
    if (!nn) return TOHIP_ERR_ARG;
    nn->nsamples = 0;
    return 0;
}

int tohip_nn_num_samples(const tohip_nn* nn) { return nn ? nn->nsamples : -1; }

int tohip_nn_add_sample(tohip_nn* nn, const double* xphys, const double* f, const double* u) {
    if (!nn || !xphys || !f || !u) return TOHIP_ERR_ARG;
    if (nn->nsamples >= nn->maxS) {
        nn->err = "tohip_nn_add_sample: dataset is full";
        return TOHIP_ERR_STATE;
    }
    hipStream_t s = nn->stream;
    const int64_t nd = nn->PN * 3;
    NCK(hipMemcpyAsync(nn->tmp_x, xphys, nn->P[0] * sizeof(double), hipMemcpyHostToDevice, s));
    NCK(hipMemcpyAsync(nn->tmp_f, f, nd * sizeof(double), hipMemcpyHostToDevice, s));
    NCK(hipMemcpyAsync(nn->tmp_u, u, nd * sizeof(double), hipMemcpyHostToDevice, s));
    int rc = nn_build_features(nn, nn->nsamples, nn->tmp_x, nn->tmp_f, nn->ds_feat);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sumsq, dim3(kVecBlocks), dim3(256), 0, s, nd, nn->tmp_u, nn->lpart);
    NCK(launch_reduce(nn->lpart, kVecBlocks, 1, nn->ldev, s));
    hipLaunchKernelGGL(k_nn_target, dim3(nblk(nd)), dim3(256), 0, s, nd, nn->tmp_u, nn->ldev,
                       nn->ds_u + (size_t)nn->nsamples * nd);
    NCK(hipStreamSynchronize(s));
    nn->nsamples += 1;
    return 0;
}

// Loss (and, when train, the gradient step) on the given dataset samples.
static int nn_batch(tohip_nn* nn, const int* idx, int B, bool train, double lr, double* loss) {
    hipStream_t s = nn->stream;
    const int64_t nd = nn->PN * 3;
    for (int b = 0; b < B; ++b) {
        NCK(hipMemcpyAsync(nn->feat + (size_t)b * nn->P[0] * kFeat,
                           nn->ds_feat + (size_t)idx[b] * nn->P[0] * kFeat,
                           nn->P[0] * kFeat * sizeof(uint16_t), hipMemcpyDeviceToDevice, s));
        NCK(hipMemcpyAsync(nn->tgt + (size_t)b * nd, nn->ds_u + (size_t)idx[b] * nd,
                           nd * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    int rc = nn_forward(nn, B, train);
    if (rc) return rc;
    const int64_t npos = (int64_t)B * nn->PN;
    hipLaunchKernelGGL(k_nn_loss, dim3(kVecBlocks), dim3(256), 0, s, npos, nn->out, nn->tgt,
                       nn->dout, 1.0 / (double)(npos * 3), nn->lpart);
    NCK(launch_reduce(nn->lpart, kVecBlocks, 1, nn->ldev, s));
    if (train) {
        if ((rc = nn_backward(nn, B))) return rc;
        nn->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)nn->step), c2 = 1.f - std::pow(b2, (float)nn->step);
        hipLaunchKernelGGL(k_adam, dim3(nblk(nn->nparam)), dim3(256), 0, s, (int64_t)nn->nparam,
                           nn->params, nn->grads, nn->adm, nn->adv, (float)lr, b1, b2, 1e-8f, c1,
                           c2);
        NCK(hipGetLastError());
        if ((rc = nn_upload_weights(nn))) return rc;
    }
    NCK(hipMemcpyAsync(nn->h_loss, nn->ldev, sizeof(double), hipMemcpyDeviceToHost, s));
    NCK(hipStreamSynchronize(s));
    if (loss) *loss = nn->h_loss[0] / (double)(npos * 3);
    return 0;
}

int tohip_nn_train(tohip_nn* nn, int first, int count, int steps, int batch, double lr,
                   uint64_t seed, double* loss_hist) {
    if (!nn || steps < 0 || batch < 1 || batch > nn->maxB || !(lr > 0)) return TOHIP_ERR_ARG;
    if (first < 0 || count < 1 || first + count > nn->nsamples) {
        nn->err = "tohip_nn_train: sample range outside the dataset";
        return TOHIP_ERR_ARG;
    }
    std::mt19937_64 rng(seed);
    std::vector<int> perm(count);
    for (int i = 0; i < count; ++i) perm[i] = first + i;
    size_t pos = perm.size();
    std::vector<int> idx(batch);
    for (int st = 0; st < steps; ++st) {
        for (int b = 0; b < batch; ++b) {
            if (pos >= perm.size()) {
                std::shuffle(perm.begin(), perm.end(), rng);
                pos = 0;
            }
            idx[b] = perm[pos++];
        }
        double l = 0;
        int rc = nn_batch(nn, idx.data(), batch, true, lr, &l);
        if (rc) return rc;
        if (loss_hist) loss_hist[st] = l;
    }
    return 0;
}

int tohip_nn_eval(tohip_nn* nn, int first, int count, double* loss) {
    if (!nn || !loss || first < 0 || count < 1 || first + count > nn->nsamples) return TOHIP_ERR_ARG;
    double tot = 0;
    for (int i = first; i < first + count;) {
        const int B = std::min(nn->maxB, first + count - i);
        std::vector<int> idx(B);
        for (int b = 0; b < B; ++b) idx[b] = i + b;
        double l = 0;
        int rc = nn_batch(nn, idx.data(), B, false, 0.0, &l);
        if (rc) return rc;
        tot += l * B;
        i += B;
    }
    *loss = tot / count;
    return 0;
}

int tohip_nn_predict(tohip_nn* nn, const double* xphys, const double* f, float* u_out) {
    if (!nn || !xphys || !f || !u_out) return TOHIP_ERR_ARG;
    hipStream_t s = nn->stream;
    NCK(hipMemcpyAsync(nn->tmp_x, xphys, nn->P[0] * sizeof(double), hipMemcpyHostToDevice, s));
    NCK(hipMemcpyAsync(nn->tmp_f, f, nn->PN * 3 * sizeof(double), hipMemcpyHostToDevice, s));
    int rc = nn_build_features(nn, 0, nn->tmp_x, nn->tmp_f, nn->feat);
    if (rc) return rc;
    if ((rc = nn_forward(nn, 1, false))) return rc;
    NCK(hipMemcpyAsync(u_out, nn->out, nn->PN * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
    NCK(hipStreamSynchronize(s));
    return 0;
}

int tohip_nn_attach(tohip_ctx* ctx, tohip_nn* nn) {
    if (!ctx) return TOHIP_ERR_ARG;
    if (nn && (ctx->g.dim != 3 || ctx->g.nx != nn->nx || ctx->g.ny != nn->ny ||
               ctx->g.nz != nn->nz)) {
        ctx->err = "tohip_nn_attach: network grid differs from the solver grid";
        return TOHIP_ERR_ARG;
    }
    if (ctx->nn == nn) return 0;
    if (ctx->nn) nn_detach_ctx(ctx->nn, ctx);
    ctx->nn = nn;
    if (nn) nn->attached.push_back(ctx);
    return 0;
}

int tohip_nn_calls(const tohip_ctx* ctx, int64_t* calls) {
    if (!ctx || !calls) return TOHIP_ERR_ARG;
    *calls = ctx->nn_calls;
    return 0;
}

// Single convolution layer forward + backward with FP32 host arrays (inputs
// rounded to BF16 on the way in), for gradient checks.  Shapes: x [B][D][H][W]
// [cin], w [cout][k^3][cin], b [cout], dy and y [B][OD][OH][OW][cout], dx like
// x, dw like w, db like b.  act: 0 identity, 1 SiLU.
int tohip_nn_conv_check(int B, const int* dims, int cin, int cout, int k, int pad, int act,
                        const float* x, const float* w, const float* b, const float* dy, float* y,
                        float* dx, float* dw, float* db) {
    if (B < 1 || !dims || cin < 8 || cin % 8 || cout < 8 || cout % 8 || k < 1 || k > 3 ||
        pad < 0 || pad > k - 1 || !x || !w || !b || !dy || !y || !dx || !dw || !db) {
        g_nn_error = "tohip_nn_conv_check: bad arguments";
        return TOHIP_ERR_ARG;
    }
    tohip_nn tmp;
    tohip_nn* nn = &tmp;
    NNLayer l{};
    l.cin = cin;
    l.cout = cout;
    l.k = k;
    l.pad = pad;
    l.T = k * k * k;
    l.cin32 = up32(cin);
    l.cout32 = up32(cout);
    ConvDesc c{};
    c.B = B;
    c.D = dims[0];
    c.H = dims[1];
    c.W = dims[2];
    c.kd = c.kh = c.kw = k;
    c.pd = c.ph = c.pw = pad;
    c.OD = c.D + 2 * pad - k + 1;
    c.OH = c.H + 2 * pad - k + 1;
    c.OW = c.W + 2 * pad - k + 1;
    c.Cin = cin;
    c.Cin32 = l.cin32;
    c.Cout = cout;
    c.out_cstride = cout;
    c.out_cvalid = cout;
    c.act = act;
    const int64_t nx = (int64_t)B * c.D * c.H * c.W * cin;
    const int64_t ny = (int64_t)B * c.OD * c.OH * c.OW * cout;
    const int64_t nw = (int64_t)cout * l.T * cin;
    auto bf = [](float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (uint16_t)(u >> 16);
    };
    std::vector<uint16_t> hx(nx), hdy(ny);
    for (int64_t i = 0; i < nx; ++i) hx[i] = bf(x[i]);
    for (int64_t i = 0; i < ny; ++i) hdy[i] = bf(dy[i]);
    uint16_t *dX = nullptr, *dDY = nullptr, *dZ = nullptr, *dZg = nullptr, *dXb = nullptr;
    float *dW = nullptr, *dB = nullptr, *dYf = nullptr, *dDX = nullptr, *part = nullptr,
          *dw32 = nullptr, *dbb = nullptr;
    hipStream_t s = nullptr;
    int rc = 0;
    const int ns = 4;
    const size_t partn = (size_t)ns * cout * l.T * l.cin32;
    auto fail = [&](hipError_t e) {
        g_nn_error = std::string("tohip_nn_conv_check: ") + hipGetErrorString(e);
        rc = TOHIP_ERR_HIP;
    };
#define QCK(call)                    \
    do {                             \
        hipError_t e_ = (call);      \
        if (e_ != hipSuccess) {      \
            fail(e_);                \
            goto done;               \
        }                            \
    } while (0)
    QCK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    QCK(hipMalloc((void**)&dX, nx * 2));
    QCK(hipMalloc((void**)&dDY, ny * 2));
    QCK(hipMalloc((void**)&dZ, ny * 2));
    QCK(hipMalloc((void**)&dZg, ny * 2));
    QCK(hipMalloc((void**)&dXb, nx * 2));
    QCK(hipMalloc((void**)&dW, nw * 4));
    QCK(hipMalloc((void**)&dB, cout * 4));
    QCK(hipMalloc((void**)&dYf, ny * 4));
    QCK(hipMalloc((void**)&dDX, nx * 4));
    QCK(hipMalloc((void**)&part, (partn + (size_t)cout * l.T * l.cin32) * 4));
    QCK(hipMalloc((void**)&dbb, cout * 4));
    QCK(hipMalloc((void**)&l.wf, (size_t)cout * l.T * l.cin32 * 2));
    QCK(hipMalloc((void**)&l.wt, (size_t)cin * l.T * l.cout32 * 2));
    dw32 = part + partn;
    QCK(hipMemcpyAsync(dX, hx.data(), nx * 2, hipMemcpyHostToDevice, s));
    QCK(hipMemcpyAsync(dDY, hdy.data(), ny * 2, hipMemcpyHostToDevice, s));
    QCK(hipMemcpyAsync(dW, w, nw * 4, hipMemcpyHostToDevice, s));
    QCK(hipMemcpyAsync(dB, b, cout * 4, hipMemcpyHostToDevice, s));
    QCK(launch_weights_to_bf16(cout, l.T, cin, l.cin32, l.cout32, dW, l.wf, l.wt, s));
    QCK(launch_conv_fwd(c, dX, l.wf, dB, nullptr, dZ, dYf, s));
    if (act) {
        QCK(launch_silu_bwd(ny, dDY, dZ, dZg, s));
    } else {
        QCK(hipMemcpyAsync(dZg, dDY, ny * 2, hipMemcpyDeviceToDevice, s));
    }
    {
        ConvDesc bd = bwd_desc(nn, l, B, 0);
        bd.D = c.OD;
        bd.H = c.OH;
        bd.W = c.OW;
        bd.OD = c.D;
        bd.OH = c.H;
        bd.OW = c.W;
        QCK(launch_conv_fwd(bd, dZg, l.wt, nullptr, nullptr, nullptr, dDX, s));
    }
    QCK(launch_conv_bwd_w(c, dX, dZg, part, ns, dw32, dbb, s));
    QCK(hipMemcpy2DAsync(dW, cin * sizeof(float), dw32, l.cin32 * sizeof(float),
                         cin * sizeof(float), (size_t)cout * l.T, hipMemcpyDeviceToDevice, s));
    QCK(hipMemcpyAsync(y, dYf, ny * 4, hipMemcpyDeviceToHost, s));
    QCK(hipMemcpyAsync(dx, dDX, nx * 4, hipMemcpyDeviceToHost, s));
    QCK(hipMemcpyAsync(dw, dW, nw * 4, hipMemcpyDeviceToHost, s));
    QCK(hipMemcpyAsync(db, dbb, cout * 4, hipMemcpyDeviceToHost, s));
    QCK(hipStreamSynchronize(s));
done:
#undef QCK
    for (void* p : {(void*)dX, (void*)dDY, (void*)dZ, (void*)dZg, (void*)dXb, (void*)dW,
                    (void*)dB, (void*)dYf, (void*)dDX, (void*)part, (void*)dbb, (void*)l.wf,
                    (void*)l.wt})
        if (p) hipFree(p);
    if (s) hipStreamDestroy(s);
    nn->stream = nullptr;
    return rc;
}

}  // extern "C"
