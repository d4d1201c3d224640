// Internal declarations of the learned warm-start network kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tohip {

// One stride-1 zero-padded 3D convolution over NDHWC BF16 tensors.
struct ConvDesc {
    int B;             // batch
    int D, H, W;       // input spatial size
    int OD, OH, OW;    // output spatial size (D + 2 pd - kd + 1, ...)
    int Cin, Cin32;    // input channels (multiple of 8); Cin rounded up to 32
    int Cout;          // output channels
    int kd, kh, kw;    // kernel size
    int pd, ph, pw;    // zero padding
    int out_cstride;   // channels of the output tensor (writes into a concat view)
    int out_coff;      // channel offset of the view
    int out_cvalid;    // FP32 output: channels stored
    int act;           // 0 identity, 1 SiLU
};

hipError_t launch_conv_fwd(const ConvDesc& c, const uint16_t* X, const uint16_t* W,
                           const float* bias, uint16_t* Y, uint16_t* Z, float* Yf,
                           hipStream_t s);
hipError_t launch_conv_bwd_w(const ConvDesc& c, const uint16_t* X, const uint16_t* dY,
                             float* part, int nsplit, float* dW, float* db, hipStream_t s);
hipError_t launch_weights_to_bf16(int Cout, int T, int Cin, int Cin32, int Cout32,
                                  const float* w, uint16_t* wf, uint16_t* wt, hipStream_t s);
hipError_t launch_silu_bwd(int64_t n, const uint16_t* dA, const uint16_t* Z, uint16_t* dZ,
                           hipStream_t s);
hipError_t launch_avgpool_fwd(int B, int D, int H, int W, int C, const uint16_t* in, int in_cs,
                              int in_co, uint16_t* out, hipStream_t s);
hipError_t launch_avgpool_bwd(int B, int D, int H, int W, int C, const uint16_t* dout,
                              uint16_t* din, int in_cs, int in_co, int accumulate,
                              hipStream_t s);
hipError_t launch_upsample_fwd(int B, int D, int H, int W, int C, const uint16_t* in,
                               uint16_t* out, int out_cs, int out_co, hipStream_t s);
hipError_t launch_upsample_bwd(int B, int D, int H, int W, int C, const uint16_t* dout,
                               int out_cs, int out_co, uint16_t* din, hipStream_t s);
hipError_t launch_slice_copy(int64_t P, int C, const uint16_t* src, int src_cs, int src_co,
                             uint16_t* dst, int accumulate, hipStream_t s);

}  // namespace tohip
