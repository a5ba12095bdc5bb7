// msx_unet.h -- internal interface of the surrogate network's GPU kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace msx {
namespace unet {

enum ConvMode { kFwd = 0, kBwdData = 1, kBwdWeight = 2 };

// One convolution layer, stride 1, zero padding KS/2, NCHW FP32 tensors.
//  kFwd      : out[n, out_c0 + co] (+)= relu?(conv(x, w) + bias)
//  kBwdData  : out[n, out_c0 + ci] (+)= conv^T(G, w)
//  kBwdWeight: out[co, ci, tap]    +=  sum_{n,p} G[n, co, p] x[n, ci, p + off(tap)]
// G = g masked by (y > 0) when relu is set (y = the layer's forward output).
struct ConvArgs {
    const float* x;     // [N, Cin, H, W]   (kFwd, kBwdWeight)
    const float* w;     // [Cout, Cin, KS*KS]
    const float* bias;  // [Cout] or null   (kFwd)
    const float* g;     // [N, Cout, H, W]  (kBwdData, kBwdWeight)
    const float* y;     // [N, Cout, H, W]  ReLU mask source (backward with relu)
    float* out;
    int N, Cin, Cout, H, W;
    int relu;
    int out_c, out_c0;  // channel count and offset of the destination tensor
    int accumulate;
    // kBwdWeight split over several slices of its reduction: per-slice
    // partial weight gradients [splits][Cout][Cin * KS * KS], summed in a
    // fixed order into out (conv_wgrad_scratch floats; unused otherwise)
    float* partial;
};

hipError_t launch_conv(int mode, int ks, const ConvArgs& a, hipStream_t s);
// floats of ConvArgs::partial that launch_conv(kBwdWeight, ks, a) needs (0: none)
size_t conv_wgrad_scratch(int ks, const ConvArgs& a);
hipError_t launch_avgpool_fwd(const float* x, float* y, long long nc, int H, int W, hipStream_t s);
hipError_t launch_avgpool_bwd(const float* gy, float* gx, long long nc, int H, int W, hipStream_t s);
hipError_t launch_upsample_fwd(const float* x, float* y, int N, int C, int H, int W, int Ct, int c0, hipStream_t s);
hipError_t launch_upsample_bwd(const float* gy, float* gx, int N, int C, int H, int W, int Ct, int c0, hipStream_t s);
hipError_t launch_copy_channels(const float* src, int Cs, int s0, float* dst, int Cd, int d0, int N, int C, int HW,
                                int accumulate, hipStream_t s);
hipError_t launch_bias_grad(const float* g, const float* y, int relu, int N, int C, int HW, float* db, hipStream_t s);
hipError_t launch_residual_mse(const float* in, const float* z, const float* t, float* out, float* g, long long n,
                               float scale, hipStream_t s);
// y[n][p] = (1/C) sum_c x[n][c][p] (FP32, channels in order)
hipError_t launch_channel_mean(const float* x, float* y, int N, int C, long long HW, hipStream_t s);
hipError_t launch_sq_diff(const float* a, const float* b, long long n, double* partial, int nblocks, hipStream_t s);
hipError_t launch_adam(float* w, const float* gw, float* m, float* v, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2, hipStream_t s);

}  // namespace unet
}  // namespace msx
