// pm_net_kernels.h -- internal interface of the plane-proposal network kernels
// (pm_net.hip): a 3x3 convolution on bf16 NHWC activations with fp32
// accumulation on the matrix cores, its weight gradient, and the small
// kernels a U-Net training step and a proposal need.  Internal to the
// library: the public boundary is include/pm_net.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "pm_engine.h"

namespace pm {
namespace net {

typedef __bf16 bf16;

// K of a 3x3 convolution over cin channels, padded to the MFMA's 32
__host__ __device__ inline int conv_kpad(int cin) {
  return (9 * cin + 31) / 32 * 32;
}
// output channels padded to the 32 of a work-group's tile
__host__ __device__ inline int conv_cpad(int cout) {
  return (cout + 31) / 32 * 32;
}

// One 3x3 convolution, stride 1, zero padding, dilation `dil`.  The input
// is the channel concatenation of x1 (c1 channels, N x H x W) and x2 (c2
// channels; N x H x W, or N x ceil(H/2) x ceil(W/2) read nearest-upsampled
// when up2 is set).  c1, c2 and cout are multiples of 8.  Weights
// w[co][k], co < conv_cpad(cout), k = tap * (c1 + c2) + ci < conv_kpad(c1 +
// c2), tap = 3 (dy + 1) + (dx + 1), bf16, zero beyond the real ones.
// Output (N x H x W, channel stride ys >= cout) with the bias (when not
// null) and an optional ReLU: bf16 into y, or fp32 into yf when set.
struct ConvArgs {
  const bf16* x1 = nullptr;
  const bf16* x2 = nullptr;
  int c1 = 0, c2 = 0, up2 = 0;
  const bf16* w = nullptr;
  const float* bias = nullptr;
  int N = 0, H = 0, W = 0;
  int cout = 0, dil = 1, relu = 0;
  bf16* y = nullptr;
  float* yf = nullptr;
  int ys = 0;
};
hipError_t launch_conv(hipStream_t st, const ConvArgs& a);

// Weight gradient of the convolution above: dw[co][k] += sum over the
// N x H x W pixels of g[p][co] * x(p, k), g bf16 with cout channels; and
// db[co] += sum of g[p][co] when db is not null.  fp32 accumulation.
hipError_t launch_conv_wgrad(hipStream_t st, const ConvArgs& a,
                             const bf16* g, float* dw, float* db);

// w'[ci][k'] for the gradient of the input: the convolution of the output
// gradient (cout channels) with the flipped, transposed kernel, k' = (8 -
// tap) * cout + co; w' has conv_cpad(cin) rows of conv_kpad(cout)
hipError_t launch_flip_weights(hipStream_t st, const bf16* w, int cin,
                               int cout, bf16* wt);

// g_bf16[i] = g[i] * (y[i] > 0) (ReLU) or g[i] (relu = 0), n elements
hipError_t launch_grad_mask(hipStream_t st, size_t n, const float* g,
                            const bf16* y, int relu, bf16* out);

// t holds the input gradient of a concatenated input (N x H x W x (c1 +
// c2), fp32): g1 += its first c1 channels; g2 += the other c2, summed over
// 2 x 2 blocks when x2 was upsampled (up2).  g2 may be null (no gradient).
hipError_t launch_route_grad(hipStream_t st, int N, int H, int W, int c1,
                             int c2, int up2, const float* t, float* g1,
                             float* g2);

// 2 x 2 average pooling, bf16 (N x H x W x C -> N x H/2 x W/2 x C, H and
// W even, C a multiple of 8), and its gradient: gin += 0.25 gout
hipError_t launch_avgpool(hipStream_t st, int N, int H, int W, int C,
                          const bf16* x, bf16* y);
hipError_t launch_avgpool_grad(hipStream_t st, int N, int H, int W, int C,
                               const float* gout, float* gin);

// Masked L1 loss of the head output o (N x H x W x os fp32, channels 0..3
// = disparity / kDispScale, nx, ny, nz) against t (N x H x W x 4; a
// disparity below 0 marks a pixel without a target): g = dloss / do
// (channels 0..3, 0 elsewhere), loss += sum / count.  count: valid pixels.
hipError_t launch_l1_loss(hipStream_t st, size_t npix, const float* o,
                          int os, const float* t, float wd, float wn,
                          float inv_count, float* g, float* loss);

// Adam on n fp32 parameters; writes their bf16 copy
hipError_t launch_adam(hipStream_t st, size_t n, float* w, const float* g,
                       float* m, float* v, float lr, float b1, float b2,
                       float eps, float c1, float c2, bf16* wb);

// fp32 -> bf16, n elements
hipError_t launch_to_bf16(hipStream_t st, size_t n, const float* x, bf16* y);

// network input of a stereo pair: N = 1, Hp x Wp x 8 bf16 = (left rgba,
// right rgba) of the engine's images (channels 1 or 4, row stride in
// floats), edge-replicated beyond W x H; mirror: the pair seen as a left
// view by flipping both images horizontally and swapping them
hipError_t launch_pack_pair(hipStream_t st, const float* left,
                            const float* right, int W, int H, int channels,
                            int stride, int Hp, int Wp, int mirror, bf16* x);

// disparity levels of the matching cost volumes: the coarse one at
// quarter resolution (quarter pixels), the fine one at half resolution
constexpr int kCvLevels = 48;
constexpr int kCvFineLevels = 64;

// 4 x 4 box average of N x H x W x 8 bf16 (H, W multiples of 4)
hipError_t launch_down4(hipStream_t st, int N, int H, int W, const bf16* x,
                        bf16* y);

// N x H4 x W4 x levels bf16 matching costs of a downsampled pair
// (N x H4 x W4 x 8: first image rgba, second image rgba); levels a
// multiple of 8, at most 64
hipError_t launch_cost_volume(hipStream_t st, int N, int H4, int W4,
                              int levels, const bf16* d4, bf16* cv);

}  // namespace net

}  // namespace pm
