// pm_net_internal.h -- what the engine's C ABI (pm_capi.cpp) uses of the
// plane-proposal network (pm_net_capi.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "pm_net.h"

namespace pm {
namespace net {

// Runs the network on one stereo pair (the engine's images: channels 1 or
// 4, row stride in floats), and with views = 2 also on the mirrored pair,
// as one batch of two: out[0] is the left view's proposal, out[1] the
// right view's (the pass over the mirrored pair, columns mirrored; null
// with views = 1).  Each is the head's output, Hp x Wp x *os fp32
// (disparity / PM_NET_DISP_SCALE, nx, ny, nz, then unused channels), valid
// until the net's next pass; Wp, Hp: W, H padded to multiples of 8.
// Ordered on stream st.
int proposal(pm_net* n, hipStream_t st, const float* left, const float* right,
             int W, int H, int channels, int stride, int views,
             const float** out, int* Wp, int* os);
// Sizes the activation buffers for every proposal over inputs of at most
// Hp x Wp pixels (multiples of 8), so that later passes allocate nothing
// (stream capture).
int reserve(pm_net* n, int Hp, int Wp);
// changes whenever the activation buffers are reallocated (a recorded
// graph holds their addresses)
long generation(const pm_net* n);
const char* error_of(const pm_net* n);
int device_of(const pm_net* n);

}  // namespace net
}  // namespace pm
