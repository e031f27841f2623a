// pm_net_internal.h -- what the engine's C ABI (pm_capi.cpp) uses of the
// plane-proposal network (pm_net_capi.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "pm_net.h"

namespace pm {
namespace net {

// Runs the network on one stereo pair (the engine's images: channels 1 or
// 4, row stride in floats), on the left view's side or, with mirror, on
// the mirrored pair (the right view's side).  *out: the head's output,
// Hp x Wp x *os fp32 (disparity / PM_NET_DISP_SCALE, nx, ny, nz, then
// unused channels), valid until the net's next pass; Wp, Hp: W, H padded
// to multiples of 8.  Ordered on stream st.
int proposal(pm_net* n, hipStream_t st, const float* left, const float* right,
             int W, int H, int channels, int stride, int mirror,
             const float** out, int* Wp, int* os);
const char* error_of(const pm_net* n);
int device_of(const pm_net* n);

}  // namespace net
}  // namespace pm
