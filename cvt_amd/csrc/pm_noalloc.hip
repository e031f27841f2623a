// pm_noalloc.hip -- pm_disparity_planes_noalloc (include/pm_io.h): the LR
// check and disparity conversion of pm_disparity_device on caller-owned
// plane fields, with the LR-checked field (the one scratch buffer of that
// sequence) in a fixed device array instead of memory the library
// allocates.  For callers that cannot allocate: code under HIP graph
// capture and serving loops.
//
// The array is a file-scope __device__ variable, one copy per GPU the
// library runs on, held until the process ends.  Host code never uses its
// name as an address: each device's copy is looked up once with
// hipGetSymbolAddress, with that device current, and kept per device.
// Every call on a device shares that device's copy, so calls that overlap
// on one GPU race; k_consistency overwrites every pixel of the W x H
// region before it is read, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <vector>

#include "pm_engine.h"
#include "pm_io.h"

namespace {

// 2^21 pixels (1920 x 1080 fits): 32 MiB of float4 per GPU
constexpr size_t kNoAllocMaxPixels = (size_t)1 << 21;

static __device__ __align__(256) float4 scratch[kNoAllocMaxPixels];

std::mutex g_mu;
std::vector<float4*> g_scratch;  // per device ordinal, null until looked up

// the device address of this device's scratch array (set-up on the
// device's first call); the device must be current
int scratch_of(int device, float4** out) {
  std::lock_guard<std::mutex> lock(g_mu);
  if ((int)g_scratch.size() <= device) g_scratch.resize(device + 1, nullptr);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      char msg[256];
      std::snprintf(msg, sizeof msg, "hipGetSymbolAddress of the scratch "
                    "array on device %d: %s", device, hipGetErrorString(e));
      return pm::io_set_error(PM_ERR_HIP, msg);
    }
    g_scratch[device] = static_cast<float4*>(p);
  }
  *out = g_scratch[device];
  return PM_OK;
}

int hip_error(hipError_t e, const char* what) {
  char msg[256];
  std::snprintf(msg, sizeof msg, "%s: %s", what, hipGetErrorString(e));
  return pm::io_set_error(PM_ERR_HIP, msg);
}

}  // namespace

extern "C" int pm_disparity_planes_noalloc(const float* d_left_planes,
                                           const float* d_right_planes,
                                           int W, int H, int pp,
                                           float maxdispdiff, float maxdeg,
                                           float* d_disparity, int device,
                                           void* stream) {
  if (!d_left_planes || !d_right_planes || !d_disparity)
    return pm::io_set_error(PM_ERR_ARG, "NULL argument");
  if (W < 1 || H < 1)
    return pm::io_set_error(PM_ERR_ARG, "invalid size");
  if ((size_t)W * (size_t)H > kNoAllocMaxPixels) {
    char msg[256];
    std::snprintf(msg, sizeof msg, "%d x %d pixels exceed the static "
                  "scratch of %zu pixels", W, H, kNoAllocMaxPixels);
    return pm::io_set_error(PM_ERR_ARG, msg);
  }
  if (pp != PM_PP_NONE && pp != PM_PP_LRCHECK && pp != PM_PP_LRCHECK_FILL)
    return pm::io_set_error(PM_ERR_ARG, "unknown post-processing");
  if (!(maxdispdiff >= 0.0f) || !(maxdeg >= 0.0f))
    return pm::io_set_error(PM_ERR_ARG,
                            "maxdispdiff and maxdeg must be >= 0");
  if (device < 0) return pm::io_set_error(PM_ERR_ARG, "negative device");
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_error(e, "hipSetDevice");
  pm::Geometry g;
  g.W = W;
  g.H = H;
  g.npix = (size_t)W * H;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const float4* L = reinterpret_cast<const float4*>(d_left_planes);
  const float4* R = reinterpret_cast<const float4*>(d_right_planes);
  // the sequence of pm_disparity_device (postprocess in pm_capi.cpp)
  if (pp == PM_PP_NONE) {
    e = pm::launch_to_disparity(st, g, L, 1.0f, 1, d_disparity);
    return e == hipSuccess ? PM_OK : hip_error(e, "to_disparity");
  }
  float4* chk = nullptr;
  int rc;
  if ((rc = scratch_of(device, &chk))) return rc;
  e = pm::launch_consistency(st, g, L, R, maxdispdiff, maxdeg, chk);
  if (e != hipSuccess) return hip_error(e, "consistency");
  e = pp == PM_PP_LRCHECK
          ? pm::launch_to_disparity(st, g, chk, 1.0f, 1, d_disparity)
          : pm::launch_fill_depthmap(st, g, chk, 1.0f, d_disparity);
  return e == hipSuccess ? PM_OK : hip_error(e, "disparity");
}
