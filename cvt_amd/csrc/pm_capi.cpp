// pm_capi.cpp -- C ABI of the PatchMatch engine (include/pm_stereo.h).
//
// Host orchestration of cvt::PMHuberStereo's PatchMatch path, plain or with
// the PM-Huber smoothing coupling (cvt/vision/PMHuberStereo.cpp:349-404,
// 406-581,583-644, cvt/vision/PDOptCL.cpp:134-155,1114-1168) over the HIP
// kernels in pm_engine.hip.  Every device buffer is owned by the context
// and sized for the current image pair.  Work is queued on the context's
// stream; pm_iterate forks onto streams of its own and joins them back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>
#include <string>

#include "pm_engine.h"
#include "pm_net_internal.h"
#include "pm_stereo.h"

namespace {

constexpr int kMaxRadius = 16;

// bands of block rows a view's propagate is split into (split schedule)
constexpr int kSplitParts = 2;

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct pm_ctx {
  pm_params prm{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;

  pm::Geometry geo;
  int channels = 0;
  bool have_images = false;
  bool ready = false;   // setup done
  int iteration = 0;
  int cur = 0;          // index of the current plane field per view
  int last_pp = -1;     // post-processing of the last pm_disparity

  DevBuf img[2];        // packed input images
  DevBuf tex[2], sgn[2];
  DevBuf state[2][2];   // [view][ping-pong]
  DevBuf vbuf[2];       // vbuf[0]: filled by the left view, vbuf[1]: right
  DevBuf checked;       // LR-checked left field
  DevBuf scratch4;      // W*H float4 (host transfers, normals)
  DevBuf scratch4b;     // W*H float4
  DevBuf scratch1;      // W*H float
  DevBuf counters;      // 2 x u64 tap counters of the counting pass

  // PM-Huber coupling (PM_SMOOTH_HUBER only)
  DevBuf dt[2];         // diffusion tensor per view
  DevBuf smooth[2];     // smoothed (normal, depth) field per view
  // ROF scratch per view (the two views' smoothings may overlap)
  struct RofBufs {
    DevBuf nd;          // (normal, depth) field being smoothed
    DevBuf u;           // primal scratch
    DevBuf p[2];        // duals p (two float4 per pixel), ping-pong
    DevBuf q[2];        // duals q, ping-pong
  } rofb[2];
  float theta = 0.0f;   // theta of the next iteration
  int schedule_n = 1;   // N of the theta schedule of the running pm_iterate

  // Split schedule of the iterations: each view's propagate runs as
  // kSplitParts bands of block rows, band 0 on `stream` and band k > 0 on
  // side[k - 1], so that one band's last work-groups overlap other bands'
  // work; with PM_SMOOTH_HUBER each view's smoothing runs on `smoothing`
  // beside the other view's propagate.  ev_band[v][k] marks the end of
  // the latest launch of view v's band k, ev_smooth[v] that of view v's
  // latest smoothing.
  hipStream_t side[kSplitParts - 1] = {};
  hipStream_t smoothing = nullptr;
  hipEvent_t ev_band[2][kSplitParts] = {};
  hipEvent_t ev_smooth[2] = {};
  // pm_setup_proposal_views: the texel work forks onto side[0] beside the
  // network's pass (ev_setup[0]) and joins back (ev_setup[1])
  hipEvent_t ev_setup[2] = {};
  // LDS the propagate's work-groups ask for at least (pm_create)
  size_t prop_lds_floor = 0;
};

namespace {

int fail(pm_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

int hip_fail(pm_ctx* c, hipError_t e, const char* what) {
  return fail(c, e == hipErrorOutOfMemory ? PM_ERR_NOMEM : PM_ERR_HIP,
              "%s: %s", what, hipGetErrorString(e));
}

#define PM_HIP(ctx, call)                                   \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #call);  \
  } while (0)

int ensure(pm_ctx* c, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.ptr) return PM_OK;
  if (b.ptr) {
    (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  }
  hipError_t e = hipMalloc(&b.ptr, bytes);
  if (e != hipSuccess) {
    b.ptr = nullptr;
    return hip_fail(c, e, "hipMalloc");
  }
  b.bytes = bytes;
  return PM_OK;
}

void release(DevBuf& b) {
  if (b.ptr) (void)hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
}

int use_device(pm_ctx* c) {
  PM_HIP(c, hipSetDevice(c->prm.device));
  return PM_OK;
}

pm::ViewBufs view_bufs(pm_ctx* c, int v) {
  pm::ViewBufs b;
  b.tex = static_cast<float4*>(c->tex[v].ptr);
  b.sgn = static_cast<uint8_t*>(c->sgn[v].ptr);
  return b;
}

float4* state_ptr(pm_ctx* c, int view, int which) {
  return static_cast<float4*>(c->state[view][which].ptr);
}

int check_params(pm_ctx* c, const pm_params* p) {
  if (!p) return fail(c, PM_ERR_ARG, "params is NULL");
  if (p->radius < 1 || p->radius > kMaxRadius)
    return fail(c, PM_ERR_ARG, "radius %d outside [1, %d]", p->radius,
                kMaxRadius);
  if (!(p->disparity_max > 0.0f) || !std::isfinite(p->disparity_max))
    return fail(c, PM_ERR_ARG, "disparity_max must be finite and > 0");
  if (!(p->color_weight >= 0.0f) || !std::isfinite(p->color_weight))
    return fail(c, PM_ERR_ARG, "color_weight must be finite and >= 0");
  if (p->cost_kind != PM_COST_CENSUS && p->cost_kind != PM_COST_COLOR_GRAD)
    return fail(c, PM_ERR_ARG, "unknown cost_kind %d", p->cost_kind);
  if (p->schedule != PM_SCHEDULE_JACOBI &&
      p->schedule != PM_SCHEDULE_RED_BLACK)
    return fail(c, PM_ERR_ARG, "unknown schedule %d", p->schedule);
  if (!(p->cost_weight >= 0.0f) || !std::isfinite(p->cost_weight))
    return fail(c, PM_ERR_ARG, "cost_weight must be finite and >= 0");
  if (p->smoothing != PM_SMOOTH_NONE && p->smoothing != PM_SMOOTH_HUBER &&
      p->smoothing != PM_SMOOTH_DEPTHMAP)
    return fail(c, PM_ERR_ARG, "unknown smoothing %d", p->smoothing);
  if (!(p->theta_min >= 0.0f) || !std::isfinite(p->theta_min))
    return fail(c, PM_ERR_ARG, "theta_min must be finite and >= 0");
  if (p->schedule_iterations < 0)
    return fail(c, PM_ERR_ARG, "schedule_iterations must be >= 0");
  if (p->device < 0)
    return fail(c, PM_ERR_ARG, "device ordinal %d is negative", p->device);
  return PM_OK;
}

int check_view(pm_ctx* c, int view) {
  if (view != 0 && view != 1)
    return fail(c, PM_ERR_ARG, "view must be 0 (left) or 1 (right)");
  return PM_OK;
}

// buffers that depend on the pair's size
int alloc_pair(pm_ctx* c, int W, int H, int channels) {
  const size_t n = (size_t)W * H;
  int rc;
  for (int v = 0; v < 2; ++v) {
    if ((rc = ensure(c, c->img[v], n * channels * sizeof(float)))) return rc;
    if ((rc = ensure(c, c->tex[v], (size_t)pm::tex_stride(W) * H * 2 *
                                       sizeof(float4))))
      return rc;
    if ((rc = ensure(c, c->sgn[v], n))) return rc;
    for (int k = 0; k < 2; ++k)
      if ((rc = ensure(c, c->state[v][k], n * sizeof(float4)))) return rc;
    if ((rc = ensure(c, c->vbuf[v], n * sizeof(unsigned long long))))
      return rc;
  }
  if ((rc = ensure(c, c->checked, n * sizeof(float4)))) return rc;
  if ((rc = ensure(c, c->scratch4, n * sizeof(float4)))) return rc;
  if ((rc = ensure(c, c->scratch4b, n * sizeof(float4)))) return rc;
  if ((rc = ensure(c, c->scratch1, n * sizeof(float)))) return rc;
  if (c->prm.smoothing != PM_SMOOTH_NONE) {
    for (int v = 0; v < 2; ++v) {
      if ((rc = ensure(c, c->dt[v], n * sizeof(float4)))) return rc;
      if ((rc = ensure(c, c->smooth[v], n * sizeof(float4)))) return rc;
      pm_ctx::RofBufs& b = c->rofb[v];
      if ((rc = ensure(c, b.nd, n * sizeof(float4))) ||
          (rc = ensure(c, b.u, n * sizeof(float4))))
        return rc;
      for (int k = 0; k < 2; ++k)
        if ((rc = ensure(c, b.p[k], 2 * n * sizeof(float4))) ||
            (rc = ensure(c, b.q[k], n * sizeof(float4))))
          return rc;
    }
  }
  c->geo.W = W;
  c->geo.H = H;
  c->geo.R = c->prm.radius;
  c->geo.TS = pm::tex_stride(W);
  c->geo.npix = n;
  c->geo.color_weight = c->prm.color_weight;
  c->geo.cost_kind = c->prm.cost_kind == PM_COST_COLOR_GRAD
                         ? pm::kCostColorGrad
                         : pm::kCostCensus;
  c->geo.cost_weight = c->prm.cost_weight;
  c->channels = channels;
  return PM_OK;
}

int check_image_args(pm_ctx* c, const void* l, const void* r, int W, int H,
                     int channels, int stride) {
  if (!c) return PM_ERR_ARG;
  if (!l || !r) return fail(c, PM_ERR_ARG, "image pointer is NULL");
  if (W < 1 || H < 1) return fail(c, PM_ERR_ARG, "invalid size %dx%d", W, H);
  if ((size_t)W * H > (size_t)1 << 28)
    return fail(c, PM_ERR_ARG, "image of %dx%d pixels is too large", W, H);
  if (channels != 1 && channels != 4)
    return fail(c, PM_ERR_ARG, "channels must be 1 (GRAY) or 4 (RGBA)");
  if (stride < W * channels)
    return fail(c, PM_ERR_ARG, "stride %d < width*channels %d", stride,
                W * channels);
  return PM_OK;
}

int set_images(pm_ctx* c, const float* l, const float* r, int W, int H,
               int channels, int stride, hipMemcpyKind kind) {
  int rc = check_image_args(c, l, r, W, H, channels, stride);
  if (rc) return rc;
  if ((rc = use_device(c))) return rc;
  if ((rc = alloc_pair(c, W, H, channels))) return rc;
  const size_t row = (size_t)W * channels * sizeof(float);
  const float* src[2] = {l, r};
  // Host rows go up as one contiguous copy per image: padded rows are
  // packed on the host first (a row-by-row copy from pageable memory is
  // several times slower).  Device rows use a 2D copy when padded.
  std::vector<float> packed;
  for (int v = 0; v < 2; ++v) {
    const float* s = src[v];
    if ((size_t)stride * 4 != row && kind == hipMemcpyHostToDevice) {
      try {
        packed.resize((size_t)W * channels * H);
      } catch (const std::bad_alloc&) {
        return fail(c, PM_ERR_NOMEM, "out of host memory");
      }
      for (int y = 0; y < H; ++y)
        std::memcpy(packed.data() + (size_t)y * W * channels,
                    s + (size_t)y * stride, row);
      s = packed.data();
      // the copy below reads `packed` before it is refilled for the
      // next image
      PM_HIP(c, hipMemcpyAsync(c->img[v].ptr, s, row * H, kind, c->stream));
      PM_HIP(c, hipStreamSynchronize(c->stream));
    } else if ((size_t)stride * 4 == row) {
      PM_HIP(c, hipMemcpyAsync(c->img[v].ptr, s, row * H, kind, c->stream));
    } else {
      PM_HIP(c, hipMemcpy2DAsync(c->img[v].ptr, row, s, (size_t)stride * 4,
                                 row, H, kind, c->stream));
    }
  }
  if (kind == hipMemcpyHostToDevice) PM_HIP(c, hipStreamSynchronize(c->stream));
  c->have_images = true;
  c->ready = false;
  c->last_pp = -1;
  return PM_OK;
}

// bilateral weights + gradients of both views (PMHuberStereo.cpp:445-450)
int prepare_views(pm_ctx* c, hipStream_t st = nullptr) {
  for (int v = 0; v < 2; ++v)
    PM_HIP(c, pm::launch_prepare(st ? st : c->stream, c->geo,
                                 static_cast<const float*>(c->img[v].ptr),
                                 c->channels, c->geo.W * c->channels,
                                 view_bufs(c, v)));
  return PM_OK;
}

int to_device(pm_ctx* c, void* dst, const void* src, size_t bytes) {
  PM_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return PM_OK;
}

int to_host(pm_ctx* c, void* dst, const void* src, size_t bytes) {
  PM_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  PM_HIP(c, hipStreamSynchronize(c->stream));
  return PM_OK;
}

int require_images(pm_ctx* c) {
  if (!c->have_images)
    return fail(c, PM_ERR_STATE, "no images: call pm_set_images first");
  return PM_OK;
}

int require_ready(pm_ctx* c) {
  if (!c->ready) return fail(c, PM_ERR_STATE, "call pm_setup first");
  return PM_OK;
}

// disparityMapAll's per-iteration sequence (PMHuberStereo.cpp:531-543):
// clear viewbuf1, propagate L (reads viewbuf2, writes viewbuf1), clear
// viewbuf2, propagate R (reads viewbuf1, writes viewbuf2)
float4* f4buf(DevBuf& b) { return static_cast<float4*>(b.ptr); }

// PM-Huber constants of disparityMapAll (PMHuberStereo.cpp:416-428)
constexpr float kHuberAlpha = 24.0f;      // diffusion tensor, beta = 1
constexpr float kThetaScaleNorm = 10.0f;  // THETASCALE_NORM = THETASCALE_D
constexpr float kThetaBias = 1.0f;        // THETABIAS_NORM = THETABIAS_D
constexpr float kThetaScale = 50.0f;      // THETASCALE
constexpr int kRofIters = 40;             // PDOPTITER
// depthMap's constants (PMHuberStereo.cpp:153-165)
constexpr float kDepthAlpha = 12.0f;      // diffusion tensor, beta = 1
constexpr float kDepthThetaScaleD = 15.0f;
constexpr float kDepthThetaBiasD = 2.0f;
constexpr float kDepthThetaScale = 10.0f;
constexpr float kDepthHuberEps = 1e-4f;   // HUBEREPS
constexpr float kDepthTheta0 = 1e-3f;

bool depthmap_mode(const pm_ctx* c) {
  return c->prm.smoothing == PM_SMOOTH_DEPTHMAP;
}

// LAMBDA (PMHuberStereo.cpp:532): the normal channels and nz take
// theta * 10 + 1, the depth channel 1
float4 huber_lambda(float theta) {
  const float l = theta * kThetaScaleNorm + kThetaBias;
  return make_float4(l, l, l, 1.0f);
}

// denoiseDiffusionTensorROF_PDD (PDOptCL.cpp:1114-1168): out = f, duals 0,
// `iters` steps ping-ponging between out and the scratch primal (one more
// when the count is odd, so that the result lands in out).  f must not
// alias out.
int rof(pm_ctx* c, int view, const float4* f, float4 lambda, int iters,
        float4* out, hipStream_t st, float hubereps = -1.0f) {
  const size_t n = c->geo.npix;
  pm_ctx::RofBufs& b = c->rofb[view];
  float4* u[2] = {f4buf(b.u), out};
  float4* p[2] = {f4buf(b.p[0]), f4buf(b.p[1])};
  float4* q[2] = {f4buf(b.q[0]), f4buf(b.q[1])};
  const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float4* dt = f4buf(c->dt[view]);
  // The reference ping-pongs one step per launch and adds a step when the
  // count is odd, so that the result lands in the output image.  Here a
  // launch runs up to kRofStepsMax steps; the launches alternate between
  // the two buffer sets, starting from the set that makes the last one
  // write the output.
  const int steps = iters + (iters & 1);
  const int launches = (steps + pm::kRofStepsMax - 1) / pm::kRofStepsMax;
  int src = (launches & 1) ? 0 : 1;
  PM_HIP(c, hipMemcpyAsync(u[src], f, n * sizeof(float4),
                           hipMemcpyDeviceToDevice, st));
  PM_HIP(c, pm::launch_fill4(st, 2 * n, zero, p[src]));
  PM_HIP(c, pm::launch_fill4(st, n, zero, q[src]));
  for (int done = 0; done < steps; done += pm::kRofStepsMax) {
    const int k = std::min(pm::kRofStepsMax, steps - done);
    PM_HIP(c, pm::launch_rof_steps(st, c->geo, k, u[src], p[src],
                                   q[src], f, dt, lambda, hubereps,
                                   u[1 - src], p[1 - src], q[1 - src]));
    src = 1 - src;
  }
  return PM_OK;
}

// (normal, depth) field of a view's current planes, then smoothed into the
// view's smooth field (PMHuberStereo.cpp:552-567)
int smooth_view(pm_ctx* c, int view, const float4* planes, float theta,
                hipStream_t st) {
  float4* nd = f4buf(c->rofb[view].nd);
  if (depthmap_mode(c)) {
    // depthMap: filled field, Huber steps, lambda theta * 15 + 2 on all
    // four channels (PMHuberStereo.cpp:249, 273-274, 295-296)
    PM_HIP(c, pm::launch_fill_state(st, c->geo, planes,
                                    1.0f / c->prm.disparity_max,
                                    view == 0 ? 1 : 0, nd));
    const float l = theta * kDepthThetaScaleD + kDepthThetaBiasD;
    return rof(c, view, nd, make_float4(l, l, l, l), kRofIters,
               f4buf(c->smooth[view]), st, kDepthHuberEps);
  }
  PM_HIP(c, pm::launch_normal_depth(st, c->geo, planes,
                                    1.0f / c->prm.disparity_max,
                                    view == 0 ? 1 : 0, nd));
  return rof(c, view, nd, huber_lambda(theta), kRofIters,
             f4buf(c->smooth[view]), st);
}

static_assert(kSplitParts >= 2, "the split schedule needs two bands");

// Rows above and below its own that a band reads of its view's previous
// field: the random neighbours' row offsets lie in [-4, 5]
// (pmhstereo.cl:981-991, NUMRNDSAMPLERADIUS 5).
constexpr int kBandReach = 5;

int block_rows(const pm_ctx* c) {
  return (c->geo.H + pm::kBlockRows - 1) / pm::kBlockRows;
}

// true when every band of the split schedule gets a block row; the
// red-black schedule runs its two passes whole, on the context's stream
bool can_split(const pm_ctx* c) {
  return block_rows(c) >= kSplitParts &&
         c->prm.schedule == PM_SCHEDULE_JACOBI;
}

// first block row of band k (band k ends where band k + 1 starts)
int band_first(const pm_ctx* c, int k) {
  return (int)((long long)block_rows(c) * k / kSplitParts);
}

// image rows [*r0, *r1) of band k
void band_rows(const pm_ctx* c, int k, int* r0, int* r1) {
  *r0 = band_first(c, k) * pm::kBlockRows;
  *r1 = std::min(c->geo.H, band_first(c, k + 1) * pm::kBlockRows);
}

hipStream_t band_stream(const pm_ctx* c, int k) {
  return k == 0 ? c->stream : c->side[k - 1];
}

// Makes everything enqueued on the context's stream so far a predecessor
// of the next band or smoothing launched on any stream (every event is
// recorded on the context's stream).  Streams and events are made on
// first use.
int split_rebase(pm_ctx* c) {
  for (hipStream_t& s : c->side)
    if (!s) PM_HIP(c, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  if (!c->smoothing)
    PM_HIP(c, hipStreamCreateWithFlags(&c->smoothing, hipStreamNonBlocking));
  for (int v = 0; v < 2; ++v) {
    for (hipEvent_t& e : c->ev_band[v]) {
      if (!e) PM_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
      PM_HIP(c, hipEventRecord(e, c->stream));
    }
    hipEvent_t& e = c->ev_smooth[v];
    if (!e) PM_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    PM_HIP(c, hipEventRecord(e, c->stream));
  }
  return PM_OK;
}

// One view's propagate as kSplitParts bands of block rows.  A band reads,
// besides its own rows, the view's previous field within kBandReach rows
// of them (the random neighbours) and writes its rows of the new field and
// of the scatter slots; the other view's field and slots it reads are on
// its own rows.  So band k waits for every other band j of the same
// view's previous propagate whose rows lie within kBandReach of its own
// (rows it reads, and rows of the buffer it overwrites that band j still
// reads: the same condition) and follows its own stream otherwise: band k
// of the right view follows band k of the left view, and so on.  All
// waits are enqueued before this propagate records its events, so they
// refer to the view's previous propagate.
int propagate_split(pm_ctx* c, const pm::PropagateArgs& a, int view,
                    unsigned long long* vb, unsigned long long* counters) {
  hipEvent_t* ev = c->ev_band[view];
  for (int k = 0; k < kSplitParts; ++k) {
    int r0, r1;
    band_rows(c, k, &r0, &r1);
    for (int j = 0; j < kSplitParts; ++j) {
      int s0, s1;
      band_rows(c, j, &s0, &s1);
      if (j != k && s0 < r1 + kBandReach && s1 > r0 - kBandReach)
        PM_HIP(c, hipStreamWaitEvent(band_stream(c, k), ev[j], 0));
    }
    if (c->prm.smoothing != PM_SMOOTH_NONE)  // the view's smooth field
      PM_HIP(c, hipStreamWaitEvent(band_stream(c, k), c->ev_smooth[view], 0));
  }
  for (int k = 0; k < kSplitParts; ++k) {
    int r0, r1;
    band_rows(c, k, &r0, &r1);
    pm::Geometry g = c->geo;
    g.by0 = band_first(c, k);
    g.nby = band_first(c, k + 1) - g.by0;
    const hipStream_t st = band_stream(c, k);
    PM_HIP(c, pm::launch_viewbuf_clear_rows(st, c->geo, vb, r0, r1 - r0));
    PM_HIP(c, pm::launch_propagate(st, g, a, counters, c->prop_lds_floor));
    PM_HIP(c, hipEventRecord(ev[k], st));
  }
  return PM_OK;
}

// One view's propagate: the whole image on the context's stream, or as
// two halves when the caller set up the split schedule (split_rebase).
// The red-black schedule runs it as two passes over the pixels of either
// colour, the second reading the first's new planes for its colour-0
// neighbours.
int propagate_view(pm_ctx* c, const pm::PropagateArgs& a, int view,
                   unsigned long long* vb, unsigned long long* counters,
                   bool split) {
  if (split) return propagate_split(c, a, view, vb, counters);
  PM_HIP(c, pm::launch_viewbuf_clear(c->stream, c->geo, vb));
  if (c->prm.schedule == PM_SCHEDULE_RED_BLACK) {
    pm::Geometry g = c->geo;
    pm::PropagateArgs b = a;
    b.fresh = a.out;
    for (int color = 0; color < 2; ++color) {
      g.color = color;
      PM_HIP(c, pm::launch_propagate(c->stream, g, b, counters,
                                     c->prop_lds_floor));
    }
    return PM_OK;
  }
  PM_HIP(c, pm::launch_propagate(c->stream, c->geo, a, counters,
                                 c->prop_lds_floor));
  return PM_OK;
}

// The smoothing of a view's new field (PMHuberStereo.cpp:552-567).  In the
// split schedule it runs on the smoothing stream once every band of the
// view's propagate is done, beside the other view's propagate; the next
// propagate of the same view, which reads the smooth field, waits for it
// (propagate_split), and so does every later user of the field buffers it
// reads, since they follow that propagate.
int smooth_after(pm_ctx* c, int view, const float4* planes, float theta,
                 bool split) {
  if (!split) return smooth_view(c, view, planes, theta, c->stream);
  for (hipEvent_t e : c->ev_band[view])
    PM_HIP(c, hipStreamWaitEvent(c->smoothing, e, 0));
  int rc;
  if ((rc = smooth_view(c, view, planes, theta, c->smoothing))) return rc;
  PM_HIP(c, hipEventRecord(c->ev_smooth[view], c->smoothing));
  return PM_OK;
}

// theta after iteration i (0-based) of a schedule of c->schedule_n:
// disparityMapAll's (i / N)^4 + theta_min (PMHuberStereo.cpp:575-576), or
// depthMap's ((i - 3) / (N - 4))^3 + 1e-3 from i = 3 on, unchanged before
// (PMHuberStereo.cpp:315-318).  The reference uses theta only for the
// iterations of its schedule (i < N - 1 for the value computed after
// iteration i), where the fraction lies in [0, 1).  Split pm_iterate calls
// can run past the schedule (i >= N - 1, or depthMap's N <= 4, whose
// 0 / 0 would give NaN): theta then holds at the schedule's end, the
// fraction clamped to 1.
float theta_after(const pm_ctx* c, int i, float theta) {
  if (depthmap_mode(c)) {
    if (i < 3) return theta;
    const float den = (float)c->schedule_n - 4.0f;
    float t = den > 0.0f ? ((float)i - 3.0f) / den : 1.0f;
    t = std::min(t, 1.0f);
    return t * t * t + kDepthTheta0;
  }
  const float t = std::min((float)i / (float)c->schedule_n, 1.0f);
  return std::pow(t, 4.0f) + c->prm.theta_min;
}

// pmhstereo_propagate_view for the left then the right view, each followed
// by its PM-Huber smoothing (PMHuberStereo.cpp:530-578)
int one_iteration(pm_ctx* c, unsigned long long* counters, bool split) {
  const int cur = c->cur, nxt = 1 - c->cur;
  unsigned long long* vbL = static_cast<unsigned long long*>(c->vbuf[0].ptr);
  unsigned long long* vbR = static_cast<unsigned long long*>(c->vbuf[1].ptr);
  const bool huber = c->prm.smoothing != PM_SMOOTH_NONE;
  const float theta = c->theta;
  pm::PropagateArgs a;
  a.dmax = c->prm.disparity_max;
  a.theta = huber ? theta * (depthmap_mode(c) ? kDepthThetaScale : kThetaScale)
                 : 0.0f;
  a.iter = c->iteration;
  int rc;

  a.self = view_bufs(c, 0);
  a.other = view_bufs(c, 1);
  a.old = state_ptr(c, 0, cur);
  a.smooth = huber ? f4buf(c->smooth[0]) : nullptr;
  a.viewin = vbR;
  a.src_other = state_ptr(c, 1, cur);
  a.viewout = vbL;
  a.out = state_ptr(c, 0, nxt);
  a.lr = 1;
  if ((rc = propagate_view(c, a, 0, vbL, counters, split))) return rc;
  if (huber && (rc = smooth_after(c, 0, state_ptr(c, 0, nxt), theta, split)))
    return rc;

  a.self = view_bufs(c, 1);
  a.other = view_bufs(c, 0);
  a.old = state_ptr(c, 1, cur);
  a.smooth = huber ? f4buf(c->smooth[1]) : nullptr;
  a.viewin = vbL;
  a.src_other = state_ptr(c, 0, nxt);
  a.viewout = vbR;
  a.out = state_ptr(c, 1, nxt);
  a.lr = 0;
  if ((rc = propagate_view(c, a, 1, vbR, counters, split))) return rc;
  if (huber && (rc = smooth_after(c, 1, state_ptr(c, 1, nxt), theta, split)))
    return rc;

  if (huber) {
    c->theta = theta_after(c, c->iteration, c->theta);
  }
  c->cur = nxt;
  c->iteration += 1;
  return PM_OK;
}

// n iterations; with the split schedule the context's stream finally
// waits for the other streams, so later work on it sees every result
int iterate_n(pm_ctx* c, int n, unsigned long long* counters) {
  const bool split = n > 0 && can_split(c);
  int rc;
  if (c->prm.smoothing != PM_SMOOTH_NONE) {
    // Capturing the PM-Huber iteration into a graph ended with the process
    // crashing inside the runtime's end of capture (plain PatchMatch
    // captures and replays correctly); refused before anything is queued.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    PM_HIP(c, hipStreamIsCapturing(c->stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
      return fail(c, PM_ERR_ARG, "pm_iterate with smoothing cannot be "
                  "recorded by stream capture");
  }
  if (split && (rc = split_rebase(c))) return rc;
  for (int i = 0; i < n; ++i)
    if ((rc = one_iteration(c, counters, split))) break;
  if (split) {
    // joined even after a failed launch, so that no work is left behind
    // unordered on the other streams: each records how far it has got and
    // the context's stream waits for that (the events are recorded afresh
    // by the next split_rebase).  The smoothing stream joins only when it
    // ran something: otherwise it never waited for the context's stream,
    // and under stream capture its record would lie outside the capture.
    const bool huber = c->prm.smoothing != PM_SMOOTH_NONE;
    for (int k = huber ? 0 : 1; k < kSplitParts; ++k) {
      const hipStream_t st = k == 0 ? c->smoothing : c->side[k - 1];
      hipEvent_t ev = k == 0 ? c->ev_smooth[1] : c->ev_band[1][k];
      hipError_t e = hipEventRecord(ev, st);
      if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, ev, 0);
      if (!rc && e != hipSuccess) rc = hip_fail(c, e, "join of the streams");
    }
  }
  return rc;
}

// N of the theta schedule for a pm_iterate call of n iterations
void start_schedule(pm_ctx* c, int n) {
  c->schedule_n = c->prm.schedule_iterations > 0
                      ? c->prm.schedule_iterations
                      : c->iteration + n;
  if (c->schedule_n < 1) c->schedule_n = 1;
}

// PM-Huber part of the setup: diffusion tensors of both images, cleared
// smooth fields, theta = theta_min (PMHuberStereo.cpp:452-458); depthMap:
// tensor alpha 12, theta 1e-3 (PMHuberStereo.cpp:153, 166, 190-195)
int huber_setup(pm_ctx* c) {
  c->theta = depthmap_mode(c) ? kDepthTheta0 : c->prm.theta_min;
  if (c->prm.smoothing == PM_SMOOTH_NONE) return PM_OK;
  const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (int v = 0; v < 2; ++v) {
    PM_HIP(c, pm::launch_diff_tensor(c->stream, c->geo,
                                     static_cast<const float*>(c->img[v].ptr),
                                     c->channels, c->geo.W * c->channels,
                                     depthmap_mode(c) ? kDepthAlpha
                                                      : kHuberAlpha,
                                     f4buf(c->dt[v])));
    PM_HIP(c, pm::launch_fill4(c->stream, c->geo.npix, zero,
                               f4buf(c->smooth[v])));
  }
  return PM_OK;
}

int postprocess(pm_ctx* c, int pp, float maxdispdiff, float maxdeg,
                float* d_out) {
  if (pp != PM_PP_NONE && pp != PM_PP_LRCHECK && pp != PM_PP_LRCHECK_FILL)
    return fail(c, PM_ERR_ARG, "unknown post-processing %d", pp);
  if (!(maxdispdiff >= 0.0f) || !(maxdeg >= 0.0f))
    return fail(c, PM_ERR_ARG, "maxdispdiff and maxdeg must be >= 0");
  const float4* L = state_ptr(c, 0, c->cur);
  const float4* R = state_ptr(c, 1, c->cur);
  float4* chk = static_cast<float4*>(c->checked.ptr);
  if (pp == PM_PP_NONE) {
    PM_HIP(c, pm::launch_to_disparity(c->stream, c->geo, L, 1.0f, 1, d_out));
  } else {
    PM_HIP(c, pm::launch_consistency(c->stream, c->geo, L, R, maxdispdiff,
                                     maxdeg, chk));
    if (pp == PM_PP_LRCHECK)
      PM_HIP(c, pm::launch_to_disparity(c->stream, c->geo, chk, 1.0f, 1,
                                        d_out));
    else
      PM_HIP(c, pm::launch_fill_depthmap(c->stream, c->geo, chk, 1.0f, d_out));
  }
  c->last_pp = pp;
  return PM_OK;
}

// With PM_SMOOTH_HUBER each view's smoothing runs beside the other view's
// propagate.  Propagate work-groups fill a CU's VGPRs (five per CU at 96
// VGPRs each) and leave no room for a k_rof_steps work-group, which then
// gets a CU only when propagate work-groups retire.  Asking for just over
// a quarter of a CU's LDS per propagate work-group caps them at three per
// CU and leaves LDS and VGPRs (3 x 96 + 120 of 512 per lane) for one
// smoothing work-group beside them.  Measured on the 720p pair: 8 coupled
// iterations 64.4 -> 61.5 ms (propagate at four per CU then).  Plain
// PatchMatch keeps the CU to itself (three per CU would cost it 4 %).
// The floor is kept only when three such work-groups and one k_rof_steps
// work-group fit the CU's LDS, allocated in 512-byte granules.
int share_cus_with_smoothing(pm_ctx* c) {
  int lds_cu = 0;
  if (hipDeviceGetAttribute(&lds_cu,
                            hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                            c->prm.device) != hipSuccess) {
    // no floor: the propagate keeps the CUs to itself.  The failed query
    // must not surface as the error of a later launch.
    (void)hipGetLastError();
    return PM_OK;
  }
  const size_t floor = (size_t)lds_cu / 4 + 64;
  const size_t granule = 512;
  const size_t held = (floor + granule - 1) / granule * granule;
  if (3 * held + pm::rof_steps_lds_bytes() <= (size_t)lds_cu)
    c->prop_lds_floor = floor;
  return PM_OK;
}

}  // namespace

extern "C" {

void pm_params_default(pm_params* p) {
  if (!p) return;
  p->radius = 5;
  p->disparity_max = 64.0f;
  p->color_weight = 12.0f;
  p->cost_kind = PM_COST_CENSUS;
  p->cost_weight = 16.0f;
  p->schedule = PM_SCHEDULE_JACOBI;
  p->smoothing = PM_SMOOTH_NONE;
  p->theta_min = 1e-4f;
  p->schedule_iterations = 0;
  p->device = 0;
  p->stream = nullptr;
}

int pm_create(const pm_params* p, pm_ctx** out) {
  if (!out) return PM_ERR_ARG;
  *out = nullptr;
  pm_ctx* c = new (std::nothrow) pm_ctx();
  if (!c) return PM_ERR_NOMEM;
  int rc = check_params(c, p);
  if (rc) {
    *out = c;  // caller may read pm_last_error, then pm_destroy
    return rc;
  }
  c->prm = *p;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    *out = c;
    return fail(c, PM_ERR_HIP, "no HIP device available (%s)",
                e == hipSuccess ? "0 devices" : hipGetErrorString(e));
  }
  if (p->device < 0 || p->device >= ndev) {
    *out = c;
    return fail(c, PM_ERR_ARG, "device %d outside [0, %d)", p->device, ndev);
  }
  *out = c;
  if ((rc = use_device(c))) return rc;
  if (p->smoothing != PM_SMOOTH_NONE && (rc = share_cus_with_smoothing(c)))
    return rc;
  if (p->stream) {
    c->stream = static_cast<hipStream_t>(p->stream);
  } else {
    PM_HIP(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = true;
  }
  return PM_OK;
}

void pm_destroy(pm_ctx* c) {
  if (!c) return;
  if (c->stream || c->geo.npix) {
    (void)hipSetDevice(c->prm.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
  }
  for (int v = 0; v < 2; ++v) {
    release(c->img[v]);
    release(c->tex[v]);
    release(c->sgn[v]);
    release(c->state[v][0]);
    release(c->state[v][1]);
    release(c->vbuf[v]);
    release(c->dt[v]);
    release(c->smooth[v]);
    release(c->rofb[v].nd);
    release(c->rofb[v].u);
    for (int k = 0; k < 2; ++k) {
      release(c->rofb[v].p[k]);
      release(c->rofb[v].q[k]);
    }
  }
  release(c->checked);
  release(c->scratch4);
  release(c->scratch4b);
  release(c->scratch1);
  release(c->counters);
  std::vector<hipStream_t> streams(c->side, c->side + kSplitParts - 1);
  streams.push_back(c->smoothing);
  for (hipStream_t st : streams)
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  for (int v = 0; v < 2; ++v) {
    for (hipEvent_t e : c->ev_band[v])
      if (e) (void)hipEventDestroy(e);
    if (c->ev_smooth[v]) (void)hipEventDestroy(c->ev_smooth[v]);
    if (c->ev_setup[v]) (void)hipEventDestroy(c->ev_setup[v]);
  }
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* pm_last_error(const pm_ctx* c) {
  return c ? c->err.c_str() : "context is NULL";
}

int pm_set_images(pm_ctx* c, const float* l, const float* r, int W, int H,
                  int channels, int stride) {
  return set_images(c, l, r, W, H, channels, stride, hipMemcpyHostToDevice);
}

int pm_set_images_device(pm_ctx* c, const float* l, const float* r, int W,
                         int H, int channels, int stride) {
  return set_images(c, l, r, W, H, channels, stride, hipMemcpyDeviceToDevice);
}

int pm_setup(pm_ctx* c) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = require_images(c)) || (rc = use_device(c))) return rc;
  if ((rc = prepare_views(c))) return rc;
  PM_HIP(c, pm::launch_init(c->stream, c->geo, view_bufs(c, 0),
                            view_bufs(c, 1), c->prm.disparity_max, 1,
                            state_ptr(c, 0, 0)));
  PM_HIP(c, pm::launch_init(c->stream, c->geo, view_bufs(c, 1),
                            view_bufs(c, 0), c->prm.disparity_max, 0,
                            state_ptr(c, 1, 0)));
  PM_HIP(c, pm::launch_viewbuf_clear(
                c->stream, c->geo,
                static_cast<unsigned long long*>(c->vbuf[1].ptr)));
  if ((rc = huber_setup(c))) return rc;
  c->cur = 0;
  c->iteration = 0;
  c->ready = true;
  c->last_pp = -1;
  return PM_OK;
}

int pm_setup_proposal(pm_ctx* c, pm_net* net) {
  return pm_setup_proposal_views(c, net, 2);
}

int pm_setup_proposal_views(pm_ctx* c, pm_net* net, int views) {
  if (!c) return PM_ERR_ARG;
  if (!net) return fail(c, PM_ERR_ARG, "net is NULL");
  if (views != 1 && views != 2)
    return fail(c, PM_ERR_ARG, "views must be 1 or 2, not %d", views);
  int rc;
  if ((rc = require_images(c)) || (rc = use_device(c))) return rc;
  if (pm::net::device_of(net) != c->prm.device)
    return fail(c, PM_ERR_ARG, "the net is on device %d, the context on %d",
                pm::net::device_of(net), c->prm.device);
  // The texels of both views on a side stream, beside the network's pass
  // on the context's stream: neither reads what the other writes, and the
  // network's deeper layers leave the GPU room.  The seeding waits for
  // both.  Everything queued on the context's stream before comes first.
  if (!c->side[0])
    PM_HIP(c, hipStreamCreateWithFlags(&c->side[0], hipStreamNonBlocking));
  for (hipEvent_t& e : c->ev_setup)
    if (!e) PM_HIP(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  const hipStream_t side = c->side[0];
  PM_HIP(c, hipEventRecord(c->ev_setup[0], c->stream));
  PM_HIP(c, hipStreamWaitEvent(side, c->ev_setup[0], 0));
  const int side_rc = prepare_views(c, side);
  const float* img[2] = {static_cast<const float*>(c->img[0].ptr),
                         static_cast<const float*>(c->img[1].ptr)};
  const int W = c->geo.W, H = c->geo.H, C = c->channels;
  // the network on the pair (left view) and, with views = 2, on the
  // mirrored pair (right view)
  const float* prop[2] = {nullptr, nullptr};
  int Wp = 0, os = 0;
  rc = side_rc ? side_rc
               : pm::net::proposal(net, c->stream, img[0], img[1], W, H, C,
                                   W * C, views, prop, &Wp, &os);
  // joined whatever failed, so that no side work is left unordered
  hipError_t e = hipEventRecord(c->ev_setup[1], side);
  if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, c->ev_setup[1], 0);
  if (side_rc) return side_rc;
  if (rc)
    return fail(c, PM_ERR_HIP, "proposal network: %s",
                pm::net::error_of(net));
  if (e != hipSuccess) return hip_fail(c, e, "join of the setup streams");
  for (int view = 0; view < views; ++view)
    PM_HIP(c, pm::launch_init_proposal(
                  c->stream, c->geo, view_bufs(c, view),
                  view_bufs(c, 1 - view), c->prm.disparity_max,
                  view == 0 ? 1 : 0, prop[view], os, Wp, PM_NET_DISP_SCALE,
                  state_ptr(c, view, 0)));
  if (views == 1) {
    // the right view from the seeded left view through the view scatter
    // (into the left view's slots, which its first propagate clears)
    unsigned long long* vbL = static_cast<unsigned long long*>(c->vbuf[0].ptr);
    PM_HIP(c, pm::launch_viewbuf_clear(c->stream, c->geo, vbL));
    PM_HIP(c, pm::launch_scatter(c->stream, c->geo, state_ptr(c, 0, 0), 1,
                                 vbL));
    PM_HIP(c, pm::launch_init_from_view(c->stream, c->geo,
                                        c->prm.disparity_max, vbL,
                                        state_ptr(c, 0, 0),
                                        state_ptr(c, 1, 0)));
  }
  PM_HIP(c, pm::launch_viewbuf_clear(
                c->stream, c->geo,
                static_cast<unsigned long long*>(c->vbuf[1].ptr)));
  if ((rc = huber_setup(c))) return rc;
  c->cur = 0;
  c->iteration = 0;
  c->ready = true;
  c->last_pp = -1;
  return PM_OK;
}

int pm_setup_prior(pm_ctx* c, const float* disparity, const float* normals) {
  if (!c) return PM_ERR_ARG;
  if (!disparity) return fail(c, PM_ERR_ARG, "disparity is NULL");
  int rc;
  if ((rc = require_images(c)) || (rc = use_device(c))) return rc;
  const size_t n = c->geo.npix;
  if ((rc = prepare_views(c))) return rc;
  float* d_disp = static_cast<float*>(c->scratch1.ptr);
  float4* d_norm = static_cast<float4*>(c->scratch4.ptr);
  if ((rc = to_device(c, d_disp, disparity, n * sizeof(float)))) return rc;
  if (normals) {
    if ((rc = to_device(c, d_norm, normals, n * sizeof(float4)))) return rc;
  } else {
    PM_HIP(c, pm::launch_disparity_to_normal(c->stream, c->geo, d_disp,
                                             d_norm));
  }
  PM_HIP(c, pm::launch_init_prior(c->stream, c->geo, view_bufs(c, 0),
                                  view_bufs(c, 1), c->prm.disparity_max, 1,
                                  d_disp, d_norm, state_ptr(c, 0, 0)));
  PM_HIP(c, pm::launch_init(c->stream, c->geo, view_bufs(c, 1),
                            view_bufs(c, 0), c->prm.disparity_max, 0,
                            state_ptr(c, 1, 0)));
  PM_HIP(c, pm::launch_viewbuf_clear(
                c->stream, c->geo,
                static_cast<unsigned long long*>(c->vbuf[1].ptr)));
  if ((rc = huber_setup(c))) return rc;
  // the seeded left planes start the left smooth field
  // (PMHuberStereo.cpp:473-477)
  if (c->prm.smoothing == PM_SMOOTH_HUBER &&
      (rc = smooth_view(c, 0, state_ptr(c, 0, 0), c->theta, c->stream)))
    return rc;
  PM_HIP(c, hipStreamSynchronize(c->stream));  // host inputs may be freed
  c->cur = 0;
  c->iteration = 0;
  c->ready = true;
  c->last_pp = -1;
  return PM_OK;
}

int pm_iterate(pm_ctx* c, int n) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = require_ready(c)) || (rc = use_device(c))) return rc;
  if (n < 0) return fail(c, PM_ERR_ARG, "negative iteration count %d", n);
  start_schedule(c, n);
  if ((rc = iterate_n(c, n, nullptr))) return rc;
  c->last_pp = -1;
  return PM_OK;
}

int pm_iteration(const pm_ctx* c) { return c ? c->iteration : -1; }

int pm_iterate_counted(pm_ctx* c, int n, uint64_t* cost_taps,
                       uint64_t* klt_taps) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = require_ready(c)) || (rc = use_device(c))) return rc;
  if (n < 0) return fail(c, PM_ERR_ARG, "negative iteration count %d", n);
  if (!cost_taps || !klt_taps) return fail(c, PM_ERR_ARG, "NULL counter");
  if ((rc = ensure(c, c->counters, 2 * sizeof(unsigned long long)))) return rc;
  unsigned long long* d = static_cast<unsigned long long*>(c->counters.ptr);
  const unsigned long long zero[2] = {0ull, 0ull};
  if ((rc = to_device(c, d, zero, sizeof zero))) return rc;
  start_schedule(c, n);
  if ((rc = iterate_n(c, n, d))) return rc;
  unsigned long long h[2];
  if ((rc = to_host(c, h, d, sizeof h))) return rc;
  *cost_taps = h[0];
  *klt_taps = h[1];
  c->last_pp = -1;
  return PM_OK;
}

int pm_get_planes(pm_ctx* c, int view, float* out) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = check_view(c, view)) || (rc = require_ready(c)) ||
      (rc = use_device(c)))
    return rc;
  if (!out) return fail(c, PM_ERR_ARG, "out is NULL");
  return to_host(c, out, state_ptr(c, view, c->cur),
                 c->geo.npix * sizeof(float4));
}

int pm_set_planes(pm_ctx* c, int view, const float* in, int iteration) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = check_view(c, view)) || (rc = require_ready(c)) ||
      (rc = use_device(c)))
    return rc;
  if (!in) return fail(c, PM_ERR_ARG, "in is NULL");
  if (iteration < 0) return fail(c, PM_ERR_ARG, "negative iteration");
  if ((rc = to_device(c, state_ptr(c, view, c->cur), in,
                      c->geo.npix * sizeof(float4))))
    return rc;
  // the buffer the next left step reads is the right field's scatter of
  // the previous iteration (empty before the first): rebuild it so that a
  // run resumes exactly from the two fields and the counter
  unsigned long long* vbR = static_cast<unsigned long long*>(c->vbuf[1].ptr);
  PM_HIP(c, pm::launch_viewbuf_clear(c->stream, c->geo, vbR));
  if (iteration > 0)
    PM_HIP(c, pm::launch_scatter(c->stream, c->geo, state_ptr(c, 1, c->cur), 0,
                                 vbR));
  PM_HIP(c, hipStreamSynchronize(c->stream));
  c->iteration = iteration;
  // PM-Huber: theta of the next iteration when the schedule length is
  // fixed (PMHuberStereo.cpp:575-576); the smooth fields stay as they are
  if (c->prm.smoothing != PM_SMOOTH_NONE && c->prm.schedule_iterations > 0) {
    c->schedule_n = c->prm.schedule_iterations;
    float theta = depthmap_mode(c) ? kDepthTheta0 : c->prm.theta_min;
    for (int i = 0; i < iteration; ++i) theta = theta_after(c, i, theta);
    c->theta = theta;
  }
  c->last_pp = -1;
  return PM_OK;
}

int pm_cost(pm_ctx* c, int view, const float* planes, float* cost) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = check_view(c, view)) || (rc = require_ready(c)) ||
      (rc = use_device(c)))
    return rc;
  if (!planes || !cost) return fail(c, PM_ERR_ARG, "NULL buffer");
  float4* d_pl = static_cast<float4*>(c->scratch4.ptr);
  float* d_cost = static_cast<float*>(c->scratch1.ptr);
  if ((rc = to_device(c, d_pl, planes, c->geo.npix * sizeof(float4))))
    return rc;
  PM_HIP(c, pm::launch_cost_field(c->stream, c->geo, view_bufs(c, view),
                                  view_bufs(c, 1 - view), d_pl, d_cost));
  return to_host(c, cost, d_cost, c->geo.npix * sizeof(float));
}

int pm_disparity(pm_ctx* c, int pp, float maxdispdiff, float maxdeg,
                 float* disparity, float* checked) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = require_ready(c)) || (rc = use_device(c))) return rc;
  if (!disparity) return fail(c, PM_ERR_ARG, "disparity is NULL");
  float* d_disp = static_cast<float*>(c->scratch1.ptr);
  if ((rc = postprocess(c, pp, maxdispdiff, maxdeg, d_disp))) return rc;
  if (checked) {
    const void* src = pp == PM_PP_NONE ? (const void*)state_ptr(c, 0, c->cur)
                                       : c->checked.ptr;
    PM_HIP(c, hipMemcpyAsync(checked, src, c->geo.npix * sizeof(float4),
                             hipMemcpyDeviceToHost, c->stream));
  }
  return to_host(c, disparity, d_disp, c->geo.npix * sizeof(float));
}

int pm_disparity_device(pm_ctx* c, int pp, float maxdispdiff, float maxdeg,
                        float* d_disparity) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = require_ready(c)) || (rc = use_device(c))) return rc;
  if (!d_disparity) return fail(c, PM_ERR_ARG, "d_disparity is NULL");
  return postprocess(c, pp, maxdispdiff, maxdeg, d_disparity);
}

int pm_normals(pm_ctx* c, int pp, float* normals) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = require_ready(c)) || (rc = use_device(c))) return rc;
  if (!normals) return fail(c, PM_ERR_ARG, "normals is NULL");
  if (pp != PM_PP_NONE && pp != PM_PP_LRCHECK && pp != PM_PP_LRCHECK_FILL)
    return fail(c, PM_ERR_ARG, "unknown post-processing %d", pp);
  if (pp != PM_PP_NONE && c->last_pp != pp)
    return fail(c, PM_ERR_STATE,
                "call pm_disparity with the same post-processing first");
  const float4* src = pp == PM_PP_NONE
                          ? state_ptr(c, 0, c->cur)
                          : static_cast<const float4*>(c->checked.ptr);
  float4* d_n = static_cast<float4*>(c->scratch4.ptr);
  PM_HIP(c, pm::launch_fill_normalmap(c->stream, c->geo, src, d_n));
  return to_host(c, normals, d_n, c->geo.npix * sizeof(float4));
}

int pm_sync(pm_ctx* c) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = use_device(c))) return rc;
  if (c->stream) PM_HIP(c, hipStreamSynchronize(c->stream));
  return PM_OK;
}

int pm_get_texels(pm_ctx* c, int view, float* out) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = check_view(c, view)) || (rc = require_ready(c)) ||
      (rc = use_device(c)))
    return rc;
  if (!out) return fail(c, PM_ERR_ARG, "out is NULL");
  // the W texels of each row without the edge padding: colour plane, then
  // gradient plane, interleaved on the host into (colour, gradient) pairs
  const int W = c->geo.W, H = c->geo.H;
  const size_t rec = sizeof(float4), npix = (size_t)W * H;
  const size_t plane = (size_t)c->geo.TS * H * rec;
  std::vector<float4> tmp;
  try {
    tmp.resize(2 * npix);
  } catch (const std::bad_alloc&) {
    return fail(c, PM_ERR_NOMEM, "out of host memory");
  }
  for (int k = 0; k < 2; ++k) {
    const char* src = static_cast<const char*>(c->tex[view].ptr) + k * plane +
                      pm::kTexPadL * rec;
    PM_HIP(c, hipMemcpy2DAsync(tmp.data() + k * npix, W * rec, src,
                               c->geo.TS * rec, W * rec, H,
                               hipMemcpyDeviceToHost, c->stream));
  }
  PM_HIP(c, hipStreamSynchronize(c->stream));
  float4* o = reinterpret_cast<float4*>(out);
  for (size_t p = 0; p < npix; ++p) {
    o[2 * p] = tmp[p];
    o[2 * p + 1] = tmp[npix + p];
  }
  return PM_OK;
}

int pm_draws(pm_ctx* c, uint64_t k, int n, float* out) {
  if (!c) return PM_ERR_ARG;
  if (n < 0 || n > (1 << 26)) return fail(c, PM_ERR_ARG, "n out of range");
  if (!out) return fail(c, PM_ERR_ARG, "out is NULL");
  if (n == 0) return PM_OK;
  int rc;
  if ((rc = use_device(c))) return rc;
  DevBuf tmp;
  if ((rc = ensure(c, tmp, (size_t)n * sizeof(float)))) return rc;
  hipError_t e = pm::launch_draws(c->stream, k, n, static_cast<float*>(tmp.ptr));
  if (e == hipSuccess) rc = to_host(c, out, tmp.ptr, (size_t)n * sizeof(float));
  else rc = hip_fail(c, e, "launch_draws");
  release(tmp);
  return rc;
}

int pm_klt(pm_ctx* c, int view, const float* planes, int passes, float* out) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = check_view(c, view)) || (rc = require_ready(c)) ||
      (rc = use_device(c)))
    return rc;
  if (!planes || !out) return fail(c, PM_ERR_ARG, "NULL buffer");
  if (passes < 0 || passes > 64) return fail(c, PM_ERR_ARG, "passes out of range");
  float4* d_in = static_cast<float4*>(c->scratch4.ptr);
  float4* d_out = static_cast<float4*>(c->scratch4b.ptr);
  if ((rc = to_device(c, d_in, planes, c->geo.npix * sizeof(float4))))
    return rc;
  PM_HIP(c, pm::launch_klt_field(c->stream, c->geo, view_bufs(c, view),
                                 view_bufs(c, 1 - view), d_in, passes, d_out));
  return to_host(c, out, d_out, c->geo.npix * sizeof(float4));
}

int pm_copy_probe(pm_ctx* c, const float* d_src, float* d_dst, size_t bytes) {
  if (!c) return PM_ERR_ARG;
  if (!d_src || !d_dst) return fail(c, PM_ERR_ARG, "NULL buffer");
  if (bytes % 16 != 0 || (uintptr_t)d_src % 16 || (uintptr_t)d_dst % 16)
    return fail(c, PM_ERR_ARG, "copy probe needs 16-byte sizes and alignment");
  int rc;
  if ((rc = use_device(c))) return rc;
  PM_HIP(c, pm::launch_copy_probe(c->stream,
                                  reinterpret_cast<const float4*>(d_src),
                                  reinterpret_cast<float4*>(d_dst),
                                  bytes / 16));
  return PM_OK;
}

static int require_huber(pm_ctx* c) {
  if (c->prm.smoothing == PM_SMOOTH_NONE)
    return fail(c, PM_ERR_STATE, "the context does not smooth (PM_SMOOTH_NONE)");
  return PM_OK;
}

int pm_get_smooth(pm_ctx* c, int view, float* out) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = check_view(c, view)) || (rc = require_ready(c)) ||
      (rc = require_huber(c)) || (rc = use_device(c)))
    return rc;
  if (!out) return fail(c, PM_ERR_ARG, "out is NULL");
  return to_host(c, out, c->smooth[view].ptr, c->geo.npix * sizeof(float4));
}

int pm_get_tensor(pm_ctx* c, int view, float* out) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = check_view(c, view)) || (rc = require_ready(c)) ||
      (rc = require_huber(c)) || (rc = use_device(c)))
    return rc;
  if (!out) return fail(c, PM_ERR_ARG, "out is NULL");
  return to_host(c, out, c->dt[view].ptr, c->geo.npix * sizeof(float4));
}

int pm_rof(pm_ctx* c, int view, const float* f, const float* lambda,
           int iters, float* out) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = check_view(c, view)) || (rc = require_ready(c)) ||
      (rc = require_huber(c)) || (rc = use_device(c)))
    return rc;
  if (!f || !lambda || !out) return fail(c, PM_ERR_ARG, "NULL argument");
  if (iters < 1) return fail(c, PM_ERR_ARG, "iters must be >= 1");
  const size_t bytes = c->geo.npix * sizeof(float4);
  float4* d_f = static_cast<float4*>(c->scratch4.ptr);
  float4* d_out = static_cast<float4*>(c->scratch4b.ptr);
  if ((rc = to_device(c, d_f, f, bytes))) return rc;
  const float4 lam = make_float4(lambda[0], lambda[1], lambda[2], lambda[3]);
  if ((rc = rof(c, view, d_f, lam, iters, d_out, c->stream))) return rc;
  return to_host(c, out, d_out, bytes);
}

int pm_rof_device(pm_ctx* c, int view, const float* d_f, const float* lambda,
                  int iters, float* d_out) {
  if (!c) return PM_ERR_ARG;
  int rc;
  if ((rc = check_view(c, view)) || (rc = require_ready(c)) ||
      (rc = require_huber(c)) || (rc = use_device(c)))
    return rc;
  if (!d_f || !lambda || !d_out) return fail(c, PM_ERR_ARG, "NULL argument");
  if (iters < 1) return fail(c, PM_ERR_ARG, "iters must be >= 1");
  const size_t bytes = c->geo.npix * sizeof(float4);
  const char* f0 = reinterpret_cast<const char*>(d_f);
  const char* o0 = reinterpret_cast<const char*>(d_out);
  if (f0 < o0 + bytes && o0 < f0 + bytes)
    return fail(c, PM_ERR_ARG, "d_f and d_out overlap");
  const float4 lam = make_float4(lambda[0], lambda[1], lambda[2], lambda[3]);
  return rof(c, view, reinterpret_cast<const float4*>(d_f), lam, iters,
             reinterpret_cast<float4*>(d_out), c->stream);
}

int pm_math_selftest(pm_ctx* c, const float* in, int n, float* out) {
  if (!c) return PM_ERR_ARG;
  if (n < 0 || n > (1 << 24)) return fail(c, PM_ERR_ARG, "n out of range");
  if (!in || !out) return fail(c, PM_ERR_ARG, "NULL buffer");
  if (n == 0) return PM_OK;
  int rc;
  if ((rc = use_device(c))) return rc;
  DevBuf din, dout;
  if ((rc = ensure(c, din, (size_t)n * sizeof(float))) ||
      (rc = ensure(c, dout, (size_t)n * 5 * sizeof(float)))) {
    release(din);
    release(dout);
    return rc;
  }
  rc = to_device(c, din.ptr, in, (size_t)n * sizeof(float));
  if (!rc) {
    hipError_t e = pm::launch_math(c->stream, static_cast<float*>(din.ptr), n,
                                   static_cast<float*>(dout.ptr));
    rc = e == hipSuccess ? to_host(c, out, dout.ptr, (size_t)n * 5 * sizeof(float))
                         : hip_fail(c, e, "launch_math");
  }
  release(din);
  release(dout);
  return rc;
}

}  // extern "C"
