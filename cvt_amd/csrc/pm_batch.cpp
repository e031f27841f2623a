// pm_batch.cpp -- C ABI of batches of stereo pairs (include/pm_batch.h).
//
// Every pair gets the buffers a single-pair context gives a view pair
// (texel planes, sign bits, two plane fields and a scatter slot buffer per
// view, the LR-checked field and the disparity), carved out of one
// allocation.  A descriptor per pair in device memory carries its
// geometry, those addresses and its place in the sequence of tile blocks,
// and every kernel covers all pairs in one launch.  The sequence of a run
// (pm_setup + pm_iterate + pm_disparity of a plain-PatchMatch context,
// PMHuberStereo.cpp:445-545) is recorded into a HIP graph on the first
// run with a given set of arguments and replayed afterwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <new>
#include <string>
#include <vector>

#include "pm_batch.h"
#include "pm_engine.h"
#include "pm_net_internal.h"

namespace {

constexpr int kMaxRadius = 16;
constexpr size_t kAlign = 256;

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// arguments a recorded graph was made for
struct RunKey {
  int iterations = -1, pp = -1;
  float maxdispdiff = 0.0f, maxdeg = 0.0f, err_abs = -1.0f, err_rel = 0.0f;
  pm::BatchShape shape;
  const pm_net* net = nullptr;
  long net_generation = 0;
  bool operator==(const RunKey& o) const {
    return net == o.net && net_generation == o.net_generation &&
           iterations == o.iterations && pp == o.pp &&
           maxdispdiff == o.maxdispdiff && maxdeg == o.maxdeg &&
           err_abs == o.err_abs && err_rel == o.err_rel &&
           shape.rows == o.shape.rows &&
           shape.workgroups == o.shape.workgroups;
  }
};

}  // namespace

struct pm_batch {
  pm_params prm{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;

  int n = 0;
  int maxW = 0, maxH = 0;
  std::vector<pm::BatchPair> host;   // descriptors (host copy)
  pm::BatchPair* dev = nullptr;      // descriptors (device)
  void* arena = nullptr;
  size_t arena_bytes = 0;
  long long nitems = 0;              // tile blocks of all pairs
  pm::BatchShape shape;
  bool ran = false;
  int cur = 0;                       // plane field of the last run's result
  bool counted = false;              // the last run counted errors

  pm_net* net = nullptr;              // proposal seeding (may be null)
  int net_views = 2;                  // views the proposal seeds (1 or 2)

  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  RunKey key;
  float tuned_ms = 0.0f;
};

namespace {

int fail(pm_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b) b->err = buf;
  return code;
}

int hip_fail(pm_batch* b, hipError_t e, const char* what) {
  return fail(b, e == hipErrorOutOfMemory ? PM_ERR_NOMEM : PM_ERR_HIP,
              "%s: %s", what, hipGetErrorString(e));
}

#define PB_HIP(b, call)                                     \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return hip_fail(b, e_, #call);    \
  } while (0)

int check_pair(pm_batch* b, int i, const pm_batch_pair& q) {
  if (q.width < 1 || q.height < 1)
    return fail(b, PM_ERR_ARG, "pair %d: size %d x %d", i, q.width, q.height);
  if (q.channels != 1 && q.channels != 4)
    return fail(b, PM_ERR_ARG, "pair %d: channels must be 1 or 4", i);
  if (q.stride < q.width * q.channels)
    return fail(b, PM_ERR_ARG, "pair %d: stride %d below width x channels",
                i, q.stride);
  if (!(q.disparity_max > 0.0f) || !std::isfinite(q.disparity_max))
    return fail(b, PM_ERR_ARG, "pair %d: disparity_max must be finite and "
                "> 0", i);
  if (!q.d_left || !q.d_right)
    return fail(b, PM_ERR_ARG, "pair %d: NULL image", i);
  if (q.d_gt && q.gt_stride < q.width)
    return fail(b, PM_ERR_ARG, "pair %d: gt_stride below width", i);
  return PM_OK;
}

int check_params(pm_batch* b, const pm_params* p) {
  if (!p) return fail(b, PM_ERR_ARG, "params is NULL");
  if (p->radius < 1 || p->radius > kMaxRadius)
    return fail(b, PM_ERR_ARG, "radius %d outside [1, %d]", p->radius,
                kMaxRadius);
  if (!(p->color_weight >= 0.0f) || !std::isfinite(p->color_weight))
    return fail(b, PM_ERR_ARG, "color_weight must be finite and >= 0");
  if (p->cost_kind != PM_COST_CENSUS && p->cost_kind != PM_COST_COLOR_GRAD)
    return fail(b, PM_ERR_ARG, "unknown cost_kind %d", p->cost_kind);
  if (!(p->cost_weight >= 0.0f) || !std::isfinite(p->cost_weight))
    return fail(b, PM_ERR_ARG, "cost_weight must be finite and >= 0");
  if (p->smoothing != PM_SMOOTH_NONE)
    return fail(b, PM_ERR_ARG, "a batch runs plain PatchMatch "
                "(smoothing must be PM_SMOOTH_NONE)");
  if (p->schedule != PM_SCHEDULE_JACOBI)
    return fail(b, PM_ERR_ARG, "a batch runs the Jacobi schedule");
  if (p->device < 0)
    return fail(b, PM_ERR_ARG, "device ordinal %d is negative", p->device);
  return PM_OK;
}

// descriptor fields that depend on the tile rows: blocks per row and each
// pair's first block in the batch sequence
void number_blocks(pm_batch* b) {
  long long item = 0;
  for (pm::BatchPair& P : b->host) {
    P.nbx = (P.g.W + 63) / 64;
    P.item0 = item;
    item += pm::batch_blocks(P.g.W, P.g.H, b->shape.rows);
  }
  b->nitems = item;
}

int upload(pm_batch* b) {
  PB_HIP(b, hipMemcpyAsync(b->dev, b->host.data(),
                           b->host.size() * sizeof(pm::BatchPair),
                           hipMemcpyHostToDevice, b->stream));
  // the host copy may change again before the copy ran
  PB_HIP(b, hipStreamSynchronize(b->stream));
  return PM_OK;
}

void drop_graph(pm_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->key = RunKey();
}

size_t tile_lds(const pm_batch* b, int rows) {
  return pm::batch_lds_bytes(b->prm.radius, b->prm.cost_kind, rows);
}

// the proposal-seeded init of every pair (pm_setup_proposal_views): the
// network's pass over pair i (and its mirror), then the init of the seeded
// views, in stream order (the next pass overwrites the network's output);
// with one seeded view the right view follows from the left through the
// view scatter (into the left view's slots, cleared again by its first
// propagate)
int enqueue_proposal_init(pm_batch* b, hipStream_t st) {
  for (int i = 0; i < b->n; ++i) {
    const pm::BatchPair& P = b->host[i];
    const float* prop[2] = {nullptr, nullptr};
    int Wp = 0, os = 0;
    if (pm::net::proposal(b->net, st, P.img[0], P.img[1], P.g.W, P.g.H,
                          P.channels, P.stride, b->net_views, prop, &Wp,
                          &os))
      return fail(b, PM_ERR_HIP, "proposal network, pair %d: %s", i,
                  pm::net::error_of(b->net));
    for (int view = 0; view < b->net_views; ++view)
      PB_HIP(b, pm::launch_init_proposal(st, P.g, P.v[view], P.v[1 - view],
                                         P.dmax, view == 0 ? 1 : 0,
                                         prop[view], os, Wp,
                                         PM_NET_DISP_SCALE,
                                         P.state[view][0]));
    if (b->net_views == 1) {
      PB_HIP(b, pm::launch_viewbuf_clear(st, P.g, P.vbuf[0]));
      PB_HIP(b, pm::launch_scatter(st, P.g, P.state[0][0], 1, P.vbuf[0]));
      PB_HIP(b, pm::launch_init_from_view(st, P.g, P.dmax, P.vbuf[0],
                                          P.state[0][0], P.state[1][0]));
    }
  }
  return PM_OK;
}

// texels and init of both views (random, or seeded by the proposal
// network), scatter slots of the right view cleared (pm_setup,
// pm_setup_proposal)
int enqueue_setup(pm_batch* b, hipStream_t st) {
  const size_t lds = tile_lds(b, b->shape.rows);
  PB_HIP(b, pm::launch_prepare_batch(st, b->dev, b->n, b->maxW, b->maxH,
                                     b->prm.radius, b->prm.color_weight));
  if (b->net) {
    int rc;
    if ((rc = enqueue_proposal_init(b, st))) return rc;
  } else {
    for (int view = 0; view < 2; ++view)
      PB_HIP(b, pm::launch_init_batch(st, b->dev, b->n, b->nitems, view,
                                      b->prm.cost_kind, lds, b->shape));
  }
  PB_HIP(b, pm::launch_viewbuf_clear_batch(st, b->dev, b->n, b->maxW,
                                           b->maxH, 1));
  return PM_OK;
}

// one view's propagate of iteration `iter` (one_iteration in pm_capi.cpp):
// its scatter slots cleared, then the propagate
int enqueue_propagate(pm_batch* b, hipStream_t st, int view, int cur,
                      int iter) {
  PB_HIP(b, pm::launch_viewbuf_clear_batch(st, b->dev, b->n, b->maxW,
                                           b->maxH, view));
  PB_HIP(b, pm::launch_propagate_batch(st, b->dev, b->n, b->nitems, view,
                                       cur, iter, b->prm.cost_kind,
                                       tile_lds(b, b->shape.rows), b->shape));
  return PM_OK;
}

int enqueue_run(pm_batch* b, hipStream_t st, const RunKey& k) {
  int rc;
  if ((rc = enqueue_setup(b, st))) return rc;
  int cur = 0;
  for (int i = 0; i < k.iterations; ++i) {
    if ((rc = enqueue_propagate(b, st, 0, cur, i))) return rc;
    if ((rc = enqueue_propagate(b, st, 1, cur, i))) return rc;
    cur = 1 - cur;
  }
  PB_HIP(b, pm::launch_postprocess_batch(st, b->dev, b->n, b->maxW, b->maxH,
                                         cur, k.pp, k.maxdispdiff, k.maxdeg));
  if (k.err_abs >= 0.0f)
    PB_HIP(b, pm::launch_eval_error_batch(st, b->dev, b->n, b->maxW, b->maxH,
                                          k.err_abs, k.err_rel));
  return PM_OK;
}

// dynamic LDS above 64 KiB is allowed per kernel outside the capture
int allow_shape(pm_batch* b, int rows) {
  int resident = 0;
  PB_HIP(b, pm::batch_resident_per_cu(b->prm.cost_kind, rows,
                                      tile_lds(b, rows), &resident));
  if (resident < 1)
    return fail(b, PM_ERR_ARG, "tile rows %d at radius %d do not fit a CU",
                rows, b->prm.radius);
  return PM_OK;
}

int record(pm_batch* b, const RunKey& k) {
  drop_graph(b);
  int rc;
  if ((rc = allow_shape(b, k.shape.rows))) return rc;
  PB_HIP(b, hipStreamBeginCapture(b->stream,
                                  hipStreamCaptureModeThreadLocal));
  rc = enqueue_run(b, b->stream, k);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(b, e, "hipStreamEndCapture");
  b->graph = g;
  PB_HIP(b, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  b->key = k;
  return PM_OK;
}

int check_index(pm_batch* b, int i) {
  if (i < 0 || i >= b->n)
    return fail(b, PM_ERR_ARG, "pair %d outside [0, %d)", i, b->n);
  return PM_OK;
}

int require_run(pm_batch* b) {
  if (!b->ran) return fail(b, PM_ERR_STATE, "no pm_batch_run yet");
  return PM_OK;
}

// the net's buffers sized for the largest pair (padded to multiples of 8)
int reserve_net(pm_batch* b, pm_net* net) {
  int hp = 8, wp = 8;
  long long best = 0;
  for (const pm::BatchPair& P : b->host) {
    const int h = (P.g.H + 7) / 8 * 8, w = (P.g.W + 7) / 8 * 8;
    if ((long long)h * w > best) {
      best = (long long)h * w;
      hp = h;
      wp = w;
    }
  }
  if (pm::net::reserve(net, hp, wp))
    return fail(b, PM_ERR_HIP, "proposal network: %s",
                pm::net::error_of(net));
  PB_HIP(b, hipSetDevice(b->prm.device));
  return PM_OK;
}

int valid_rows(int rows) {
  for (int r : pm::kBatchRowChoices)
    if (r == rows) return 1;
  return 0;
}

}  // namespace

extern "C" {

int pm_batch_create(const pm_params* p, const pm_batch_pair* pairs, int n,
                    pm_batch** out) {
  if (!out) return PM_ERR_ARG;
  *out = nullptr;
  pm_batch* b = new (std::nothrow) pm_batch();
  if (!b) return PM_ERR_NOMEM;
  *out = b;
  int rc;
  if ((rc = check_params(b, p))) return rc;
  if (n < 1 || !pairs) return fail(b, PM_ERR_ARG, "no pairs");
  for (int i = 0; i < n; ++i)
    if ((rc = check_pair(b, i, pairs[i]))) return rc;
  b->prm = *p;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0)
    return fail(b, PM_ERR_HIP, "no HIP device available (%s)",
                e == hipSuccess ? "0 devices" : hipGetErrorString(e));
  if (p->device >= ndev)
    return fail(b, PM_ERR_ARG, "device %d outside [0, %d)", p->device, ndev);
  PB_HIP(b, hipSetDevice(p->device));
  if (p->stream) {
    b->stream = static_cast<hipStream_t>(p->stream);
  } else {
    PB_HIP(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->own_stream = true;
  }

  // one allocation: every pair's buffers, 256-byte aligned
  b->n = n;
  b->host.resize(n);
  std::vector<size_t> off(n);
  size_t total = align_up(n * sizeof(pm::BatchPair));
  for (int i = 0; i < n; ++i) {
    const size_t W = pairs[i].width, H = pairs[i].height, np = W * H;
    off[i] = total;
    total += 2 * align_up(pm::tex_stride((int)W) * H * 2 * sizeof(float4));
    total += 2 * align_up(np);                        // sign bits
    total += 4 * align_up(np * sizeof(float4));       // plane fields
    total += 2 * align_up(np * sizeof(unsigned long long));
    total += align_up(np * sizeof(float4));           // checked
    total += align_up(np * sizeof(float));            // disparity
    total += align_up(3 * sizeof(unsigned long long));
  }
  e = hipMalloc(&b->arena, total);
  if (e != hipSuccess) {
    b->arena = nullptr;
    return hip_fail(b, e, "hipMalloc of the batch");
  }
  b->arena_bytes = total;
  char* base = static_cast<char*>(b->arena);
  b->dev = reinterpret_cast<pm::BatchPair*>(base);
  for (int i = 0; i < n; ++i) {
    const pm_batch_pair& q = pairs[i];
    pm::BatchPair& P = b->host[i];
    const size_t W = q.width, H = q.height, np = W * H;
    char* at = base + off[i];
    auto take = [&at](size_t bytes) {
      char* r = at;
      at += align_up(bytes);
      return r;
    };
    P.g.W = q.width;
    P.g.H = q.height;
    P.g.R = p->radius;
    P.g.TS = pm::tex_stride(q.width);
    P.g.npix = np;
    P.g.color_weight = p->color_weight;
    P.g.cost_kind = p->cost_kind;
    P.g.cost_weight = p->cost_weight;
    for (int v = 0; v < 2; ++v)
      P.v[v].tex = reinterpret_cast<float4*>(
          take(pm::tex_stride(q.width) * H * 2 * sizeof(float4)));
    for (int v = 0; v < 2; ++v)
      P.v[v].sgn = reinterpret_cast<uint8_t*>(take(np));
    for (int v = 0; v < 2; ++v)
      for (int k = 0; k < 2; ++k)
        P.state[v][k] = reinterpret_cast<float4*>(take(np * sizeof(float4)));
    for (int v = 0; v < 2; ++v)
      P.vbuf[v] = reinterpret_cast<unsigned long long*>(
          take(np * sizeof(unsigned long long)));
    P.checked = reinterpret_cast<float4*>(take(np * sizeof(float4)));
    P.disp = reinterpret_cast<float*>(take(np * sizeof(float)));
    P.counts = reinterpret_cast<unsigned long long*>(
        take(3 * sizeof(unsigned long long)));
    P.img[0] = q.d_left;
    P.img[1] = q.d_right;
    P.gt = q.d_gt;
    P.gt_stride = q.gt_stride;
    P.channels = q.channels;
    P.stride = q.stride;
    P.dmax = q.disparity_max;
    b->maxW = std::max(b->maxW, q.width);
    b->maxH = std::max(b->maxH, q.height);
  }
  number_blocks(b);
  return upload(b);
}

void pm_batch_destroy(pm_batch* b) {
  if (!b) return;
  if (b->stream || b->arena) {
    (void)hipSetDevice(b->prm.device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
  }
  drop_graph(b);
  if (b->arena) (void)hipFree(b->arena);
  if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char* pm_batch_last_error(const pm_batch* b) {
  return b ? b->err.c_str() : "NULL batch";
}

int pm_batch_size(const pm_batch* b) { return b ? b->n : -1; }

int pm_batch_set_pair(pm_batch* b, int i, const pm_batch_pair* q) {
  if (!b) return PM_ERR_ARG;
  int rc;
  if ((rc = check_index(b, i))) return rc;
  if (!q) return fail(b, PM_ERR_ARG, "pair is NULL");
  if ((rc = check_pair(b, i, *q))) return rc;
  pm::BatchPair& P = b->host[i];
  if (q->width != P.g.W || q->height != P.g.H || q->channels != P.channels ||
      q->disparity_max != P.dmax)
    return fail(b, PM_ERR_ARG, "pair %d: size, channels and disparity_max "
                "are fixed at creation", i);
  PB_HIP(b, hipSetDevice(b->prm.device));
  P.img[0] = q->d_left;
  P.img[1] = q->d_right;
  P.stride = q->stride;
  P.gt = q->d_gt;
  P.gt_stride = q->gt_stride;
  return upload(b);
}

int pm_batch_run(pm_batch* b, int iterations, int pp, float maxdispdiff,
                 float maxdeg, float err_abs, float err_rel) {
  if (!b) return PM_ERR_ARG;
  if (iterations < 0)
    return fail(b, PM_ERR_ARG, "negative iteration count %d", iterations);
  if (pp != PM_PP_NONE && pp != PM_PP_LRCHECK && pp != PM_PP_LRCHECK_FILL)
    return fail(b, PM_ERR_ARG, "unknown post-processing %d", pp);
  if (!(maxdispdiff >= 0.0f) || !(maxdeg >= 0.0f))
    return fail(b, PM_ERR_ARG, "maxdispdiff and maxdeg must be >= 0");
  if (std::isnan(err_abs) || !(err_rel == err_rel))
    return fail(b, PM_ERR_ARG, "error thresholds must not be NaN");
  PB_HIP(b, hipSetDevice(b->prm.device));
  RunKey k;
  k.iterations = iterations;
  k.pp = pp;
  k.maxdispdiff = maxdispdiff;
  k.maxdeg = maxdeg;
  k.err_abs = err_abs >= 0.0f ? err_abs : -1.0f;
  k.err_rel = err_abs >= 0.0f ? err_rel : 0.0f;
  k.shape = b->shape;
  int rc;
  if (b->net) {
    // the net may have been used on larger inputs since: size it again,
    // and record again when that moved its buffers
    if ((rc = reserve_net(b, b->net))) return rc;
    k.net = b->net;
    k.net_generation = pm::net::generation(b->net);
  }
  if (!b->exec || !(b->key == k))
    if ((rc = record(b, k))) return rc;
  PB_HIP(b, hipGraphLaunch(b->exec, b->stream));
  b->ran = true;
  b->cur = iterations & 1;
  b->counted = k.err_abs >= 0.0f;
  return PM_OK;
}

int pm_batch_get_disparity(pm_batch* b, int i, float* out) {
  if (!b) return PM_ERR_ARG;
  int rc;
  if ((rc = check_index(b, i)) || (rc = require_run(b))) return rc;
  if (!out) return fail(b, PM_ERR_ARG, "out is NULL");
  PB_HIP(b, hipSetDevice(b->prm.device));
  const pm::BatchPair& P = b->host[i];
  PB_HIP(b, hipMemcpyAsync(out, P.disp, P.g.npix * sizeof(float),
                           hipMemcpyDeviceToHost, b->stream));
  PB_HIP(b, hipStreamSynchronize(b->stream));
  return PM_OK;
}

int pm_batch_get_planes(pm_batch* b, int i, int view, float* out) {
  if (!b) return PM_ERR_ARG;
  int rc;
  if ((rc = check_index(b, i)) || (rc = require_run(b))) return rc;
  if (view != 0 && view != 1)
    return fail(b, PM_ERR_ARG, "view must be 0 (left) or 1 (right)");
  if (!out) return fail(b, PM_ERR_ARG, "out is NULL");
  PB_HIP(b, hipSetDevice(b->prm.device));
  const pm::BatchPair& P = b->host[i];
  PB_HIP(b, hipMemcpyAsync(out, P.state[view][b->cur],
                           P.g.npix * sizeof(float4), hipMemcpyDeviceToHost,
                           b->stream));
  PB_HIP(b, hipStreamSynchronize(b->stream));
  return PM_OK;
}

int pm_batch_get_error_counts(pm_batch* b, uint64_t* counts) {
  if (!b) return PM_ERR_ARG;
  int rc;
  if ((rc = require_run(b))) return rc;
  if (!counts) return fail(b, PM_ERR_ARG, "counts is NULL");
  if (!b->counted)
    return fail(b, PM_ERR_STATE, "the last run counted no errors "
                "(err_abs < 0)");
  PB_HIP(b, hipSetDevice(b->prm.device));
  for (int i = 0; i < b->n; ++i) {
    if (!b->host[i].gt) {
      counts[3 * i] = counts[3 * i + 1] = counts[3 * i + 2] = 0;
      continue;
    }
    PB_HIP(b, hipMemcpyAsync(counts + 3 * i, b->host[i].counts,
                             3 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                             b->stream));
  }
  PB_HIP(b, hipStreamSynchronize(b->stream));
  return PM_OK;
}

int pm_batch_disparity_ptr(pm_batch* b, int i, const float** d_out) {
  if (!b) return PM_ERR_ARG;
  int rc;
  if ((rc = check_index(b, i))) return rc;
  if (!d_out) return fail(b, PM_ERR_ARG, "d_out is NULL");
  *d_out = b->host[i].disp;
  return PM_OK;
}

int pm_batch_get_tuning(const pm_batch* b, pm_batch_tuning* out) {
  if (!b || !out) return PM_ERR_ARG;
  out->tile_rows = b->shape.rows;
  out->workgroups = b->shape.workgroups;
  out->ms = b->tuned_ms;
  return PM_OK;
}

int pm_batch_set_tuning(pm_batch* b, const pm_batch_tuning* t) {
  if (!b) return PM_ERR_ARG;
  if (!t) return fail(b, PM_ERR_ARG, "tuning is NULL");
  if (!valid_rows(t->tile_rows))
    return fail(b, PM_ERR_ARG, "tile rows %d: 2, 4 or 8", t->tile_rows);
  if (t->workgroups < 0)
    return fail(b, PM_ERR_ARG, "negative work-group count");
  PB_HIP(b, hipSetDevice(b->prm.device));
  int rc;
  if ((rc = allow_shape(b, t->tile_rows))) return rc;
  if (t->tile_rows != b->shape.rows || t->workgroups != b->shape.workgroups) {
    b->shape.rows = t->tile_rows;
    b->shape.workgroups = t->workgroups;
    b->tuned_ms = 0.0f;
    drop_graph(b);
    number_blocks(b);
    b->ran = false;  // the descriptors changed under any earlier result
    if ((rc = upload(b))) return rc;
  }
  return PM_OK;
}

int pm_batch_autotune(pm_batch* b, int reps, pm_batch_tuning* table,
                      int table_size, int* count) {
  if (!b) return PM_ERR_ARG;
  if (reps < 1) return fail(b, PM_ERR_ARG, "reps must be >= 1");
  if (table_size < 0 || (table_size > 0 && !table))
    return fail(b, PM_ERR_ARG, "table is NULL");
  PB_HIP(b, hipSetDevice(b->prm.device));
  int cus = 0;
  PB_HIP(b, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                  b->prm.device));
  std::vector<pm_batch_tuning> cand;
  for (int rows : pm::kBatchRowChoices) {
    int resident = 0;
    if (pm::batch_resident_per_cu(b->prm.cost_kind, rows, tile_lds(b, rows),
                                  &resident) != hipSuccess ||
        resident < 1)
      continue;
    const int full = cus * resident;
    for (int wg : {full, 2 * full, 0}) {
      pm_batch_tuning t{};
      t.tile_rows = rows;
      t.workgroups = wg;
      cand.push_back(t);
    }
  }
  if (cand.empty())
    return fail(b, PM_ERR_ARG, "no tile shape fits a CU at radius %d",
                b->prm.radius);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PB_HIP(b, hipEventCreate(&e0));
  hipError_t he = hipEventCreate(&e1);
  if (he != hipSuccess) {
    (void)hipEventDestroy(e0);
    return hip_fail(b, he, "hipEventCreate");
  }
  int rc = PM_OK;
  int best = -1;
  for (size_t c = 0; c < cand.size() && !rc; ++c) {
    pm_batch_tuning t = cand[c];
    t.ms = 0.0f;
    if ((rc = pm_batch_set_tuning(b, &t))) break;
    if ((rc = enqueue_setup(b, b->stream))) break;
    float bestms = INFINITY;
    for (int r = 0; r < reps && !rc; ++r) {
      he = hipEventRecord(e0, b->stream);
      if (he == hipSuccess && (rc = enqueue_propagate(b, b->stream, 0, 0, 0)))
        break;
      if (he == hipSuccess) he = hipEventRecord(e1, b->stream);
      if (he == hipSuccess) he = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
      if (he != hipSuccess) rc = hip_fail(b, he, "timing a candidate");
      bestms = std::min(bestms, ms);
    }
    cand[c].ms = bestms;
    if (!rc && (best < 0 || bestms < cand[best].ms)) best = (int)c;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->ran = false;
  if (rc) return rc;
  for (int c = 0; c < (int)cand.size() && c < table_size; ++c)
    table[c] = cand[c];
  if (count) *count = (int)cand.size();
  if ((rc = pm_batch_set_tuning(b, &cand[best]))) return rc;
  b->tuned_ms = cand[best].ms;
  return PM_OK;
}

int pm_batch_set_proposal(pm_batch* b, pm_net* net) {
  return pm_batch_set_proposal_views(b, net, 2);
}

int pm_batch_set_proposal_views(pm_batch* b, pm_net* net, int views) {
  if (!b) return PM_ERR_ARG;
  if (views != 1 && views != 2)
    return fail(b, PM_ERR_ARG, "views must be 1 or 2, not %d", views);
  PB_HIP(b, hipSetDevice(b->prm.device));
  if (net) {
    if (pm::net::device_of(net) != b->prm.device)
      return fail(b, PM_ERR_ARG, "the net is on device %d, the batch on %d",
                  pm::net::device_of(net), b->prm.device);
    int rc;
    if ((rc = reserve_net(b, net))) return rc;
  }
  if (net != b->net || views != b->net_views) {
    b->net = net;
    b->net_views = views;
    drop_graph(b);
    b->ran = false;
  }
  return PM_OK;
}

int pm_batch_sync(pm_batch* b) {
  if (!b) return PM_ERR_ARG;
  PB_HIP(b, hipSetDevice(b->prm.device));
  PB_HIP(b, hipStreamSynchronize(b->stream));
  return PM_OK;
}

}  // extern "C"
