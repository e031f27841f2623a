// pm_net_capi.cpp -- the plane-proposal U-Net (include/pm_net.h): its
// topology, parameters, activation buffers, forward pass, backward pass
// and Adam step, over the kernels of pm_net.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "pm_net.h"
#include "pm_net_internal.h"
#include "pm_net_kernels.h"

namespace {

using pm::net::bf16;
using pm::net::conv_cpad;
using pm::net::conv_kpad;

// ---------------------------------------------------------------------
// topology: tensors (level = resolution 2^-level, channels) and the
// operations in forward order
// ---------------------------------------------------------------------
enum Tensor {
  X0, E0A, E0, P1, E1A, E1, P2, E2A, E2, P3, B0, B1, B2, B3,
  D2A, D2, D1A, D1, D0A, D0, OUT, X4, CV, X2, CV1, kTensors
};
struct TensorDesc {
  int level, ch;
};
// X4, X2: the input box-averaged to quarter and half resolution; CV,
// CV1: the coarse and fine matching cost volumes computed from them (no
// parameters, no gradient)
constexpr TensorDesc kT[kTensors] = {
    {0, 8},   {0, 32},  {0, 32},  {1, 32},  {1, 64},  {1, 64},  {2, 64},
    {2, 96},  {2, 96},  {3, 96},  {3, 128}, {3, 128}, {3, 128}, {3, 128},
    {2, 96},  {2, 96},  {1, 64},  {1, 64},  {0, 32},  {0, 32},  {0, 8},
    {2, 8},   {2, pm::net::kCvLevels}, {1, 8},  {1, pm::net::kCvFineLevels}};

// conv: in1 (+ in2, upsampled when up2) -> out; OUT is the fp32 head
struct ConvDesc {
  int in1, in2, up2, out, dil, relu;
};
constexpr ConvDesc kC[] = {
    {X0, -1, 0, E0A, 1, 1},  {E0A, -1, 0, E0, 1, 1},
    {P1, CV1, 0, E1A, 1, 1}, {E1A, -1, 0, E1, 1, 1},
    {P2, CV, 0, E2A, 1, 1},  {E2A, -1, 0, E2, 1, 1},
    {P3, -1, 0, B0, 1, 1},   {B0, -1, 0, B1, 2, 1},
    {B1, -1, 0, B2, 4, 1},   {B2, -1, 0, B3, 8, 1},
    {E2, B3, 1, D2A, 1, 1},  {D2A, -1, 0, D2, 1, 1},
    {E1, D2, 1, D1A, 1, 1},  {D1A, -1, 0, D1, 1, 1},
    {E0, D1, 1, D0A, 1, 1},  {D0A, -1, 0, D0, 1, 1},
    {D0, -1, 0, OUT, 1, 0}};
constexpr int kConvs = sizeof(kC) / sizeof(kC[0]);
struct PoolDesc {
  int in, out;
};
constexpr PoolDesc kP[] = {{E0, P1}, {E1, P2}, {E2, P3}};
// forward order: conv k (k >= 0) or pool -(k + 1)
constexpr int kOrder[] = {0, 1, -1, 2, 3, -2, 4, 5, -3, 6, 7, 8, 9,
                          10, 11, 12, 13, 14, 15, 16};

// tensors computed from the input alone: no gradient is formed for them
bool input_only(int t) { return t == X4 || t == CV || t == X2 || t == CV1; }

int conv_cin(int l) {
  return kT[kC[l].in1].ch + (kC[l].in2 >= 0 ? kT[kC[l].in2].ch : 0);
}
int conv_cout(int l) { return kT[kC[l].out].ch; }

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct pm_net {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;

  // parameters: per conv the padded weights [cpad(cout)][kpad(cin)], then
  // the bias [cpad(cout)]; fp32 master, bf16 copy, gradient, Adam moments
  std::vector<size_t> woff, boff;
  size_t nparams = 0;
  DevBuf w32, w16, grad, m, v;
  DevBuf wflip[kConvs];  // flipped, transposed bf16 weights (input grads)
  long step = 0;

  // activations (bf16; OUT fp32) and their gradients (fp32) for the
  // largest N x H x W seen
  long cap_pix = 0;  // N * H * W the buffers hold
  long generation = 0;  // count of reallocations of the buffers
  DevBuf act[kTensors], gact[kTensors];
  DevBuf tmp;        // input gradient of a concatenated input (fp32)
  DevBuf gz;         // masked output gradient (bf16)
  DevBuf loss;       // one float
  int N = 0, H = 0, W = 0;
};

namespace {

int fail(pm_net* n, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (n) n->err = buf;
  return code;
}

int hip_fail(pm_net* n, hipError_t e, const char* what) {
  return fail(n, e == hipErrorOutOfMemory ? PM_ERR_NOMEM : PM_ERR_HIP,
              "%s: %s", what, hipGetErrorString(e));
}

#define NET_HIP(net, call)                                  \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return hip_fail(net, e_, #call);  \
  } while (0)

int ensure(pm_net* n, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes) return PM_OK;
  if (b.ptr) NET_HIP(n, hipFree(b.ptr));
  b.ptr = nullptr;
  b.bytes = 0;
  NET_HIP(n, hipMalloc(&b.ptr, bytes));
  b.bytes = bytes;
  return PM_OK;
}

template <typename T>
T* as(DevBuf& b) {
  return static_cast<T*>(b.ptr);
}

long level_pix(int N, int H, int W, int level) {
  return (long)N * (H >> level) * (W >> level);
}

// buffers for a batch of N pairs of H x W (multiples of 8)
int alloc_batch(pm_net* n, int N, int H, int W) {
  const long pix = (long)N * H * W;
  int rc;
  if (pix > n->cap_pix) {
    for (int t = 0; t < kTensors; ++t) {
      const long lp = level_pix(N, H, W, kT[t].level);
      const size_t el = (size_t)lp * kT[t].ch;
      if ((rc = ensure(n, n->act[t], el * (t == OUT ? 4 : 2)))) return rc;
      if (t != X0 && !input_only(t) &&
          (rc = ensure(n, n->gact[t], el * 4)))
        return rc;
    }
    size_t tmp = 0, gz = 0;
    for (int l = 0; l < kConvs; ++l) {
      const long lp = level_pix(N, H, W, kT[kC[l].out].level);
      tmp = std::max(tmp, (size_t)lp * conv_cin(l));
      gz = std::max(gz, (size_t)lp * conv_cout(l));
    }
    if ((rc = ensure(n, n->tmp, tmp * 4)) || (rc = ensure(n, n->gz, gz * 2)))
      return rc;
    n->cap_pix = pix;
    ++n->generation;
  }
  n->N = N;
  n->H = H;
  n->W = W;
  return PM_OK;
}

pm::net::ConvArgs conv_args(pm_net* n, int l) {
  const ConvDesc& d = kC[l];
  const int lev = kT[d.out].level;
  pm::net::ConvArgs a;
  a.x1 = as<bf16>(n->act[d.in1]);
  a.c1 = kT[d.in1].ch;
  if (d.in2 >= 0) {
    a.x2 = as<bf16>(n->act[d.in2]);
    a.c2 = kT[d.in2].ch;
    a.up2 = d.up2;
  }
  a.w = as<bf16>(n->w16) + n->woff[l];
  a.bias = as<float>(n->w32) + n->boff[l];
  a.N = n->N;
  a.H = n->H >> lev;
  a.W = n->W >> lev;
  a.cout = conv_cout(l);
  a.dil = d.dil;
  a.relu = d.relu;
  a.ys = a.cout;
  if (d.out == OUT)
    a.yf = as<float>(n->act[OUT]);
  else
    a.y = as<bf16>(n->act[d.out]);
  return a;
}

// out: when set, the head writes its 4 real outputs there (N x H x W x 4
// fp32) instead of into act[OUT]
int forward(pm_net* n, hipStream_t st, float* out = nullptr) {
  NET_HIP(n, pm::net::launch_down4(st, n->N, n->H, n->W, as<bf16>(n->act[X0]),
                                   as<bf16>(n->act[X4])));
  NET_HIP(n, pm::net::launch_cost_volume(st, n->N, n->H >> 2, n->W >> 2,
                                         kT[CV].ch, as<bf16>(n->act[X4]),
                                         as<bf16>(n->act[CV])));
  NET_HIP(n, pm::net::launch_avgpool(st, n->N, n->H, n->W, kT[X0].ch,
                                     as<bf16>(n->act[X0]),
                                     as<bf16>(n->act[X2])));
  NET_HIP(n, pm::net::launch_cost_volume(st, n->N, n->H >> 1, n->W >> 1,
                                         kT[CV1].ch, as<bf16>(n->act[X2]),
                                         as<bf16>(n->act[CV1])));
  for (int op : kOrder) {
    if (op >= 0) {
      pm::net::ConvArgs a = conv_args(n, op);
      if (out && kC[op].out == OUT) {
        a.yf = out;
        a.cout = PM_NET_OUT_CHANNELS;
        a.ys = PM_NET_OUT_CHANNELS;
      }
      NET_HIP(n, pm::net::launch_conv(st, a));
    } else {
      const PoolDesc& p = kP[-op - 1];
      const int lev = kT[p.in].level;
      NET_HIP(n, pm::net::launch_avgpool(st, n->N, n->H >> lev, n->W >> lev,
                                         kT[p.in].ch, as<bf16>(n->act[p.in]),
                                         as<bf16>(n->act[p.out])));
    }
  }
  return PM_OK;
}

int refresh_flipped(pm_net* n) {
  for (int l = 1; l < kConvs; ++l)
    NET_HIP(n, pm::net::launch_flip_weights(
                   n->stream, as<bf16>(n->w16) + n->woff[l], conv_cin(l),
                   conv_cout(l), as<bf16>(n->wflip[l])));
  return PM_OK;
}

// backward from the head's gradient (already in gact[OUT])
int backward(pm_net* n, hipStream_t st) {
  for (int i = (int)(sizeof(kOrder) / sizeof(kOrder[0])) - 1; i >= 0; --i) {
    const int op = kOrder[i];
    if (op < 0) {
      const PoolDesc& p = kP[-op - 1];
      const int lev = kT[p.in].level;
      NET_HIP(n, pm::net::launch_avgpool_grad(
                     st, n->N, n->H >> lev, n->W >> lev, kT[p.in].ch,
                     as<float>(n->gact[p.out]), as<float>(n->gact[p.in])));
      continue;
    }
    const int l = op;
    const ConvDesc& d = kC[l];
    pm::net::ConvArgs a = conv_args(n, l);
    const size_t nout = (size_t)a.N * a.H * a.W * a.cout;
    bf16* gz = as<bf16>(n->gz);
    NET_HIP(n, pm::net::launch_grad_mask(
                   st, nout, as<float>(n->gact[d.out]),
                   d.relu ? as<bf16>(n->act[d.out]) : nullptr, d.relu, gz));
    NET_HIP(n, pm::net::launch_conv_wgrad(st, a, gz,
                                          as<float>(n->grad) + n->woff[l],
                                          as<float>(n->grad) + n->boff[l]));
    if (d.in1 == X0) continue;  // the input needs no gradient
    pm::net::ConvArgs t;
    t.x1 = gz;
    t.c1 = a.cout;
    t.w = as<bf16>(n->wflip[l]);
    t.N = a.N;
    t.H = a.H;
    t.W = a.W;
    t.cout = a.c1 + a.c2;
    t.dil = a.dil;
    t.yf = as<float>(n->tmp);
    t.ys = t.cout;
    NET_HIP(n, pm::net::launch_conv(st, t));
    // the cost volume takes no gradient
    NET_HIP(n, pm::net::launch_route_grad(
                   st, a.N, a.H, a.W, a.c1, a.c2, a.up2, as<float>(n->tmp),
                   as<float>(n->gact[d.in1]),
                   d.in2 >= 0 && !input_only(d.in2)
                       ? as<float>(n->gact[d.in2])
                       : nullptr));
  }
  return PM_OK;
}

int check_dims(pm_net* n, int N, int H, int W) {
  if (N < 1 || H < 8 || W < 8 || H % 8 || W % 8)
    return fail(n, PM_ERR_ARG, "batch %d of %d x %d: H and W must be "
                "positive multiples of 8", N, H, W);
  if ((long)N * H * W > (1L << 27))
    return fail(n, PM_ERR_ARG, "batch of %d x %d x %d pixels is too large",
                N, H, W);
  return PM_OK;
}

}  // namespace

namespace pm {
namespace net {

int proposal(pm_net* n, hipStream_t st, const float* left, const float* right,
             int W, int H, int channels, int stride, int views,
             const float** out, int* Wp, int* os) {
  if (views != 1 && views != 2)
    return fail(n, PM_ERR_ARG, "views must be 1 or 2, not %d", views);
  const int Hp = (H + 7) / 8 * 8, Wq = (W + 7) / 8 * 8;
  int rc;
  if ((rc = alloc_batch(n, views, Hp, Wq))) return rc;
  // batch entry 0: the pair; entry 1: the mirrored pair.  Every output
  // pixel depends on its own batch entry only, so each entry's result is
  // what a pass over it alone gives, bit for bit.
  const size_t plane = (size_t)Hp * Wq;
  for (int mirror = 0; mirror < views; ++mirror)
    NET_HIP(n, launch_pack_pair(st, left, right, W, H, channels, stride, Hp,
                                Wq, mirror,
                                as<bf16>(n->act[X0]) +
                                    mirror * plane * kT[X0].ch));
  if ((rc = forward(n, st))) return rc;
  out[0] = as<float>(n->act[OUT]);
  out[1] = views == 2 ? out[0] + plane * kT[OUT].ch : nullptr;
  *Wp = Wq;
  *os = kT[OUT].ch;
  return PM_OK;
}

int reserve(pm_net* n, int Hp, int Wp) {
  if (Hp < 8 || Wp < 8 || Hp % 8 || Wp % 8)
    return fail(n, PM_ERR_ARG, "reserve: %d x %d is not a positive multiple "
                "of 8", Hp, Wp);
  NET_HIP(n, hipSetDevice(n->device));
  return alloc_batch(n, 2, Hp, Wp);
}

long generation(const pm_net* n) { return n->generation; }

const char* error_of(const pm_net* n) { return n->err.c_str(); }
int device_of(const pm_net* n) { return n->device; }

}  // namespace net
}  // namespace pm

extern "C" {

int pm_net_create(int device, void* stream, uint64_t seed, pm_net** out) {
  if (!out) return PM_ERR_ARG;
  *out = nullptr;
  pm_net* n = new (std::nothrow) pm_net();
  if (!n) return PM_ERR_NOMEM;
  *out = n;  // the caller reads pm_net_last_error on failure, then destroys
  n->device = device;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    return fail(n, PM_ERR_HIP, "no HIP device");
  }
  if (device < 0 || device >= ndev)
    return fail(n, PM_ERR_ARG, "device %d of %d", device, ndev);
  NET_HIP(n, hipSetDevice(device));
  if (stream) {
    n->stream = static_cast<hipStream_t>(stream);
  } else {
    NET_HIP(n, hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
    n->own_stream = true;
  }
  // parameter layout, He-initialised weights (std sqrt(2 / fan_in)); the
  // head starts small with the mean disparity / 128 and a fronto-parallel
  // normal as its biases
  size_t off = 0;
  for (int l = 0; l < kConvs; ++l) {
    n->woff.push_back(off);
    off += (size_t)conv_cpad(conv_cout(l)) * conv_kpad(conv_cin(l));
    off = (off + 63) / 64 * 64;
    n->boff.push_back(off);
    off += conv_cpad(conv_cout(l));
    off = (off + 63) / 64 * 64;
  }
  n->nparams = off;
  std::vector<float> w(off, 0.0f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < kConvs; ++l) {
    const int cin = conv_cin(l), cout = conv_cout(l), kpad = conv_kpad(cin);
    const bool head = kC[l].out == OUT;
    std::normal_distribution<float> nd(
        0.0f, std::sqrt(2.0f / (9.0f * cin)) * (head ? 0.1f : 1.0f));
    const int real = head ? PM_NET_OUT_CHANNELS : cout;
    for (int co = 0; co < real; ++co)
      for (int k = 0; k < 9 * cin; ++k)
        w[n->woff[l] + (size_t)co * kpad + k] = nd(rng);
    if (head) {
      w[n->boff[l] + 0] = 0.35f;
      w[n->boff[l] + 3] = 1.0f;
    }
  }
  int rc;
  if ((rc = ensure(n, n->w32, off * 4)) || (rc = ensure(n, n->w16, off * 2)) ||
      (rc = ensure(n, n->grad, off * 4)) || (rc = ensure(n, n->m, off * 4)) ||
      (rc = ensure(n, n->v, off * 4)) || (rc = ensure(n, n->loss, 4)))
    return rc;
  for (int l = 1; l < kConvs; ++l)
    if ((rc = ensure(n, n->wflip[l], (size_t)conv_cpad(conv_cin(l)) *
                                         conv_kpad(conv_cout(l)) * 2)))
      return rc;
  NET_HIP(n, hipMemsetAsync(n->m.ptr, 0, off * 4, n->stream));
  NET_HIP(n, hipMemsetAsync(n->v.ptr, 0, off * 4, n->stream));
  return pm_net_set_params(n, w.data(), off);
}

void pm_net_destroy(pm_net* n) {
  if (!n) return;
  (void)hipSetDevice(n->device);
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  DevBuf* bufs[] = {&n->w32, &n->w16, &n->grad, &n->m, &n->v, &n->tmp, &n->gz,
                    &n->loss};
  for (DevBuf* b : bufs)
    if (b->ptr) (void)hipFree(b->ptr);
  for (int t = 0; t < kTensors; ++t) {
    if (n->act[t].ptr) (void)hipFree(n->act[t].ptr);
    if (n->gact[t].ptr) (void)hipFree(n->gact[t].ptr);
  }
  for (int l = 0; l < kConvs; ++l)
    if (n->wflip[l].ptr) (void)hipFree(n->wflip[l].ptr);
  if (n->own_stream && n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

const char* pm_net_last_error(const pm_net* n) {
  return n ? n->err.c_str() : "net is NULL";
}

size_t pm_net_num_params(const pm_net* n) { return n ? n->nparams : 0; }

int pm_net_get_params(pm_net* n, float* out, size_t count) {
  if (!n || !out) return PM_ERR_ARG;
  if (count != n->nparams)
    return fail(n, PM_ERR_ARG, "%zu parameters, not %zu", n->nparams, count);
  NET_HIP(n, hipSetDevice(n->device));
  NET_HIP(n, hipMemcpyAsync(out, n->w32.ptr, count * 4, hipMemcpyDeviceToHost,
                            n->stream));
  NET_HIP(n, hipStreamSynchronize(n->stream));
  return PM_OK;
}

int pm_net_get_grads(pm_net* n, float* out, size_t count) {
  if (!n || !out) return PM_ERR_ARG;
  if (count != n->nparams)
    return fail(n, PM_ERR_ARG, "%zu parameters, not %zu", n->nparams, count);
  NET_HIP(n, hipSetDevice(n->device));
  NET_HIP(n, hipMemcpyAsync(out, n->grad.ptr, count * 4, hipMemcpyDeviceToHost,
                            n->stream));
  NET_HIP(n, hipStreamSynchronize(n->stream));
  return PM_OK;
}

int pm_net_set_params(pm_net* n, const float* in, size_t count) {
  if (!n || !in) return PM_ERR_ARG;
  if (count != n->nparams)
    return fail(n, PM_ERR_ARG, "%zu parameters, not %zu", n->nparams, count);
  NET_HIP(n, hipSetDevice(n->device));
  NET_HIP(n, hipMemcpyAsync(n->w32.ptr, in, count * 4, hipMemcpyHostToDevice,
                            n->stream));
  NET_HIP(n, pm::net::launch_to_bf16(n->stream, count, as<float>(n->w32),
                                     as<bf16>(n->w16)));
  int rc;
  if ((rc = refresh_flipped(n))) return rc;
  NET_HIP(n, hipStreamSynchronize(n->stream));  // `in` may be freed
  return PM_OK;
}

int pm_net_forward(pm_net* n, const float* d_x, int N, int H, int W,
                   float* d_out) {
  if (!n) return PM_ERR_ARG;
  if (!d_x || !d_out) return fail(n, PM_ERR_ARG, "NULL buffer");
  int rc;
  if ((rc = check_dims(n, N, H, W))) return rc;
  NET_HIP(n, hipSetDevice(n->device));
  if ((rc = alloc_batch(n, N, H, W))) return rc;
  const size_t npix = (size_t)N * H * W;
  NET_HIP(n, pm::net::launch_to_bf16(n->stream, npix * PM_NET_IN_CHANNELS, d_x,
                                     as<bf16>(n->act[X0])));
  return forward(n, n->stream, d_out);
}

int pm_net_train_step(pm_net* n, const float* d_x, const float* d_target,
                      int N, int H, int W, float lr, float w_disp,
                      float w_normal, float* loss) {
  if (!n) return PM_ERR_ARG;
  if (!d_x || !d_target) return fail(n, PM_ERR_ARG, "NULL buffer");
  if (!(lr > 0.0f) || !std::isfinite(lr))
    return fail(n, PM_ERR_ARG, "learning rate must be finite and > 0");
  int rc;
  if ((rc = check_dims(n, N, H, W))) return rc;
  NET_HIP(n, hipSetDevice(n->device));
  if ((rc = alloc_batch(n, N, H, W))) return rc;
  hipStream_t st = n->stream;
  const size_t npix = (size_t)N * H * W;
  NET_HIP(n, pm::net::launch_to_bf16(st, npix * PM_NET_IN_CHANNELS, d_x,
                                     as<bf16>(n->act[X0])));
  if ((rc = forward(n, st))) return rc;
  // gradients start at zero: activations (accumulated by their consumers)
  // and parameters (accumulated over the pixels)
  for (int t = 0; t < kTensors; ++t)
    if (t != X0 && t != OUT && !input_only(t))
      NET_HIP(n, hipMemsetAsync(n->gact[t].ptr, 0,
                                (size_t)level_pix(N, H, W, kT[t].level) *
                                    kT[t].ch * 4,
                                st));
  NET_HIP(n, hipMemsetAsync(n->grad.ptr, 0, n->nparams * 4, st));
  NET_HIP(n, hipMemsetAsync(n->loss.ptr, 0, 4, st));
  NET_HIP(n, pm::net::launch_l1_loss(st, npix, as<float>(n->act[OUT]),
                                     kT[OUT].ch, d_target, w_disp, w_normal,
                                     1.0f / (float)npix,
                                     as<float>(n->gact[OUT]),
                                     as<float>(n->loss)));
  if ((rc = backward(n, st))) return rc;
  n->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)n->step);
  const float c2 = 1.0f - std::pow(b2, (float)n->step);
  NET_HIP(n, pm::net::launch_adam(st, n->nparams, as<float>(n->w32),
                                  as<float>(n->grad), as<float>(n->m),
                                  as<float>(n->v), lr, b1, b2, 1e-8f, c1, c2,
                                  as<bf16>(n->w16)));
  if ((rc = refresh_flipped(n))) return rc;
  if (loss) {
    NET_HIP(n, hipMemcpyAsync(loss, n->loss.ptr, 4, hipMemcpyDeviceToHost, st));
    NET_HIP(n, hipStreamSynchronize(st));
  }
  return PM_OK;
}

}  // extern "C"
