// pm_smooth.hip -- gfx950 kernels of the PM-Huber smoothing coupling:
// diffusion tensor of an image, (normal, depth) field of a plane field and
// the diffusion-tensor ROF primal-dual step (cvt/cl/kernel/PDDiffTensor.cl,
// PDDROFDiffTensor.cl, pmhstereo.cl:1384-1402).  One thread per pixel,
// 64 x 4 work-groups; every step is a bandwidth-bound stencil.
#include <hip/hip_runtime.h>

#include <climits>

#include "pm_device.h"
#include "pm_engine.h"

namespace pm {

constexpr int kBX = 64, kBY = 4;

// gray value of the diffusion tensor: the 4-component dot with
// (0.2126, 0.7152, 0.0722, 0) for RGBA, the intensity for gray images,
// 0 outside (CLK_ADDRESS_CLAMP), PDDiffTensor.cl:40-46,82-87
__device__ __forceinline__ float dt_gray(const float* __restrict__ img, int W,
                                         int H, int C, int stride, int x,
                                         int y) {
  if (x < 0 || x >= W || y < 0 || y >= H) return 0.0f;
  if (C == 1) return img[(size_t)y * stride + x];
  return gray_at(img, W, H, C, stride, x, y);
}

// PDDiffTensorColor / PDDiffTensorGray (PDDiffTensor.cl:25-116), beta = 1
__global__ void __launch_bounds__(256)
k_diff_tensor(const float* __restrict__ img, int W, int H, int C, int stride,
              float alpha, float4* __restrict__ dt) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  const float g = dt_gray(img, W, H, C, stride, x, y);
  float dx = dt_gray(img, W, H, C, stride, x + 1, y) - g;
  float dy = dt_gray(img, W, H, C, stride, x, y + 1) - g;
  const float len = sqrtf(fmaf(dx, dx, dy * dy));
  const float w = exp_nonpos(-alpha * len);
  float4 o = mk4(1.0f, 0.0f, 0.0f, 1.0f);
  if (!(len <= 0.001f)) {
    dx = dx / len;
    dy = dy / len;
    const float dxo = dy, dyo = -dx;
    o = mk4(fmaf(w * dx, dx, dxo * dxo), fmaf(w * dx, dy, dxo * dyo),
            fmaf(w * dy, dx, dyo * dxo), fmaf(w * dy, dy, dyo * dyo));
  }
  dt[(size_t)y * W + x] = o;
}

// pmhstereo_normal_depth (pmhstereo.cl:1384-1402)
__global__ void __launch_bounds__(256)
k_normal_depth(int W, int H, const float4* __restrict__ states,
               float dispscale, int lr, float4* __restrict__ out) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  const size_t p = (size_t)y * W + x;
  float4 v = to_ref_normal_depth(states[p], (float)x, (float)y, lr);
  v.w = v.w * dispscale;
  out[p] = v;
}

// float to int, truncating towards zero, with NaN -> 0 and values
// outside the int range saturated (the conversion OpenCL leaves undefined
// there, fixed so that every device and the host agree)
__device__ __forceinline__ int f2i_trunc(float v) {
  if (v != v) return 0;
  if (v >= 2147483648.0f) return INT_MAX;
  if (v <= -2147483648.0f) return INT_MIN;
  return (int)v;
}

// pmhstereo_fill_state (pmhstereo.cl:1131-1185): a plane whose cost is
// negative or whose disparity lies outside [0, W] is replaced by the
// nearest plane to its left or right whose cost is not negative (the
// search tests only the cost); of two, the one with the smaller positive
// disparity at this pixel (<= 0 counts as W, ties to the left).  Then
// non-finite a, b become 0 and the plane is converted as k_normal_depth
// converts it.
__global__ void __launch_bounds__(256)
k_fill_state(int W, int H, const float4* __restrict__ states,
             float dispscale, int lr, float4* __restrict__ out) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  const float4* row = states + (size_t)y * W;
  const float xf = (float)x, yf = (float)y;
  const float sidesign = lr ? -1.0f : 1.0f;
  const float4 state = row[x];
  float4 val = state;
  const float disp = (transform(state, xf, yf) - xf) * sidesign;
  if (state.w < 0.0f || disp < 0.0f || disp > (float)W) {
    float4 left = mk4(-1.0f, -1.0f, -1.0f, -1.0f);
    for (int i = x - 1; left.w < 0.0f && i >= 0; --i) left = row[i];
    float4 right = mk4(-1.0f, -1.0f, -1.0f, -1.0f);
    for (int i = x + 1; right.w < 0.0f && i < W; ++i) right = row[i];
    if (left.w >= 0.0f && right.w >= 0.0f) {
      int dl = f2i_trunc(sidesign * (transform(left, xf, yf) - xf));
      int dr = f2i_trunc(sidesign * (transform(right, xf, yf) - xf));
      dl = dl <= 0 ? W : dl;
      dr = dr <= 0 ? W : dr;
      val = dl <= dr ? left : right;
    } else {
      val = right.w < 0.0f ? left : right;
    }
  }
  if (!(finite_f(val.x) && finite_f(val.y))) {
    val.x = 0.0f;
    val.y = 0.0f;
  }
  val = to_ref_normal_depth(val, xf, yf, lr);
  val.w = val.w * dispscale;
  out[(size_t)y * W + x] = val;
}

constexpr float kRofSigma = 4.0f;  // PDDROFDiffTensor.cl:36-39
constexpr float kRofTau = 0.05f;
constexpr float kRofEta = 8.0f;
constexpr float kRofTheta = 0.5f;  // PDDROFDiffTensor.cl:41

__device__ __forceinline__ float comp(const float4& v, int k) {
  return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

struct Dual {
  float lo[4], hi[4];
};

// tensor-applied projected dual of one step (PDDROFDiffTensor.cl:69-98)
// from u at the pixel, its right and its lower neighbour (clamped to the
// edge), the tensor D and the old dual (both 0 outside the image)
// kHuber: the dual of PDDHuberDiffTensor.cl:64-92 instead, the ascent
// divided by shrink = 1 + SIGMA eps and one norm over all four channels
template <bool kHuber>
__device__ __forceinline__ Dual dual_of(const float4& u0, const float4& ux,
                                        const float4& uy, const float4& D,
                                        const float4& plo4,
                                        const float4& phi4, float shrink) {
  Dual r;
  float n[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float dx = comp(u0, k) - comp(ux, k), dy = comp(u0, k) - comp(uy, k);
    const float DTdx = fmaf(D.x, dx, D.z * dy);
    const float DTdy = fmaf(D.y, dx, D.w * dy);
    r.lo[k] = fmaf(kRofSigma, DTdx, comp(plo4, k));
    r.hi[k] = fmaf(kRofSigma, DTdy, comp(phi4, k));
    if (kHuber) {
      r.lo[k] = r.lo[k] / shrink;
      r.hi[k] = r.hi[k] / shrink;
    }
    n[k] = fmaf(r.lo[k], r.lo[k], r.hi[k] * r.hi[k]);
  }
  const float pxyz =
      kHuber ? fmaxf(1.0f, sqrtf(((n[0] + n[1]) + n[2]) + n[3]))
             : fmaxf(1.0f, sqrtf((n[0] + n[1]) + n[2]));
  const float pw = kHuber ? pxyz : fmaxf(1.0f, sqrtf(n[3]));
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float pr = k < 3 ? pxyz : pw;
    // x / 1 is x exactly: the division only runs for a projected dual
    float lo = r.lo[k], hi = r.hi[k];
    if (pr != 1.0f) {
      lo = lo / pr;
      hi = hi / pr;
    }
    r.lo[k] = fmaf(D.x, lo, D.z * hi);
    r.hi[k] = fmaf(D.y, lo, D.w * hi);
  }
  return r;
}

// primal and q update of one pixel (PDDROFDiffTensor.cl:100-124) from its
// dual P and the duals Lf, Up of its left and upper neighbours
__device__ __forceinline__ void rof_primal(const Dual& P, const Dual& Lf,
                                           const Dual& Up, const float4& fi,
                                           const float4& ui, const float4& qi,
                                           const float4& lambda, float4* uo,
                                           float4* qo) {
  float un[4], qn[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float lam = comp(lambda, k);
    const float div = ((P.lo[k] - Lf.lo[k]) + P.hi[k]) - Up.hi[k];
    const float img = comp(fi, k), qq = comp(qi, k), u0 = comp(ui, k);
    float v = fmaf(kRofTau, fmaf(lam, img, -div) - qq, u0) /
              fmaf(kRofTau, lam, 1.0f);
    v = fmaf(kRofTheta, v - u0, v);
    // 1 + 4 ETA / lambda: a quotient, nothing to fuse
    float w = fmaf(kRofEta, v - img, qq) / (1.0f + 4.0f * kRofEta / lam);
    w = fmaf(kRofTheta, w - qq, w);
    un[k] = v;
    qn[k] = w;
  }
  *uo = mk4(un[0], un[1], un[2], un[3]);
  *qo = mk4(qn[0], qn[1], qn[2], qn[3]);
}

// Up to kRofK PDDROFDiffTensor steps (PDDROFDiffTensor.cl:43-125) in one
// launch (temporal tiling).  The reference runs one step per launch over
// the whole image; a block here owns a kRofTX x kRofTY output tile and
// works on the region around it, kRofK pixels wider on every side: one
// wavefront row of 64 columns by kRofRH rows.  Each step forms the duals
// on the shrinking area that the later steps still need (a pixel's dual
// from u at the pixel and at its right and lower neighbours), then the
// primal and q there (from the pixel's dual and those of its left and
// upper neighbours); after nsteps <= kRofK steps the tile holds what
// nsteps whole-image steps leave.  The dual stored is the tensor-applied
// one, as the reference stores it.  u and the duals sit in LDS (the steps
// read them at neighbours); q, f and the tensor of a thread's own pixels
// stay in registers.  The halo pixels repeat work with the same
// operations, so every value is the one a whole-image step forms.
constexpr int kRofK = kRofStepsMax;
constexpr int kRofRW = 64;
constexpr int kRofTX = kRofRW - 2 * kRofK;
constexpr int kRofTY = 8;
constexpr int kRofRH = kRofTY + 2 * kRofK;
constexpr int kRofRows = kRofRH / 4;  // region rows per wavefront
static_assert(kRofRH % 4 == 0, "four wavefronts share the region rows");

template <bool kHuber>
__global__ void __launch_bounds__(256)
k_rof_steps(int W, int H, int nsteps, float eps, const float4* __restrict__ u,
            const float4* __restrict__ p, const float4* __restrict__ q,
            const float4* __restrict__ f, const float4* __restrict__ dt,
            float4 lambda, float4* __restrict__ u_out,
            float4* __restrict__ p_out, float4* __restrict__ q_out) {
  __shared__ float4 su[kRofRH * kRofRW];
  __shared__ float4 sp[2 * kRofRH * kRofRW];
  const float shrink = fmaf(kRofSigma, eps, 1.0f);  // Huber dual only
  const int lx = threadIdx.x;
  const int rx0 = blockIdx.x * kRofTX - kRofK;
  const int ry0 = blockIdx.y * kRofTY - kRofK;
  const int gx = rx0 + lx;
  const float4 zero = mk4(0.0f, 0.0f, 0.0f, 0.0f);
  float4 fr[kRofRows], dr[kRofRows], qr[kRofRows];
#pragma unroll
  for (int r = 0; r < kRofRows; ++r) {
    const int ly = threadIdx.y * kRofRows + r, gy = ry0 + ly;
    const int l = ly * kRofRW + lx;
    fr[r] = dr[r] = qr[r] = zero;
    if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
      const size_t o = (size_t)gy * W + gx;
      su[l] = u[o];
      sp[2 * l] = p[2 * o];
      sp[2 * l + 1] = p[2 * o + 1];
      qr[r] = q[o];
      fr[r] = f[o];
      dr[r] = dt[o];
    }
  }
  __syncthreads();
  for (int s = 1; s <= nsteps; ++s) {
    // duals on [s-1, kRofRW-s) x [s-1, kRofRH-s): the primal area of this
    // step and its left and upper neighbours
#pragma unroll
    for (int r = 0; r < kRofRows; ++r) {
      const int ly = threadIdx.y * kRofRows + r, gy = ry0 + ly;
      if (lx < s - 1 || lx >= kRofRW - s || ly < s - 1 || ly >= kRofRH - s)
        continue;
      const bool inside = gx >= 0 && gx < W && gy >= 0 && gy < H;
      // u at the edge-clamped pixel and neighbours, all inside the region
      const int ci = clampi(gx, 0, W - 1) - rx0;
      const int ci1 = clampi(gx + 1, 0, W - 1) - rx0;
      const int cj = (clampi(gy, 0, H - 1) - ry0) * kRofRW;
      const int cj1 = (clampi(gy + 1, 0, H - 1) - ry0) * kRofRW;
      const int l = ly * kRofRW + lx;
      const Dual P =
          dual_of<kHuber>(su[cj + ci], su[cj + ci1], su[cj1 + ci], dr[r],
                          inside ? sp[2 * l] : zero,
                          inside ? sp[2 * l + 1] : zero, shrink);
      sp[2 * l] = mk4(P.lo[0], P.lo[1], P.lo[2], P.lo[3]);
      sp[2 * l + 1] = mk4(P.hi[0], P.hi[1], P.hi[2], P.hi[3]);
    }
    __syncthreads();
    // primal and q on [s, kRofRW-s) x [s, kRofRH-s), pixels in the image
#pragma unroll
    for (int r = 0; r < kRofRows; ++r) {
      const int ly = threadIdx.y * kRofRows + r, gy = ry0 + ly;
      if (lx < s || lx >= kRofRW - s || ly < s || ly >= kRofRH - s) continue;
      if (gx < 0 || gx >= W || gy < 0 || gy >= H) continue;
      const int l = ly * kRofRW + lx;
      Dual P, Lf, Up;
      const float4 a = sp[2 * l], b = sp[2 * l + 1];
      const float4 c = sp[2 * (l - 1)], d = sp[2 * (l - kRofRW) + 1];
      P.lo[0] = a.x; P.lo[1] = a.y; P.lo[2] = a.z; P.lo[3] = a.w;
      P.hi[0] = b.x; P.hi[1] = b.y; P.hi[2] = b.z; P.hi[3] = b.w;
      Lf.lo[0] = c.x; Lf.lo[1] = c.y; Lf.lo[2] = c.z; Lf.lo[3] = c.w;
      Up.hi[0] = d.x; Up.hi[1] = d.y; Up.hi[2] = d.z; Up.hi[3] = d.w;
      float4 un, qn;
      rof_primal(P, Lf, Up, fr[r], su[l], qr[r], lambda, &un, &qn);
      su[l] = un;
      qr[r] = qn;
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < kRofRows; ++r) {
    const int ly = threadIdx.y * kRofRows + r, gy = ry0 + ly;
    if (lx < kRofK || lx >= kRofK + kRofTX || ly < kRofK ||
        ly >= kRofK + kRofTY)
      continue;
    if (gx >= W || gy >= H) continue;
    const int l = ly * kRofRW + lx;
    const size_t o = (size_t)gy * W + gx;
    u_out[o] = su[l];
    q_out[o] = qr[r];
    p_out[2 * o] = sp[2 * l];
    p_out[2 * o + 1] = sp[2 * l + 1];
  }
}

// n float4 set to v (graph-safe replacement of a memset)
__global__ void __launch_bounds__(256)
k_fill4(size_t n, float4 v, float4* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = v;
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
// Launches and returns the launch's own error.  The runtime's last-error
// slot is cleared first: an earlier failed call, already returned to its
// caller, is not reported again here.
#define PM_LAUNCH(kernel, grid, block, lds, st, ...)                \
  do {                                                              \
    (void)hipGetLastError();                                        \
    hipLaunchKernelGGL(kernel, grid, block, lds, st, __VA_ARGS__);  \
    return hipGetLastError();                                       \
  } while (0)

static const dim3 kBlock2D(kBX, kBY);

hipError_t launch_diff_tensor(hipStream_t st, const Geometry& g,
                              const float* img, int channels, int stride,
                              float alpha, float4* dt) {
  PM_LAUNCH(k_diff_tensor,
            dim3((g.W + kBX - 1) / kBX, (g.H + kBY - 1) / kBY), kBlock2D, 0,
            st, img, g.W, g.H, channels, stride, alpha, dt);
}

hipError_t launch_normal_depth(hipStream_t st, const Geometry& g,
                               const float4* states, float dispscale, int lr,
                               float4* out) {
  PM_LAUNCH(k_normal_depth,
            dim3((g.W + kBX - 1) / kBX, (g.H + kBY - 1) / kBY), kBlock2D, 0,
            st, g.W, g.H, states, dispscale, lr, out);
}

size_t rof_steps_lds_bytes() {
  return (size_t)3 * kRofRH * kRofRW * sizeof(float4);  // su and sp
}

hipError_t launch_rof_steps(hipStream_t st, const Geometry& g, int nsteps,
                            const float4* u, const float4* p, const float4* q,
                            const float4* f, const float4* dt, float4 lambda,
                            float hubereps, float4* u_out, float4* p_out,
                            float4* q_out) {
  if (nsteps < 1 || nsteps > kRofK) return hipErrorInvalidValue;
  const dim3 grid((g.W + kRofTX - 1) / kRofTX, (g.H + kRofTY - 1) / kRofTY);
  if (hubereps >= 0.0f)
    PM_LAUNCH(k_rof_steps<true>, grid, dim3(kRofRW, 4), 0, st, g.W, g.H,
              nsteps, hubereps, u, p, q, f, dt, lambda, u_out, p_out, q_out);
  PM_LAUNCH(k_rof_steps<false>, grid, dim3(kRofRW, 4), 0, st, g.W, g.H,
            nsteps, 0.0f, u, p, q, f, dt, lambda, u_out, p_out, q_out);
}

hipError_t launch_fill_state(hipStream_t st, const Geometry& g,
                             const float4* states, float dispscale, int lr,
                             float4* out) {
  PM_LAUNCH(k_fill_state,
            dim3((g.W + kBX - 1) / kBX, (g.H + kBY - 1) / kBY), kBlock2D, 0,
            st, g.W, g.H, states, dispscale, lr, out);
}

hipError_t launch_fill4(hipStream_t st, size_t n, float4 v, float4* out) {
  PM_LAUNCH(k_fill4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
            v, out);
}

}  // namespace pm
