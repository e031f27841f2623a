// pm_smooth.hip -- gfx950 kernels of the PM-Huber smoothing coupling:
// diffusion tensor of an image, (normal, depth) field of a plane field and
// the diffusion-tensor ROF primal-dual step (cvt/cl/kernel/PDDiffTensor.cl,
// PDDROFDiffTensor.cl, pmhstereo.cl:1384-1402).  One thread per pixel,
// 64 x 4 work-groups; every step is a bandwidth-bound stencil.
#include <hip/hip_runtime.h>

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
  const float len = sqrtf(dx * dx + dy * dy);
  const float w = exp_nonpos(-alpha * len);
  float4 o = mk4(1.0f, 0.0f, 0.0f, 1.0f);
  if (!(len <= 0.001f)) {
    dx = dx / len;
    dy = dy / len;
    const float dxo = dy, dyo = -dx;
    o = mk4((w * dx) * dx + dxo * dxo, (w * dx) * dy + dxo * dyo,
            (w * dy) * dx + dyo * dxo, (w * dy) * dy + dyo * dyo);
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

// tensor-applied projected dual of one step at (i, j) (PDDROFDiffTensor.cl:
// 69-98): u clamped to the edge, tensor and old dual 0 outside the image
__device__ __forceinline__ Dual rof_dual(const float4* __restrict__ u,
                                         const float4* __restrict__ p,
                                         const float4* __restrict__ dt, int W,
                                         int H, int i, int j) {
  const bool inside = i >= 0 && i < W && j >= 0 && j < H;
  const int ci = clampi(i, 0, W - 1), cj = clampi(j, 0, H - 1);
  const int ci1 = clampi(i + 1, 0, W - 1), cj1 = clampi(j + 1, 0, H - 1);
  const float4 u0 = u[(size_t)cj * W + ci];
  const float4 ux = u[(size_t)cj * W + ci1];
  const float4 uy = u[(size_t)cj1 * W + ci];
  float4 D = mk4(0.0f, 0.0f, 0.0f, 0.0f);
  float4 plo4 = D, phi4 = D;
  if (inside) {
    const size_t q = (size_t)j * W + i;
    D = dt[q];
    plo4 = p[2 * q];
    phi4 = p[2 * q + 1];
  }
  Dual r;
  float n[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float dx = comp(u0, k) - comp(ux, k), dy = comp(u0, k) - comp(uy, k);
    const float DTdx = D.x * dx + D.z * dy;
    const float DTdy = D.y * dx + D.w * dy;
    r.lo[k] = comp(plo4, k) + kRofSigma * DTdx;
    r.hi[k] = comp(phi4, k) + kRofSigma * DTdy;
    n[k] = r.lo[k] * r.lo[k] + r.hi[k] * r.hi[k];
  }
  const float pxyz = fmaxf(1.0f, sqrtf((n[0] + n[1]) + n[2]));
  const float pw = fmaxf(1.0f, sqrtf(n[3]));
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float pr = k < 3 ? pxyz : pw;
    const float lo = r.lo[k] / pr, hi = r.hi[k] / pr;
    r.lo[k] = D.x * lo + D.z * hi;
    r.hi[k] = D.y * lo + D.w * hi;
  }
  return r;
}

// one PDDROFDiffTensor step (PDDROFDiffTensor.cl:43-125).  Each pixel forms
// its own dual and those of its left and upper neighbours (the kernel's
// LDS tile shares them; the values are the same).  The dual written back
// is the tensor-applied one, as the kernel stores it.
__global__ void __launch_bounds__(256)
k_rof_step(int W, int H, const float4* __restrict__ u,
           const float4* __restrict__ p, const float4* __restrict__ q,
           const float4* __restrict__ f, const float4* __restrict__ dt,
           float4 lambda, float4* __restrict__ u_out,
           float4* __restrict__ p_out, float4* __restrict__ q_out) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  const size_t o = (size_t)y * W + x;
  const Dual P = rof_dual(u, p, dt, W, H, x, y);
  const Dual Lf = rof_dual(u, p, dt, W, H, x - 1, y);
  const Dual Up = rof_dual(u, p, dt, W, H, x, y - 1);
  const float4 fi = f[o], qi = q[o], ui = u[o];
  float un[4], qn[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float lam = comp(lambda, k);
    const float div = ((P.lo[k] - Lf.lo[k]) + P.hi[k]) - Up.hi[k];
    const float img = comp(fi, k), qq = comp(qi, k), uo = comp(ui, k);
    float v = (uo + kRofTau * ((lam * img - div) - qq)) /
              (1.0f + kRofTau * lam);
    v = v + kRofTheta * (v - uo);
    float w = (qq + kRofEta * (v - img)) / (1.0f + 4.0f * kRofEta / lam);
    w = w + kRofTheta * (w - qq);
    un[k] = v;
    qn[k] = w;
  }
  u_out[o] = mk4(un[0], un[1], un[2], un[3]);
  q_out[o] = mk4(qn[0], qn[1], qn[2], qn[3]);
  p_out[2 * o] = mk4(P.lo[0], P.lo[1], P.lo[2], P.lo[3]);
  p_out[2 * o + 1] = mk4(P.hi[0], P.hi[1], P.hi[2], P.hi[3]);
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
#define PM_LAUNCH(kernel, grid, block, lds, st, ...)                \
  do {                                                              \
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

hipError_t launch_rof_step(hipStream_t st, const Geometry& g, const float4* u,
                           const float4* p, const float4* q, const float4* f,
                           const float4* dt, float4 lambda, float4* u_out,
                           float4* p_out, float4* q_out) {
  PM_LAUNCH(k_rof_step,
            dim3((g.W + kBX - 1) / kBX, (g.H + kBY - 1) / kBY), kBlock2D, 0,
            st, g.W, g.H, u, p, q, f, dt, lambda, u_out, p_out, q_out);
}

hipError_t launch_fill4(hipStream_t st, size_t n, float4 v, float4* out) {
  PM_LAUNCH(k_fill4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
            v, out);
}

}  // namespace pm
