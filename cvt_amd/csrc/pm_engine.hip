// pm_engine.hip -- gfx950 kernels of the PatchMatch stereo engine.
//
// One work-item per pixel, 64 x 4 work-groups: a wavefront covers 64
// consecutive pixels of one row, so per-tap loads of the weight planes and
// of the tap's own texel row are contiguous across the wave, and the
// other view's samples (same row, x' = a x + b y + c) are nearly
// contiguous for coherent plane fields.
//
// Per-pixel work follows cvt/cl/kernel/pmhstereo.cl (cited per kernel);
// the arithmetic contract is in pm_device.h.
#include "pm_device.h"
#include "pm_engine.h"

namespace pm {

// ---------------------------------------------------------------------
// A^(2^i) mod M, folded at compile time
// ---------------------------------------------------------------------
struct ApowTable {
  uint64_t v[64];
};

constexpr uint64_t cx_mulmod(uint64_t a, uint64_t b) {
  unsigned __int128 p = (unsigned __int128)a * b;
  while ((p >> 64) != 0)
    p = (unsigned __int128)(uint64_t)(p >> 64) * kMwcR + (uint64_t)p;
  uint64_t r = (uint64_t)p;
  return r >= kMwcM ? r - kMwcM : r;
}

constexpr ApowTable make_apow() {
  ApowTable t{};
  uint64_t sq = kMwcA;
  for (int i = 0; i < 64; ++i) {
    t.v[i] = sq;
    sq = cx_mulmod(sq, sq);
  }
  return t;
}

__constant__ ApowTable c_apow_table = make_apow();

// state of draw k: decode(BASE * A^k mod M), x = v / A, c = v % A
__device__ __forceinline__ Mwc mwc_jump(uint64_t k) {
  uint64_t acc = 1;
  for (int i = 0; k != 0; ++i, k >>= 1)
    if (k & 1) acc = mulmod_m(acc, c_apow_table.v[i]);
  uint64_t v = mulmod_m(kMwcBase, acc);
  Mwc s;
  s.x = (uint32_t)(v / kMwcA);
  s.c = (uint32_t)(v % kMwcA);
  return s;
}

constexpr int kBX = 64, kBY = 4;

__host__ __forceinline__ dim3 grid_for(int W, int H) {
  return dim3((unsigned)((W + kBX - 1) / kBX), (unsigned)((H + kBY - 1) / kBY));
}

struct ViewDev {
  const float4* __restrict__ tex1;
  const float4* __restrict__ tex2;
  const uint8_t* __restrict__ sgn1;
  const float* __restrict__ wts1;
  int W, H, R;
  size_t npix;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Linear sample of a texel row at column u (CLK_FILTER_LINEAR with
// CLK_ADDRESS_CLAMP_TO_EDGE, pmhstereo.cl:71): colour rgb and gradient.
struct Sample {
  float r, g, b;
  float4 grad;
};

__device__ __forceinline__ Sample sample_row(const float4* __restrict__ row,
                                             int W, float u) {
  u = fmaxf(fminf(u, (float)W), -1.0f);
  const float fl = floorf(u);
  const float a = u - fl;
  int i0 = (int)fl;
  int i1 = i0 + 1;
  i0 = clampi(i0, 0, W - 1);
  i1 = clampi(i1, 0, W - 1);
  const float4 A0 = row[2 * i0], G0 = row[2 * i0 + 1];
  const float4 A1 = row[2 * i1], G1 = row[2 * i1 + 1];
  Sample s;
  s.r = fmaf(a, A1.x - A0.x, A0.x);
  s.g = fmaf(a, A1.y - A0.y, A0.y);
  s.b = fmaf(a, A1.z - A0.z, A0.z);
  s.grad.x = fmaf(a, G1.x - G0.x, G0.x);
  s.grad.y = fmaf(a, G1.y - G0.y, G0.y);
  s.grad.z = fmaf(a, G1.z - G0.z, G0.z);
  s.grad.w = fmaf(a, G1.w - G0.w, G0.w);
  return s;
}

// ---------------------------------------------------------------------
// patch_eval_census (pmhstereo.cl:661-736).  The reference tests
// census/bilw >= mincost after every tap; the census terms are >= 0 and
// accumulate monotonically, so testing at the end of each row (and after
// the last tap) returns 1.0 exactly when the per-tap test would.
// kCount: also count the taps the reference's per-tap test evaluates
// (the algorithmic work; used by an untimed counting pass only).
// ---------------------------------------------------------------------
template <bool kCount = false>
__device__ float cost_eval(const ViewDev& v, int x, int y, float4 s,
                           float mincost, uint32_t* ntaps = nullptr) {
  const float xf = (float)x, yf = (float)y;
  const float Wf = (float)v.W;
  const float xp = transform(s, xf, yf);
  if (!(finite_f(s.x) && finite_f(s.y) && finite_f(s.z)) ||
      to_normal(s).z < kNormalZMin || xp < 0.0f || xp >= Wf)
    return 1e5f;
  const int R = v.R, W = v.W, T = 2 * R + 1;
  const size_t p = (size_t)y * W + x;
  const float bilw = v.tex1[2 * p].w;
  const int dy0 = max(-R, -y), dy1 = min(R, v.H - 1 - y);
  const int dx0 = max(-R, -x), dx1 = min(R, W - 1 - x);
  float census = 0.0f, wsum = 0.0f;
  for (int dy = dy0; dy <= dy1; ++dy) {
    const int qy = y + dy;
    const float rowc = fmaf(s.y, (float)qy, s.z);
    const float4* __restrict__ row1 = v.tex1 + 2 * (size_t)qy * W;
    const float4* __restrict__ row2 = v.tex2 + 2 * (size_t)qy * W;
    const uint8_t* __restrict__ srow = v.sgn1 + (size_t)qy * W;
    const float* __restrict__ wrow = v.wts1 + (size_t)((dy + R) * T) * v.npix + p;
    for (int dx = dx0; dx <= dx1; ++dx) {
      const int qx = x + dx;
      const float xq = fmaf(s.x, (float)qx, rowc);
      if (xq < 0.0f || xq >= Wf) continue;
      const float w = wrow[(size_t)(dx + R) * v.npix];
      const float4 t1 = row1[2 * qx];
      const uint32_t m1 = srow[qx];
      const Sample s2 = sample_row(row2, W, xq);
      const float c0 = fminf(fabsf(s2.r - t1.x), 0.05f) * 20.0f;
      const float c1 = fminf(fabsf(s2.g - t1.y), 0.05f) * 20.0f;
      const float c2 = fminf(fabsf(s2.b - t1.z), 0.05f) * 20.0f;
      const float cad = 0.33f * ((c0 + c1) + c2);
      float cost = 0.125f * cad;
      const int nd = __popc(m1 ^ signmask4(s2.grad));
      const float cc = 0.25f * (float)nd;
      cost += 0.875f * cc;
      census = fmaf(w, cost, census);
      wsum += w;
      if (kCount) {
        ++*ntaps;
        if (census / bilw >= mincost) return 1.0f;
      }
    }
    if (census / bilw >= mincost) return 1.0f;
  }
  if (wsum <= 1.0f) return 1.1f;
  return census / wsum;
}

// one census term of patch_census_klt2 (pmhstereo.cl:360-392); v0..v2 is
// the centred colour I1(q) - I1(p) the reference compares against
__device__ __forceinline__ float klt_term(const float4* __restrict__ row2,
                                          int W, float u, float v0, float v1,
                                          float v2, uint32_t ga, float w1) {
  const Sample s = sample_row(row2, W, u);
  const float c0 = fminf(fabsf(s.r - v0), 0.05f) * 20.0f;
  const float c1 = fminf(fabsf(s.g - v1), 0.05f) * 20.0f;
  const float c2 = fminf(fabsf(s.b - v2), 0.05f) * 20.0f;
  const float C = ((w1 * 0.125f) * 0.33f) * ((c0 + c1) + c2);
  const int nd = __popc(ga ^ signmask4(s.grad));
  // C += w1 * (1 - alpha) * 0.25 * n, 0.25 a double literal
  // (pmhstereo.cl:370): exact product, one rounding
  return fmaf(w1 * 0.875f, 0.25f * (float)nd, C);
}

// patch_census_klt2 (pmhstereo.cl:326-410) + mat3f_lusolve
// (cvt/cl/kernel/Matrix3.cl:286-311)
__device__ float4 klt_pass(const ViewDev& v, int x, int y, float4 s) {
  const int R = v.R, W = v.W, T = 2 * R + 1;
  const size_t p = (size_t)y * W + x;
  const float4 tc = v.tex1[2 * p];
  float J0 = 0.0f, J1 = 0.0f, J2 = 0.0f;
  float H00 = 0.0f, H01 = 0.0f, H02 = 0.0f, H11 = 0.0f, H12 = 0.0f,
        H22 = 0.0f;
  const int dy0 = max(-R, -y), dy1 = min(R, v.H - 1 - y);
  // taps need 1 <= x+dx < W-1 (the test is on the tap, pmhstereo.cl:358)
  const int dx0 = max(-R, 1 - x), dx1 = min(R, W - 2 - x);
  for (int dy = dy0; dy <= dy1; ++dy) {
    const int qy = y + dy;
    const float py = (float)qy;
    const float rowc = fmaf(s.y, py, s.z);
    const float4* __restrict__ row1 = v.tex1 + 2 * (size_t)qy * W;
    const float4* __restrict__ row2 = v.tex2 + 2 * (size_t)qy * W;
    const uint8_t* __restrict__ srow = v.sgn1 + (size_t)qy * W;
    const float* __restrict__ wrow = v.wts1 + (size_t)((dy + R) * T) * v.npix + p;
    for (int dx = dx0; dx <= dx1; ++dx) {
      const int qx = x + dx;
      const float px = (float)qx;
      const float4 t1 = row1[2 * qx];
      const float v0 = t1.x - tc.x, v1 = t1.y - tc.y, v2 = t1.z - tc.z;
      const uint32_t ga = srow[qx];
      const float w1 = wrow[(size_t)(dx + R) * v.npix];
      const float xq = fmaf(s.x, px, rowc);
      const float C = klt_term(row2, W, xq, v0, v1, v2, ga, w1);
      const float Cxp = klt_term(row2, W, xq + 1.0f, v0, v1, v2, ga, w1);
      const float Cxm = klt_term(row2, W, xq - 1.0f, v0, v1, v2, ga, w1);
      const float curv = (Cxp + Cxm) - 2.0f * C;
      if (curv > 0.0f) {
        const float gj = (w1 * (Cxm - Cxp)) * 0.5f;
        J0 = J0 + gj * px;
        J1 = J1 + gj * py;
        J2 = J2 + gj * 1.0f;
        const float h = w1 * curv;
        H00 = H00 + h * (px * px);
        H01 = H01 + h * (px * py);
        H02 = H02 + h * (px * 1.0f);
        H11 = H11 + h * (py * py);
        H12 = H12 + h * (py * 1.0f);
        H22 = H22 + h * (1.0f * 1.0f);
      }
    }
  }
  // LU without pivoting; H is symmetric (H10 = H01, ...)
  float m00 = H00, m01 = H01, m02 = H02;
  float m10 = H01, m11 = H11, m12 = H12;
  float m20 = H02, m21 = H12, m22 = H22;
  m10 /= m00;
  m20 /= m00;
  m11 -= m10 * m01;
  m12 -= m10 * m02;
  m21 -= m20 * m01;
  m22 -= m20 * m02;
  m21 /= m11;
  m22 -= m21 * m12;
  float b0 = J0, b1 = J1, b2 = J2;
  b1 -= m10 * b0;
  b2 -= m20 * b0 + m21 * b1;
  const float r2 = b2 / m22;
  const float r1 = (b1 - m12 * r2) / m11;
  const float r0 = (b0 - (m01 * r1 + m02 * r2)) / m00;
  return mk4(s.x - r0, s.y - r1, s.z - r2, s.w - 1e5f);
}

// ---------------------------------------------------------------------
// precompute: pmhstereo_bilateral_weight_to_alpha (pmhstereo.cl:1462-1498)
// and pmhstereo_gradxy (pmhstereo.cl:1549-1602); the per-tap support
// weights it sums are kept as planes for the cost and KLT loops
// ---------------------------------------------------------------------
__device__ __forceinline__ float4 fetch_rgba(const float* __restrict__ img,
                                             int C, int stride, int x, int y) {
  const float* r = img + (size_t)y * stride;
  if (C == 1) {
    const float v = r[x];
    return mk4(v, v, v, v);  // CL_INTENSITY
  }
  return *reinterpret_cast<const float4*>(r + 4 * (size_t)x);
}

__device__ __forceinline__ float gray_at(const float* __restrict__ img, int W,
                                         int H, int C, int stride, int x,
                                         int y) {
  if (x < 0 || x >= W || y < 0 || y >= H) return 0.0f;  // CLK_ADDRESS_CLAMP
  const float4 v = fetch_rgba(img, C, stride, x, y);
  return ((v.x * 0.2126f + v.y * 0.7152f) + v.z * 0.0722f) + v.w * 0.0f;
}

__global__ void __launch_bounds__(256)
k_prepare(const float* __restrict__ img, int W, int H, int C, int stride,
          int R, float cw, size_t npix, float4* __restrict__ tex,
          uint8_t* __restrict__ sgn, float* __restrict__ wts) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  const size_t p = (size_t)y * W + x;
  const float4 c = fetch_rgba(img, C, stride, x, y);
  const int T = 2 * R + 1;
  float wsum = 0.0f;
  for (int dy = -R; dy <= R; ++dy) {
    const int qy = y + dy;
    for (int dx = -R; dx <= R; ++dx) {
      const int qx = x + dx;
      float w = 0.0f;
      if (qx >= 0 && qx < W && qy >= 0 && qy < H) {
        const float4 q = fetch_rgba(img, C, stride, qx, qy);
        w = support_weight(q.x - c.x, q.y - c.y, q.z - c.z, cw);
        wsum += w;
      }
      wts[(size_t)((dy + R) * T + (dx + R)) * npix + p] = w;
    }
  }
  const float g = gray_at(img, W, H, C, stride, x, y);
  const float gxp = gray_at(img, W, H, C, stride, x + 1, y);
  const float gxm = gray_at(img, W, H, C, stride, x - 1, y);
  const float gyp = gray_at(img, W, H, C, stride, x, y + 1);
  const float gym = gray_at(img, W, H, C, stride, x, y - 1);
  const float4 grad = mk4(gxp - g, gyp - g, gxp - gxm, gyp - gym);
  tex[2 * p] = mk4(c.x, c.y, c.z, wsum);
  tex[2 * p + 1] = grad;
  sgn[p] = (uint8_t)signmask4(grad);
}

__device__ __forceinline__ ViewDev make_view(const Geometry& g, ViewBufs self,
                                             ViewBufs other) {
  ViewDev v;
  v.tex1 = self.tex;
  v.tex2 = other.tex;
  v.sgn1 = self.sgn;
  v.wts1 = self.wts;
  v.W = g.W;
  v.H = g.H;
  v.R = g.R;
  v.npix = g.npix;
  return v;
}

// pmhstereo_init (pmhstereo.cl:755-773): RNG_init(p, 3), three draws
__global__ void __launch_bounds__(256)
k_init(Geometry g, ViewBufs self, ViewBufs other, float dmax, int lr,
       float4* __restrict__ states) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= g.W || y >= g.H) return;
  const ViewDev v = make_view(g, self, other);
  const size_t p = (size_t)y * g.W + x;
  Mwc rng = mwc_jump((uint64_t)p * 3u);
  const float u0 = mwc_next(&rng), u1 = mwc_next(&rng), u2 = mwc_next(&rng);
  float4 s = state_init(u0, u1, u2, (float)x, (float)y, lr, dmax);
  s.w = cost_eval(v, x, y, s, 1e10f);
  states[p] = s;
}

// pmhstereo_disparity_to_normal (pmhstereo.cl:1604-1637)
__global__ void __launch_bounds__(256)
k_disparity_to_normal(int W, int H, const float* __restrict__ disp,
                      float4* __restrict__ normals) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  const size_t p = (size_t)y * W + x;
  const float d = disp[p];
  const float dxp = x + 1 < W ? disp[p + 1] : 0.0f;
  const float dyp = y + 1 < H ? disp[p + W] : 0.0f;
  const float dx = dxp - d, dy = dyp - d;
  float3 n = make_float3(0.0f * dy - dx * 1.0f, dx * 0.0f - 1.0f * dy,
                         1.0f * 1.0f - 0.0f * 0.0f);
  n = normalize3(n);
  normals[p] = mk4(n.x, n.y, n.z, 0.0f);
}

// pmhstereo_init_disparity_normal (pmhstereo.cl:802-828)
__global__ void __launch_bounds__(256)
k_init_prior(Geometry g, ViewBufs self, ViewBufs other, float dmax, int lr,
             const float* __restrict__ disp,
             const float4* __restrict__ normals, float4* __restrict__ states) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= g.W || y >= g.H) return;
  const ViewDev v = make_view(g, self, other);
  const size_t p = (size_t)y * g.W + x;
  Mwc rng = mwc_jump((uint64_t)p * 3u);
  const float d = disp[p];
  float4 s;
  if (d > 0.0f) {
    const float4 nn = normals[p];
    const float3 n = normalize3(make_float3(nn.x, nn.y, nn.z));
    s = from_disparity_normal((float)x, (float)y, lr ? d : -d, n);
  } else {
    const float u0 = mwc_next(&rng), u1 = mwc_next(&rng), u2 = mwc_next(&rng);
    s = state_init(u0, u1, u2, (float)x, (float)y, lr, dmax);
  }
  s.w = cost_eval(v, x, y, s, 1e10f);
  states[p] = s;
}

__global__ void __launch_bounds__(256)
k_viewbuf_clear(size_t n, unsigned long long* __restrict__ vbuf) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) vbuf[i] = ~0ull;
}

// ---------------------------------------------------------------------
// pmhstereo_propagate_view (pmhstereo.cl:880-1071), Jacobi schedule
// ---------------------------------------------------------------------
struct PropCtx {
  float xf, yf, dmax, theta;
  int lr;
  float4 smooth;
};

// candidate test with minimalCost (pmhstereo.cl:873-877) and the select
// rule (pmhstereo.cl:976)
template <bool kCount>
__device__ __forceinline__ void try_candidate(const ViewDev& v,
                                              const PropCtx& c, int x, int y,
                                              float4& self, float4 cand,
                                              uint32_t* ntaps) {
  const float ds = smooth_dist(self, c.smooth, c.xf, c.yf, c.dmax, c.lr);
  const float dc = smooth_dist(cand, c.smooth, c.xf, c.yf, c.dmax, c.lr);
  const float mincost = (self.w + c.theta * ds) - c.theta * dc;
  cand.w = cost_eval<kCount>(v, x, y, cand, mincost, ntaps);
  if (cand.w + c.theta * dc <= self.w + c.theta * ds) self = cand;
}

// keep the two smallest keys of a slot: the two nearest surfaces win the
// view scatter, independent of the order of arrival (the reference's
// atomic_inc race, pmhstereo.cl:1057-1067, keeps an arbitrary two)
__device__ __forceinline__ void slot_insert(unsigned long long* slot,
                                            uint32_t key) {
  unsigned long long old = *reinterpret_cast<volatile unsigned long long*>(slot);
  while (true) {
    const uint32_t k0 = (uint32_t)(old & 0xFFFFFFFFull);
    const uint32_t k1 = (uint32_t)(old >> 32);
    unsigned long long nw;
    if (key < k0)
      nw = ((unsigned long long)k0 << 32) | key;
    else if (key < k1)
      nw = ((unsigned long long)key << 32) | k0;
    else
      return;
    const unsigned long long prev = atomicCAS(slot, old, nw);
    if (prev == old) return;
    old = prev;
  }
}

// view scatter of a final state (pmhstereo.cl:1060-1067): slot at
// rte(x') of the other view when the integer disparity is > 0
__device__ __forceinline__ void scatter(float4 self, int x, int y, int W,
                                        int lr,
                                        unsigned long long* viewout) {
  const float xf = (float)x, yf = (float)y;
  const float xs = transform(self, xf, yf);
  const float xr = rintf(xs);
  const float disp = (xs - xf) * (lr ? -1.0f : 1.0f);
  if (xr >= 0.0f && xr < (float)W && disp >= 1.0f) {
    const uint32_t key = (uint32_t)(lr ? W - 1 - x : x);
    slot_insert(&viewout[(size_t)y * W + (int)xr], key);
  }
}

// kCount: accumulate the cost taps (counters[0]) and KLT taps
// (counters[1]) this launch evaluates; the results are identical
template <bool kCount>
__global__ void __launch_bounds__(256)
k_propagate(Geometry g, PropagateArgs a, unsigned long long* counters) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= g.W || y >= g.H) return;
  uint32_t ncost = 0;
  uint32_t* nc = kCount ? &ncost : nullptr;
  const ViewDev v = make_view(g, a.self, a.other);
  const int W = g.W, H = g.H;
  const size_t p = (size_t)y * W + x;
  PropCtx c;
  c.xf = (float)x;
  c.yf = (float)y;
  c.dmax = a.dmax;
  c.theta = a.theta;
  c.lr = a.lr;
  float4 sm = a.smooth ? a.smooth[p] : mk4(0.0f, 0.0f, 0.0f, 0.0f);
  sm.w = fminf(fmaxf(sm.w, 0.0f), 1.0f) * a.dmax;
  c.smooth = sm;

  float4 self = a.old[p];
  self.w = cost_eval<kCount>(v, x, y, self, 1e10f, nc);
  Mwc rng = mwc_jump((uint64_t)(p + (size_t)a.iter) * kPropDraws);

  // smooth proposal (pmhstereo.cl:968-976)
  {
    float4 cand = from_disparity_normal(
        c.xf, c.yf, sm.w * (a.lr ? -1.0f : 1.0f),
        normalize3(make_float3(sm.x, sm.y, sm.z)));
    if (!a.lr) cand = viewprop(cand);
    try_candidate<kCount>(v, c, x, y, self, cand, nc);
  }
  // random neighbours (pmhstereo.cl:981-991), x offset drawn first
  for (int i = 0; i < kNumRndSample; ++i) {
    const float ux = mwc_next(&rng);
    const float uy = mwc_next(&rng);
    const int ox = (int)(((ux - 0.5f) * 2.0f) * kRndSampleRadius + 0.5f);
    const int oy = (int)(((uy - 0.5f) * 2.0f) * kRndSampleRadius + 0.5f);
    const int nx = clampi(x + ox, 0, W - 1), ny = clampi(y + oy, 0, H - 1);
    try_candidate<kCount>(v, c, x, y, self, a.old[(size_t)ny * W + nx],
                          nc);
  }
  // random plane (pmhstereo.cl:996-1002)
  {
    const float u0 = mwc_next(&rng), u1 = mwc_next(&rng), u2 = mwc_next(&rng);
    try_candidate<kCount>(v, c, x, y, self,
                          state_init(u0, u1, u2, c.xf, c.yf, a.lr, a.dmax),
                          nc);
  }
  // other view (pmhstereo.cl:1007-1016)
  if (a.viewin) {
    const unsigned long long slot = a.viewin[p];
    const uint32_t keys[2] = {(uint32_t)(slot & 0xFFFFFFFFull),
                              (uint32_t)(slot >> 32)};
    for (int i = 0; i < kViewSamples; ++i) {
      if (keys[i] == kSlotEmpty) break;
      const int xs = a.lr ? (int)keys[i] : W - 1 - (int)keys[i];
      try_candidate<kCount>(v, c, x, y, self,
                            viewprop(a.src_other[(size_t)y * W + xs]), nc);
    }
  }
  // randomized refinement (pmhstereo.cl:1021-1028)
  {
    const float u0 = mwc_next(&rng), u1 = mwc_next(&rng), u2 = mwc_next(&rng);
    try_candidate<kCount>(
        v, c, x, y, self,
        state_refine(u0, u1, u2, self, c.xf, c.yf, a.dmax, a.lr), nc);
  }
  // KLT step (pmhstereo.cl:1034-1053)
  {
    float4 k = self;
    for (int it = 0; it < kKltPasses; ++it) k = klt_pass(v, x, y, k);
    try_candidate<kCount>(v, c, x, y, self, k, nc);
  }
  if (a.viewout) scatter(self, x, y, W, a.lr, a.viewout);
  a.out[p] = self;
  if (kCount) {
    // KLT taps: rows in the image times columns with 1 <= x+dx < W-1
    const int R = g.R;
    const int rows = min(R, H - 1 - y) - max(-R, -y) + 1;
    const int cols = min(R, W - 2 - x) - max(-R, 1 - x) + 1;
    const unsigned long long nklt =
        cols > 0 ? (unsigned long long)rows * cols * kKltPasses : 0ull;
    atomicAdd(&counters[0], (unsigned long long)ncost);
    atomicAdd(&counters[1], nklt);
  }
}

// view scatter alone, from a stored field: rebuilds the buffer the field's
// last propagate left behind
__global__ void __launch_bounds__(256)
k_scatter(int W, int H, const float4* __restrict__ states, int lr,
          unsigned long long* __restrict__ viewout) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  scatter(states[(size_t)y * W + x], x, y, W, lr, viewout);
}

__global__ void __launch_bounds__(256)
k_cost_field(Geometry g, ViewBufs self, ViewBufs other,
             const float4* __restrict__ planes, float* __restrict__ cost) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= g.W || y >= g.H) return;
  const ViewDev v = make_view(g, self, other);
  const size_t p = (size_t)y * g.W + x;
  cost[p] = cost_eval(v, x, y, planes[p], 1e10f);
}

__global__ void __launch_bounds__(256)
k_klt_field(Geometry g, ViewBufs self, ViewBufs other,
            const float4* __restrict__ planes, int passes,
            float4* __restrict__ out) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= g.W || y >= g.H) return;
  const ViewDev v = make_view(g, self, other);
  const size_t p = (size_t)y * g.W + x;
  float4 s = planes[p];
  for (int i = 0; i < passes; ++i) s = klt_pass(v, x, y, s);
  out[p] = s;
}

// pmhstereo_consistency (pmhstereo.cl:1095-1129)
__global__ void __launch_bounds__(256)
k_consistency(int W, int H, const float4* __restrict__ left,
              const float4* __restrict__ right, float maxdispdiff,
              float maxdeg, float4* __restrict__ out) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  const size_t p = (size_t)y * W + x;
  const float athr = maxdeg / 180.0f;
  const float4 sl = left[p];
  const float xf = (float)x, yf = (float)y;
  const float x2 = transform(sl, xf, yf);
  float4 val = mk4(0.0f, 0.0f, 0.0f, -1.0f);
  const float fl = floorf(x2), ce = ceilf(x2);
  if (ce >= 0.0f && fl < (float)W) {
    const int i0 = clampi((int)fl, 0, W - 1), i1 = clampi((int)ce, 0, W - 1);
    const float4 r1 = right[(size_t)y * W + i0];
    const float4 r2 = right[(size_t)y * W + i1];
    const float3 nl = to_normal(sl);
    const float3 n1 = to_normal(viewprop(r1));
    const float3 n2 = to_normal(viewprop(r2));
    const float nd1 = (nl.x * n1.x + nl.y * n1.y) + nl.z * n1.z;
    const float nd2 = (nl.x * n2.x + nl.y * n2.y) + nl.z * n2.z;
    const float dm1 = fabsf(xf - transform(r1, x2, yf));
    const float dm2 = fabsf(xf - transform(r2, x2, yf));
    const float a1 = acospi_f(fminf(fmaxf(nd1, -1.0f), 1.0f));
    const float a2 = acospi_f(fminf(fmaxf(nd2, -1.0f), 1.0f));
    if (!(dm1 >= maxdispdiff || a1 >= athr) ||
        !(dm2 >= maxdispdiff || a2 >= athr))
      val = sl;
  }
  out[p] = val;
}

// pmhstereo_to_disparity (pmhstereo.cl:1232-1257)
__global__ void __launch_bounds__(256)
k_to_disparity(int W, int H, const float4* __restrict__ states, float scale,
               int lr, float* __restrict__ disp) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  const size_t p = (size_t)y * W + x;
  const float4 s = states[p];
  const float d = (transform(s, (float)x, (float)y) - (float)x) *
                  (lr ? -1.0f : 1.0f);
  float val;
  if (len3(s) < 1e-4f || d < 0.0f || d > (float)W)
    val = 0.0f;
  else
    val = fmaxf(d, 0.0f);
  disp[p] = val * scale;
}

// row scans of pmhstereo_fill_depthmap / pmhstereo_fill_normalmap
__device__ __forceinline__ void row_fill_pair(const float4* __restrict__ in,
                                              int W, int x, int y, float4* l,
                                              float4* r) {
  const float4* row = in + (size_t)y * W;
  float4 a = mk4(0.0f, 0.0f, 0.0f, 0.0f), b = a;
  for (int i = x - 1; len3(a) < 1e-1f && i >= 0; --i) a = row[i];
  for (int i = x + 1; len3(b) < 1e-1f && i < W; ++i) b = row[i];
  if (len3(a) < 1e-1f) a.x = a.y = a.z = -1e5f;
  if (len3(b) < 1e-1f) b.x = b.y = b.z = -1e5f;
  const float xf = (float)x, yf = (float)y;
  a.w = -(transform(a, xf, yf) - xf);
  b.w = -(transform(b, xf, yf) - xf);
  *l = a;
  *r = b;
}

// pmhstereo_fill_depthmap (pmhstereo.cl:1187-1230)
__global__ void __launch_bounds__(256)
k_fill_depthmap(int W, int H, const float4* __restrict__ states, float scale,
                float* __restrict__ disp) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  const size_t p = (size_t)y * W + x;
  const float4 s = states[p];
  float val;
  if (len3(s) < 1e-4f) {
    float4 l, r;
    row_fill_pair(states, W, x, y, &l, &r);
    val = fminf(l.w, r.w);
  } else {
    val = fabsf(transform(s, (float)x, (float)y) - (float)x);
  }
  disp[p] = val * scale;
}

// pmhstereo_fill_normalmap (pmhstereo.cl:1259-1305)
__global__ void __launch_bounds__(256)
k_fill_normalmap(int W, int H, const float4* __restrict__ states,
                 float4* __restrict__ normals) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  const size_t p = (size_t)y * W + x;
  const float4 s = states[p];
  float3 n;
  if (len3(s) < 1e-1f) {
    float4 l, r;
    row_fill_pair(states, W, x, y, &l, &r);
    n = (l.w < r.w) ? to_normal(l) : to_normal(r);
  } else {
    n = to_normal(s);
  }
  normals[p] = mk4(n.x, n.y, n.z, 0.0f);
}

// draws [k + 16 t, k + 16 t + 16) by one jump and sixteen steps
__global__ void __launch_bounds__(256)
k_draws(uint64_t k, int n, float* __restrict__ out) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int i0 = t * 16;
  if (i0 >= n) return;
  Mwc s = mwc_jump(k + (uint64_t)i0);
  for (int j = 0; j < 16 && i0 + j < n; ++j) out[i0 + j] = mwc_next(&s);
}

// plain streaming copy of n float4: the known-size workload that the HBM
// byte counters are calibrated with
__global__ void __launch_bounds__(256)
k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst,
             size_t n) {
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

__global__ void __launch_bounds__(256)
k_math(const float* __restrict__ in, int n, float* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float x = in[i];
  float s, c;
  sincos_f(x, &s, &c);
  out[5 * i + 0] = exp_f(x);
  out[5 * i + 1] = s;
  out[5 * i + 2] = c;
  out[5 * i + 3] = acospi_f(x);
  out[5 * i + 4] = sqrtf(x);
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
#define PM_LAUNCH(kernel, grid, block, st, ...)                 \
  do {                                                          \
    hipLaunchKernelGGL(kernel, grid, block, 0, st, __VA_ARGS__); \
    return hipGetLastError();                                   \
  } while (0)

static const dim3 kBlock2D(kBX, kBY);

hipError_t launch_prepare(hipStream_t st, const Geometry& g, const float* img,
                          int channels, int stride, ViewBufs v) {
  PM_LAUNCH(k_prepare, grid_for(g.W, g.H), kBlock2D, st, img, g.W, g.H,
            channels, stride, g.R, g.color_weight, g.npix, v.tex, v.sgn,
            v.wts);
}

hipError_t launch_init(hipStream_t st, const Geometry& g, ViewBufs self,
                       ViewBufs other, float dmax, int lr, float4* states) {
  PM_LAUNCH(k_init, grid_for(g.W, g.H), kBlock2D, st, g, self, other, dmax,
            lr, states);
}

hipError_t launch_disparity_to_normal(hipStream_t st, const Geometry& g,
                                      const float* disp, float4* normals) {
  PM_LAUNCH(k_disparity_to_normal, grid_for(g.W, g.H), kBlock2D, st, g.W, g.H,
            disp, normals);
}

hipError_t launch_init_prior(hipStream_t st, const Geometry& g, ViewBufs self,
                             ViewBufs other, float dmax, int lr,
                             const float* disp, const float4* normals,
                             float4* states) {
  PM_LAUNCH(k_init_prior, grid_for(g.W, g.H), kBlock2D, st, g, self, other,
            dmax, lr, disp, normals, states);
}

hipError_t launch_viewbuf_clear(hipStream_t st, const Geometry& g,
                                unsigned long long* vbuf) {
  const unsigned blocks = (unsigned)((g.npix + 255) / 256);
  PM_LAUNCH(k_viewbuf_clear, dim3(blocks), dim3(256), st, g.npix, vbuf);
}

hipError_t launch_propagate(hipStream_t st, const Geometry& g,
                            const PropagateArgs& a,
                            unsigned long long* counters) {
  if (counters)
    PM_LAUNCH(k_propagate<true>, grid_for(g.W, g.H), kBlock2D, st, g, a,
              counters);
  PM_LAUNCH(k_propagate<false>, grid_for(g.W, g.H), kBlock2D, st, g, a,
            counters);
}

hipError_t launch_scatter(hipStream_t st, const Geometry& g,
                          const float4* states, int lr,
                          unsigned long long* viewout) {
  PM_LAUNCH(k_scatter, grid_for(g.W, g.H), kBlock2D, st, g.W, g.H, states, lr,
            viewout);
}

hipError_t launch_cost_field(hipStream_t st, const Geometry& g, ViewBufs self,
                             ViewBufs other, const float4* planes,
                             float* cost) {
  PM_LAUNCH(k_cost_field, grid_for(g.W, g.H), kBlock2D, st, g, self, other,
            planes, cost);
}

hipError_t launch_klt_field(hipStream_t st, const Geometry& g, ViewBufs self,
                            ViewBufs other, const float4* planes, int passes,
                            float4* out) {
  PM_LAUNCH(k_klt_field, grid_for(g.W, g.H), kBlock2D, st, g, self, other,
            planes, passes, out);
}

hipError_t launch_consistency(hipStream_t st, const Geometry& g,
                              const float4* left, const float4* right,
                              float maxdispdiff, float maxdeg, float4* out) {
  PM_LAUNCH(k_consistency, grid_for(g.W, g.H), kBlock2D, st, g.W, g.H, left,
            right, maxdispdiff, maxdeg, out);
}

hipError_t launch_to_disparity(hipStream_t st, const Geometry& g,
                               const float4* states, float scale, int lr,
                               float* disp) {
  PM_LAUNCH(k_to_disparity, grid_for(g.W, g.H), kBlock2D, st, g.W, g.H,
            states, scale, lr, disp);
}

hipError_t launch_fill_depthmap(hipStream_t st, const Geometry& g,
                                const float4* states, float scale,
                                float* disp) {
  PM_LAUNCH(k_fill_depthmap, grid_for(g.W, g.H), kBlock2D, st, g.W, g.H,
            states, scale, disp);
}

hipError_t launch_fill_normalmap(hipStream_t st, const Geometry& g,
                                 const float4* states, float4* normals) {
  PM_LAUNCH(k_fill_normalmap, grid_for(g.W, g.H), kBlock2D, st, g.W, g.H,
            states, normals);
}

hipError_t launch_draws(hipStream_t st, uint64_t k, int n, float* out) {
  const int threads = (n + 15) / 16;
  PM_LAUNCH(k_draws, dim3((unsigned)((threads + 255) / 256)), dim3(256), st,
            k, n, out);
}

hipError_t launch_copy_probe(hipStream_t st, const float4* src, float4* dst,
                             size_t n) {
  PM_LAUNCH(k_copy_probe, dim3(4096), dim3(256), st, src, dst, n);
}

hipError_t launch_math(hipStream_t st, const float* in, int n, float* out) {
  PM_LAUNCH(k_math, dim3((unsigned)((n + 255) / 256)), dim3(256), st, in, n,
            out);
}

}  // namespace pm
