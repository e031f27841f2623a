// pm_device.h -- device-side scalar math and plane algebra of the
// PatchMatch engine (gfx950).
//
// Arithmetic contract (the whole engine is built with -ffp-contract=off and
// correctly rounded float32 division / sqrt): every expression below is
// evaluated as written, left to right, one rounding per operation, except
// the fused operations written as fmaf().  Those are the contractions an
// OpenCL compiler with FP_CONTRACT ON forms in the reference's source
// expressions: a product operand of a sum or difference fused into it,
// the left one first (a * b + c, a + b * c, a - b * c, a -= b * c), and a
// built-in dot as the chain fma(x2, y2, fma(x1, y1, x0 * y0)); never
// across statements or helper calls.  exp, sin/cos and acospi are the
// engine's own Cody-Waite + Taylor implementations, so every result is a
// defined function of the inputs, independent of the vendor math library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pm {

constexpr float kNormalZMin = 1e-4f;        // NORMAL_Z_MIN, pmhstereo.cl:50
constexpr float kNormalCompMax = 0.95f;     // NORMALCOMPMAX, pmhstereo.cl:32
constexpr float kDepthRefineMul = 2.0f;     // DEPTHREFINEMUL, pmhstereo.cl:30
constexpr float kNormalRefineMul = 0.05f;   // NORMALREFINEMUL, pmhstereo.cl:31
constexpr int kNumRndSample = 4;            // NUMRNDSAMPLE, pmhstereo.cl:34
constexpr float kRndSampleRadius = 5.0f;    // NUMRNDSAMPLERADIUS, :35
constexpr int kViewSamples = 2;             // VIEWSAMPLES, pmhstereo.cl:46
constexpr int kKltPasses = 6;               // pmhstereo.cl:1036-1038
constexpr uint32_t kSlotEmpty = 0xFFFFFFFFu;

// ---------------------------------------------------------------------
// scalar math
// ---------------------------------------------------------------------

// e^x: k = rint(x log2 e), r = x - k ln2 (two-part ln2), degree-7 Taylor
// polynomial for e^r, result scaled by 2^k.  x < -87 gives 0.
__device__ __forceinline__ float exp_f(float x) {
  if (x != x) return x;
  if (x < -87.0f) return 0.0f;
  if (x > 88.5f) return __int_as_float(0x7f800000);
  float k = rintf(x * 0x1.715476p+0f);
  float r = fmaf(-k, 0x1.62e400p-1f, x);
  r = fmaf(-k, 0x1.7f7d1cp-20f, r);
  float p = 0x1.a01a02p-13f;
  p = fmaf(p, r, 0x1.6c16c2p-10f);
  p = fmaf(p, r, 0x1.111112p-7f);
  p = fmaf(p, r, 0x1.555556p-5f);
  p = fmaf(p, r, 0x1.555556p-3f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)k);
}

// sin and cos: three-part pi/2 reduction, Taylor polynomials on
// |r| <= pi/4, quadrant selection.
__device__ __forceinline__ void sincos_f(float x, float* s, float* c) {
  float k = rintf(x * 0x1.45f306p-1f);
  float r = fmaf(-k, 0x1.920000p+0f, x);
  r = fmaf(-k, 0x1.fb4000p-12f, r);
  r = fmaf(-k, 0x1.4442d2p-24f, r);
  float r2 = r * r;
  float ps = -0x1.ae6456p-26f;
  ps = fmaf(ps, r2, 0x1.71de3ap-19f);
  ps = fmaf(ps, r2, -0x1.a01a02p-13f);
  ps = fmaf(ps, r2, 0x1.111112p-7f);
  ps = fmaf(ps, r2, -0x1.555556p-3f);
  float sr = fmaf(r * r2, ps, r);
  float pc = 0x1.1eed8ep-29f;
  pc = fmaf(pc, r2, -0x1.27e4fcp-22f);
  pc = fmaf(pc, r2, 0x1.a01a02p-16f);
  pc = fmaf(pc, r2, -0x1.6c16c2p-10f);
  pc = fmaf(pc, r2, 0x1.555556p-5f);
  pc = fmaf(pc, r2, -0.5f);
  float cr = fmaf(pc, r2, 1.0f);
  int q = ((int)k) & 3;
  float ss = (q & 1) ? cr : sr;
  float cc = (q & 1) ? sr : cr;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

// asin on |t| <= 1/2, Taylor series through t^27
__device__ __forceinline__ float asin_small(float t) {
  float t2 = t * t;
  float p = 0x1.782ddap-8f;
  p = fmaf(p, t2, 0x1.a6863ep-8f);
  p = fmaf(p, t2, 0x1.df3bd4p-8f);
  p = fmaf(p, t2, 0x1.12ef3cp-7f);
  p = fmaf(p, t2, 0x1.3fde50p-7f);
  p = fmaf(p, t2, 0x1.7a8788p-7f);
  p = fmaf(p, t2, 0x1.c9999ap-7f);
  p = fmaf(p, t2, 0x1.1c4ec4p-6f);
  p = fmaf(p, t2, 0x1.6e8ba2p-6f);
  p = fmaf(p, t2, 0x1.f1c71cp-6f);
  p = fmaf(p, t2, 0x1.6db6dcp-5f);
  p = fmaf(p, t2, 0x1.333334p-4f);
  p = fmaf(p, t2, 0x1.555556p-3f);
  return fmaf(t * t2, p, t);
}

// acos(x)/pi on [-1, 1]
__device__ __forceinline__ float acospi_f(float x) {
  if (x != x) return x;
  float ax = fabsf(x);
  float a;
  if (ax <= 0.5f) {
    a = (0x1.921fb6p+0f - asin_small(x)) + -0x1.777a5cp-25f;
  } else {
    float z = (1.0f - ax) * 0.5f;
    float h = 2.0f * asin_small(sqrtf(z));
    a = (x > 0.0f) ? h : (0x1.921fb6p+1f - h) + -0x1.777a5cp-24f;
  }
  return a * 0x1.45f306p-2f;
}

__device__ __forceinline__ bool finite_f(float v) {
  return (__float_as_uint(v) & 0x7f800000u) != 0x7f800000u;
}

// ---------------------------------------------------------------------
// MWC64X stream (cvt/cl/kernel/MWC64X.cl, RNG.cl): draw k of the single
// global stream is decode(BASE * A^k mod M), x = v / A, c = v % A.
// ---------------------------------------------------------------------
constexpr uint64_t kMwcA = 4294883355ULL;                 // MWC64X.cl:106
constexpr uint64_t kMwcM = 18446383549859758079ULL;       // MWC64X.cl:107
constexpr uint64_t kMwcBase = 4077358422479273989ULL;     // MWC64X.cl:94
constexpr uint64_t kMwcR = 360523849793537ULL;            // 2^64 - M

// a*b mod M: 128-bit product, high word folded with 2^64 = R (mod M)
__device__ __forceinline__ uint64_t mulmod_m(uint64_t a, uint64_t b) {
  uint64_t lo = a * b;
  uint64_t hi = __umul64hi(a, b);
  while (hi != 0) {
    uint64_t l2 = hi * kMwcR;
    uint64_t h2 = __umul64hi(hi, kMwcR);
    uint64_t s = lo + l2;
    h2 += (s < lo) ? 1u : 0u;
    lo = s;
    hi = h2;
  }
  if (lo >= kMwcM) lo -= kMwcM;
  return lo;
}

struct Mwc {
  uint32_t x, c;
};

// Stream index of the first draw of a propagate at pixel p, iteration
// iter.  The reference calls RNG_init(&rng, p + iter, <numdraws text>)
// (pmhstereo.cl:942); the macro expands to Skip((linpos) * numdraws)
// (RNG.cl:54-57) with the argument text unparenthesised, which with
// PROPSIZE 0, NUMRNDTRIES 2 and NUMRNDSAMPLE 4 evaluates to
// (p + iter) * 2 * 3 + 2 * 4.  A propagate consumes 14 draws, so the
// windows of neighbouring pixels overlap.
__device__ __forceinline__ uint64_t prop_draw_start(uint64_t p, int iter) {
  return (p + (uint64_t)iter) * 6u + 8u;
}

// MWC64X_NextFloat (RNG.cl:30-35) + MWC64X_Step (MWC64X.cl:109-120)
__device__ __forceinline__ float mwc_next(Mwc* s) {
  uint32_t res = s->x ^ s->c;
  uint32_t X = s->x, C = s->c;
  uint32_t Xn = (uint32_t)kMwcA * X + C;
  uint32_t carry = (Xn < C) ? 1u : 0u;
  uint32_t Cn = __umulhi((uint32_t)kMwcA, X) + carry;
  s->x = Xn;
  s->c = Cn;
  return 0x1.0p-32f * (float)res;
}

// ---------------------------------------------------------------------
// plane algebra (pmhstereo.cl:77-224); state = (a, b, c, cost)
// ---------------------------------------------------------------------
__device__ __forceinline__ float4 mk4(float x, float y, float z, float w) {
  return make_float4(x, y, z, w);
}

// nd_state_viewprop, pmhstereo.cl:174-181
__device__ __forceinline__ float4 viewprop(float4 n) {
  float4 r = mk4(1.0f / n.x, -n.y / n.x, -n.z / n.x, 0.0f);
  if (fabsf(n.x) < 1e-8f) r = mk4(0.0f, -n.y, -n.z, 0.0f);
  return r;
}

// nd_state_transform, pmhstereo.cl:183-186: (a x + b y) + c with the
// product a x fused into the first sum (the contraction OpenCL's default
// FP_CONTRACT permits for this expression)
__device__ __forceinline__ float transform(float4 s, float x, float y) {
  return fmaf(s.x, x, s.y * y) + s.z;
}

__device__ __forceinline__ float4 one_minus(float4 s) {
  return mk4(1.0f - s.x, 0.0f - s.y, 0.0f - s.z, 0.0f - s.w);
}

// OpenCL normalize(): v / |v|, zero vector unchanged
__device__ __forceinline__ float3 normalize3(float3 v) {
  float l2 = fmaf(v.z, v.z, fmaf(v.y, v.y, v.x * v.x));
  if (l2 == 0.0f) return v;
  float l = sqrtf(l2);
  return make_float3(v.x / l, v.y / l, v.z / l);
}

// nd_state_to_normal, pmhstereo.cl:201-210 (native_rsqrt as 1/sqrt)
__device__ __forceinline__ float3 to_normal(float4 s0) {
  float4 s = one_minus(s0);
  float3 n;
  n.z = 1.0f / sqrtf(fmaf(s.x, s.x, s.y * s.y) + 1.0f);
  n.x = -s.x * n.z;
  n.y = -s.y * n.z;
  n.z = sqrtf(fmaf(-n.y, n.y, fmaf(-n.x, n.x, 1.0f)));
  return n;
}

__device__ __forceinline__ float4 plane_from_nd(float3 n, float x, float y,
                                                float d) {
  float4 t = mk4(-n.x / n.z, -n.y / n.z, fmaf(n.x, x, n.y * y) / n.z + d,
                 0.0f);
  return mk4(1.0f - t.x, 0.0f - t.y, 0.0f - t.z, 0.0f - t.w);
}

// nd_state_from_disparity_normal, pmhstereo.cl:99-106
__device__ __forceinline__ float4 from_disparity_normal(float x, float y,
                                                        float d, float3 n) {
  n.z = fmaxf(n.z, kNormalZMin);
  n = normalize3(n);
  return plane_from_nd(n, x, y, d);
}

// 2*M_PI*u with M_PI a double constant (pmhstereo.cl:82,144)
__device__ __forceinline__ float azimuth_of(float u) {
  return (float)(2.0 * 3.14159265358979323846 * (double)u);
}

// nd_state_init, pmhstereo.cl:77-97
__device__ __forceinline__ float4 state_init(float u0, float u1, float u2,
                                             float x, float y, int lr,
                                             float dmax) {
  float z = u0 * dmax;
  float az = azimuth_of(u1);
  float u = u2 * kNormalCompMax;
  float w = sqrtf(fmaf(-u, u, 1.0f));
  float sn, cs;
  sincos_f(az, &sn, &cs);
  float3 n = make_float3(w * cs, w * sn, u);
  n.z = fmaxf(n.z, kNormalZMin);
  n = normalize3(n);
  float4 r = plane_from_nd(n, x, y, z);
  if (!lr) r = viewprop(r);
  return r;
}

// nd_state_refine, pmhstereo.cl:129-172 (right view returns viewprop of
// the unperturbed converted state, pmhstereo.cl:169-170)
__device__ __forceinline__ float4 state_refine(float u0, float u1, float u2,
                                               float4 s0, float x, float y,
                                               float dmax, int lr) {
  float4 st = s0;
  if (!lr) st = viewprop(st);
  st = one_minus(st);
  float z = fmaf(st.x, x, st.y * y) + st.z;
  float3 n;
  n.z = 1.0f / sqrtf(fmaf(st.x, st.x, st.y * st.y) + 1.0f);
  n.x = -st.x * n.z;
  n.y = -st.y * n.z;
  z = fmaf((u0 - 0.5f) * 2.0f, kDepthRefineMul, z);
  z = fminf(fmaxf(z, 0.0f), dmax);
  float az = azimuth_of(u1);
  float u = u2 * (1.0f - kNormalRefineMul);
  float w = sqrtf(fmaf(-u, u, 1.0f));
  float sn, cs;
  sincos_f(az, &sn, &cs);
  float3 n2 = make_float3(w * cs, w * sn, u);
  n.x = (n.x + n2.x) * 0.5f;
  n.y = (n.y + n2.y) * 0.5f;
  n.z = (n.z + n2.z) * 0.5f;
  n.z = fmaxf(n.z, kNormalZMin);
  n = normalize3(n);
  float4 r = plane_from_nd(n, x, y, z);
  if (!lr) return viewprop(st);
  return r;
}

// nd_state_to_ref_normal_depth, pmhstereo.cl:108-127
__device__ __forceinline__ float4 to_ref_normal_depth(float4 s0, float x,
                                                      float y, int lr) {
  float4 st = s0;
  float sidesign = lr ? -1.0f : 1.0f;
  if (!lr) st = viewprop(s0);
  st = one_minus(st);
  float4 r;
  r.w = (transform(s0, x, y) - x) * sidesign;
  r.z = 1.0f / sqrtf(fmaf(st.x, st.x, st.y * st.y) + 1.0f);
  r.x = -st.x * r.z;
  r.y = -st.y * r.z;
  if (!(finite_f(r.x) && finite_f(r.y) && finite_f(r.z)))
    return mk4(0.0f, 0.0f, 1.0f, 0.0f);
  return r;
}

// smoothDistance, pmhstereo.cl:846-871
__device__ __forceinline__ float smooth_dist(float4 s, float4 sm, float x,
                                             float y, float dmax, int lr) {
  float4 a = to_ref_normal_depth(s, x, y, lr);
  float id2 = 1.0f / (dmax * dmax);
  float den = id2 + 0.3f;
  float dg = 0.1f / den, dw = id2 / den;
  float dx = a.x - sm.x, dy = a.y - sm.y, dz = a.z - sm.z, dd = a.w - sm.w;
  return fmaf(dd, dw * dd, fmaf(dz, dg * dz, fmaf(dy, dg * dy, dx * (dg * dx))));
}

// gray value used by the support weight: the built-in
// dot(v, (0.2126, 0.7152, 0.0722)) as a chain of fused multiply-adds
__device__ __forceinline__ float gray3(float x, float y, float z) {
  return fmaf(z, 0.0722f, fmaf(y, 0.7152f, x * 0.2126f));
}

// exp_f for x <= 0 without branches: the same operation sequence, the
// x < -87 case selected afterwards (equal to exp_f for every x <= 0)
__device__ __forceinline__ float exp_nonpos(float x) {
  const float k = rintf(x * 0x1.715476p+0f);
  float r = fmaf(-k, 0x1.62e400p-1f, x);
  r = fmaf(-k, 0x1.7f7d1cp-20f, r);
  float p = 0x1.a01a02p-13f;
  p = fmaf(p, r, 0x1.6c16c2p-10f);
  p = fmaf(p, r, 0x1.111112p-7f);
  p = fmaf(p, r, 0x1.555556p-5f);
  p = fmaf(p, r, 0x1.555556p-3f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  const float v = ldexpf(p, (int)k);
  return x < -87.0f ? 0.0f : v;
}

// support weight exp(-pow(|dot(d, gray)|, 1) * cw), pmhstereo.cl:701; the
// argument is <= 0 (cw >= 0 is checked at context creation)
__device__ __forceinline__ float support_weight(float dx, float dy, float dz,
                                                float cw) {
  return exp_nonpos(-fabsf(gray3(dx, dy, dz)) * cw);
}

// ---------------------------------------------------------------------
// image access
// ---------------------------------------------------------------------
__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

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
  return fmaf(v.w, 0.0f, fmaf(v.z, 0.0722f, fmaf(v.y, 0.7152f, v.x * 0.2126f)));
}

__device__ __forceinline__ uint32_t signmask4(float4 g) {
  return (__float_as_uint(g.x) >> 31) | ((__float_as_uint(g.y) >> 31) << 1) |
         ((__float_as_uint(g.z) >> 31) << 2) |
         ((__float_as_uint(g.w) >> 31) << 3);
}

__device__ __forceinline__ float len3(float4 s) {
  return sqrtf(fmaf(s.z, s.z, fmaf(s.y, s.y, s.x * s.x)));
}

}  // namespace pm
