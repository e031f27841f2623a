// pm_engine.hip -- gfx950 kernels of the PatchMatch stereo engine.
//
// Work decomposition (kernels that evaluate plane costs):
//  - a work-group is a 64 x 4 pixel block, each wavefront one row of 64
//    consecutive pixels; blocks are renumbered so that each XCD walks one
//    contiguous band of the image (its L2 then holds the rows the band
//    reads);
//  - the block stages its own-view tile (pixels plus an r halo: colour and
//    gradient sign bits) in LDS once; support weights are recomputed per
//    tap from that tile instead of being streamed from HBM;
//  - the other view is sampled straight from its colour and gradient
//    planes (two adjacent 16-byte records of each per linear sample), so
//    a wavefront's gather of one plane spans half the cache lines that
//    interleaved 32-byte texels would.
// The cost loops run wave-uniformly over patch rows (one image row per
// wavefront) and every lane stays resident (no early return): lanes are
// masked with flags, and kernels reach the block barrier with all threads.
//
// Per-pixel work follows cvt/cl/kernel/pmhstereo.cl (cited per kernel);
// the arithmetic contract is in pm_device.h.
#include <algorithm>
#include <climits>

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

// ---------------------------------------------------------------------
// block geometry and LDS layout
// ---------------------------------------------------------------------
constexpr int kTX = kBlockCols;  // block width = one wavefront
constexpr int kTY = kBlockRows;  // wavefronts (rows) per block
constexpr int kXcds = 8;

// image columns a block covers: one per lane, or every other column in a
// red-black pass (the lanes take the pass's colour only)
__host__ __device__ __forceinline__ int block_cols(int color) {
  return color >= 0 ? 2 * kTX : kTX;
}

// the blocks of a launch: every block column, the launch's block rows
__host__ __forceinline__ dim3 grid_for(const Geometry& g) {
  const int nby = g.nby > 0 ? g.nby : (g.H + kTY - 1) / kTY;
  const int cols = block_cols(g.color);
  return dim3((unsigned)((g.W + cols - 1) / cols), (unsigned)nby);
}

// LDS bytes of the block kernels: the own-view tile (colour and sign
// bits), followed for the colour+gradient cost by a tile of the own view's
// gradients
__host__ __forceinline__ size_t block_lds_bytes(int R, int kind,
                                                int color = -1) {
  const size_t tile = (size_t)(block_cols(color) + 2 * R) * (kTY + 2 * R) *
                      sizeof(float4);
  return kind == kCostColorGrad ? 2 * tile : tile;
}

// sign bits of a gradient as one byte per component: byte k is 0xff when
// component k has its sign bit set (the tile's .w holds the own view's)
__device__ __forceinline__ uint32_t expand_signs(uint32_t m) {
  return ((m & 1u) ? 0x000000ffu : 0u) | ((m & 2u) ? 0x0000ff00u : 0u) |
         ((m & 4u) ? 0x00ff0000u : 0u) | ((m & 8u) ? 0xff000000u : 0u);
}

// Work-group (bx, by) of the image for the hardware work-group id.
// Consecutive work-groups are dispatched round-robin over the XCDs; this
// numbering gives XCD k the k-th contiguous eighth of the blocks.
__device__ __forceinline__ void xcd_block(int* bx, int* by) {
  const int nbx = (int)gridDim.x;
  const int n = nbx * (int)gridDim.y;
  const int hw = (int)blockIdx.y * nbx + (int)blockIdx.x;
  const int xcd = hw % kXcds, j = hw / kXcds;
  const int per = n / kXcds, rem = n % kXcds;
  const int logical = xcd * per + min(xcd, rem) + j;
  *bx = logical % nbx;
  *by = logical / nbx;
}

// Everything one lane needs to evaluate planes for its pixel.
struct Lane {
  int W, H, R;
  float Wf, cw;
  int x, y;      // y is the same for the whole wavefront
  bool valid;
  float xf, yf;
  size_t p;
  float4 tc;     // own pixel colour
  float bilw;    // sum of its support weights
  const float4* trow0;  // LDS tile at this lane: tap (dx, dy) is
                        // trow0[(dy + R) * TW + dx + R]
  int TW;
  int GT;               // offset of the gradient tile (records), when
                        // staged: tap (dx, dy)'s gradient is at
                        // trow0[GT + (dy + R) * TW + dx + R]
  float kw;             // colour+gradient cost's support-weight factor
  int TS;               // texel row stride (records, with the edge pads)
  size_t GP;            // offset of the gradient plane (records)
  const float4* __restrict__ tex2;  // other view's colour record (0, 0)
};

// Block prologue shared by the cost-evaluating kernels, for the block
// (bx, by) of tile blocks of 64 x kRows pixels: stages the own-view tile
// (r, g, b, gradient sign bits), and for kGrad the tile of its gradients
// after it, and returns the lane context.  For kRB (a red-black pass,
// g.color = 0 or 1) the block covers 2 x 64 columns and lane i takes
// column 2 i + ((y + color) & 1) of its row, the pixel of the pass's
// colour.  Must be reached by every thread of the block; the block is
// 64 x kRows threads.
template <bool kGrad, bool kRB, int kRows>
__device__ Lane block_setup_at(const Geometry& g, const ViewBufs& self,
                               const ViewBufs& other, float4* smem, int bx,
                               int by) {
  const int R = g.R, W = g.W, H = g.H;
  const int cols = kRB ? 2 * kTX : kTX;
  const int TW = cols + 2 * R, TH = kRows + 2 * R;
  const int x0 = bx * cols, y0 = by * kRows;
  const int tid = threadIdx.y * kTX + threadIdx.x;
  const size_t gplane = (size_t)g.TS * H;
  for (int i = tid; i < TW * TH; i += kTX * kRows) {
    const int gx = x0 - R + i % TW, gy = y0 - R + i / TW;
    float4 v = mk4(0.0f, 0.0f, 0.0f, 0.0f), gv = v;
    if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
      const size_t rec = (size_t)gy * g.TS + kTexPadL + gx;
      const float4 a = self.tex[rec];
      v = mk4(a.x, a.y, a.z,
              __uint_as_float(expand_signs(self.sgn[(size_t)gy * W + gx])));
      if (kGrad) gv = self.tex[gplane + rec];
    }
    smem[i] = v;
    if (kGrad) smem[TW * TH + i] = gv;
  }
  __syncthreads();
  // the row index is uniform per wavefront: keep it scalar
  const int ly = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const int lx = kRB ? 2 * (int)threadIdx.x + ((y0 + ly + g.color) & 1)
                     : (int)threadIdx.x;
  Lane L;
  L.W = W;
  L.H = H;
  L.R = R;
  L.Wf = (float)W;
  L.cw = g.color_weight;
  L.x = x0 + lx;
  L.y = y0 + ly;
  L.valid = L.x < W && L.y < H;
  L.xf = (float)L.x;
  L.yf = (float)L.y;
  L.p = L.valid ? (size_t)L.y * W + L.x : 0;
  L.TW = TW;
  L.GT = TW * TH;
  L.kw = g.cost_weight;
  L.TS = g.TS;
  L.GP = (size_t)g.TS * g.H;
  L.trow0 = smem + ly * TW + lx;
  L.tex2 = other.tex + kTexPadL;
  L.tc = L.trow0[R * TW + R];
  L.bilw = L.valid
               ? self.tex[(size_t)L.y * g.TS + kTexPadL + L.x].w
               : 1.0f;
  return L;
}

// The same for the single-image launches: 64 x kTY blocks, the block from
// the hardware work-group id in XCD order, offset by the launch's first
// block row.
template <bool kGrad, bool kRB = false>
__device__ Lane block_setup(const Geometry& g, const ViewBufs& self,
                            const ViewBufs& other, float4* smem) {
  int bx, by;
  xcd_block(&bx, &by);
  return block_setup_at<kGrad, kRB, kTY>(g, self, other, smem, bx,
                                         by + g.by0);
}

// Linear sample (CLK_FILTER_LINEAR, CLK_ADDRESS_CLAMP_TO_EDGE,
// pmhstereo.cl:71) of the other view at column u of a texel row:
// t0 + a*(t1 - t0), from adjacent records of the colour and the gradient
// plane.  The edge-padded rows stand in for the coordinate clamp: every
// column the kernels sample at lies in [-1, W], and the records it reads
// there are the image's or copies of its edge texels.
struct Sample {
  float r, g, b;
  float4 grad;
};

struct Texel {
  float4 A, G;  // (r, g, b, bilw), gradient
};

__device__ __forceinline__ Sample lerp_texels(const Texel& t0,
                                              const Texel& t1, float a) {
  Sample s;
  s.r = fmaf(a, t1.A.x - t0.A.x, t0.A.x);
  s.g = fmaf(a, t1.A.y - t0.A.y, t0.A.y);
  s.b = fmaf(a, t1.A.z - t0.A.z, t0.A.z);
  s.grad.x = fmaf(a, t1.G.x - t0.G.x, t0.G.x);
  s.grad.y = fmaf(a, t1.G.y - t0.G.y, t0.G.y);
  s.grad.z = fmaf(a, t1.G.z - t0.G.z, t0.G.z);
  s.grad.w = fmaf(a, t1.G.w - t0.G.w, t0.G.w);
  return s;
}

// texel at record p of the colour plane and p + gp of the gradient plane
__device__ __forceinline__ Texel texel_at(const float4* __restrict__ p,
                                          size_t gp) {
  Texel t;
  t.A = p[0];
  t.G = p[gp];
  return t;
}

// 8 x the gradient sign mismatches between expanded sign bytes m1x and
// the signs of g.  v_perm_b32 selectors 9 and 11 replicate the sign bit of
// the low and the high source operand into a whole byte, so the signs of
// g arrive in the layout of m1x; xor and count.
__device__ __forceinline__ int sign_mismatches8(uint32_t m1x, float4 g) {
  const uint32_t lo = __builtin_amdgcn_perm(__float_as_uint(g.y),
                                            __float_as_uint(g.x), 0x0c0c0b09u);
  const uint32_t hi = __builtin_amdgcn_perm(__float_as_uint(g.w),
                                            __float_as_uint(g.z), 0x0b090c0cu);
  return __popc((lo | hi) ^ m1x);
}

// ---------------------------------------------------------------------
// patch_eval_census (pmhstereo.cl:661-736), one plane per lane (lanes with
// active == false idle), patch rows wave-uniform.  The reference tests
// census/bilw >= mincost after every tap; the census terms are >= 0 and
// accumulate monotonically, so testing at the end of each row returns 1.0
// exactly when the per-tap test would.
// kCount: count the taps the reference's per-tap test evaluates (the
// algorithmic work; used by an untimed counting pass only).
// ---------------------------------------------------------------------
template <bool kCount>
__device__ float cost_eval(const Lane& L, float4 s, float mincost,
                           bool active, uint32_t* ntaps) {
  const int R = L.R, W = L.W;
  bool live = active && L.valid;
  float res = 0.0f;
  if (live) {
    const float xp = transform(s, L.xf, L.yf);
    if (!(finite_f(s.x) && finite_f(s.y) && finite_f(s.z)) ||
        to_normal(s).z < kNormalZMin || xp < 0.0f || xp >= L.Wf) {
      res = 1e5f;
      live = false;
    }
  }
  const int dx0 = max(-R, -L.x), dx1 = min(R, W - 1 - L.x);
  float census = 0.0f, wsum = 0.0f;
  for (int dy = -R; dy <= R; ++dy) {
    if (!__any(live)) break;
    const int qy = L.y + dy;  // uniform
    if (qy < 0 || qy >= L.H) continue;
    if (!live) continue;
    const float rowb = s.y * (float)qy;  // transform(), b y per row
    const float4* __restrict__ srow = L.tex2 + (size_t)qy * L.TS;
    const float4* trow = L.trow0 + (dy + R) * L.TW + R;
    // tap column as a float, stepped by 1 (exact below 2^24)
    float qx = L.xf + (float)dx0;
    for (int dx = dx0; dx <= dx1; ++dx, qx += 1.0f) {
      const float xq = fmaf(s.x, qx, rowb) + s.z;
      if (xq < 0.0f || xq >= L.Wf) continue;
      const float4 t1 = trow[dx];
      const uint32_t m1 = __float_as_uint(t1.w);
      const float w = support_weight(t1.x - L.tc.x, t1.y - L.tc.y,
                                     t1.z - L.tc.z, L.cw);
      // 0 <= xq < W: no coordinate clamp; the second texel of xq in
      // [W-1, W) is the right edge pad, a copy of texel W-1
      const float fq = floorf(xq);
      const float4* __restrict__ tq = srow + (int)fq;
      const Sample s2 =
          lerp_texels(texel_at(tq, L.GP), texel_at(tq + 1, L.GP), xq - fq);
      const float c0 = fminf(fabsf(s2.r - t1.x), 0.05f) * 20.0f;
      const float c1 = fminf(fabsf(s2.g - t1.y), 0.05f) * 20.0f;
      const float c2 = fminf(fabsf(s2.b - t1.z), 0.05f) * 20.0f;
      const float cad = 0.33f * ((c0 + c1) + c2);
      // cost = 0.125 cad, then cost += 0.875 * (0.25 n): both products
      // are exact, so one fma with 0.21875 n = (7/256) (8 n) rounds the
      // same sum
      const float cost = fmaf(0.02734375f,
                              (float)sign_mismatches8(m1, s2.grad),
                              0.125f * cad);
      census = fmaf(w, cost, census);
      wsum += w;
      if (kCount) {
        ++*ntaps;
        if (census / L.bilw >= mincost) break;
      }
    }
    if (census / L.bilw >= mincost) {
      res = 1.0f;
      live = false;
    }
  }
  if (live) res = (wsum <= 1.0f) ? 1.1f : census / wsum;
  return res;
}

// patch_eval_color_grad_weighted (pmhstereo.cl:510-565), one plane per lane
// (lanes with active == false idle), patch rows wave-uniform:
// w = exp(-|dot(I1(p) - I1(q), gray)| * kw), C = 0.05 sum_3 min(|I1(q) -
// I2(x')|, 0.05) + 0.95 sum_4 min(|G1(q) - G2(x')|, 0.05), result
// sum(w C) / sum(w).  No early exit and no invalid-plane test: a tap counts
// when its own pixel is in the image and x' lies in [0, W) (a non-finite x'
// fails the test; the reference samples at it, which OpenCL leaves
// undefined).  No qualifying tap gives 0 / 0 = NaN, which the accept rule
// never takes.  Needs the gradient tile (block_setup<true>).
// kCount: count the taps evaluated.
template <bool kCount>
__device__ float cost_eval_cg(const Lane& L, float4 s, bool active,
                              uint32_t* ntaps) {
  const int R = L.R, W = L.W;
  const bool live = active && L.valid;
  const int dx0 = max(-R, -L.x), dx1 = min(R, W - 1 - L.x);
  float ret = 0.0f, wsum = 0.0f;
  for (int dy = -R; dy <= R; ++dy) {
    if (!__any(live)) break;
    const int qy = L.y + dy;  // uniform
    if (qy < 0 || qy >= L.H) continue;
    if (!live) continue;
    const float rowb = s.y * (float)qy;  // transform(), b y per row
    const float4* __restrict__ srow = L.tex2 + (size_t)qy * L.TS;
    const float4* trow = L.trow0 + (dy + R) * L.TW + R;
    const float4* grow = trow + L.GT;
    float qx = L.xf + (float)dx0;
    for (int dx = dx0; dx <= dx1; ++dx, qx += 1.0f) {
      const float xq = fmaf(s.x, qx, rowb) + s.z;
      if (!(xq >= 0.0f && xq < L.Wf)) continue;
      const float4 t1 = trow[dx];
      const float4 g1 = grow[dx];
      // dot(valcenter - val1, gray): the negated difference gives the
      // negated dot bit for bit, so |.| is the same
      const float w = support_weight(t1.x - L.tc.x, t1.y - L.tc.y,
                                     t1.z - L.tc.z, L.kw);
      const float fq = floorf(xq);
      const float4* __restrict__ tq = srow + (int)fq;
      const Sample s2 =
          lerp_texels(texel_at(tq, L.GP), texel_at(tq + 1, L.GP), xq - fq);
      const float sc = (fminf(fabsf(t1.x - s2.r), 0.05f) +
                        fminf(fabsf(t1.y - s2.g), 0.05f)) +
                       fminf(fabsf(t1.z - s2.b), 0.05f);
      const float sg = ((fminf(fabsf(g1.x - s2.grad.x), 0.05f) +
                         fminf(fabsf(g1.y - s2.grad.y), 0.05f)) +
                        fminf(fabsf(g1.z - s2.grad.z), 0.05f)) +
                       fminf(fabsf(g1.w - s2.grad.w), 0.05f);
      // alpha dot(.) + (1 - alpha) dot(.), the left product fused
      const float C = fmaf(0.05f, sc, (1.0f - 0.05f) * sg);
      wsum += w;
      ret = fmaf(w, C, ret);
      if (kCount) ++*ntaps;
    }
  }
  return ret / wsum;
}

// the candidate cost the kernels' patch_eval selects
template <int kKind, bool kCount>
__device__ __forceinline__ float cost_any(const Lane& L, float4 s,
                                          float mincost, bool active,
                                          uint32_t* ntaps) {
  if (kKind == kCostColorGrad) return cost_eval_cg<kCount>(L, s, active, ntaps);
  return cost_eval<kCount>(L, s, mincost, active, ntaps);
}

// one census term of patch_census_klt2 (pmhstereo.cl:360-392); v0..v2 is
// the centred colour I1(q) - I1(p) the reference compares against;
// k1 = (w1 * alpha) * 0.33, k2 = w1 * (1 - alpha)
__device__ __forceinline__ float klt_term(const Sample& s, float v0, float v1,
                                          float v2, uint32_t ga, float k1,
                                          float k2) {
  const float c0 = fminf(fabsf(s.r - v0), 0.05f) * 20.0f;
  const float c1 = fminf(fabsf(s.g - v1), 0.05f) * 20.0f;
  const float c2 = fminf(fabsf(s.b - v2), 0.05f) * 20.0f;
  const float C = k1 * ((c0 + c1) + c2);
  // C += w1 * (1 - alpha) * 0.25 * n, 0.25 a double literal
  // (pmhstereo.cl:370): 0.25 n = 0.03125 (8 n) is exact, one rounding
  return fmaf(k2, 0.03125f * (float)sign_mismatches8(ga, s.grad), C);
}

// patch_census_klt2 (pmhstereo.cl:326-410) + mat3f_lusolve
// (cvt/cl/kernel/Matrix3.cl:286-311), one plane per lane
__device__ float4 klt_pass(const Lane& L, float4 s) {
  const int R = L.R, W = L.W;
  float J0 = 0.0f, J1 = 0.0f, J2 = 0.0f;
  float H00 = 0.0f, H01 = 0.0f, H02 = 0.0f, H11 = 0.0f, H12 = 0.0f,
        H22 = 0.0f;
  // taps need 1 <= x+dx < W-1 (the test is on the tap, pmhstereo.cl:358)
  const int dx0 = max(-R, 1 - L.x), dx1 = min(R, W - 2 - L.x);
  if (L.valid) {
    for (int dy = -R; dy <= R; ++dy) {
      const int qy = L.y + dy;  // uniform
      if (qy < 0 || qy >= L.H) continue;
      const float py = (float)qy;
      const float rowb = s.y * py;  // transform(), b y per row
      const float4* __restrict__ srow = L.tex2 + (size_t)qy * L.TS;
      const float4* trow = L.trow0 + (dy + R) * L.TW + R;
      // texel window srow[kf - 1 .. kf + 2] of the previous tap; a slope
      // near 1 slides it by one texel per tap
      int kf = INT_MIN;
      Texel tw0, tw1, tw2, tw3;
      float px = L.xf + (float)dx0;  // stepped by 1, exact below 2^24
      for (int dx = dx0; dx <= dx1; ++dx, px += 1.0f) {
        const float4 t1 = trow[dx];
        const float v0 = t1.x - L.tc.x, v1 = t1.y - L.tc.y,
                    v2 = t1.z - L.tc.z;
        const uint32_t ga = __float_as_uint(t1.w);
        const float w1 = support_weight(v0, v1, v2, L.cw);
        const float k1 = (w1 * 0.125f) * 0.33f, k2 = w1 * 0.875f;
        const float xq = fmaf(s.x, px, rowb) + s.z;
        // The samples at x'-1, x' and x'+1 read the four texels from
        // floor(xc)-1 on, xc = x' clamped to [-1, W] (a NaN x' clamps to W):
        //  - x'-1 is exact for x' >= 0.5 and shares the floor and fraction
        //    of x'; below that it and its neighbour both read texel 0 (the
        //    left pads), whatever the fraction;
        //  - past either edge every texel involved is a copy of the edge
        //    texel, which is what the clamped sampler returns;
        //  - x'+1 is rounded and has its own fraction; its floor is
        //    floor(xc)+1, or floor(xc)+2 when the sum rounded up to the
        //    next integer (xc < W, so texel floor(xc)+3 is at most the
        //    last right pad).
        const float xc = fmaxf(fminf(xq, L.Wf), -1.0f);
        const float xp1 = xc + 1.0f;
        const float fl = floorf(xc);
        const float flp = floorf(xp1);
        const float a = xc - fl;
        const float ap = xp1 - flp;
        const int f = (int)fl;  // -1 <= f <= W
        if (f != kf) {
          if (f == kf + 1) {
            tw0 = tw1;
            tw1 = tw2;
            tw2 = tw3;
          } else {
            tw0 = texel_at(srow + (f - 1), L.GP);
            tw1 = texel_at(srow + f, L.GP);
            tw2 = texel_at(srow + (f + 1), L.GP);
          }
          tw3 = texel_at(srow + (f + 2), L.GP);
          kf = f;
        }
        const Sample sm = lerp_texels(tw0, tw1, a);
        const Sample s0 = lerp_texels(tw1, tw2, a);
        Sample sp = lerp_texels(tw2, tw3, ap);
        if (flp != fl + 1.0f)
          sp = lerp_texels(tw3, texel_at(srow + (f + 3), L.GP), ap);
        const float C = klt_term(s0, v0, v1, v2, ga, k1, k2);
        const float Cxp = klt_term(sp, v0, v1, v2, ga, k1, k2);
        const float Cxm = klt_term(sm, v0, v1, v2, ga, k1, k2);
        // (Cxp + Cxm) - 2 C: the product is exact, so the fma rounds the
        // same difference
        const float curv = fmaf(-2.0f, C, Cxp + Cxm);
        if (curv > 0.0f) {
          // J = J + w1 (Cxm - Cxp) 0.5 p, the last product fused
          // (pmhstereo.cl:394)
          const float gj = (w1 * (Cxm - Cxp)) * 0.5f;
          J0 = fmaf(gj, px, J0);
          J1 = fmaf(gj, py, J1);
          J2 = J2 + gj;
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
  }
  // LU without pivoting; H is symmetric (H10 = H01, ...)
  float m00 = H00, m01 = H01, m02 = H02;
  float m10 = H01, m11 = H11, m12 = H12;
  float m20 = H02, m21 = H12, m22 = H22;
  m10 /= m00;
  m20 /= m00;
  // x -= a * b contracts to fma(-a, b, x); the built-in dot of two
  // components is fma(y0, y1, x0 * x1)
  m11 = fmaf(-m10, m01, m11);
  m12 = fmaf(-m10, m02, m12);
  m21 = fmaf(-m20, m01, m21);
  m22 = fmaf(-m20, m02, m22);
  m21 /= m11;
  m22 = fmaf(-m21, m12, m22);
  float b0 = J0, b1 = J1, b2 = J2;
  b1 = fmaf(-m10, b0, b1);
  b2 -= fmaf(m21, b1, m20 * b0);
  const float r2 = b2 / m22;
  const float r1 = fmaf(-m12, r2, b1) / m11;
  const float r0 = (b0 - fmaf(m02, r2, m01 * r1)) / m00;
  return mk4(s.x - r0, s.y - r1, s.z - r2, s.w - 1e5f);
}

// ---------------------------------------------------------------------
// precompute: pmhstereo_bilateral_weight_to_alpha (pmhstereo.cl:1462-1498)
// and pmhstereo_gradxy (pmhstereo.cl:1549-1602) into one texel record,
// plus the gradient sign bits
// ---------------------------------------------------------------------
constexpr int kBX = 64, kBY = 4;  // simple per-pixel kernels

__device__ __forceinline__ void prepare_pixel(const float* __restrict__ img,
                                              int W, int H, int C, int stride,
                                              int R, float cw,
                                              float4* __restrict__ tex,
                                              uint8_t* __restrict__ sgn,
                                              int x, int y) {
  const size_t p = (size_t)y * W + x;
  const float4 c = fetch_rgba(img, C, stride, x, y);
  float wsum = 0.0f;
  for (int dy = -R; dy <= R; ++dy) {
    const int qy = y + dy;
    if (qy < 0 || qy >= H) continue;
    for (int dx = -R; dx <= R; ++dx) {
      const int qx = x + dx;
      if (qx < 0 || qx >= W) continue;
      const float4 q = fetch_rgba(img, C, stride, qx, qy);
      wsum += support_weight(q.x - c.x, q.y - c.y, q.z - c.z, cw);
    }
  }
  const float g = gray_at(img, W, H, C, stride, x, y);
  const float gxp = gray_at(img, W, H, C, stride, x + 1, y);
  const float gxm = gray_at(img, W, H, C, stride, x - 1, y);
  const float gyp = gray_at(img, W, H, C, stride, x, y + 1);
  const float gym = gray_at(img, W, H, C, stride, x, y - 1);
  const float4 grad = mk4(gxp - g, gyp - g, gxp - gxm, gyp - gym);
  const float4 col = mk4(c.x, c.y, c.z, wsum);
  const size_t gp = (size_t)tex_stride(W) * H;  // gradient plane offset
  float4* row = tex + (size_t)y * tex_stride(W) + kTexPadL;
  row[x] = col;
  row[gp + x] = grad;
  // edge pads: copies of the first and last texel of the row
  if (x == 0)
    for (int i = 1; i <= kTexPadL; ++i) {
      row[-i] = col;
      row[gp - i] = grad;
    }
  if (x == W - 1)
    for (int i = 1; i <= kTexPadR; ++i) {
      row[x + i] = col;
      row[gp + x + i] = grad;
    }
  sgn[p] = (uint8_t)signmask4(grad);
}

__global__ void __launch_bounds__(256)
k_prepare(const float* __restrict__ img, int W, int H, int C, int stride,
          int R, float cw, float4* __restrict__ tex,
          uint8_t* __restrict__ sgn) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  prepare_pixel(img, W, H, C, stride, R, cw, tex, sgn, x, y);
}

// pmhstereo_init (pmhstereo.cl:755-773): RNG_init(p, 3), three draws; one
// tile block of 64 x kRows pixels, reached by every thread of the block
template <int kKind, int kRows>
__device__ __forceinline__ void init_block(const Geometry& g, ViewBufs self,
                                           ViewBufs other, float dmax, int lr,
                                           float4* __restrict__ states,
                                           float4* smem, int bx, int by) {
  const Lane L = block_setup_at<kKind == kCostColorGrad, false, kRows>(
      g, self, other, smem, bx, by);
  float4 s = mk4(0.0f, 0.0f, 0.0f, 0.0f);
  if (L.valid) {
    Mwc rng = mwc_jump((uint64_t)L.p * 3u);
    const float u0 = mwc_next(&rng), u1 = mwc_next(&rng), u2 = mwc_next(&rng);
    s = state_init(u0, u1, u2, L.xf, L.yf, lr, dmax);
  }
  const float w = cost_any<kKind, false>(L, s, 1e10f, true, nullptr);
  if (L.valid) {
    s.w = w;
    states[L.p] = s;
  }
}

template <int kKind>
__global__ void __launch_bounds__(256)
k_init(Geometry g, ViewBufs self, ViewBufs other, float dmax, int lr,
       float4* __restrict__ states) {
  extern __shared__ float4 smem[];
  int bx, by;
  xcd_block(&bx, &by);
  init_block<kKind, kTY>(g, self, other, dmax, lr, states, smem, bx,
                         by + g.by0);
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
template <int kKind>
__global__ void __launch_bounds__(256)
k_init_prior(Geometry g, ViewBufs self, ViewBufs other, float dmax, int lr,
             const float* __restrict__ disp,
             const float4* __restrict__ normals, float4* __restrict__ states) {
  extern __shared__ float4 smem[];
  const Lane L = block_setup<kKind == kCostColorGrad>(g, self, other, smem);
  float4 s = mk4(0.0f, 0.0f, 0.0f, 0.0f);
  if (L.valid) {
    Mwc rng = mwc_jump((uint64_t)L.p * 3u);
    const float d = disp[L.p];
    if (d > 0.0f) {
      const float4 nn = normals[L.p];
      const float3 n = normalize3(make_float3(nn.x, nn.y, nn.z));
      s = from_disparity_normal(L.xf, L.yf, lr ? d : -d, n);
    } else {
      const float u0 = mwc_next(&rng), u1 = mwc_next(&rng),
                  u2 = mwc_next(&rng);
      s = state_init(u0, u1, u2, L.xf, L.yf, lr, dmax);
    }
  }
  const float w = cost_any<kKind, false>(L, s, 1e10f, true, nullptr);
  if (L.valid) {
    s.w = w;
    states[L.p] = s;
  }
}

// Init of a view from a plane proposal, with a per-pixel fallback: the
// proposed plane (disparity d = o.x * dscale and normal (o.y, o.z, o.w) at
// the pixel, from_disparity_normal as pmhstereo_init_disparity_normal
// builds it, pmhstereo.cl:802-828) competes with the random plane of
// pmhstereo_init (draws 3p..3p+2, pmhstereo.cl:755-773), and the random
// plane takes the pixel only when its cost is lower (a NaN cost never
// wins).  The right view's proposal (lr = 0) comes from the network's pass
// over the mirrored pair: it is read at column W - 1 - x, and its normal
// maps to (nx, -ny, nz) (the mirror flips the sign of the disparity's x
// slope, the right view's plane convention flips both slopes).  A
// non-finite or non-positive proposed disparity leaves the random plane.
template <int kKind>
__global__ void __launch_bounds__(256)
k_init_proposal(Geometry g, ViewBufs self, ViewBufs other, float dmax, int lr,
                const float* __restrict__ prop, int os, int Wp, float dscale,
                float4* __restrict__ states) {
  extern __shared__ float4 smem[];
  const Lane L = block_setup<kKind == kCostColorGrad>(g, self, other, smem);
  float4 q = mk4(0.0f, 0.0f, 0.0f, 0.0f), sp = q;
  bool have = false;
  if (L.valid) {
    Mwc rng = mwc_jump((uint64_t)L.p * 3u);
    const float u0 = mwc_next(&rng), u1 = mwc_next(&rng), u2 = mwc_next(&rng);
    q = state_init(u0, u1, u2, L.xf, L.yf, lr, dmax);
    const int xs = lr ? L.x : L.W - 1 - L.x;
    const float* o = prop + ((size_t)L.y * Wp + xs) * os;
    const float d = o[0] * dscale;
    float3 n = make_float3(o[1], lr ? o[2] : -o[2], o[3]);
    if (d > 0.0f && finite_f(d) && finite_f(n.x) && finite_f(n.y) &&
        finite_f(n.z)) {
      n = normalize3(n);
      sp = from_disparity_normal(L.xf, L.yf, lr ? d : -d, n);
      have = true;
    }
  }
  const float cq = cost_any<kKind, false>(L, q, 1e10f, true, nullptr);
  const float cp = cost_any<kKind, false>(L, sp, 1e10f, have, nullptr);
  if (L.valid) {
    q.w = cq;
    sp.w = cp;
    const bool keep = have && (cp <= cq || cq != cq);
    states[L.p] = keep ? sp : q;
  }
}

// The right view's start when only the left view is seeded
// (pm_setup_proposal_views, views = 1): at every pixel the nearest left
// surface the view scatter put there (slot key 0, the propagate's first
// view candidate, pmhstereo.cl:1007-1016) carried over by viewprop, else
// the random plane of pmhstereo_init (draws 3p..3p+2, pmhstereo.cl:
// 755-773).  No cost is evaluated: the cost channel is NaN, and every
// propagate re-costs its own planes before it uses them
// (pmhstereo.cl:940).
__global__ void __launch_bounds__(256)
k_init_from_view(int W, int H, float dmax,
                 const unsigned long long* __restrict__ slots,
                 const float4* __restrict__ src, float4* __restrict__ states) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  const size_t p = (size_t)y * W + x;
  const uint32_t k0 = (uint32_t)(slots[p] & 0xFFFFFFFFull);
  float4 s;
  if (k0 != kSlotEmpty) {
    // the left view's key is W - 1 - x (slot_insert's order: nearest first)
    s = viewprop(src[(size_t)y * W + (W - 1 - (int)k0)]);
  } else {
    Mwc rng = mwc_jump((uint64_t)p * 3u);
    const float u0 = mwc_next(&rng), u1 = mwc_next(&rng), u2 = mwc_next(&rng);
    s = state_init(u0, u1, u2, (float)x, (float)y, 0, dmax);
  }
  s.w = __int_as_float(0x7fc00000);
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
  float dmax, theta;
  int lr;
  float4 smooth;
};

// candidate test with minimalCost (pmhstereo.cl:873-877) and the select
// rule (pmhstereo.cl:976); called by the whole wavefront
template <int kKind, bool kCount>
__device__ __forceinline__ void try_candidate(const Lane& L, const PropCtx& c,
                                              float4& self, float4 cand,
                                              bool active, uint32_t* ntaps) {
  active = active && L.valid;
  float ds = 0.0f, dc = 0.0f, mincost = 0.0f;
  if (active) {
    ds = smooth_dist(self, c.smooth, L.xf, L.yf, c.dmax, c.lr);
    dc = smooth_dist(cand, c.smooth, L.xf, L.yf, c.dmax, c.lr);
    mincost = fmaf(-c.theta, dc, fmaf(c.theta, ds, self.w));
  }
  const float w = cost_any<kKind, kCount>(L, cand, mincost, active, ntaps);
  if (active) {
    cand.w = w;
    if (fmaf(c.theta, dc, cand.w) <= fmaf(c.theta, ds, self.w)) self = cand;
  }
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

// One tile block of pmhstereo_propagate_view: the pixels of the block
// (bx, by) of 64 x kRows tiles (kRB: of 128 x kRows, one colour).  Must be
// reached by every thread of the block.
// kCount: accumulate the cost taps (counters[0]) and KLT taps
// (counters[1]) this block evaluates; the results are identical.
// kRB: one pass of the red-black schedule (g.color); pass 1 reads the
// neighbours of colour 0 from a.fresh, the planes pass 0 just wrote.
template <int kKind, bool kCount, bool kRB, int kRows>
__device__ __forceinline__ void propagate_block(
    const Geometry& g, const PropagateArgs& a, unsigned long long* counters,
    float4* smem, int bx, int by) {
  const Lane L = block_setup_at<kKind == kCostColorGrad, kRB, kRows>(
      g, a.self, a.other, smem, bx, by);
  const int W = g.W, H = g.H;
  uint32_t ncost = 0;
  uint32_t* nc = kCount ? &ncost : nullptr;
  PropCtx c;
  c.dmax = a.dmax;
  c.theta = a.theta;
  c.lr = a.lr;
  float4 sm = mk4(0.0f, 0.0f, 0.0f, 0.0f);
  float4 self = sm;
  if (L.valid) {
    if (a.smooth) sm = a.smooth[L.p];
    self = a.old[L.p];
  }
  sm.w = fminf(fmaxf(sm.w, 0.0f), 1.0f) * a.dmax;
  c.smooth = sm;

  {
    const float w = cost_any<kKind, kCount>(L, self, 1e10f, true, nc);
    if (L.valid) self.w = w;
  }
  Mwc rng = mwc_jump(prop_draw_start(L.p, a.iter));

  // smooth proposal (pmhstereo.cl:968-976)
  {
    float4 cand = from_disparity_normal(
        L.xf, L.yf, sm.w * (a.lr ? -1.0f : 1.0f),
        normalize3(make_float3(sm.x, sm.y, sm.z)));
    if (!a.lr) cand = viewprop(cand);
    try_candidate<kKind, kCount>(L, c, self, cand, true, nc);
  }
  // random neighbours (pmhstereo.cl:981-991), x offset drawn first
  for (int i = 0; i < kNumRndSample; ++i) {
    const float ux = mwc_next(&rng);
    const float uy = mwc_next(&rng);
    const int ox = (int)fmaf((ux - 0.5f) * 2.0f, kRndSampleRadius, 0.5f);
    const int oy = (int)fmaf((uy - 0.5f) * 2.0f, kRndSampleRadius, 0.5f);
    float4 cand = mk4(0.0f, 0.0f, 0.0f, 0.0f);
    if (L.valid) {
      const int nx = clampi(L.x + ox, 0, W - 1), ny = clampi(L.y + oy, 0, H - 1);
      const float4* src = a.old;
      if (kRB && g.color == 1 && ((nx + ny) & 1) == 0) src = a.fresh;
      cand = src[(size_t)ny * W + nx];
    }
    try_candidate<kKind, kCount>(L, c, self, cand, true, nc);
  }
  // random plane (pmhstereo.cl:996-1002)
  {
    const float u0 = mwc_next(&rng), u1 = mwc_next(&rng), u2 = mwc_next(&rng);
    try_candidate<kKind, kCount>(L, c, self,
                          state_init(u0, u1, u2, L.xf, L.yf, a.lr, a.dmax),
                          true, nc);
  }
  // other view (pmhstereo.cl:1007-1016)
  {
    uint32_t keys[2] = {kSlotEmpty, kSlotEmpty};
    if (a.viewin && L.valid) {
      const unsigned long long slot = a.viewin[L.p];
      keys[0] = (uint32_t)(slot & 0xFFFFFFFFull);
      keys[1] = (uint32_t)(slot >> 32);
    }
    for (int i = 0; i < kViewSamples; ++i) {
      const bool have = keys[0] != kSlotEmpty && keys[i] != kSlotEmpty;
      float4 cand = mk4(0.0f, 0.0f, 0.0f, 0.0f);
      if (have) {
        const int xs = a.lr ? (int)keys[i] : W - 1 - (int)keys[i];
        cand = viewprop(a.src_other[(size_t)L.y * W + xs]);
      }
      try_candidate<kKind, kCount>(L, c, self, cand, have, nc);
    }
  }
  // randomized refinement (pmhstereo.cl:1021-1028)
  {
    const float u0 = mwc_next(&rng), u1 = mwc_next(&rng), u2 = mwc_next(&rng);
    try_candidate<kKind, kCount>(
        L, c, self, state_refine(u0, u1, u2, self, L.xf, L.yf, a.dmax, a.lr),
        true, nc);
  }
  // KLT step (pmhstereo.cl:1034-1053)
  {
    float4 k = self;
    for (int it = 0; it < kKltPasses; ++it) k = klt_pass(L, k);
    try_candidate<kKind, kCount>(L, c, self, k, true, nc);
  }
  if (!L.valid) return;
  if (a.viewout) scatter(self, L.x, L.y, W, a.lr, a.viewout);
  a.out[L.p] = self;
  if (kCount) {
    // KLT taps: rows in the image times columns with 1 <= x+dx < W-1
    const int R = g.R;
    const int rows = min(R, H - 1 - L.y) - max(-R, -L.y) + 1;
    const int cols = min(R, W - 2 - L.x) - max(-R, 1 - L.x) + 1;
    const unsigned long long nklt =
        cols > 0 ? (unsigned long long)rows * cols * kKltPasses : 0ull;
    atomicAdd(&counters[0], (unsigned long long)ncost);
    atomicAdd(&counters[1], nklt);
  }
}

// The single-image launch: one tile block per work-group.  Built for five
// wavefronts per SIMD: 96 VGPRs, with a few values spilled to scratch.
// Measured on the 720p pair against the unconstrained build (107 VGPRs,
// four wavefronts): 8 iterations 52.1 -> 51.8 ms; six wavefronts (80
// VGPRs, more spills) gain nothing.
template <int kKind, bool kCount, bool kRB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 5)))
k_propagate(Geometry g, PropagateArgs a, unsigned long long* counters) {
  extern __shared__ float4 smem[];
  int bx, by;
  xcd_block(&bx, &by);
  propagate_block<kKind, kCount, kRB, kTY>(g, a, counters, smem, bx,
                                           by + g.by0);
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

template <int kKind>
__global__ void __launch_bounds__(256)
k_cost_field(Geometry g, ViewBufs self, ViewBufs other,
             const float4* __restrict__ planes, float* __restrict__ cost) {
  extern __shared__ float4 smem[];
  const Lane L = block_setup<kKind == kCostColorGrad>(g, self, other, smem);
  const float4 s = L.valid ? planes[L.p] : mk4(0.0f, 0.0f, 0.0f, 0.0f);
  const float w = cost_any<kKind, false>(L, s, 1e10f, true, nullptr);
  if (L.valid) cost[L.p] = w;
}

__global__ void __launch_bounds__(256)
k_klt_field(Geometry g, ViewBufs self, ViewBufs other,
            const float4* __restrict__ planes, int passes,
            float4* __restrict__ out) {
  extern __shared__ float4 smem[];
  const Lane L = block_setup<false>(g, self, other, smem);
  float4 s = L.valid ? planes[L.p] : mk4(0.0f, 0.0f, 0.0f, 0.0f);
  for (int i = 0; i < passes; ++i) s = klt_pass(L, s);
  if (L.valid) out[L.p] = s;
}

// pmhstereo_consistency (pmhstereo.cl:1095-1129) at pixel (x, y)
__device__ __forceinline__ void consistency_pixel(
    int W, const float4* __restrict__ left, const float4* __restrict__ right,
    float maxdispdiff, float maxdeg, float4* __restrict__ out, int x, int y) {
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
    const float nd1 = fmaf(nl.z, n1.z, fmaf(nl.y, n1.y, nl.x * n1.x));
    const float nd2 = fmaf(nl.z, n2.z, fmaf(nl.y, n2.y, nl.x * n2.x));
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

__global__ void __launch_bounds__(256)
k_consistency(int W, int H, const float4* __restrict__ left,
              const float4* __restrict__ right, float maxdispdiff,
              float maxdeg, float4* __restrict__ out) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  consistency_pixel(W, left, right, maxdispdiff, maxdeg, out, x, y);
}

// pmhstereo_to_disparity (pmhstereo.cl:1232-1257) at pixel (x, y)
__device__ __forceinline__ void to_disparity_pixel(
    int W, const float4* __restrict__ states, float scale, int lr,
    float* __restrict__ disp, int x, int y) {
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

__global__ void __launch_bounds__(256)
k_to_disparity(int W, int H, const float4* __restrict__ states, float scale,
               int lr, float* __restrict__ disp) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  to_disparity_pixel(W, states, scale, lr, disp, x, y);
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

// pmhstereo_fill_depthmap (pmhstereo.cl:1187-1230) at pixel (x, y)
__device__ __forceinline__ void fill_depthmap_pixel(
    int W, const float4* __restrict__ states, float scale,
    float* __restrict__ disp, int x, int y) {
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

__global__ void __launch_bounds__(256)
k_fill_depthmap(int W, int H, const float4* __restrict__ states, float scale,
                float* __restrict__ disp) {
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= W || y >= H) return;
  fill_depthmap_pixel(W, states, scale, disp, x, y);
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

// Disparity::evaluateError (cvt/vision/Disparity.cpp:201-233) on the
// device: over the pixels with gt > 0, counts[2] += 1, counts[1] += 1 when
// the estimate is > 0, counts[0] += 1 when |gt - est| > abs_thres and
// |gt - est| / |gt| > rel_thres.  A block sums its pixels' counts, one
// atomic per counter.
// Called by every lane of the wavefront (pixels outside the image count
// nothing).
__device__ __forceinline__ void eval_error_pixel(
    const float* __restrict__ disp, int sd, const float* __restrict__ gt,
    int sg, int W, int H, float abs_thres, float rel_thres,
    unsigned long long* __restrict__ counts, int x, int y) {
  unsigned e = 0, v = 0, a = 0;
  if (x < W && y < H) {
    const float g = gt[(size_t)y * sg + x];
    if (g > 0.0f) {
      const float d = disp[(size_t)y * sd + x];
      const float diff = fabsf(g - d);
      e = (diff > abs_thres && diff / fabsf(g) > rel_thres) ? 1u : 0u;
      v = d > 0.0f ? 1u : 0u;
      a = 1u;
    }
  }
  // wave sums by ballot, then one atomic per wave and counter
  const unsigned long long be = __ballot(e), bv = __ballot(v),
                           ba = __ballot(a);
  if ((threadIdx.x & 63) == 0) {
    if (be) atomicAdd(&counts[0], (unsigned long long)__popcll(be));
    if (bv) atomicAdd(&counts[1], (unsigned long long)__popcll(bv));
    if (ba) atomicAdd(&counts[2], (unsigned long long)__popcll(ba));
  }
}

__global__ void __launch_bounds__(256)
k_eval_error(const float* __restrict__ disp, int sd, const float* __restrict__ gt,
             int sg, int W, int H, float abs_thres, float rel_thres,
             unsigned long long* __restrict__ counts) {
  eval_error_pixel(disp, sd, gt, sg, W, H, abs_thres, rel_thres, counts,
                   blockIdx.x * kBX + threadIdx.x,
                   blockIdx.y * kBY + threadIdx.y);
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
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
// batches of pairs of uneven sizes (BatchPair descriptors in device
// memory).  The tile kernels number the tile blocks of all pairs in one
// sequence, pair after pair, row-major within a pair; the per-pixel
// kernels take the pair (and view) from the grid's z index.
// ---------------------------------------------------------------------

// Persistent walk over the tile blocks: the sequence is cut into kXcds
// contiguous eighths, and the work-groups of XCD k (hardware ids congruent
// to k, dispatched round-robin) take the k-th eighth with a stride of
// their number, so each XCD's L2 serves a contiguous run of rows.  Each
// work-group calls body(pair, bx, by) for its blocks in increasing order;
// every block of [0, nitems) is visited exactly once for any grid of at
// least kXcds work-groups.  The barrier before each block keeps the
// previous block's LDS tile intact until every wavefront is done with it.
template <int kRows, typename Body>
__device__ __forceinline__ void batch_walk(const BatchPair* __restrict__ pairs,
                                           int npairs, long long nitems,
                                           Body body) {
  const int G = (int)gridDim.x, w = (int)blockIdx.x;
  const int xcd = w % kXcds, j = w / kXcds;
  const int g8 = (G - xcd + kXcds - 1) / kXcds;
  const long long i0 = nitems * xcd / kXcds, i1 = nitems * (xcd + 1) / kXcds;
  int pi = 0;
  for (long long it = i0 + j; it < i1; it += g8) {
    __syncthreads();
    while (pi + 1 < npairs && pairs[pi + 1].item0 <= it) ++pi;
    const BatchPair& P = pairs[pi];
    const long long li = it - P.item0;
    body(P, (int)(li % P.nbx), (int)(li / P.nbx));
  }
}

template <int kKind, int kRows>
__global__ void __launch_bounds__(64 * kRows)
k_init_batch(const BatchPair* __restrict__ pairs, int npairs,
             long long nitems, int view) {
  extern __shared__ float4 smem[];
  batch_walk<kRows>(pairs, npairs, nitems,
                    [&](const BatchPair& P, int bx, int by) {
    init_block<kKind, kRows>(P.g, P.v[view], P.v[1 - view], P.dmax,
                             view == 0 ? 1 : 0, P.state[view][0], smem, bx,
                             by);
  });
}

// one view's propagate over all pairs: the single-image launch's
// arguments (one_iteration in pm_capi.cpp) from each pair's descriptor
template <int kKind, int kRows>
__global__ void __launch_bounds__(64 * kRows)
__attribute__((amdgpu_waves_per_eu(5, 5)))
k_propagate_batch(const BatchPair* __restrict__ pairs, int npairs,
                  long long nitems, int view, int cur, int iter) {
  extern __shared__ float4 smem[];
  batch_walk<kRows>(pairs, npairs, nitems,
                    [&](const BatchPair& P, int bx, int by) {
    PropagateArgs a;
    a.self = P.v[view];
    a.other = P.v[1 - view];
    a.old = P.state[view][cur];
    a.out = P.state[view][1 - cur];
    a.smooth = nullptr;
    a.viewin = P.vbuf[1 - view];
    a.viewout = P.vbuf[view];
    // the left view reads the right view's field of the previous
    // iteration, the right view the left view's new one
    a.src_other = view == 0 ? P.state[1][cur] : P.state[0][1 - cur];
    a.dmax = P.dmax;
    a.theta = 0.0f;
    a.lr = view == 0 ? 1 : 0;
    a.iter = iter;
    propagate_block<kKind, false, false, kRows>(P.g, a, nullptr, smem, bx,
                                                by);
  });
}

__global__ void __launch_bounds__(256)
k_prepare_batch(const BatchPair* __restrict__ pairs, int R, float cw) {
  const BatchPair& P = pairs[blockIdx.z >> 1];
  const int view = blockIdx.z & 1;
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= P.g.W || y >= P.g.H) return;
  prepare_pixel(P.img[view], P.g.W, P.g.H, P.channels, P.stride, R, cw,
                P.v[view].tex, P.v[view].sgn, x, y);
}

__global__ void __launch_bounds__(256)
k_viewbuf_clear_batch(const BatchPair* __restrict__ pairs, int view) {
  const BatchPair& P = pairs[blockIdx.z];
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= P.g.W || y >= P.g.H) return;
  P.vbuf[view][(size_t)y * P.g.W + x] = ~0ull;
}

__global__ void __launch_bounds__(256)
k_consistency_batch(const BatchPair* __restrict__ pairs, int cur,
                    float maxdispdiff, float maxdeg) {
  const BatchPair& P = pairs[blockIdx.z];
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= P.g.W || y >= P.g.H) return;
  consistency_pixel(P.g.W, P.state[0][cur], P.state[1][cur], maxdispdiff,
                    maxdeg, P.checked, x, y);
}

// pp 0: the left field's disparity; 1: the LR-checked field's; 2: the
// LR-checked field's, holes filled
__global__ void __launch_bounds__(256)
k_disparity_batch(const BatchPair* __restrict__ pairs, int cur, int pp) {
  const BatchPair& P = pairs[blockIdx.z];
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  if (x >= P.g.W || y >= P.g.H) return;
  if (pp == 2)
    fill_depthmap_pixel(P.g.W, P.checked, 1.0f, P.disp, x, y);
  else
    to_disparity_pixel(P.g.W, pp == 1 ? P.checked : P.state[0][cur], 1.0f,
                       1, P.disp, x, y);
}

__global__ void __launch_bounds__(256)
k_zero_counts_batch(const BatchPair* __restrict__ pairs, int npairs) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < 3 * npairs && pairs[i / 3].counts) pairs[i / 3].counts[i % 3] = 0ull;
}

__global__ void __launch_bounds__(256)
k_eval_error_batch(const BatchPair* __restrict__ pairs, float abs_thres,
                   float rel_thres) {
  const BatchPair& P = pairs[blockIdx.z];
  if (!P.gt || !P.counts) return;   // uniform over the work-group
  const int x = blockIdx.x * kBX + threadIdx.x;
  const int y = blockIdx.y * kBY + threadIdx.y;
  eval_error_pixel(P.disp, P.g.W, P.gt, P.gt_stride, P.g.W, P.g.H, abs_thres,
                   rel_thres, P.counts, x, y);
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
static const dim3 kBlockTile(kTX, kTY);

// dynamic LDS above 64 KiB has to be allowed per kernel
template <typename K>
static hipError_t allow_lds(K kernel, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes);
}

#define PM_LAUNCH_TILE(kernel, kind, g, st, ...)                        \
  do {                                                                  \
    const size_t lds_ = block_lds_bytes((g).R, kind);                   \
    hipError_t e_ = allow_lds(kernel, lds_);                            \
    if (e_ != hipSuccess) return e_;                                    \
    PM_LAUNCH(kernel, grid_for(g), kBlockTile, lds_, st,                \
              __VA_ARGS__);                                             \
  } while (0)

hipError_t launch_prepare(hipStream_t st, const Geometry& g, const float* img,
                          int channels, int stride, ViewBufs v) {
  const dim3 grid((g.W + kBX - 1) / kBX, (g.H + kBY - 1) / kBY);
  PM_LAUNCH(k_prepare, grid, kBlock2D, 0, st, img, g.W, g.H, channels, stride,
            g.R, g.color_weight, v.tex, v.sgn);
}

hipError_t launch_init(hipStream_t st, const Geometry& g, ViewBufs self,
                       ViewBufs other, float dmax, int lr, float4* states) {
  if (g.cost_kind == kCostColorGrad)
    PM_LAUNCH_TILE(k_init<kCostColorGrad>, kCostColorGrad, g, st, g, self,
                   other, dmax, lr, states);
  PM_LAUNCH_TILE(k_init<kCostCensus>, kCostCensus, g, st, g, self, other,
                 dmax, lr, states);
}

hipError_t launch_disparity_to_normal(hipStream_t st, const Geometry& g,
                                      const float* disp, float4* normals) {
  PM_LAUNCH(k_disparity_to_normal,
            dim3((g.W + kBX - 1) / kBX, (g.H + kBY - 1) / kBY), kBlock2D, 0,
            st, g.W, g.H, disp, normals);
}

hipError_t launch_init_prior(hipStream_t st, const Geometry& g, ViewBufs self,
                             ViewBufs other, float dmax, int lr,
                             const float* disp, const float4* normals,
                             float4* states) {
  if (g.cost_kind == kCostColorGrad)
    PM_LAUNCH_TILE(k_init_prior<kCostColorGrad>, kCostColorGrad, g, st, g,
                   self, other, dmax, lr, disp, normals, states);
  PM_LAUNCH_TILE(k_init_prior<kCostCensus>, kCostCensus, g, st, g, self,
                 other, dmax, lr, disp, normals, states);
}

hipError_t launch_init_proposal(hipStream_t st, const Geometry& g,
                                ViewBufs self, ViewBufs other, float dmax,
                                int lr, const float* prop, int os, int Wp,
                                float dscale, float4* states) {
  if (g.cost_kind == kCostColorGrad)
    PM_LAUNCH_TILE(k_init_proposal<kCostColorGrad>, kCostColorGrad, g, st, g,
                   self, other, dmax, lr, prop, os, Wp, dscale, states);
  PM_LAUNCH_TILE(k_init_proposal<kCostCensus>, kCostCensus, g, st, g, self,
                 other, dmax, lr, prop, os, Wp, dscale, states);
}

hipError_t launch_viewbuf_clear_rows(hipStream_t st, const Geometry& g,
                                     unsigned long long* vbuf, int row0,
                                     int rows) {
  if (row0 < 0 || rows < 0 || row0 + rows > g.H) return hipErrorInvalidValue;
  const size_t n = (size_t)rows * g.W;
  if (n == 0) return hipSuccess;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  PM_LAUNCH(k_viewbuf_clear, dim3(blocks), dim3(256), 0, st, n,
            vbuf + (size_t)row0 * g.W);
}

hipError_t launch_viewbuf_clear(hipStream_t st, const Geometry& g,
                                unsigned long long* vbuf) {
  const unsigned blocks = (unsigned)((g.npix + 255) / 256);
  PM_LAUNCH(k_viewbuf_clear, dim3(blocks), dim3(256), 0, st, g.npix, vbuf);
}

template <int kKind, bool kRB>
static hipError_t launch_propagate_kind(hipStream_t st, const Geometry& g,
                                        const PropagateArgs& a,
                                        unsigned long long* counters,
                                        size_t lds_floor) {
  const size_t lds =
      std::max(block_lds_bytes(g.R, kKind, kRB ? g.color : -1), lds_floor);
  if (counters) {
    const hipError_t e = allow_lds(k_propagate<kKind, true, kRB>, lds);
    if (e != hipSuccess) return e;
    PM_LAUNCH((k_propagate<kKind, true, kRB>), grid_for(g), kBlockTile, lds,
              st, g, a, counters);
  }
  const hipError_t e = allow_lds(k_propagate<kKind, false, kRB>, lds);
  if (e != hipSuccess) return e;
  PM_LAUNCH((k_propagate<kKind, false, kRB>), grid_for(g), kBlockTile, lds,
            st, g, a, counters);
}

hipError_t launch_propagate(hipStream_t st, const Geometry& g,
                            const PropagateArgs& a,
                            unsigned long long* counters, size_t lds_floor) {
  if (g.color != -1 && g.color != 0 && g.color != 1)
    return hipErrorInvalidValue;
  if (g.color >= 0 && g.color == 1 && !a.fresh) return hipErrorInvalidValue;
  const bool rb = g.color >= 0;
  if (g.cost_kind == kCostColorGrad)
    return rb ? launch_propagate_kind<kCostColorGrad, true>(st, g, a, counters,
                                                           lds_floor)
              : launch_propagate_kind<kCostColorGrad, false>(st, g, a,
                                                            counters,
                                                            lds_floor);
  return rb ? launch_propagate_kind<kCostCensus, true>(st, g, a, counters,
                                                      lds_floor)
            : launch_propagate_kind<kCostCensus, false>(st, g, a, counters,
                                                       lds_floor);
}

hipError_t launch_init_from_view(hipStream_t st, const Geometry& g,
                                 float dmax,
                                 const unsigned long long* slots,
                                 const float4* src, float4* states) {
  PM_LAUNCH(k_init_from_view,
            dim3((g.W + kBX - 1) / kBX, (g.H + kBY - 1) / kBY), kBlock2D, 0,
            st, g.W, g.H, dmax, slots, src, states);
}

hipError_t launch_scatter(hipStream_t st, const Geometry& g,
                          const float4* states, int lr,
                          unsigned long long* viewout) {
  PM_LAUNCH(k_scatter, dim3((g.W + kBX - 1) / kBX, (g.H + kBY - 1) / kBY),
            kBlock2D, 0, st, g.W, g.H, states, lr, viewout);
}

hipError_t launch_cost_field(hipStream_t st, const Geometry& g, ViewBufs self,
                             ViewBufs other, const float4* planes,
                             float* cost) {
  if (g.cost_kind == kCostColorGrad)
    PM_LAUNCH_TILE(k_cost_field<kCostColorGrad>, kCostColorGrad, g, st, g,
                   self, other, planes, cost);
  PM_LAUNCH_TILE(k_cost_field<kCostCensus>, kCostCensus, g, st, g, self,
                 other, planes, cost);
}

hipError_t launch_klt_field(hipStream_t st, const Geometry& g, ViewBufs self,
                            ViewBufs other, const float4* planes, int passes,
                            float4* out) {
  PM_LAUNCH_TILE(k_klt_field, kCostCensus, g, st, g, self, other, planes,
                 passes, out);
}

hipError_t launch_consistency(hipStream_t st, const Geometry& g,
                              const float4* left, const float4* right,
                              float maxdispdiff, float maxdeg, float4* out) {
  PM_LAUNCH(k_consistency, dim3((g.W + kBX - 1) / kBX, (g.H + kBY - 1) / kBY),
            kBlock2D, 0, st, g.W, g.H, left, right, maxdispdiff, maxdeg, out);
}

hipError_t launch_to_disparity(hipStream_t st, const Geometry& g,
                               const float4* states, float scale, int lr,
                               float* disp) {
  PM_LAUNCH(k_to_disparity,
            dim3((g.W + kBX - 1) / kBX, (g.H + kBY - 1) / kBY), kBlock2D, 0,
            st, g.W, g.H, states, scale, lr, disp);
}

hipError_t launch_fill_depthmap(hipStream_t st, const Geometry& g,
                                const float4* states, float scale,
                                float* disp) {
  PM_LAUNCH(k_fill_depthmap,
            dim3((g.W + kBX - 1) / kBX, (g.H + kBY - 1) / kBY), kBlock2D, 0,
            st, g.W, g.H, states, scale, disp);
}

hipError_t launch_fill_normalmap(hipStream_t st, const Geometry& g,
                                 const float4* states, float4* normals) {
  PM_LAUNCH(k_fill_normalmap,
            dim3((g.W + kBX - 1) / kBX, (g.H + kBY - 1) / kBY), kBlock2D, 0,
            st, g.W, g.H, states, normals);
}

hipError_t launch_eval_error(hipStream_t st, const float* disp, int sd,
                             const float* gt, int sg, int W, int H,
                             float abs_thres, float rel_thres,
                             unsigned long long* counts) {
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, counts, 3);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  PM_LAUNCH(k_eval_error, dim3((W + kBX - 1) / kBX, (H + kBY - 1) / kBY),
            kBlock2D, 0, st, disp, sd, gt, sg, W, H, abs_thres, rel_thres,
            counts);
}

// ---------------------------------------------------------------------
// batch launchers
// ---------------------------------------------------------------------
size_t batch_lds_bytes(int R, int cost_kind, int rows) {
  const size_t tile = (size_t)(kTX + 2 * R) * (rows + 2 * R) * sizeof(float4);
  return cost_kind == kCostColorGrad ? 2 * tile : tile;
}

static dim3 batch_pixel_grid(int maxW, int maxH, int z) {
  return dim3((unsigned)((maxW + kBX - 1) / kBX),
              (unsigned)((maxH + kBY - 1) / kBY), (unsigned)z);
}

// persistent grid: at least one work-group per XCD (batch_walk's
// requirement), at most one per tile block
static unsigned batch_tile_grid(long long nitems, BatchShape shape) {
  long long g = shape.workgroups > 0 ? shape.workgroups : nitems;
  g = std::min(g, nitems);
  return (unsigned)std::max<long long>(g, kXcds);
}

hipError_t launch_prepare_batch(hipStream_t st, const BatchPair* pairs,
                                int npairs, int maxW, int maxH, int R,
                                float cw) {
  if (npairs < 1) return hipSuccess;
  PM_LAUNCH(k_prepare_batch, batch_pixel_grid(maxW, maxH, 2 * npairs),
            kBlock2D, 0, st, pairs, R, cw);
}

hipError_t launch_viewbuf_clear_batch(hipStream_t st, const BatchPair* pairs,
                                      int npairs, int maxW, int maxH,
                                      int view) {
  if (npairs < 1) return hipSuccess;
  PM_LAUNCH(k_viewbuf_clear_batch, batch_pixel_grid(maxW, maxH, npairs),
            kBlock2D, 0, st, pairs, view);
}

template <int kKind, int kRows>
static hipError_t init_batch_rows(hipStream_t st, const BatchPair* pairs,
                                  int npairs, long long nitems, int view,
                                  size_t lds, BatchShape shape) {
  const hipError_t e = allow_lds(k_init_batch<kKind, kRows>, lds);
  if (e != hipSuccess) return e;
  PM_LAUNCH((k_init_batch<kKind, kRows>), dim3(batch_tile_grid(nitems, shape)),
            dim3(kTX, kRows), lds, st, pairs, npairs, nitems, view);
}

template <int kKind>
static hipError_t init_batch_kind(hipStream_t st, const BatchPair* pairs,
                                  int npairs, long long nitems, int view,
                                  size_t lds, BatchShape shape) {
  switch (shape.rows) {
    case 2:
      return init_batch_rows<kKind, 2>(st, pairs, npairs, nitems, view, lds,
                                       shape);
    case 4:
      return init_batch_rows<kKind, 4>(st, pairs, npairs, nitems, view, lds,
                                       shape);
    case 8:
      return init_batch_rows<kKind, 8>(st, pairs, npairs, nitems, view, lds,
                                       shape);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_init_batch(hipStream_t st, const BatchPair* pairs,
                             int npairs, long long nitems, int view,
                             int cost_kind, size_t lds, BatchShape shape) {
  if (nitems < 1) return hipSuccess;
  if (cost_kind == kCostColorGrad)
    return init_batch_kind<kCostColorGrad>(st, pairs, npairs, nitems, view,
                                           lds, shape);
  return init_batch_kind<kCostCensus>(st, pairs, npairs, nitems, view, lds,
                                      shape);
}

template <int kKind, int kRows>
static hipError_t propagate_batch_rows(hipStream_t st, const BatchPair* pairs,
                                       int npairs, long long nitems, int view,
                                       int cur, int iter, size_t lds,
                                       BatchShape shape) {
  const hipError_t e = allow_lds(k_propagate_batch<kKind, kRows>, lds);
  if (e != hipSuccess) return e;
  PM_LAUNCH((k_propagate_batch<kKind, kRows>),
            dim3(batch_tile_grid(nitems, shape)), dim3(kTX, kRows), lds, st,
            pairs, npairs, nitems, view, cur, iter);
}

template <int kKind>
static hipError_t propagate_batch_kind(hipStream_t st, const BatchPair* pairs,
                                       int npairs, long long nitems, int view,
                                       int cur, int iter, size_t lds,
                                       BatchShape shape) {
  switch (shape.rows) {
    case 2:
      return propagate_batch_rows<kKind, 2>(st, pairs, npairs, nitems, view,
                                            cur, iter, lds, shape);
    case 4:
      return propagate_batch_rows<kKind, 4>(st, pairs, npairs, nitems, view,
                                            cur, iter, lds, shape);
    case 8:
      return propagate_batch_rows<kKind, 8>(st, pairs, npairs, nitems, view,
                                            cur, iter, lds, shape);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_propagate_batch(hipStream_t st, const BatchPair* pairs,
                                  int npairs, long long nitems, int view,
                                  int cur, int iter, int cost_kind,
                                  size_t lds, BatchShape shape) {
  if (nitems < 1) return hipSuccess;
  if (cost_kind == kCostColorGrad)
    return propagate_batch_kind<kCostColorGrad>(st, pairs, npairs, nitems,
                                                view, cur, iter, lds, shape);
  return propagate_batch_kind<kCostCensus>(st, pairs, npairs, nitems, view,
                                           cur, iter, lds, shape);
}

template <int kKind, int kRows>
static hipError_t resident_rows(size_t lds, int* out) {
  hipError_t e = allow_lds(k_propagate_batch<kKind, kRows>, lds);
  if (e != hipSuccess) return e;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(
      out, k_propagate_batch<kKind, kRows>, kTX * kRows, lds);
}

hipError_t batch_resident_per_cu(int cost_kind, int rows, size_t lds,
                                 int* out) {
  const bool cg = cost_kind == kCostColorGrad;
  switch (rows) {
    case 2:
      return cg ? resident_rows<kCostColorGrad, 2>(lds, out)
                : resident_rows<kCostCensus, 2>(lds, out);
    case 4:
      return cg ? resident_rows<kCostColorGrad, 4>(lds, out)
                : resident_rows<kCostCensus, 4>(lds, out);
    case 8:
      return cg ? resident_rows<kCostColorGrad, 8>(lds, out)
                : resident_rows<kCostCensus, 8>(lds, out);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_postprocess_batch(hipStream_t st, const BatchPair* pairs,
                                    int npairs, int maxW, int maxH, int cur,
                                    int pp, float maxdispdiff, float maxdeg) {
  if (npairs < 1) return hipSuccess;
  if (pp < 0 || pp > 2) return hipErrorInvalidValue;
  const dim3 grid = batch_pixel_grid(maxW, maxH, npairs);
  if (pp != 0) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_consistency_batch, grid, kBlock2D, 0, st, pairs,
                       cur, maxdispdiff, maxdeg);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  PM_LAUNCH(k_disparity_batch, grid, kBlock2D, 0, st, pairs, cur, pp);
}

hipError_t launch_eval_error_batch(hipStream_t st, const BatchPair* pairs,
                                   int npairs, int maxW, int maxH,
                                   float abs_thres, float rel_thres) {
  if (npairs < 1) return hipSuccess;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_zero_counts_batch,
                     dim3((unsigned)((3 * npairs + 255) / 256)), dim3(256), 0,
                     st, pairs, npairs);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  PM_LAUNCH(k_eval_error_batch, batch_pixel_grid(maxW, maxH, npairs),
            kBlock2D, 0, st, pairs, abs_thres, rel_thres);
}

hipError_t launch_draws(hipStream_t st, uint64_t k, int n, float* out) {
  const int threads = (n + 15) / 16;
  PM_LAUNCH(k_draws, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
            st, k, n, out);
}

hipError_t launch_copy_probe(hipStream_t st, const float4* src, float4* dst,
                             size_t n) {
  PM_LAUNCH(k_copy_probe, dim3(4096), dim3(256), 0, st, src, dst, n);
}

hipError_t launch_math(hipStream_t st, const float* in, int n, float* out) {
  PM_LAUNCH(k_math, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in,
            n, out);
}

}  // namespace pm
