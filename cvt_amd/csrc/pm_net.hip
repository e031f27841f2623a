// pm_net.hip -- gfx950 kernels of the plane-proposal network: a 3x3
// convolution on bf16 NHWC activations as an implicit GEMM on the matrix
// cores (v_mfma_f32_16x16x32_bf16, fp32 accumulation), its weight gradient
// (fp32 VALU), and the pooling, gradient routing, loss, Adam and input
// packing kernels of a U-Net training step.
#include <algorithm>

#include "pm_net_kernels.h"

namespace pm {
namespace net {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

#define NET_LAUNCH(kernel, grid, block, lds, st, ...)               \
  do {                                                              \
    (void)hipGetLastError();                                        \
    hipLaunchKernelGGL(kernel, grid, block, lds, st, __VA_ARGS__);  \
    return hipGetLastError();                                       \
  } while (0)

__device__ __forceinline__ bf16x8 load8(const bf16* p) {
  return *reinterpret_cast<const bf16x8*>(p);
}

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int i = 0; i < 8; ++i) z[i] = (bf16)0.0f;
  return z;
}

// ---------------------------------------------------------------------
// convolution: a work-group is one wavefront computing 64 output pixels x
// 16 kNT output channels as 4 x kNT tiles of 16 x 16.  Per K step of 32 a lane
// holds the A operand of its tile row (pixel) and k-group (8 consecutive
// k: one tap, 8 channels, one 16-byte load from NHWC) and the B operand
// of its tile column (output channel: 8 consecutive k of its weight row).
// The accumulator of a lane holds rows 4 g .. 4 g + 3, column lane % 16.
// ---------------------------------------------------------------------
constexpr int kConvPix = 64;

// kNT column tiles of 16 output channels per wavefront: 2 (32 channels)
// or 4 (64, when the padded output count is a multiple of 64, which
// halves the A loads per matrix-core op)
template <int kNT>
__global__ void __launch_bounds__(64) k_conv(ConvArgs a) {
  const int lane = threadIdx.x;
  const int row = lane & 15, grp = lane >> 4;
  const int cin = a.c1 + a.c2;
  const int kpad = conv_kpad(cin), kreal = 9 * cin;
  const long npix = (long)a.N * a.H * a.W;
  const long p0 = (long)blockIdx.x * kConvPix;
  const int co0 = blockIdx.y * (16 * kNT);
  const int H2 = a.up2 ? (a.H + 1) >> 1 : a.H;
  const int W2 = a.up2 ? (a.W + 1) >> 1 : a.W;
  int pn[4], py[4], px[4];
  bool pv[4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const long p = p0 + mt * 16 + row;
    pv[mt] = p < npix;
    const long q = pv[mt] ? p : 0;
    px[mt] = (int)(q % a.W);
    const long t = q / a.W;
    py[mt] = (int)(t % a.H);
    pn[mt] = (int)(t / a.H);
  }
  f32x4 acc[4][kNT];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16* wrow = a.w + (size_t)(co0 + row) * kpad;
  // this lane's k = k0 + 8 grp as (tap, channel), advanced by 32 per step
  int tap = (8 * grp) / cin, ci = (8 * grp) % cin;
  for (int k0 = 0; k0 < kpad; k0 += 32) {
    const int k = k0 + 8 * grp;
    bf16x8 A[4];
    if (k < kreal) {
      const int dy = (tap / 3 - 1) * a.dil, dx = (tap % 3 - 1) * a.dil;
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        const int yy = py[mt] + dy, xx = px[mt] + dx;
        A[mt] = zero8();
        if (pv[mt] && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
          if (ci < a.c1) {
            A[mt] = load8(a.x1 +
                          (((size_t)pn[mt] * a.H + yy) * a.W + xx) * a.c1 +
                          ci);
          } else {
            const int y2 = a.up2 ? yy >> 1 : yy, x2 = a.up2 ? xx >> 1 : xx;
            A[mt] = load8(a.x2 +
                          (((size_t)pn[mt] * H2 + y2) * W2 + x2) * a.c2 +
                          (ci - a.c1));
          }
        }
      }
    } else {
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) A[mt] = zero8();
    }
    bf16x8 B[kNT];
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt)
      B[nt] = load8(wrow + (size_t)nt * 16 * kpad + k);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt)
        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            A[mt], B[nt], acc[mt][nt], 0, 0, 0);
    ci += 32;
    while (ci >= cin) {
      ci -= cin;
      ++tap;
    }
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt) {
      const int co = co0 + nt * 16 + row;
      if (co >= a.cout) continue;
      const float b = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const long p = p0 + mt * 16 + 4 * grp + j;
        if (p >= npix) continue;
        float v = acc[mt][nt][j] + b;
        if (a.relu) v = fmaxf(v, 0.0f);
        if (a.yf)
          a.yf[p * a.ys + co] = v;
        else
          a.y[p * a.ys + co] = (bf16)v;
      }
    }
}

// ---------------------------------------------------------------------
// LDS-tiled convolution (dilation 1, c1 and c2 multiples of 32): a
// work-group of four wavefronts computes a 64 x 4 pixel tile of one image
// (wavefront w: tile row w) x 16 kNT output channels.  The input channels
// are taken 32 at a time: the chunk's (4 + 2) x (64 + 2) pixel window and
// its weights for the 9 taps are staged in LDS, and every tap is one K
// step of 32 read from there, so each input pixel is fetched from memory
// 1.5 times per chunk instead of 9.  The next chunk's global loads are in
// flight during the current chunk's matrix-core work.  LDS rows of 64
// bytes (32 channels) are read 16 bytes per lane by 16 consecutive
// pixels; the 16-byte slot of channel group g of window column x is
// g ^ ((x >> 1) & 3), so any 8 consecutive columns hit distinct banks.
// K is summed chunk by chunk, tap by tap within a chunk.
// ---------------------------------------------------------------------
constexpr int kTileX = 64, kTileY = 4;
constexpr int kWinX = kTileX + 2, kWinY = kTileY + 2;
constexpr int kWinPieces = kWinY * kWinX * 4;   // 16-byte pieces

__device__ __forceinline__ int lds_slot(int g, int col) {
  return g ^ ((col >> 1) & 3);
}

template <int kNT>
__device__ __forceinline__ void conv_lds_block(const ConvArgs& a, int tiles_x,
                                               int tiles_y) {
  constexpr int kCO = 16 * kNT;
  constexpr int kWPieces = 9 * kCO * 4;
  constexpr int kAIters = (kWinPieces + 255) / 256;
  constexpr int kBIters = (kWPieces + 255) / 256;
  __shared__ uint4 sA[kWinPieces];
  __shared__ uint4 sB[kWPieces];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int row = lane & 15, grp = lane >> 4;
  int t = blockIdx.x;
  const int tx = t % tiles_x;
  t /= tiles_x;
  const int ty = t % tiles_y, n = t / tiles_y;
  const int x0 = tx * kTileX, y0 = ty * kTileY;
  const int co0 = blockIdx.y * kCO;
  const int cin = a.c1 + a.c2;
  const int kpad = conv_kpad(cin);
  const int H2 = a.up2 ? (a.H + 1) >> 1 : a.H;
  const int W2 = a.up2 ? (a.W + 1) >> 1 : a.W;
  const int nchunks = cin / 32;

  uint4 ra[kAIters], rb[kBIters];
  // global loads of chunk c (channels 32 c .. 32 c + 31) into ra, rb
  auto fetch = [&](int c) {
    const int cc = 32 * c;
#pragma unroll
    for (int i = 0; i < kAIters; ++i) {
      const int q = tid + 256 * i;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (q < kWinPieces) {
        const int r = q / (kWinX * 4), rem = q - r * (kWinX * 4);
        const int xt = rem >> 2, g = rem & 3;
        const int yy = y0 - 1 + r, xx = x0 - 1 + xt;
        if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
          if (cc < a.c1) {
            v = *reinterpret_cast<const uint4*>(
                a.x1 + (((size_t)n * a.H + yy) * a.W + xx) * a.c1 + cc +
                8 * g);
          } else {
            const int y2 = a.up2 ? yy >> 1 : yy, x2 = a.up2 ? xx >> 1 : xx;
            v = *reinterpret_cast<const uint4*>(
                a.x2 + (((size_t)n * H2 + y2) * W2 + x2) * a.c2 +
                (cc - a.c1) + 8 * g);
          }
        }
      }
      ra[i] = v;
    }
#pragma unroll
    for (int i = 0; i < kBIters; ++i) {
      const int q = tid + 256 * i;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (q < kWPieces) {
        const int tap = q / (kCO * 4), rem = q - tap * (kCO * 4);
        const int co = rem >> 2, g = rem & 3;
        v = *reinterpret_cast<const uint4*>(
            a.w + (size_t)(co0 + co) * kpad + tap * cin + cc + 8 * g);
      }
      rb[i] = v;
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int i = 0; i < kAIters; ++i) {
      const int q = tid + 256 * i;
      if (q < kWinPieces) {
        const int r = q / (kWinX * 4), rem = q - r * (kWinX * 4);
        const int xt = rem >> 2, g = rem & 3;
        sA[(r * kWinX + xt) * 4 + lds_slot(g, xt)] = ra[i];
      }
    }
#pragma unroll
    for (int i = 0; i < kBIters; ++i) {
      const int q = tid + 256 * i;
      if (q < kWPieces) {
        const int tap = q / (kCO * 4), rem = q - tap * (kCO * 4);
        const int co = rem >> 2, g = rem & 3;
        sB[(tap * kCO + co) * 4 + lds_slot(g, co)] = rb[i];
      }
    }
  };

  f32x4 acc[4][kNT];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  fetch(0);
  for (int c = 0; c < nchunks; ++c) {
    __syncthreads();  // every wavefront is done reading the previous chunk
    stage();
    __syncthreads();
    if (c + 1 < nchunks) fetch(c + 1);
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int dy = tap / 3, dx = tap % 3;  // window offsets 0..2
      bf16x8 A[4], B[kNT];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        const int xt = mt * 16 + row + dx;
        A[mt] = __builtin_bit_cast(
            bf16x8, sA[((wave + dy) * kWinX + xt) * 4 + lds_slot(grp, xt)]);
      }
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt) {
        const int co = nt * 16 + row;
        B[nt] = __builtin_bit_cast(
            bf16x8, sB[(tap * kCO + co) * 4 + lds_slot(grp, co)]);
      }
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < kNT; ++nt)
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
              A[mt], B[nt], acc[mt][nt], 0, 0, 0);
    }
  }
  const int y = y0 + wave;
  if (y >= a.H) return;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < kNT; ++nt) {
      const int co = co0 + nt * 16 + row;
      if (co >= a.cout) continue;
      const float b = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int x = x0 + mt * 16 + 4 * grp + j;
        if (x >= a.W) continue;
        const size_t p = ((size_t)n * a.H + y) * a.W + x;
        float v = acc[mt][nt][j] + b;
        if (a.relu) v = fmaxf(v, 0.0f);
        if (a.yf)
          a.yf[p * a.ys + co] = v;
        else
          a.y[p * a.ys + co] = (bf16)v;
      }
    }
}

// 32 output channels per work-group.  With one input chunk (32 input
// channels) nothing overlaps a work-group's fetch, so that form is built
// for three wavefronts per SIMD (168 VGPRs, a few spilled); with more
// chunks the spills cost more than the extra wavefront gains and the
// unconstrained build (two per SIMD) is used.  At 720p: the single-chunk
// layers 143 -> 115 us, the others 3-12 % slower at three.  The
// 64-channel form cannot reach three and keeps two.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_conv_lds32_1chunk(ConvArgs a, int tiles_x, int tiles_y) {
  conv_lds_block<2>(a, tiles_x, tiles_y);
}

__global__ void __launch_bounds__(256)
k_conv_lds32(ConvArgs a, int tiles_x, int tiles_y) {
  conv_lds_block<2>(a, tiles_x, tiles_y);
}

__global__ void __launch_bounds__(256)
k_conv_lds64(ConvArgs a, int tiles_x, int tiles_y) {
  conv_lds_block<4>(a, tiles_x, tiles_y);
}

hipError_t launch_conv(hipStream_t st, const ConvArgs& a) {
  const long npix = (long)a.N * a.H * a.W;
  const int cpad = conv_cpad(a.cout);
  if (a.dil == 1 && a.c1 % 32 == 0 && a.c2 % 32 == 0 && a.c1 + a.c2 > 0) {
    const int tiles_x = (a.W + kTileX - 1) / kTileX;
    const int tiles_y = (a.H + kTileY - 1) / kTileY;
    const unsigned g = (unsigned)((long)tiles_x * tiles_y * a.N);
    if (cpad % 64 == 0)
      NET_LAUNCH(k_conv_lds64, dim3(g, (unsigned)(cpad / 64)), dim3(256), 0,
                 st, a, tiles_x, tiles_y);
    if (a.c1 + a.c2 == 32)
      NET_LAUNCH(k_conv_lds32_1chunk, dim3(g, (unsigned)(cpad / 32)),
                 dim3(256), 0, st, a, tiles_x, tiles_y);
    NET_LAUNCH(k_conv_lds32, dim3(g, (unsigned)(cpad / 32)), dim3(256), 0,
               st, a, tiles_x, tiles_y);
  }
  const unsigned gx = (unsigned)((npix + kConvPix - 1) / kConvPix);
  if (cpad % 64 == 0)
    NET_LAUNCH(k_conv<4>, dim3(gx, (unsigned)(cpad / 64)), dim3(64), 0, st,
               a);
  NET_LAUNCH(k_conv<2>, dim3(gx, (unsigned)(cpad / 32)), dim3(64), 0, st, a);
}

// ---------------------------------------------------------------------
// weight gradient: a work-group owns 32 k x 32 output channels and a chunk
// of pixels; 32-pixel slices of the input patches and of the output
// gradient are staged in LDS, each thread accumulates 4 channels of one k
// ---------------------------------------------------------------------
constexpr int kWgChunk = 1024;

__global__ void __launch_bounds__(256)
k_conv_wgrad(ConvArgs a, const bf16* __restrict__ g, float* __restrict__ dw,
             float* __restrict__ db) {
  __shared__ float sx[32][33];
  __shared__ float sg[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int cin = a.c1 + a.c2, kpad = conv_kpad(cin), kreal = 9 * cin;
  const int k = blockIdx.x * 32 + tx;
  const int co0 = blockIdx.y * 32;
  const long npix = (long)a.N * a.H * a.W;
  const long pb0 = (long)blockIdx.z * kWgChunk;
  const long pe = std::min<long>(npix, pb0 + kWgChunk);
  const int H2 = a.up2 ? (a.H + 1) >> 1 : a.H;
  const int W2 = a.up2 ? (a.W + 1) >> 1 : a.W;
  // this thread stages k (its tx) of the patches and channel co0 + tx of g
  const bool kin = k < kreal;
  const int tap = kin ? k / cin : 0, ci = kin ? k - tap * cin : 0;
  const int dy = (tap / 3 - 1) * a.dil, dx = (tap % 3 - 1) * a.dil;
  const int gco = co0 + tx;
  float acc[4] = {0.f, 0.f, 0.f, 0.f}, accb[4] = {0.f, 0.f, 0.f, 0.f};
  for (long pb = pb0; pb < pe; pb += 32) {
    for (int pp = ty; pp < 32; pp += 8) {
      const long p = pb + pp;
      float xv = 0.0f, gv = 0.0f;
      if (p < pe) {
        const int xq = (int)(p % a.W);
        const long t = p / a.W;
        const int yq = (int)(t % a.H), nq = (int)(t / a.H);
        const int yy = yq + dy, xx = xq + dx;
        if (kin && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
          if (ci < a.c1) {
            xv = (float)a.x1[(((size_t)nq * a.H + yy) * a.W + xx) * a.c1 + ci];
          } else {
            const int y2 = a.up2 ? yy >> 1 : yy, x2 = a.up2 ? xx >> 1 : xx;
            xv = (float)a.x2[(((size_t)nq * H2 + y2) * W2 + x2) * a.c2 +
                             (ci - a.c1)];
          }
        }
        if (gco < a.cout) gv = (float)g[p * a.cout + gco];
      }
      sx[pp][tx] = xv;
      sg[pp][tx] = gv;
    }
    __syncthreads();
#pragma unroll 8
    for (int pp = 0; pp < 32; ++pp) {
      const float xv = sx[pp][tx];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float gv = sg[pp][ty * 4 + j];
        acc[j] = fmaf(gv, xv, acc[j]);
        accb[j] += gv;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int co = co0 + ty * 4 + j;
    if (co >= a.cout) continue;
    if (kin) atomicAdd(&dw[(size_t)co * kpad + k], acc[j]);
    if (db && blockIdx.x == 0 && tx == 0) atomicAdd(&db[co], accb[j]);
  }
}

hipError_t launch_conv_wgrad(hipStream_t st, const ConvArgs& a, const bf16* g,
                             float* dw, float* db) {
  const long npix = (long)a.N * a.H * a.W;
  const dim3 grid((unsigned)(conv_kpad(a.c1 + a.c2) / 32),
                  (unsigned)(conv_cpad(a.cout) / 32),
                  (unsigned)((npix + kWgChunk - 1) / kWgChunk));
  NET_LAUNCH(k_conv_wgrad, grid, dim3(256), 0, st, a, g, dw, db);
}

// ---------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_flip_weights(const bf16* __restrict__ w, int cin, int cout,
               bf16* __restrict__ wt) {
  const int kpt = conv_kpad(cout), rows = conv_cpad(cin);
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)rows * kpt) return;
  const int ci = (int)(i / kpt), kk = (int)(i % kpt);
  bf16 v = (bf16)0.0f;
  if (ci < cin && kk < 9 * cout) {
    const int t = kk / cout, co = kk - t * cout;
    v = w[(size_t)co * conv_kpad(cin) + (size_t)(8 - t) * cin + ci];
  }
  wt[i] = v;
}

hipError_t launch_flip_weights(hipStream_t st, const bf16* w, int cin,
                               int cout, bf16* wt) {
  const long n = (long)conv_cpad(cin) * conv_kpad(cout);
  NET_LAUNCH(k_flip_weights, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
             st, w, cin, cout, wt);
}

__global__ void __launch_bounds__(256)
k_grad_mask(size_t n, const float* __restrict__ g, const bf16* __restrict__ y,
            int relu, bf16* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float v = g[i];
  if (relu && !((float)y[i] > 0.0f)) v = 0.0f;
  out[i] = (bf16)v;
}

hipError_t launch_grad_mask(hipStream_t st, size_t n, const float* g,
                            const bf16* y, int relu, bf16* out) {
  NET_LAUNCH(k_grad_mask, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
             n, g, y, relu, out);
}

__global__ void __launch_bounds__(256)
k_route_first(size_t npix, int c1, int c2, const float* __restrict__ t,
              float* __restrict__ g1) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= npix * c1) return;
  const size_t p = i / c1, c = i % c1;
  g1[i] += t[p * (c1 + c2) + c];
}

__global__ void __launch_bounds__(256)
k_route_second(int N, int H, int W, int c1, int c2, int up2,
               const float* __restrict__ t, float* __restrict__ g2) {
  const int H2 = up2 ? (H + 1) >> 1 : H, W2 = up2 ? (W + 1) >> 1 : W;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)N * H2 * W2 * c2) return;
  const int c = (int)(i % c2);
  const size_t q = i / c2;
  const int x2 = (int)(q % W2);
  const size_t r = q / W2;
  const int y2 = (int)(r % H2), n = (int)(r / H2);
  const int ct = c1 + c2;
  float s = 0.0f;
  if (up2) {
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const int y = 2 * y2 + dy, x = 2 * x2 + dx;
        if (y < H && x < W)
          s += t[(((size_t)n * H + y) * W + x) * ct + c1 + c];
      }
  } else {
    s = t[(((size_t)n * H + y2) * W + x2) * ct + c1 + c];
  }
  g2[i] += s;
}

hipError_t launch_route_grad(hipStream_t st, int N, int H, int W, int c1,
                             int c2, int up2, const float* t, float* g1,
                             float* g2) {
  const size_t npix = (size_t)N * H * W;
  if (g1 && c1 > 0) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_route_first,
                       dim3((unsigned)((npix * c1 + 255) / 256)), dim3(256),
                       0, st, npix, c1, c2, t, g1);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (!g2 || c2 == 0) return hipSuccess;
  const size_t H2 = up2 ? (H + 1) / 2 : H, W2 = up2 ? (W + 1) / 2 : W;
  const size_t n2 = (size_t)N * H2 * W2 * c2;
  NET_LAUNCH(k_route_second, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0,
             st, N, H, W, c1, c2, up2, t, g2);
}

__global__ void __launch_bounds__(256)
k_avgpool(int N, int H, int W, int C, const bf16* __restrict__ x,
          bf16* __restrict__ y) {
  // a thread: 8 consecutive channels of one output pixel (C % 8 == 0)
  const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)N * Ho * Wo * C8) return;
  const int c = (int)(i % C8) * 8;
  const size_t q = i / C8;
  const int xo = (int)(q % Wo);
  const size_t r = q / Wo;
  const int yo = (int)(r % Ho), n = (int)(r / Ho);
  const size_t b = (((size_t)n * H + 2 * yo) * W + 2 * xo) * C + c;
  const bf16x8 v00 = load8(x + b), v01 = load8(x + b + C);
  const bf16x8 v10 = load8(x + b + (size_t)W * C);
  const bf16x8 v11 = load8(x + b + (size_t)W * C + C);
  bf16x8 o;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float s = ((float)v00[k] + (float)v01[k]) +
                    ((float)v10[k] + (float)v11[k]);
    o[k] = (bf16)(0.25f * s);
  }
  *reinterpret_cast<bf16x8*>(y + q * C + c) = o;
}

hipError_t launch_avgpool(hipStream_t st, int N, int H, int W, int C,
                          const bf16* x, bf16* y) {
  if (C % 8) return hipErrorInvalidValue;
  const size_t n = (size_t)N * (H / 2) * (W / 2) * (C / 8);
  NET_LAUNCH(k_avgpool, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
             N, H, W, C, x, y);
}

__global__ void __launch_bounds__(256)
k_avgpool_grad(int N, int H, int W, int C, const float* __restrict__ gout,
               float* __restrict__ gin) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)N * H * W * C) return;
  const int c = (int)(i % C);
  const size_t q = i / C;
  const int x = (int)(q % W);
  const size_t r = q / W;
  const int y = (int)(r % H), n = (int)(r / H);
  const int Ho = H / 2, Wo = W / 2;
  gin[i] += 0.25f * gout[(((size_t)n * Ho + (y >> 1)) * Wo + (x >> 1)) * C + c];
}

hipError_t launch_avgpool_grad(hipStream_t st, int N, int H, int W, int C,
                               const float* gout, float* gin) {
  const size_t n = (size_t)N * H * W * C;
  NET_LAUNCH(k_avgpool_grad, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
             st, N, H, W, C, gout, gin);
}

__global__ void __launch_bounds__(256)
k_l1_loss(size_t npix, const float* __restrict__ o, int os,
          const float4* __restrict__ t, float wd, float wn, float inv_count,
          float* __restrict__ g, float* __restrict__ loss) {
  __shared__ float part[256];
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  float l = 0.0f;
  if (p < npix) {
    const float4 tv = t[p];
    const float* op = o + p * os;
    float* gp = g + p * os;
    for (int c = 0; c < os; ++c) gp[c] = 0.0f;
    if (tv.x >= 0.0f) {
      const float d[4] = {op[0] - tv.x, op[1] - tv.y, op[2] - tv.z,
                          op[3] - tv.w};
      const float w[4] = {wd, wn, wn, wn};
      for (int c = 0; c < 4; ++c) {
        l += w[c] * fabsf(d[c]);
        gp[c] = d[c] > 0.0f ? w[c] * inv_count
                            : (d[c] < 0.0f ? -w[c] * inv_count : 0.0f);
      }
    }
  }
  part[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, part[0] * inv_count);
}

hipError_t launch_l1_loss(hipStream_t st, size_t npix, const float* o, int os,
                          const float* t, float wd, float wn, float inv_count,
                          float* g, float* loss) {
  NET_LAUNCH(k_l1_loss, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0,
             st, npix, o, os, reinterpret_cast<const float4*>(t), wd, wn,
             inv_count, g, loss);
}

__global__ void __launch_bounds__(256)
k_adam(size_t n, float* __restrict__ w, const float* __restrict__ g,
       float* __restrict__ m, float* __restrict__ v, float lr, float b1,
       float b2, float eps, float c1, float c2, bf16* __restrict__ wb) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  wb[i] = (bf16)wi;
}

hipError_t launch_adam(hipStream_t st, size_t n, float* w, const float* g,
                       float* m, float* v, float lr, float b1, float b2,
                       float eps, float c1, float c2, bf16* wb) {
  NET_LAUNCH(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
             w, g, m, v, lr, b1, b2, eps, c1, c2, wb);
}

__global__ void __launch_bounds__(256)
k_to_bf16(size_t n, const float* __restrict__ x, bf16* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = (bf16)x[i];
}

hipError_t launch_to_bf16(hipStream_t st, size_t n, const float* x, bf16* y) {
  NET_LAUNCH(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
             n, x, y);
}

// ---------------------------------------------------------------------
// matching cost volumes: the network's explicit stereo evidence.  The
// packed pair is box-averaged (4 x 4 for the coarse volume, 2 x 2 for the
// fine one), and at every pixel of that resolution channel d of the
// volume is the mean over the 3 x 3 window of the summed |r, g, b|
// differences between the first image at (x + i, y + j) and the second at
// (x + i - d, y + j), d = 0 .. levels - 1 (coarse: 48 quarter pixels,
// 0 .. 188 px at full resolution; fine: 64 half pixels, 0 .. 126 px);
// coordinates are clamped to the image.  No parameters: the network
// learns what follows it.
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_down4(int N, int H, int W, const bf16* __restrict__ x,
        bf16* __restrict__ y) {
  const int Ho = H / 4, Wo = W / 4;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)N * Ho * Wo) return;
  const int xo = (int)(i % Wo);
  const size_t r = i / Wo;
  const int yo = (int)(r % Ho), n = (int)(r / Ho);
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int dy = 0; dy < 4; ++dy)
    for (int dx = 0; dx < 4; ++dx) {
      const bf16x8 v = load8(
          x + (((size_t)n * H + 4 * yo + dy) * W + 4 * xo + dx) * 8);
#pragma unroll
      for (int c = 0; c < 8; ++c) s[c] += (float)v[c];
    }
  bf16x8 o;
#pragma unroll
  for (int c = 0; c < 8; ++c) o[c] = (bf16)(s[c] * 0.0625f);
  *reinterpret_cast<bf16x8*>(y + i * 8) = o;
}

hipError_t launch_down4(hipStream_t st, int N, int H, int W, const bf16* x,
                        bf16* y) {
  const size_t n = (size_t)N * (H / 4) * (W / 4);
  NET_LAUNCH(k_down4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, N,
             H, W, x, y);
}

// A work-group computes every level of 64 consecutive pixels of one row.
// It first stages, as float4 (r, g, b, 0) with clamped coordinates, the
// three window rows of the first image over columns x0 - 1 .. x0 + 64 and
// of the second image over x0 - levels .. x0 + 64 (the columns any level
// of the segment reads); then an item is a (pixel, 8 consecutive levels):
// levels 8 g .. 8 g + 7 over the taps x - 1 .. x + 1 read the second image
// at x + 1 - 8 g - 9 .. x + 1 - 8 g, ten pixels per window row.  levels: a
// multiple of 8, at most kCvMaxLevels.
constexpr int kCvSeg = 64;
constexpr int kCvMaxLevels = 64;
constexpr int kCvAW = kCvSeg + 2;                  // first-image columns
constexpr int kCvBW = kCvSeg + 1 + kCvMaxLevels;   // second-image columns

__global__ void __launch_bounds__(256)
k_cost_volume(int N, int H4, int W4, int levels, int segs,
              const bf16* __restrict__ d4, bf16* __restrict__ cv) {
  __shared__ float4 sa[3][kCvAW];
  __shared__ float4 sb[3][kCvBW];
  int t = blockIdx.x;
  const int seg = t % segs;
  t /= segs;
  const int y = t % H4, n = t / H4;
  const int x0 = seg * kCvSeg;
  const int bw = kCvSeg + 1 + levels;  // second-image columns x0 - levels ..
  for (int i = threadIdx.x; i < 3 * (kCvAW + bw); i += 256) {
    const int j = i / (kCvAW + bw), c = i - j * (kCvAW + bw);
    const int yy = min(max(y + j - 1, 0), H4 - 1);
    const bf16* row = d4 + ((size_t)n * H4 + yy) * W4 * 8;
    if (c < kCvAW) {
      const bf16x8 v = load8(row + (size_t)min(max(x0 - 1 + c, 0), W4 - 1) * 8);
      sa[j][c] = make_float4((float)v[0], (float)v[1], (float)v[2], 0.0f);
    } else {
      const int cb = c - kCvAW;
      const bf16x8 v =
          load8(row + (size_t)min(max(x0 - levels + cb, 0), W4 - 1) * 8);
      sb[j][cb] = make_float4((float)v[4], (float)v[5], (float)v[6], 0.0f);
    }
  }
  __syncthreads();
  const int groups = levels / 8;
  for (int item = threadIdx.x; item < kCvSeg * groups; item += 256) {
    const int xl = item % kCvSeg, grp = item / kCvSeg;
    const int x = x0 + xl;
    if (x >= W4) continue;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    // second-image column of t = 1, d = 8 g, as an index into sb
    const int b0 = xl + 1 - 8 * grp + levels;
    for (int j = 0; j < 3; ++j) {
      float4 ar[3], br[10];
#pragma unroll
      for (int tt = 0; tt < 3; ++tt) ar[tt] = sa[j][xl + tt];
      // br[m] = second image at column x + 1 - 8 g - m
#pragma unroll
      for (int m = 0; m < 10; ++m) br[m] = sb[j][b0 - m];
      // tap tt (x + tt - 1) at level 8 g + k reads br[2 - tt + k]
#pragma unroll
      for (int tt = 0; tt < 3; ++tt)
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const float4 bb = br[2 - tt + k];
          acc[k] += (fabsf(ar[tt].x - bb.x) + fabsf(ar[tt].y - bb.y)) +
                    fabsf(ar[tt].z - bb.z);
        }
    }
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = (bf16)(acc[k] * (1.0f / 9.0f));
    const size_t p = ((size_t)n * H4 + y) * W4 + x;
    *reinterpret_cast<bf16x8*>(cv + p * levels + grp * 8) = o;
  }
}

hipError_t launch_cost_volume(hipStream_t st, int N, int H4, int W4,
                              int levels, const bf16* d4, bf16* cv) {
  if (levels < 8 || levels % 8 || levels > kCvMaxLevels)
    return hipErrorInvalidValue;
  const int segs = (W4 + kCvSeg - 1) / kCvSeg;
  const size_t blocks = (size_t)N * H4 * segs;
  NET_LAUNCH(k_cost_volume, dim3((unsigned)blocks), dim3(256), 0, st, N, H4,
             W4, levels, segs, d4, cv);
}

__global__ void __launch_bounds__(256)
k_pack_pair(const float* __restrict__ left, const float* __restrict__ right,
            int W, int H, int C, int stride, int Hp, int Wp, int mirror,
            bf16* __restrict__ x) {
  const int xo = blockIdx.x * 64 + (threadIdx.x & 63);
  const int yo = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (xo >= Wp || yo >= Hp) return;
  const int ys = min(yo, H - 1);
  int xs = min(xo, W - 1);
  const float* a = left;
  const float* b = right;
  if (mirror) {
    xs = W - 1 - xs;
    a = right;
    b = left;
  }
  bf16x8 v;
  const float* pa = a + (size_t)ys * stride;
  const float* pb = b + (size_t)ys * stride;
  for (int c = 0; c < 4; ++c) {
    v[c] = (bf16)(C == 1 ? pa[xs] : pa[4 * (size_t)xs + c]);
    v[4 + c] = (bf16)(C == 1 ? pb[xs] : pb[4 * (size_t)xs + c]);
  }
  *reinterpret_cast<bf16x8*>(x + ((size_t)yo * Wp + xo) * 8) = v;
}

hipError_t launch_pack_pair(hipStream_t st, const float* left,
                            const float* right, int W, int H, int channels,
                            int stride, int Hp, int Wp, int mirror, bf16* x) {
  const dim3 grid((unsigned)((Wp + 63) / 64), (unsigned)((Hp + 3) / 4));
  NET_LAUNCH(k_pack_pair, grid, dim3(256), 0, st, left, right, W, H, channels,
             stride, Hp, Wp, mirror, x);
}

}  // namespace net
}  // namespace pm
