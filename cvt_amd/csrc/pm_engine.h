// pm_engine.h -- host-side launchers of the PatchMatch kernels
// (pm_engine.hip).  Internal to the library: the public boundary is the C
// ABI in include/pm_stereo.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace pm {

// Device buffers of one view.  The texels live in two planes of
// tex_stride(W) * H float4 records each: the colour plane (r, g, b, bilw)
// and, right after it, the gradient plane (gx, gy, gxy, gyx).  Texel rows
// are padded with copies of their edge texels, kTexPadL on the left and
// kTexPadR on the right, so that every linear sample the kernels form
// (coordinates clamped to [-1, W]) reads existing records without
// clamping its indices: texel (x, y) is record y * tex_stride(W) +
// kTexPadL + x of each plane.
// sgn[y * W + x] = sign bits of the gradient (bit k = signbit of
// component k).
struct ViewBufs {
  float4* tex = nullptr;
  uint8_t* sgn = nullptr;
};

constexpr int kTexPadL = 2;
constexpr int kTexPadR = 3;

__host__ __device__ inline int tex_stride(int W) {
  return W + kTexPadL + kTexPadR;
}

// work-group tiles of the cost-evaluating kernels: kBlockCols x kBlockRows
// pixels
constexpr int kBlockCols = 64;
constexpr int kBlockRows = 4;

// candidate cost of the plane-evaluating kernels: what the reference's
// `#define patch_eval` (pmhstereo.cl:753) selects
enum CostKind {
  kCostCensus = 0,     // patch_eval_census, pmhstereo.cl:661-736
  kCostColorGrad = 1   // patch_eval_color_grad_weighted, pmhstereo.cl:510-565
};

struct Geometry {
  int W = 0, H = 0, R = 0;
  int TS = 0;  // texel row stride in records, tex_stride(W)
  size_t npix = 0;
  // a launch covers the block rows [by0, by0 + nby) of the image
  // (nby = 0: all of them)
  int by0 = 0, nby = 0;
  // CENSUS_COLOR_WEIGHT: bilateral weight sum, census cost and KLT step
  float color_weight = 12.0f;
  int cost_kind = kCostCensus;
  // support-weight factor of the colour+gradient cost (16 in the
  // reference, pmhstereo.cl:531)
  float cost_weight = 16.0f;
  // red-black propagate pass: -1 = every pixel (the reference's Jacobi
  // sweep), 0 / 1 = only the pixels with (x + y) & 1 equal to it
  int color = -1;
};

hipError_t launch_prepare(hipStream_t st, const Geometry& g, const float* img,
                          int channels, int stride, ViewBufs v);

hipError_t launch_init(hipStream_t st, const Geometry& g, ViewBufs self,
                       ViewBufs other, float dmax, int lr, float4* states);

hipError_t launch_disparity_to_normal(hipStream_t st, const Geometry& g,
                                      const float* disp, float4* normals);

hipError_t launch_init_prior(hipStream_t st, const Geometry& g, ViewBufs self,
                             ViewBufs other, float dmax, int lr,
                             const float* disp, const float4* normals,
                             float4* states);

// init of a view from a plane proposal (Hp x Wp x os fp32 per pixel:
// disparity / dscale, normal; the right view's from the mirrored pass)
// with a per-pixel fallback to the random plane of launch_init
hipError_t launch_init_proposal(hipStream_t st, const Geometry& g,
                                ViewBufs self, ViewBufs other, float dmax,
                                int lr, const float* prop, int os, int Wp,
                                float dscale, float4* states);

// clear the view-scatter slots of image rows [row0, row0 + rows)
hipError_t launch_viewbuf_clear_rows(hipStream_t st, const Geometry& g,
                                     unsigned long long* vbuf, int row0,
                                     int rows);
hipError_t launch_viewbuf_clear(hipStream_t st, const Geometry& g,
                                unsigned long long* vbuf);

struct PropagateArgs {
  ViewBufs self, other;
  const float4* old = nullptr;
  const float4* smooth = nullptr;       // may be null (cleared field)
  const unsigned long long* viewin = nullptr;
  const float4* src_other = nullptr;    // states the viewin keys refer to
  unsigned long long* viewout = nullptr;
  float4* out = nullptr;
  float dmax = 0.0f, theta = 0.0f;
  int lr = 1, iter = 0;
  // red-black pass 1 (Geometry::color == 1): the pass-0 pixels' new planes,
  // read instead of `old` for neighbours of colour 0 (may alias `out`)
  const float4* fresh = nullptr;
};

// counters (optional, 2 x u64): cost taps and KLT taps evaluated, counted
// with the reference's per-tap early exit.  lds_floor: the work-groups ask
// for at least this much LDS (only the tile's share is used), which caps
// how many share a CU.
hipError_t launch_propagate(hipStream_t st, const Geometry& g,
                            const PropagateArgs& a,
                            unsigned long long* counters = nullptr,
                            size_t lds_floor = 0);

hipError_t launch_scatter(hipStream_t st, const Geometry& g,
                          const float4* states, int lr,
                          unsigned long long* viewout);

// right-view planes (lr = 0) from the left view's field src through the
// slots a left-view scatter (lr = 1) filled: the nearest surface where one
// landed, pmhstereo_init's random plane elsewhere; cost channel NaN
hipError_t launch_init_from_view(hipStream_t st, const Geometry& g,
                                 float dmax,
                                 const unsigned long long* slots,
                                 const float4* src, float4* states);

hipError_t launch_cost_field(hipStream_t st, const Geometry& g, ViewBufs self,
                             ViewBufs other, const float4* planes,
                             float* cost);

hipError_t launch_klt_field(hipStream_t st, const Geometry& g, ViewBufs self,
                            ViewBufs other, const float4* planes, int passes,
                            float4* out);

hipError_t launch_consistency(hipStream_t st, const Geometry& g,
                              const float4* left, const float4* right,
                              float maxdispdiff, float maxdeg, float4* out);

hipError_t launch_to_disparity(hipStream_t st, const Geometry& g,
                               const float4* states, float scale, int lr,
                               float* disp);

hipError_t launch_fill_depthmap(hipStream_t st, const Geometry& g,
                                const float4* states, float scale,
                                float* disp);

hipError_t launch_fill_normalmap(hipStream_t st, const Geometry& g,
                                 const float4* states, float4* normals);

// PM-Huber coupling: diffusion tensor of an image (alpha, beta = 1),
// (normal, depth) field of a plane field, diffusion-tensor ROF steps, and
// a float4 fill
hipError_t launch_diff_tensor(hipStream_t st, const Geometry& g,
                              const float* img, int channels, int stride,
                              float alpha, float4* dt);

hipError_t launch_normal_depth(hipStream_t st, const Geometry& g,
                               const float4* states, float dispscale, int lr,
                               float4* out);

// nsteps <= kRofStepsMax diffusion-tensor ROF steps in one launch (u, q,
// f, the tensor: one float4 per pixel; p: two)
constexpr int kRofStepsMax = 2;
// LDS of one k_rof_steps work-group (bytes)
size_t rof_steps_lds_bytes();
// hubereps >= 0: diffusion-tensor Huber steps with that epsilon instead
// (PDDHuberDiffTensor.cl:38-109)
hipError_t launch_rof_steps(hipStream_t st, const Geometry& g, int nsteps,
                            const float4* u, const float4* p, const float4* q,
                            const float4* f, const float4* dt, float4 lambda,
                            float hubereps, float4* u_out, float4* p_out,
                            float4* q_out);

// pmhstereo_fill_state: filled (normal, disparity * dispscale) field
hipError_t launch_fill_state(hipStream_t st, const Geometry& g,
                             const float4* states, float dispscale, int lr,
                             float4* out);

hipError_t launch_fill4(hipStream_t st, size_t n, float4 v, float4* out);

// Disparity::evaluateError on the device: counts[0..2] = error, valid,
// all (zeroed first by a kernel, so the pair can be recorded in a graph)
hipError_t launch_eval_error(hipStream_t st, const float* disp, int sd,
                             const float* gt, int sg, int W, int H,
                             float abs_thres, float rel_thres,
                             unsigned long long* counts);

// ---------------------------------------------------------------------
// Batches of stereo pairs of uneven sizes (pm_batch.cpp).  Every pair is
// laid out as a single-pair context lays out its views; a launch covers
// all pairs, reading each pair's geometry and buffers from a descriptor
// in device memory.
// ---------------------------------------------------------------------
struct BatchPair {
  Geometry g;               // full image: by0 = 0, nby = 0, color = -1
  ViewBufs v[2];
  float4* state[2][2];      // [view][ping-pong]
  unsigned long long* vbuf[2];
  float4* checked;          // LR-checked left field
  float* disp;              // disparity out (W x H)
  const float* img[2];      // input images, `stride` floats per row
  const float* gt;          // ground truth for the error counts (may be null)
  unsigned long long* counts;  // 3 x u64: error, valid, all
  int channels, stride, gt_stride;
  float dmax;
  int nbx;                  // tile blocks per row (64 columns each)
  long long item0;          // first tile block of the pair in the batch
};

// tile blocks of 64 x rows pixels of a W x H image
__host__ __device__ inline long long batch_blocks(int W, int H, int rows) {
  return (long long)((W + 63) / 64) * ((H + rows - 1) / rows);
}

// Launch shape of the batch's tile kernels: tile rows (2, 4 or 8) and the
// work-group count of the persistent grid (each work-group walks tile
// blocks of its XCD's eighth of the batch; 0 = one work-group per block).
struct BatchShape {
  int rows = 4;
  int workgroups = 0;
};

constexpr int kBatchRowChoices[3] = {2, 4, 8};

// per-pixel kernels over all pairs: grid z = pair (and view); maxW / maxH
// bound the pairs' sizes
hipError_t launch_prepare_batch(hipStream_t st, const BatchPair* pairs,
                                int npairs, int maxW, int maxH, int R,
                                float cw);
hipError_t launch_viewbuf_clear_batch(hipStream_t st, const BatchPair* pairs,
                                      int npairs, int maxW, int maxH,
                                      int view);
hipError_t launch_init_batch(hipStream_t st, const BatchPair* pairs,
                             int npairs, long long nitems, int view,
                             int cost_kind, size_t lds, BatchShape shape);
// one view's propagate of iteration `iter` over all pairs, reading
// state[view][cur] and writing state[view][1 - cur]
hipError_t launch_propagate_batch(hipStream_t st, const BatchPair* pairs,
                                  int npairs, long long nitems, int view,
                                  int cur, int iter, int cost_kind,
                                  size_t lds, BatchShape shape);
// pp: 0 none, 1 LR check, 2 LR check + fill (pm_postproc)
hipError_t launch_postprocess_batch(hipStream_t st, const BatchPair* pairs,
                                    int npairs, int maxW, int maxH, int cur,
                                    int pp, float maxdispdiff, float maxdeg);
hipError_t launch_eval_error_batch(hipStream_t st, const BatchPair* pairs,
                                   int npairs, int maxW, int maxH,
                                   float abs_thres, float rel_thres);
// LDS bytes of a tile work-group of `rows` rows at patch radius R
size_t batch_lds_bytes(int R, int cost_kind, int rows);
// propagate work-groups of that shape one CU holds at once
hipError_t batch_resident_per_cu(int cost_kind, int rows, size_t lds,
                                 int* out);

// pm_io_last_error's message on this thread (pm_io.cpp); returns code
int io_set_error(int code, const char* msg);

hipError_t launch_draws(hipStream_t st, uint64_t k, int n, float* out);

hipError_t launch_math(hipStream_t st, const float* in, int n, float* out);

hipError_t launch_copy_probe(hipStream_t st, const float4* src, float4* dst,
                             size_t n);

}  // namespace pm
