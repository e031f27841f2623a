/*
 * pm_oracle.h -- CPU restatement of cvt's slanted-plane PatchMatch stereo
 * (cvt::PMHuberStereo driving cvt/cl/kernel/pmhstereo.cl).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine
 * in cvt_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product never links or calls it.
 *
 * Layouts (all row-major, one view):
 *   image    : W x H float32, 1 channel (GRAY, read as (I,I,I,I) like
 *              CL_INTENSITY, cvt/gfx/IFormat.h:308) or 4 channels (RGBA),
 *              row stride in floats.
 *   tex      : W x H x 8 float32 texel records
 *              (r, g, b, bilw, gx, gy, gxy, gyx):
 *              the reference's bilateral-weight image
 *              (pmhstereo_bilateral_weight_to_alpha, pmhstereo.cl:1462-1498)
 *              followed by its gradient image (pmhstereo_gradxy,
 *              pmhstereo.cl:1549-1602).
 *   state    : W x H x 4 float32 plane field (a, b, c, cost); the matching
 *              x-coordinate in the other view is a*x + b*y + c
 *              (nd_state_transform, pmhstereo.cl:183-186).
 *   viewbuf  : W x H uint64 view-propagation slots (deterministic
 *              replacement of VIEWPROP_t, pmhstereo.cl:830-833): low word
 *              = smallest key, high word = second smallest, 0xFFFFFFFF =
 *              empty; key = (lr ? W-1-x : x) of the scattering pixel.
 *
 * Arithmetic contract shared with the HIP engine: IEEE float32, no
 * contraction by the compiler (built with -ffp-contract=off); the fused
 * operations are written as fmaf(), by one rule: the contractions an
 * OpenCL compiler with FP_CONTRACT ON forms in the reference's source
 * expressions (a product operand of a sum or difference fused into it,
 * the left one first; a built-in dot as the chain fma(x2, y2, fma(x1, y1,
 * x0 * y0)); never across statements or helper calls).  Division and sqrt
 * correctly rounded; exp, sin, cos and acospi are this file's own
 * Cody-Waite + Taylor implementations.
 */
#ifndef PM_ORACLE_H
#define PM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scalar math (own implementations, pinned against libm in tests) --- */
float pmo_expf(float x);
void pmo_sincosf(float x, float* s, float* c);
float pmo_acospif(float x);

/* ---- MWC64X stream (cvt/cl/kernel/MWC64X.cl, RNG.cl) ----------------- */
/* state (x, c) of draw k of the global stream */
void pmo_mwc_state(uint64_t k, uint32_t* x, uint32_t* c);
/* n consecutive draws starting at stream index k (RNG_float semantics) */
void pmo_draws(uint64_t k, int n, float* out);
/* n draws starting at k produced by Skip(k) then n Step()s */
void pmo_draws_stepped(uint64_t k, int n, float* out);
/* stream index of the first draw of a propagate at pixel p, iteration
 * iter: 6 (p + iter) + 8, the expansion of RNG_init at pmhstereo.cl:942 */
uint64_t pmo_prop_draw_start(uint64_t p, int iter);

/* ---- plane algebra (pmhstereo.cl:77-224) -------------------------------*/
void pmo_viewprop(const float* s, float* out);
float pmo_transform(const float* s, float x, float y);
void pmo_to_normal(const float* s, float* n3);
void pmo_from_disparity_normal(float x, float y, float d, const float* n3,
                               float* out);
void pmo_state_init(float u0, float u1, float u2, float x, float y, int lr,
                    float dmax, float* out);
void pmo_state_refine(float u0, float u1, float u2, const float* s, float x,
                      float y, float dmax, int lr, float* out);
void pmo_to_ref_normal_depth(const float* s, float x, float y, int lr,
                             float* out);

/* ---- per-image precompute ----------------------------------------------*/
/* gradient + bilateral weight sum -> tex (W*H*8) */
void pmo_prepare(const float* img, int W, int H, int channels, int stride,
                 int radius, float color_weight, float* tex);

/* ---- per-pixel kernels -------------------------------------------------
 * Candidate costs: cost_kind selects what the reference's
 * `#define patch_eval` (pmhstereo.cl:753) names.  color_weight is
 * CENSUS_COLOR_WEIGHT (pmhstereo.cl:47: the bilateral weight sum, the
 * census cost and the KLT step), cost_weight the colour+gradient cost's
 * support-weight factor (16 in pmhstereo.cl:531). */
#define PMO_COST_CENSUS 0      /* patch_eval_census, pmhstereo.cl:661-736 */
#define PMO_COST_COLOR_GRAD 1  /* patch_eval_color_grad_weighted, :510-565 */
/* cost of plane s at (x, y); mincost is the census early-exit bound */
float pmo_cost_at(const float* tex1, const float* tex2, int W, int H,
                  int radius, float color_weight, int cost_kind,
                  float cost_weight, int x, int y, const float* s,
                  float mincost);
/* cost of every pixel's plane with mincost = 1e10 (no early exit) */
void pmo_cost_field(const float* tex1, const float* tex2, int W, int H,
                    int radius, float color_weight, int cost_kind,
                    float cost_weight, const float* planes, float* cost);
/* one patch_census_klt2 pass (pmhstereo.cl:326-410) */
void pmo_klt_at(const float* tex1, const float* tex2, int W, int H,
                int radius, float color_weight, int x, int y,
                const float* s, float* out);
/* pmhstereo_init over the whole view (pmhstereo.cl:755-773) */
void pmo_init(const float* tex1, const float* tex2, int W, int H, int radius,
              float color_weight, int cost_kind, float cost_weight,
              float dmax, int lr, float* states);
/* pmhstereo_init_disparity_normal (pmhstereo.cl:802-828) */
void pmo_init_prior(const float* tex1, const float* tex2, int W, int H,
                    int radius, float color_weight, int cost_kind,
                    float cost_weight, float dmax, int lr,
                    const float* disparity, const float* normals,
                    float* states);
void pmo_viewbuf_clear(uint64_t* vbuf, int W, int H);
/* pmhstereo_propagate_view, Jacobi schedule (pmhstereo.cl:880-1071).
 * src_other: the other view's state field the keys in viewin refer to. */
/* color >= 0: one pass of the red-black schedule (not in the reference):
 * only the pixels with (x + y) & 1 == color; pass 1 reads its colour-0
 * random neighbours from `fresh` (pass 0's output), the rest from `old`.
 * color = -1: every pixel (fresh unused). */
void pmo_propagate(const float* tex1, const float* tex2, int W, int H,
                   int radius, float color_weight, int cost_kind,
                   float cost_weight, float dmax, float theta, int lr,
                   int iter, const float* old, const float* smooth,
                   const uint64_t* viewin, const float* src_other,
                   uint64_t* viewout, float* out, int color,
                   const float* fresh);
/* the view scatter alone (inserts into vbuf, which is not cleared) */
void pmo_scatter(const float* states, int W, int H, int lr, uint64_t* vbuf);
/* pmhstereo_consistency (pmhstereo.cl:1095-1129) */
void pmo_consistency(const float* left, const float* right, int W, int H,
                     float maxdispdiff, float maxdeg, float* out);
/* pmhstereo_to_disparity (pmhstereo.cl:1232-1257) */
void pmo_to_disparity(const float* states, int W, int H, float scale, int lr,
                      float* disp);

/* pmhstereo_fill_depthmap (pmhstereo.cl:1187-1230), .x channel */
void pmo_fill_depthmap(const float* states, int W, int H, float scale,
                       float* disp);
/* pmhstereo_fill_normalmap (pmhstereo.cl:1259-1305), W x H x 4 */
void pmo_fill_normalmap(const float* states, int W, int H, float* normals);
/* pmhstereo_disparity_to_normal (pmhstereo.cl:1604-1637), W x H x 4 */
void pmo_disparity_to_normal(const float* disp, int W, int H,
                             float* normals);

/* ---- whole plain-PatchMatch pipeline (PMHuberStereo.cpp:406-581 with
 * theta = 0, cleared smooth field, no PD smoothing, exactly `iters`
 * iterations), followed by the LR check and disparity conversion.
 * stateL/stateR receive the final plane fields, disp the left disparity,
 * cost_trace (may be NULL) the per-iteration per-pixel cost of both views:
 * layout [iter][view][H][W]. */
#define PMO_SCHEDULE_JACOBI 0     /* the reference's sweep */
#define PMO_SCHEDULE_RED_BLACK 1  /* two checkerboard passes per propagate */
int pmo_run(const float* left, const float* right, int W, int H,
            int channels, int stride, int radius, float dmax, int iters,
            int cost_kind, float cost_weight, int schedule,
            float maxdispdiff, float maxdeg, float* stateL, float* stateR,
            float* disp, float* cost_trace);

/* ---- PM-Huber smoothing coupling (theta > 0) ---------------------------
 * PDDiffTensorColor / PDDiffTensorGray (cvt/cl/kernel/PDDiffTensor.cl:
 * 25-116) with beta = 1: dt = W x H x 4 tensor (a, b, c, d) from the
 * forward differences of the gray image (zero outside). */
void pmo_diffusion_tensor(const float* img, int W, int H, int channels,
                          int stride, float alpha, float* dt);
/* pmhstereo_normal_depth (pmhstereo.cl:1384-1402): (nx, ny, nz,
 * disparity * dispscale) of every plane, W x H x 4 */
void pmo_normal_depth(const float* states, int W, int H, float dispscale,
                      int lr, float* out);
/* one PDDROFDiffTensor step (cvt/cl/kernel/PDDROFDiffTensor.cl:43-125):
 * u, q, f, dt are W x H x 4, p is W x H x 8 (lo, hi); writes u_out,
 * p_out, q_out (p_out holds the tensor-applied projected dual, as the
 * kernel stores it) */
void pmo_rof_step(const float* u, const float* p, const float* q,
                  const float* f, const float* dt, const float lambda[4],
                  int W, int H, float* u_out, float* p_out, float* q_out);
/* denoiseDiffusionTensorROF_PDD (cvt/vision/PDOptCL.cpp:1114-1168): u = f,
 * p = q = 0, then `iters` steps; u_out W x H x 4 */
int pmo_rof(const float* f, const float* dt, const float lambda[4],
            int iters, int W, int H, float* u_out);

/* one PDDHuberDiffTensor step (cvt/cl/kernel/PDDHuberDiffTensor.cl:
 * 38-109): the ROF step with the Huber dual (ascent divided by
 * 1 + 4 eps, one norm over the four channels) */
void pmo_huber_step(const float* u, const float* p, const float* q,
                    const float* f, const float* dt, const float lambda[4],
                    float eps, int W, int H, float* u_out, float* p_out,
                    float* q_out);
/* denoiseDiffusionTensorHuber_PDD (cvt/vision/PDOptCL.cpp:1170-1225),
 * eps >= 0; u_out W x H x 4 */
int pmo_huber(const float* f, const float* dt, const float lambda[4],
              float eps, int iters, int W, int H, float* u_out);
/* pmhstereo_fill_state (pmhstereo.cl:1131-1185): planes whose cost is
 * negative or whose disparity is outside [0, W] replaced by a row
 * neighbour, then (normal, disparity * dispscale), W x H x 4 */
void pmo_fill_state(const float* states, int W, int H, float dispscale,
                    int lr, float* out);

/* ---- whole PM-Huber pipeline: disparityMapAll (PMHuberStereo.cpp:
 * 406-581) with one pyramid level: exactly `iters` iterations (the caller
 * makes the count odd as disparityMap does), theta from theta_min,
 * diffusion-tensor ROF smoothing of both views after every iteration;
 * then the LR check and disparity conversion.  smoothL / smoothR (may be
 * NULL) receive the final smoothed (normal, depth) fields. */
int pmo_run_huber(const float* left, const float* right, int W, int H,
                  int channels, int stride, int radius, float dmax,
                  int iters, float theta_min, float maxdispdiff,
                  float maxdeg, float* stateL, float* stateR, float* disp,
                  float* smoothL, float* smoothR);

/* ---- the depthMap pipeline (PMHuberStereo.cpp:147-347): the coupling
 * with diffusion-tensor Huber smoothing of filled fields, its own theta
 * schedule and constants; outputs as pmo_run_huber. */
int pmo_run_depthmap(const float* left, const float* right, int W, int H,
                     int channels, int stride, int radius, float dmax,
                     int iters, float maxdispdiff, float maxdeg,
                     float* stateL, float* stateR, float* disp,
                     float* smoothL, float* smoothR);

int pmo_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
