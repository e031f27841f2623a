/*
 * pm_stereo.h -- C ABI of the MI355X PatchMatch stereo engine.
 *
 * Drop-in for the setup/iterate path of cvt::PMHuberStereo
 * (cvt/vision/PMHuberStereo.h:42-110): images in, plane fields and a
 * disparity map out.  Every entry point is extern "C", takes plain
 * pointers and sizes, and returns a pm_status; no exception crosses the
 * boundary (the reference throws CVTException / CLException,
 * cvt/util/Exception.h:35, cvt/cl/CLKernel.h:82-83).  The message of the
 * last failure is available from pm_last_error().
 *
 * Conventions
 *  - image: W x H float32, channels 1 (GRAY_FLOAT, read as (I,I,I,I)) or
 *    4 (RGBA_FLOAT), row stride in floats (>= W * channels).
 *  - plane field ("state"): W x H x 4 float32 (a, b, c, cost), the
 *    reference's RGBA_FLOAT match image (PMHuberStereo.cpp:460-466): the
 *    corresponding x in the other view is a*x + b*y + c.
 *  - view 0 = left (the reference's LR = true), view 1 = right.
 *  - host-pointer entry points synchronise before returning; the *_device
 *    variants and pm_iterate are ordered on the context's stream and
 *    return without waiting (pm_sync waits).
 *  - contexts are independent: one per thread or stream.  Calls on one
 *    context must not overlap.
 */
#ifndef PM_STEREO_H
#define PM_STEREO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pm_ctx pm_ctx;

typedef enum pm_status {
    PM_OK = 0,
    PM_ERR_ARG = 1,     /* invalid argument (size, pointer, parameter)   */
    PM_ERR_STATE = 2,   /* call out of order (e.g. iterate before setup) */
    PM_ERR_HIP = 3,     /* HIP runtime failure or no device              */
    PM_ERR_NOMEM = 4    /* device allocation failed                      */
} pm_status;

/* Post-processing, cvt::DisparityPostProcessing (PMHuberStereo.h:36-40) */
typedef enum pm_postproc {
    PM_PP_NONE = 0,
    PM_PP_LRCHECK = 1,
    PM_PP_LRCHECK_FILL = 2
} pm_postproc;

/* Coupling of the plane fields to a smoothed (normal, depth) field */
typedef enum pm_smoothing {
    /* plain PatchMatch: theta = 0, no smoothing */
    PM_SMOOTH_NONE = 0,
    /* PM-Huber, disparityMapAll's pipeline (PMHuberStereo.cpp:406-581):
     * theta rises from theta_min, the accept rule weighs theta * 50 times
     * the distance to the smoothed field (pmhstereo.cl:846-877), and after
     * every iteration each view's (normal, depth) field is smoothed by 40
     * diffusion-tensor ROF steps (PDOptCL.cpp:1114-1168) */
    PM_SMOOTH_HUBER = 1,
    /* depthMap's pipeline (PMHuberStereo.cpp:147-347): theta starts at
     * 1e-3 and follows ((i - 3) / (N - 4))^3 + 1e-3 after iteration
     * i >= 3, the accept rule weighs theta * 10 times the distance, and
     * each view's new field, filled by pmhstereo_fill_state, is smoothed
     * by 40 diffusion-tensor Huber steps (PDOptCL.cpp:1170-1225, eps 1e-4,
     * lambda theta * 15 + 2); theta_min is not used */
    PM_SMOOTH_DEPTHMAP = 2
} pm_smoothing;

/* Candidate cost: what the reference's `#define patch_eval`
 * (pmhstereo.cl:753) selects.  The KLT step is patch_census_klt2 with
 * either (the propagate calls it by name, pmhstereo.cl:1036-1038). */
typedef enum pm_cost_kind {
    /* patch_eval_census (pmhstereo.cl:661-736), the reference's choice:
     * support-weighted truncated colour difference and gradient-sign
     * census, with the early exit */
    PM_COST_CENSUS = 0,
    /* patch_eval_color_grad_weighted (pmhstereo.cl:510-565): weight
     * exp(-|dot(I1(p) - I1(q), gray)| * cost_weight), tap cost 0.05 x the
     * colour differences + 0.95 x the gradient differences, each
     * truncated at 0.05; no early exit and no invalid-plane test; a patch
     * with no tap inside the other image costs NaN, which is never
     * accepted */
    PM_COST_COLOR_GRAD = 1
} pm_cost_kind;

/* Order in which a view's propagate visits its pixels */
typedef enum pm_schedule {
    /* the reference's Jacobi sweep: every pixel reads the previous field
     * (pmhstereo_propagate_view reads `old`, writes `output`,
     * pmhstereo.cl:880-1071) */
    PM_SCHEDULE_JACOBI = 0,
    /* not in the reference: two passes over the checkerboard colours
     * ((x + y) even, then odd); the second pass's pixels see the first
     * pass's new planes at their colour-0 random neighbours, so good planes
     * spread twice per iteration.  Same candidates, draws, cost and accept
     * rule; deterministic; runs without the split schedule */
    PM_SCHEDULE_RED_BLACK = 1
} pm_schedule;

typedef struct pm_params {
    int radius;          /* patch half-width r: the reference's `patchsize`
                            (PMHuberStereo.h:47; loops -r..r,
                            pmhstereo.cl:685-687).  1..16, default 5 */
    float disparity_max; /* `depthmax`, > 0, default 64 */
    float color_weight;  /* support-weight gain, CENSUS_COLOR_WEIGHT = 12
                            (pmhstereo.cl:47): bilateral weight sum, census
                            cost and KLT step */
    int cost_kind;       /* pm_cost_kind, default PM_COST_CENSUS */
    float cost_weight;   /* PM_COST_COLOR_GRAD's support-weight gain: 16 in
                            the reference (pmhstereo.cl:531); a gamma of
                            Bleyer's exp(-dI / gamma) on 0-255 data is
                            255 / gamma on [0, 1] data (gamma 10: 25.5) */
    int schedule;        /* pm_schedule, default PM_SCHEDULE_JACOBI */
    int smoothing;       /* pm_smoothing, default PM_SMOOTH_NONE */
    float theta_min;     /* PM-Huber: first theta, 1e-4 as disparityMap
                            passes it (PMHuberStereo.cpp:92) */
    int schedule_iterations; /* PM-Huber: N of theta = (i / N)^4 + theta_min
                            after iteration i (PMHuberStereo.cpp:575-576),
                            depthMap: of ((i - 3) / (N - 4))^3 + 1e-3
                            (:315-318); 0 = the iterations run so far plus
                            the count of the pm_iterate call (one call
                            after setup gives the reference's schedule).
                            Iterations past the schedule's end (split calls)
                            keep theta at its end value */
    int device;          /* HIP device ordinal, default 0 */
    void* stream;        /* hipStream_t to run on, NULL = own stream */
} pm_params;

/* Fill *p with the reference defaults. */
void pm_params_default(pm_params* p);

/* Create a context.  Replaces PMHuberStereo::PMHuberStereo()
 * (PMHuberStereo.cpp:46-67: builds the 20 CL kernels, throws on
 * failure).  Fails with PM_ERR_HIP when no HIP device is present. */
int pm_create(const pm_params* p, pm_ctx** out);
void pm_destroy(pm_ctx* ctx);
const char* pm_last_error(const pm_ctx* ctx);

/* Upload a stereo pair from host memory.  Replaces the host->CL copy of
 * Image::assign / ImageAllocatorCL (PMHuberStereo.cpp:83-87,
 * cvt/gfx/ImageAllocatorCL.cpp:58-86) and the equal-size check
 * (PMHuberStereo.cpp:412-414). */
int pm_set_images(pm_ctx* ctx, const float* left, const float* right, int W,
                  int H, int channels, int stride);
/* Same from device memory (stream-ordered copy into the context). */
int pm_set_images_device(pm_ctx* ctx, const float* d_left,
                         const float* d_right, int W, int H, int channels,
                         int stride);

/* Precompute and random init, the first half of disparityMapAll
 * (PMHuberStereo.cpp:445-487): bilateral weights and gradients of both
 * images (bilateralWeightToAlpha :678-684, gradient :646-651), smooth
 * fields cleared, init of both views (init :349-360 ->
 * pmhstereo_init), view buffer cleared; iteration counter = 0.  With
 * PM_SMOOTH_HUBER also the diffusion tensors of both images
 * (PDOptCL::diffusionTensor, PDOptCL.cpp:134-155, alpha 24, beta 1) and
 * theta = theta_min. */
int pm_setup(pm_ctx* ctx);
/* Setup seeding the left view from a prior disparity (> 0 where known)
 * and normal map, PMHuberStereo::init with initdisparity/initnormalmap
 * (PMHuberStereo.cpp:361-378 -> pmhstereo_init_disparity_normal,
 * pmhstereo.cl:802-828).  normals may be NULL: they are then derived from
 * the disparity (pmhstereo_disparity_to_normal, pmhstereo.cl:1604-1637).
 * With PM_SMOOTH_HUBER the left smooth field starts as the smoothed
 * (normal, depth) field of the seeded planes (PMHuberStereo.cpp:473-477).
 * Host arrays: disparity W x H, normals W x H x 4. */
int pm_setup_prior(pm_ctx* ctx, const float* disparity, const float* normals);

/* Run n PatchMatch iterations (left view, then right view), the loop of
 * disparityMapAll (PMHuberStereo.cpp:491-580) with propagate() :381-404
 * -> pmhstereo_propagate_view; with PM_SMOOTH_HUBER each iteration also
 * smooths both views' (normal, depth) fields (normalDepth :668-675, then
 * denoiseDiffusionTensorROF_PDD) and advances theta.  Exactly n iterations
 * (the reference's `iterations |= 1` lives in the C++ wrapper's
 * disparityMap).  Internally the call forks work to streams the context
 * owns (each view's propagate runs as two halves; PM-Huber smoothing runs
 * beside the other view's propagate) and joins them back into the
 * context's stream before it returns, so the call stays ordered on that
 * stream like any other.  Consecutive propagates overlap only within one
 * call, so run the iterations as one call where possible (8 iterations of
 * a 720p pair: 52.0 ms as one call, 56.9 ms as eight calls of one).  With
 * PM_SMOOTH_NONE the call can be recorded by stream capture on the
 * context's stream; with PM_SMOOTH_HUBER or PM_SMOOTH_DEPTHMAP a
 * capturing stream gets PM_ERR_ARG before anything is queued. */
int pm_iterate(pm_ctx* ctx, int n);
int pm_iteration(const pm_ctx* ctx);

/* Plane field of a view (W x H x 4 floats). */
int pm_get_planes(pm_ctx* ctx, int view, float* out);
/* Replace a view's plane field; `iteration` sets the counter that seeds
 * the next iteration's draws.  The view-propagation buffer the next left
 * step reads is rebuilt from the right field (empty when iteration is 0),
 * so a run resumes exactly from the two fields and the counter.  With
 * PM_SMOOTH_HUBER the smooth fields are kept and theta follows the counter
 * when schedule_iterations is set. */
int pm_set_planes(pm_ctx* ctx, int view, const float* in, int iteration);

/* Cost of the given planes at every pixel of a view with the context's
 * cost kind, mincost = 1e10 (patch_eval_census, pmhstereo.cl:661-736, or
 * patch_eval_color_grad_weighted, :510-565).  planes W x H x 4, cost W x H
 * (host).  Does not modify the context's plane fields. */
int pm_cost(pm_ctx* ctx, int view, const float* planes, float* cost);

/* Left-view disparity with post-processing: checkConsistencyLR
 * (PMHuberStereo.cpp:594-604 -> pmhstereo_consistency) then
 * convertDisparity (:626-634 -> pmhstereo_to_disparity) or, for
 * PM_PP_LRCHECK_FILL, convertFillDisparity (:615-621 ->
 * pmhstereo_fill_depthmap).  maxdispdiff in px, maxdeg in degrees
 * (the reference uses 1.0 and 5.0, PMHuberStereo.cpp:79-80).
 * disparity: W x H floats; checked (may be NULL): W x H x 4 floats, the
 * checked left plane field ((0,0,0,-1) where rejected). */
int pm_disparity(pm_ctx* ctx, int pp, float maxdispdiff, float maxdeg,
                 float* disparity, float* checked);
/* Same into device memory, stream-ordered (no host synchronisation). */
int pm_disparity_device(pm_ctx* ctx, int pp, float maxdispdiff,
                        float maxdeg, float* d_disparity);
/* Left-view normal map with the row fill of convertFillNormal
 * (PMHuberStereo.cpp:636-644 -> pmhstereo_fill_normalmap,
 * pmhstereo.cl:1259-1305) applied to the field of the last
 * pm_disparity call (or the raw left field for PM_PP_NONE).
 * normals: W x H x 4 floats (host). */
int pm_normals(pm_ctx* ctx, int pp, float* normals);

/* Wait for all work queued on the context. */
int pm_sync(pm_ctx* ctx);

/* PM_SMOOTH_HUBER, PM_SMOOTH_DEPTHMAP: the smoothed (nx, ny, nz,
 * disparity / disparity_max)
 * field a view's accept rule currently pulls towards (the reference's
 * leftsmooth / rightsmooth, PMHuberStereo.cpp:439-440), W x H x 4. */
int pm_get_smooth(pm_ctx* ctx, int view, float* out);

/* ---- diagnostics (no reference counterpart: they expose intermediate
 * results so the engine can be checked piece by piece) ----------------- */
/* texel records of a view: W x H x 8 floats (r, g, b, bilw, gx, gy, gxy,
 * gyx) = bilateral-weight image + gradient image */
int pm_get_texels(pm_ctx* ctx, int view, float* out);
/* n draws of the MWC64X stream from index k, computed on the device */
int pm_draws(pm_ctx* ctx, uint64_t k, int n, float* out);
/* `passes` KLT steps (patch_census_klt2, pmhstereo.cl:326-410) applied
 * to every pixel's plane of a view; planes and out W x H x 4 (host) */
int pm_klt(pm_ctx* ctx, int view, const float* planes, int passes,
           float* out);
/* pm_iterate that also counts the work: the cost taps evaluated (with
 * the reference's per-tap early exit) and the KLT taps, summed over both
 * views and the n iterations; results are identical to pm_iterate */
int pm_iterate_counted(pm_ctx* ctx, int n, uint64_t* cost_taps,
                       uint64_t* klt_taps);
/* stream-ordered plain copy of `bytes` (multiple of 16, 16-byte aligned
 * device buffers): the known-size kernel HBM byte counters are
 * calibrated against */
int pm_copy_probe(pm_ctx* ctx, const float* d_src, float* d_dst,
                  size_t bytes);
/* PM_SMOOTH_HUBER: a view's diffusion tensor (a, b, c, d), W x H x 4 */
int pm_get_tensor(pm_ctx* ctx, int view, float* out);
/* PM_SMOOTH_HUBER: `iters` diffusion-tensor ROF steps on the host field f
 * (W x H x 4) with a view's tensor and the per-channel lambda[4], as
 * denoiseDiffusionTensorROF_PDD runs them; out W x H x 4 (host) */
int pm_rof(pm_ctx* ctx, int view, const float* f, const float* lambda,
           int iters, float* out);
/* pm_rof on device buffers d_f and d_out (W x H x 4 in HBM, not
 * overlapping; lambda[4] on the host), stream-ordered: returns after
 * enqueueing on the context's stream */
int pm_rof_device(pm_ctx* ctx, int view, const float* d_f,
                  const float* lambda, int iters, float* d_out);
/* device evaluation of the engine's scalar math for n inputs:
 * out[5*i .. 5*i+4] = exp, sin, cos, acospi, sqrt of in[i] */
int pm_math_selftest(pm_ctx* ctx, const float* in, int n, float* out);

#ifdef __cplusplus
}
#endif
#endif
