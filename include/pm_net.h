/*
 * pm_net.h -- C ABI of the learned plane proposal: a small U-Net that
 * predicts a plane (disparity and normal) per pixel of the left view from
 * the stacked left/right RGBA pair, and seeds the PatchMatch engine with it.
 *
 * No counterpart in the reference: this is the seeding hook of
 * PMHuberStereo::init with an initial disparity and normal map
 * (cvt/vision/PMHuberStereo.cpp:361-378 -> pmhstereo_init_disparity_normal,
 * cvt/cl/kernel/pmhstereo.cl:802-828) fed by a network instead of a
 * caller's map, with a per-pixel fallback to the random init
 * (pmhstereo_init, :755-773) wherever the proposal's cost loses.
 * PatchMatch still refines every pixel.
 *
 * Network: 4 levels of 3x3 convolutions with ReLU (32, 64, 96, 128
 * channels; 2 per level, 2 x 2 average pooling between levels, the
 * bottleneck 4 convolutions with dilations 1, 2, 4, 8), a decoder of
 * nearest upsampling, concatenation with the encoder's level and 2
 * convolutions per level, and a 3x3 head of 4 outputs: disparity / 128
 * and the normal (nx, ny, nz).  The first convolutions of the second and
 * third levels also read fixed matching cost volumes: 64 levels at half
 * resolution and 48 at quarter resolution of the 3x3-window mean of
 * absolute colour differences between the 2 x 2 (4 x 4) box-averaged
 * images at horizontal shifts of 0..63 half (0..47 quarter) pixels.  bf16 weights and activations, fp32
 * accumulation on the matrix cores; fp32 master weights and Adam moments
 * for training.  Inputs are padded to multiples of 8 (edge replicated).
 *
 * All device pointers are HIP device memory of the net's device; calls are
 * ordered on the net's stream.  Status codes are pm_status
 * (include/pm_stereo.h).
 */
#ifndef PM_NET_H
#define PM_NET_H

#include <stddef.h>
#include <stdint.h>

#include "pm_stereo.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pm_net pm_net;

/* disparity scale of output channel 0: disparity = 128 * out[0] */
#define PM_NET_DISP_SCALE 128.0f
/* channels of the network input (left rgba, right rgba) and output */
#define PM_NET_IN_CHANNELS 8
#define PM_NET_OUT_CHANNELS 4

/* Create a network with He-initialised weights from `seed` on `device`,
 * ordered on `stream` (a hipStream_t, NULL = own stream). */
int pm_net_create(int device, void* stream, uint64_t seed, pm_net** out);
void pm_net_destroy(pm_net* net);
const char* pm_net_last_error(const pm_net* net);

/* Number of fp32 parameters (weights and biases, padded layout). */
size_t pm_net_num_params(const pm_net* net);
/* Parameters to / from host memory (n = pm_net_num_params floats). */
int pm_net_get_params(pm_net* net, float* out, size_t n);
int pm_net_set_params(pm_net* net, const float* in, size_t n);

/* Forward pass: d_x N x H x W x 8 fp32 (left rgba, right rgba), H and W
 * multiples of 8; d_out N x H x W x 4 fp32 (disparity / 128, nx, ny, nz). */
int pm_net_forward(pm_net* net, const float* d_x, int N, int H, int W,
                   float* d_out);

/* One training step: forward, masked L1 loss against d_target (N x H x W x
 * 4 fp32: disparity / 128, nx, ny, nz; a disparity below 0 marks a pixel
 * without a target) with weights w_disp and w_normal, backward, Adam
 * (beta 0.9, 0.999, eps 1e-8) at learning rate lr.  *loss (may be NULL)
 * receives the loss before the update (the call then waits for it). */
int pm_net_train_step(pm_net* net, const float* d_x, const float* d_target,
                      int N, int H, int W, float lr, float w_disp,
                      float w_normal, float* loss);

/* Diagnostic: the parameter gradient of the last pm_net_train_step
 * (n = pm_net_num_params floats, host), before its Adam update. */
int pm_net_get_grads(pm_net* net, float* out, size_t n);

/* Setup of a context seeded by the network, the counterpart of pm_setup
 * (PMHuberStereo.cpp:445-487) with init replaced by the proposal: the
 * network runs on the context's pair for the left view and on the
 * mirrored pair for the right view, and each pixel keeps the proposed
 * plane unless the random plane of pmhstereo_init costs less.  The net and
 * the context must be on the same device; the net runs on the context's
 * stream. */
int pm_setup_proposal(pm_ctx* ctx, pm_net* net);

/* pm_setup_proposal with a choice of seeded views (no counterpart in the
 * reference): views = 2 is pm_setup_proposal; views = 1 runs the network
 * on the pair only and seeds the left view as above; the right view takes,
 * at every pixel, the nearest left surface that the view scatter of the
 * seeded left field puts there (the plane the propagate's first view
 * candidate would read), and pmhstereo_init's random plane where none
 * lands, with no cost evaluated (NaN: every propagate re-costs its planes
 * first).  The network's pass over the mirrored pair is saved. */
int pm_setup_proposal_views(pm_ctx* ctx, pm_net* net, int views);

#ifdef __cplusplus
}
#endif
#endif
