/*
 * pm_batch.h -- many stereo pairs of uneven sizes through the PatchMatch
 * engine at once.
 *
 * No counterpart in the reference: cvt::PMHuberStereo runs one pair per
 * call, with fixed 16 x 16 work-groups and a host wait after every kernel
 * (cvt/vision/PMHuberStereo.cpp:35-36, 381-404).  A batch runs the plain
 * PatchMatch sequence of pm_setup + pm_iterate + pm_disparity
 * (PMHuberStereo.cpp:445-545, theta = 0, smooth field cleared) for every
 * pair, and each pair's planes and disparity are bit for bit what a
 * single-pair context with the same parameters computes.  Each kernel of
 * the sequence is one launch over all pairs; the cost-evaluating kernels
 * (init, propagate) are persistent: a fixed grid of work-groups walks the
 * tile blocks of every pair.  The whole sequence is recorded once into a
 * HIP graph and replayed by every pm_batch_run with the same settings.
 * pm_batch_autotune picks the tile shape (tile rows, hence the LDS tile
 * and its halo of r texels on every side) and the persistent grid size by
 * timing the candidates on the batch itself.
 *
 * Status codes are pm_status (pm_stereo.h); on an error,
 * pm_batch_last_error describes it.
 */
#ifndef PM_BATCH_H
#define PM_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "pm_net.h"
#include "pm_stereo.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pm_batch pm_batch;

/* One pair of a batch.  Images are HIP device memory of the batch's
 * device, owned by the caller and read by every pm_batch_run (the same
 * layout as pm_set_images_device: `channels` 1 or 4 floats per pixel,
 * `stride` floats per row).  d_gt (may be NULL): ground-truth disparity of
 * the left view, `gt_stride` floats per row, for the error counts. */
typedef struct pm_batch_pair {
    int width, height;
    int channels;
    int stride;
    float disparity_max;   /* the pair's `depthmax`, > 0 */
    const float* d_left;
    const float* d_right;
    const float* d_gt;
    int gt_stride;
} pm_batch_pair;

/* Launch shape of the persistent tile kernels. */
typedef struct pm_batch_tuning {
    int tile_rows;       /* 2, 4 or 8: work-groups of 64 x tile_rows pixels */
    int workgroups;      /* persistent grid; 0 = one work-group per tile */
    float ms;            /* pm_batch_autotune: time of one view's propagate
                            over the batch with this shape (ms); else 0 */
} pm_batch_tuning;

/* Create a batch of n pairs.  p supplies radius, color_weight, cost_kind,
 * cost_weight, device and stream (p->disparity_max is unused: each pair
 * brings its own); p->smoothing must be PM_SMOOTH_NONE and p->schedule
 * PM_SCHEDULE_JACOBI.  Allocates every pair's buffers in one block of
 * device memory.  On failure *out may still be set: read
 * pm_batch_last_error, then pm_batch_destroy. */
int pm_batch_create(const pm_params* p, const pm_batch_pair* pairs, int n,
                    pm_batch** out);
void pm_batch_destroy(pm_batch* b);
const char* pm_batch_last_error(const pm_batch* b);
int pm_batch_size(const pm_batch* b);

/* Replace pair i's image and ground-truth pointers (same size, channels
 * and disparity_max as at creation; the recorded graph stays valid). */
int pm_batch_set_pair(pm_batch* b, int i, const pm_batch_pair* pair);

/* Stream-ordered: for every pair, setup (texels, random init of both
 * views), `iterations` PatchMatch iterations, post-processing pp
 * (pm_postproc) with the LR thresholds, the disparity, and, with
 * err_abs >= 0, Disparity::evaluateError counts (Disparity.cpp:201-233)
 * against each pair's d_gt.  The first call with a set of arguments and
 * the current tuning records the sequence into a HIP graph; every call
 * launches that graph. */
int pm_batch_run(pm_batch* b, int iterations, int pp, float maxdispdiff,
                 float maxdeg, float err_abs, float err_rel);

/* Pair i's results of the last run, copied to host memory (waits):
 * disparity W x H; plane field of a view W x H x 4; error counts
 * (error, valid, all) of every pair, 3 n values. */
int pm_batch_get_disparity(pm_batch* b, int i, float* out);
int pm_batch_get_planes(pm_batch* b, int i, int view, float* out);
int pm_batch_get_error_counts(pm_batch* b, uint64_t* counts);
/* Device address of pair i's disparity (W x H floats, packed rows), valid
 * until pm_batch_destroy. */
int pm_batch_disparity_ptr(pm_batch* b, int i, const float** d_out);

/* Time every candidate shape (tile rows 2, 4, 8 x persistent grids of 1
 * and 2 rounds of resident work-groups and one work-group per tile) on
 * one left-view propagate over the batch after a setup, best of `reps`,
 * and keep the fastest.  table (may be NULL) receives up to table_size
 * candidates with their times, *count (may be NULL) their number.  Waits;
 * the results of earlier runs are overwritten. */
int pm_batch_autotune(pm_batch* b, int reps, pm_batch_tuning* table,
                      int table_size, int* count);
int pm_batch_get_tuning(const pm_batch* b, pm_batch_tuning* out);
int pm_batch_set_tuning(pm_batch* b, const pm_batch_tuning* t);

/* Seed every pair from the plane-proposal network (include/pm_net.h)
 * instead of the random init: each run's setup does, per pair, what
 * pm_setup_proposal does for a context (the network on the pair and on
 * the mirrored pair, each pixel keeping the proposed plane unless the
 * random init plane costs less), inside the recorded graph.  The net must
 * be on the batch's device and stay alive while the batch uses it; its
 * activation buffers are sized here for the largest pair.  net = NULL
 * restores the random init. */
int pm_batch_set_proposal(pm_batch* b, pm_net* net);

/* pm_batch_set_proposal with a choice of seeded views, as
 * pm_setup_proposal_views (include/pm_net.h): views = 1 seeds only the
 * left view of every pair (the network on the pair alone) and carries it
 * over to the right view through the view scatter; views = 2 is
 * pm_batch_set_proposal. */
int pm_batch_set_proposal_views(pm_batch* b, pm_net* net, int views);

int pm_batch_sync(pm_batch* b);

#ifdef __cplusplus
}
#endif
#endif
