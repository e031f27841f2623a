/*
 * pm_io.h -- disparity-map interchange and error metrics of the MI355X
 * PatchMatch engine (host code in libcvt_amd_pm.so).
 *
 * PFM files carry disparity maps and ground truth between this engine and
 * cvt users; the error metrics are cvt's Disparity ones.  Every entry point
 * is extern "C", takes plain pointers and sizes and returns a pm_status
 * (include/pm_stereo.h).  These functions need no GPU.  The message of the
 * last failure on the calling thread is available from pm_io_last_error().
 *
 * Image buffers are W x H float32, row-major and top row first, with a
 * row stride in floats (>= W * channels).
 */
#ifndef PM_IO_H
#define PM_IO_H

#include <stddef.h>
#include <stdint.h>

#include "pm_stereo.h"

#ifdef __cplusplus
extern "C" {
#endif

/* pm_status value for unreadable, unwritable or malformed files */
#define PM_ERR_IO 5

/* Message of the last failed pm_io call on this thread ("" if none). */
const char* pm_io_last_error(void);

/* Header of a PFM file: size and channel count (1 for "Pf", 4 for "PF",
 * whose RGB is returned as RGBA).  Follows PFMLoader::parseHeader
 * (plugins/iloader/PFM/PFMLoader.cpp:58-105): whitespace-separated tokens
 * per line, lines starting with '#' skipped, the type on the first line,
 * "<width> <height>" on the next, the scale on the next. */
int pm_pfm_info(const char* path, int* W, int* H, int* channels);

/* Read a PFM file (PFMLoader::load and parseBinary / parseBinarySwap,
 * PFMLoader.cpp:107-205): the rows are
 * stored bottom to top in the file and returned top row first; a negative
 * scale means little-endian samples, a positive one big-endian; "PF" RGB
 * samples become RGBA with alpha 1 (SIMD::Conv_XXXf_to_XXXAf,
 * cvt/util/SIMD.cpp:2384-2391).  out holds H rows of `stride` floats
 * (stride >= W * channels of the file). */
int pm_pfm_read(const char* path, float* out, int W, int H, int channels,
                int stride);

/* Write a PFM file (PFMSaver::save, plugins/isaver/PFM/PFMSaver.cpp:
 * 117-181): channels 1 -> "Pf", 4 -> "PF" with the alpha channel dropped;
 * scale -1.0 (little-endian samples), rows bottom to top. */
int pm_pfm_write(const char* path, const float* data, int W, int H,
                 int channels, int stride);

/* Disparity::evaluateError (cvt/vision/Disparity.cpp:201-233): over the
 * pixels with gt > 0, `all` counts them, `valid` those with an estimate
 * > 0 and `error` those with |gt - est| > abs_thres and
 * |gt - est| / |gt| > rel_thres. */
int pm_eval_error(const float* disparity, int stride_d, const float* gt,
                  int stride_gt, int W, int H, float abs_thres,
                  float rel_thres, uint64_t* error, uint64_t* valid,
                  uint64_t* all);

/* pm_eval_error on the device: d_disparity, d_gt in HIP device memory of
 * the current device, d_counts = 3 uint64 (error, valid, all) written in
 * stream order on `stream` (a hipStream_t, NULL = the null stream): two
 * kernels, no host synchronisation and no memset, so the call can be
 * recorded by stream capture. */
int pm_eval_error_device(const float* d_disparity, int stride_d,
                         const float* d_gt, int stride_gt, int W, int H,
                         float abs_thres, float rel_thres,
                         uint64_t* d_counts, void* stream);

/* pm_disparity_device on caller-owned plane fields, allocating nothing.
 * No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops).
 * d_left_planes, d_right_planes: W x H x 4 floats (a, b, c, cost) in HIP
 * device memory of `device`, the two views' fields as pm_get_planes
 * returns them; d_disparity: W x H floats.  pp, maxdispdiff and maxdeg as
 * pm_disparity.  Stream-ordered on `stream` (a hipStream_t of `device`,
 * NULL = its null stream), no host synchronisation.  The LR-checked field
 * lives in a fixed array of 2^21 float4 per GPU (32 MiB, held until the
 * process ends): W x H above 2^21 pixels (1920 x 1080 fits) is refused
 * with PM_ERR_ARG.  Calls that overlap on one GPU share that array and
 * race.  Failures leave their message in pm_io_last_error(). */
int pm_disparity_planes_noalloc(const float* d_left_planes,
                                const float* d_right_planes, int W, int H,
                                int pp, float maxdispdiff, float maxdeg,
                                float* d_disparity, int device, void* stream);

/* Disparity::histogramErrorAbsolute / histogramErrorRelative
 * (Disparity.cpp:236-318): over pixels where both maps are finite and
 * > 0, the fraction whose error e = |test - gt| (relative: divided by
 * |gt|) falls in (bins[i-1], bins[i]] with bins[-1] = 0, so an error of
 * exactly 0 lands in no bin; counts are float sums divided by the number
 * of such pixels (NaN when there are none).  out[nbins]. */
int pm_error_histogram(const float* gt, int stride_gt, const float* test,
                       int stride_test, int W, int H, int relative,
                       const float* bins, int nbins, float* out);

/* Disparity::defaultHistogramAbsolute / defaultHistogramRelative
 * (Disparity.cpp:320-355): the bin edges (15 absolute, 10 relative).
 * bins may be NULL to query the count. */
int pm_default_histogram(int relative, float* bins, int* nbins);

#ifdef __cplusplus
}
#endif
#endif
